// Radar postprocessing with the reference's semantics (src/graph/RadarPostprocessPointsNode.cpp,
// src/gpu/nodeKernels.cu:82-101,209-320): per-point physical-optics scattering terms, greedy clustering of the
// detections by distance scope and separation thresholds, per-cluster RCS, received power, noise and SNR.
//
// The reference runs the clustering on the host after copying every input field back (a synchronisation inside the
// graph). Here it stays on the device: the clustering is sequential in point order (a point joins the first
// cluster, in creation order, that accepts it), so one wave walks the points and tests 64 clusters per step, with
// cluster bounds and list ends in LDS; the merge phase tests 64 cluster pairs per step in the reference's pair
// order. Per-cluster sums then run with one lane per cluster.
#include <algorithm>
#include <cfloat>

#include "device_types.hpp"
#include "kernels.hpp"
#include "noise_rng.hpp"

namespace rgl {
namespace gpu {

namespace {

constexpr int kBlock = 256;
constexpr float kPi = 3.14159265358979323846f;

int blocksFor(int n) { return (n + kBlock - 1) / kBlock; }

struct Cf
{
	float re, im;
};
__device__ __forceinline__ Cf operator+(Cf a, Cf b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cf operator-(Cf a) { return {-a.re, -a.im}; }
__device__ __forceinline__ Cf operator*(Cf a, Cf b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cf operator*(Cf a, float s) { return {a.re * s, a.im * s}; }

// Complex 3-vector (the field amplitudes) and its products with real vectors.
struct Cv
{
	Cf x, y, z;
};
__device__ __forceinline__ Cv operator+(Cv a, Cv b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ Cv operator-(Cv a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ Cv scale(Cf s, Vec3f v) { return {s * Cf{v.x, 0.0f}, s * Cf{v.y, 0.0f}, s * Cf{v.z, 0.0f}}; }
__device__ __forceinline__ Cv cross(Cv a, Vec3f b)
{
	return {a.y * b.z + -(a.z * b.y), a.z * b.x + -(a.x * b.z), a.x * b.y + -(a.y * b.x)};
}
__device__ __forceinline__ Cf dot(Cv a, Vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Polarisation of the reflected wave (src/gpu/nodeKernels.cu:82-101).
__device__ Vec3f reflectPolarization(Vec3f pol, Vec3f n, Vec3f d)
{
	if (fabsf(dot(d, n)) == 1.0f) {
		return Vec3f{-pol.x, -pol.y, -pol.z};
	}
	const Vec3f polU = normalized(cross(d, n));
	const Vec3f polR = normalized(cross(d, polU));
	const Vec3f refDir = normalized(d - n * (2.0f * dot(d, n)));
	const Vec3f refPolU = polU * -1.0f;
	const Vec3f refPolR = cross(refDir, refPolU);
	return refPolR * -dot(pol, polR) + refPolU * dot(pol, polU);
}

// Scattered-field contributions BU * factor and BR * factor of detection i (src/gpu/nodeKernels.cu:209-320).
__device__ void energyOne(const RadarItem& it, int i)
{
	const RadarPostParams& p = it.p;
	const float waveLen = 299792458.0f / p.frequency;
	const float waveNum = 2.0f * kPi / waveLen;
	const Mat3x4f pose = p.lookAt * it.rayPose[i];
	const Vec3f rayDir = pose * Vec3f{0.0f, 0.0f, 1.0f};
	const Vec3f rayPol = pose * Vec3f{1.0f, 0.0f, 0.0f};
	const Vec3f nrm = p.lookAt.rotation() * it.normal[i];
	const float d = it.distance[i];
	const float rayArea = d * d * sinf(p.elevationStep) * p.azimuthStep;
	const float phi = atan2f(rayDir.y, rayDir.z);
	const float the = acosf(rayDir.x / length(rayDir));
	const float cp = cosf(phi), sp = sinf(phi), ct = cosf(the), st = sinf(the);
	const Vec3f dirP{0.0f, cp, -sp};
	const Vec3f dirT{-st, sp * ct, cp * ct};
	const Vec3f refl = normalized(rayDir - nrm * (2.0f * dot(rayDir, nrm)));
	const Vec3f reflPol = reflectPolarization(rayPol, nrm, rayDir);
	const float kr = waveNum * d;
	const Cv apE = scale(Cf{cosf(kr), sinf(kr)}, reflPol);
	const Cv apH = -cross(apE, refl);
	const Cv refField1 = -(cross(apE, Vec3f{-dirP.x, -dirP.y, -dirP.z}) + cross(apH, dirT));
	const Cv refField2 = -(cross(apE, dirT) + cross(apH, dirP));
	const Cf bu = dot(refField1, refl);
	const Cf br = dot(refField2, refl);
	const float phase = -waveNum * d;
	const Cf factor = Cf{0.0f, (waveNum * rayArea * dot(refl, nrm)) / (4.0f * kPi)} * Cf{cosf(phase), sinf(phase)};
	const Cf au = bu * factor, ar = br * factor;
	reinterpret_cast<float4*>(it.terms)[i] = make_float4(au.re, au.im, ar.re, ar.im);
}

// Bounds of a cluster: distance, azimuth, radial speed, elevation (min, max each).
enum Bound
{
	kDMin,
	kDMax,
	kAMin,
	kAMax,
	kRMin,
	kRMax,
	kEMin,
	kEMax,
	kBounds
};

__device__ __forceinline__ float fmin2(float a, float b) { return b < a ? b : a; } // std::min
__device__ __forceinline__ float fmax2(float a, float b) { return a < b ? b : a; } // std::max
__device__ __forceinline__ float lanef(float v, int j) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j)); }
__device__ __forceinline__ int lanei(int v, int j) { return __builtin_amdgcn_readlane(v, j); }

// The node's distance scopes, scope s held by lane s. Must be used by all 64 lanes together.
struct LaneScopes
{
	float begin, end, dt, rt, at;
	int count;
	// First scope containing d (src/graph/RadarPostprocessPointsNode.cpp:21-30), or -1; thresholds of that scope.
	__device__ int find(float d, float& distT, float& radialT, float& azT) const
	{
		int s = -1;
		for (int k = 0; k < count; ++k) {
			if (s < 0 && lanef(begin, k) <= d && d <= lanef(end, k)) {
				s = k;
			}
		}
		const int src = s < 0 ? 0 : s;
		distT = __shfl(dt, src);
		radialT = __shfl(rt, src);
		azT = __shfl(at, src);
		return s;
	}
};

// One wave (a 64-thread block) clusters one node's detections.
// Phase 1: every point, in order, joins the first accepting cluster or opens a new one
// (src/graph/RadarPostprocessPointsNode.cpp:74-98, RadarCluster::isCandidate :324-348). Points are fetched 64 at a
// time with coalesced loads, their scopes found in parallel, and broadcast from registers; the bounds and sizes of
// clusters 0..63 live in the registers of lane c, later clusters in LDS. Each point records its cluster and its
// rank in it.
// Phase 2: the reference merges the first mergeable pair in (i, j) order and rescans from (0, 1) (:100-117,
// canMergeWith :350-377, takeIndicesFrom :379-393). Pairs of untouched clusters that were found unmergeable stay so,
// so after merging (ci, cj) the next first mergeable pair is either (a, ci) with a < ci or in row ci or later: the
// scan resumes there and finds the same pair. A cluster's member order is the reference's: the phase-1 clusters it
// absorbed form a chain, each contributing its points in index order.
// Phase 3: per cluster (one lane each) the scattered-field sums in member order, the directional centre (:379-398
// of RadarCluster), RCS, power, noise and SNR (:128-174).
__device__ void clusterOne(const RadarItem& it)
{
	__shared__ float b[kBounds][kRadarMaxClusters];
	__shared__ int32_t segHead[kRadarMaxClusters], segTail[kRadarMaxClusters]; // per current cluster
	__shared__ int32_t segNext[kRadarMaxClusters], size[kRadarMaxClusters];    // per phase-1 cluster
	__shared__ int32_t offset[kRadarMaxClusters];
	const int lane = threadIdx.x;
	const int capacity = it.capacity;
	// Uniform across the wave; readfirstlane keeps the loops below on scalar branches.
	const int n = __builtin_amdgcn_readfirstlane(it.count ? *it.count : capacity);
	const float* distance = it.distance;
	const float* azimuth = it.azimuth;
	const float* elevation = it.elevation;
	const float* radialSpeed = it.radialSpeed;
	int32_t* pointCluster = it.work;
	int32_t* pointRank = it.work + capacity;
	int32_t* members = it.work + 2 * static_cast<size_t>(capacity);

	LaneScopes scopes{};
	scopes.count = it.p.scopeCount;
	if (lane < scopes.count) {
		const RadarScope sc = it.p.scopes[lane];
		scopes.begin = sc.begin;
		scopes.end = sc.end;
		scopes.dt = sc.distanceThreshold;
		scopes.rt = sc.radialSpeedThreshold;
		scopes.at = sc.azimuthThreshold;
	}

	int clusters = 0;
	int32_t errors = 0;
	float rb[kBounds]; // bounds of cluster `lane` while it is one of the first 64
	int rsize = 0;     // its size
	for (int q = 0; q < kBounds; ++q) {
		rb[q] = 0.0f;
	}

	for (int base = 0; base < n && errors == 0; base += 64) {
		const int pi = base + lane;
		float pd = 0.0f, pa = 0.0f, pr = 0.0f, pe = 0.0f;
		if (pi < n) {
			pd = distance[pi];
			pa = azimuth[pi];
			pr = radialSpeed[pi];
			pe = elevation[pi];
		}
		float pdt, prt, pat;
		const int ps = scopes.find(pd, pdt, prt, pat);
		int32_t myCluster = -1, myRank = 0; // lane j records point base + j
		const int m = min(64, n - base);
		for (int j = 0; j < m; ++j) {
			if (lanei(ps, j) < 0) {
				continue;
			}
			const float d = lanef(pd, j), a = lanef(pa, j), r = lanef(pr, j), e = lanef(pe, j);
			const float dt = lanef(pdt, j), rt = lanef(prt, j), at = lanef(pat, j);
			auto accepts = [&](float dlo, float dhi, float alo, float ahi, float rlo, float rhi) {
				const bool dOk = d + dt >= dlo && d - dt <= dhi;
				const bool aOk = a + at >= alo && a - at <= ahi;
				const bool rOk = (isnan(rlo) || isnan(rhi)) || (r + rt >= rlo && r - rt <= rhi);
				return dOk && aOk && rOk;
			};
			int found = -1;
			{
				const bool cand = lane < clusters && accepts(rb[kDMin], rb[kDMax], rb[kAMin], rb[kAMax], rb[kRMin], rb[kRMax]);
				const uint64_t mask = __ballot(cand);
				if (mask != 0) {
					found = __ffsll(static_cast<unsigned long long>(mask)) - 1;
				}
			}
			for (int cb = 64; cb < clusters && found < 0; cb += 64) {
				const int c = cb + lane;
				const int cc = c < clusters ? c : 0;
				const bool cand = c < clusters && accepts(b[kDMin][cc], b[kDMax][cc], b[kAMin][cc], b[kAMax][cc],
				                                          b[kRMin][cc], b[kRMax][cc]);
				const uint64_t mask = __ballot(cand);
				if (mask != 0) {
					found = cb + __ffsll(static_cast<unsigned long long>(mask)) - 1;
				}
			}
			if (found < 0 && clusters == kRadarMaxClusters) {
				errors |= kRadarErrTooManyClusters;
				break;
			}
			const int c = found >= 0 ? found : clusters;
			int rank;
			if (c < 64) {
				rank = found >= 0 ? lanei(rsize, c) : 0;
				if (lane == c) {
					if (found >= 0) {
						rb[kDMin] = fmin2(rb[kDMin], d);
						rb[kDMax] = fmax2(rb[kDMax], d);
						rb[kAMin] = fmin2(rb[kAMin], a);
						rb[kAMax] = fmax2(rb[kAMax], a);
						rb[kRMin] = isnan(rb[kRMin]) ? r : fmin2(rb[kRMin], r);
						rb[kRMax] = isnan(rb[kRMax]) ? r : fmax2(rb[kRMax], r);
						rb[kEMin] = fmin2(rb[kEMin], e);
						rb[kEMax] = fmax2(rb[kEMax], e);
					} else {
						rb[kDMin] = rb[kDMax] = d;
						rb[kAMin] = rb[kAMax] = a;
						rb[kRMin] = rb[kRMax] = r;
						rb[kEMin] = rb[kEMax] = e;
					}
					rsize = rank + 1;
				}
			} else {
				rank = found >= 0 ? __builtin_amdgcn_readfirstlane(lane == 0 ? size[c] : 0) : 0;
				if (lane == 0) {
					if (found >= 0) {
						b[kDMin][c] = fmin2(b[kDMin][c], d);
						b[kDMax][c] = fmax2(b[kDMax][c], d);
						b[kAMin][c] = fmin2(b[kAMin][c], a);
						b[kAMax][c] = fmax2(b[kAMax][c], a);
						b[kRMin][c] = isnan(b[kRMin][c]) ? r : fmin2(b[kRMin][c], r);
						b[kRMax][c] = isnan(b[kRMax][c]) ? r : fmax2(b[kRMax][c], r);
						b[kEMin][c] = fmin2(b[kEMin][c], e);
						b[kEMax][c] = fmax2(b[kEMax][c], e);
					} else {
						b[kDMin][c] = b[kDMax][c] = d;
						b[kAMin][c] = b[kAMax][c] = a;
						b[kRMin][c] = b[kRMax][c] = r;
						b[kEMin][c] = b[kEMax][c] = e;
					}
					size[c] = rank + 1;
				}
			}
			if (found < 0) {
				if (lane == 0) {
					segHead[c] = segTail[c] = c;
					segNext[c] = -1;
				}
				++clusters;
			}
			if (lane == j) {
				myCluster = c;
				myRank = rank;
			}
			__builtin_amdgcn_wave_barrier();
		}
		if (pi < n) {
			pointCluster[pi] = myCluster;
			pointRank[pi] = myRank;
		}
	}
	if (lane < clusters) {
		for (int q = 0; q < kBounds; ++q) {
			b[q][lane] = rb[q];
		}
		size[lane] = rsize;
	}
	__syncthreads();

	// Members grouped by phase-1 cluster, each group in point order: offsets by an exclusive scan of the sizes.
	if (errors == 0) {
		int running = 0;
		for (int cb = 0; cb < clusters; cb += 64) {
			const int c = cb + lane;
			const int v = c < clusters ? size[c] : 0;
			int incl = v;
			for (int k = 1; k < 64; k <<= 1) {
				const int t = __shfl_up(incl, k);
				if (lane >= k) {
					incl += t;
				}
			}
			if (c < clusters) {
				offset[c] = running + incl - v;
			}
			running += __shfl(incl, 63);
		}
		__syncthreads();
		// Each lane reads back the point records it wrote itself above.
		for (int i = lane; i < n; i += 64) {
			const int c = pointCluster[i];
			if (c >= 0) {
				members[offset[c] + pointRank[i]] = i;
			}
		}
	}

	// Whether pair (ci, cj) may merge; every lane calls it (scope lookup is wave-wide), `valid` masks the result.
	auto mergeable = [&](bool valid, int ci, int cj) {
		ci = valid ? ci : 0;
		cj = valid ? cj : 0;
		const float lo = fmax2(b[kDMin][ci], b[kDMin][cj]);
		const float hi = fmin2(b[kDMax][ci], b[kDMax][cj]);
		float dt, rt, at;
		const int s = scopes.find(fmax2(lo, hi), dt, rt, at);
		auto good = [&](int lo0, int hi0, float t) {
			const float a0 = b[lo0][ci], a1 = b[hi0][ci], b0 = b[lo0][cj], b1 = b[hi0][cj];
			const bool overlap = (b0 <= a1 && b1 >= a0) || (a0 <= b1 && a1 >= b0);
			const bool near = fabsf(a1 - b0) <= t || fabsf(b1 - a0) <= t;
			return overlap || near;
		};
		const bool rNan = isnan(b[kRMin][ci]) || isnan(b[kRMax][ci]) || isnan(b[kRMin][cj]) || isnan(b[kRMax][cj]);
		return valid && s >= 0 && good(kDMin, kDMax, dt) && good(kAMin, kAMax, at) && (good(kRMin, kRMax, rt) || rNan);
	};
	// Merges cluster cj into ci (ci < cj) and erases cj.
	auto merge = [&](int ci, int cj) {
		if (lane == 0) {
			b[kDMin][ci] = fmin2(b[kDMin][ci], b[kDMin][cj]);
			b[kDMax][ci] = fmax2(b[kDMax][ci], b[kDMax][cj]);
			b[kAMin][ci] = fmin2(b[kAMin][ci], b[kAMin][cj]);
			b[kAMax][ci] = fmax2(b[kAMax][ci], b[kAMax][cj]);
			// Both radial-speed bounds take the smaller value, as the reference does (:380-383).
			b[kRMin][ci] = isnan(b[kRMin][cj]) ? b[kRMin][ci] : fmin2(b[kRMin][cj], b[kRMin][ci]);
			b[kRMax][ci] = isnan(b[kRMax][cj]) ? b[kRMax][ci] : fmin2(b[kRMax][cj], b[kRMax][ci]);
			b[kEMin][ci] = fmin2(b[kEMin][ci], b[kEMin][cj]);
			b[kEMax][ci] = fmax2(b[kEMax][ci], b[kEMax][cj]);
			segNext[segTail[ci]] = segHead[cj];
			segTail[ci] = segTail[cj];
		}
		__syncthreads();
		// Clusters after cj move down by one, 64 at a time (all reads of a step before its writes).
		for (int kb = cj; kb < clusters - 1; kb += 64) {
			const int k = kb + lane;
			const bool act = k < clusters - 1;
			float v[kBounds];
			int32_t h = 0, t = 0;
			if (act) {
				for (int q = 0; q < kBounds; ++q) {
					v[q] = b[q][k + 1];
				}
				h = segHead[k + 1];
				t = segTail[k + 1];
			}
			__syncthreads();
			if (act) {
				for (int q = 0; q < kBounds; ++q) {
					b[q][k] = v[q];
				}
				segHead[k] = h;
				segTail[k] = t;
			}
			__syncthreads();
		}
		--clusters;
	};
	int row = 0;      // rows before `row` hold no mergeable pair except pairs with `changed`
	int changed = -1; // the cluster that absorbed another in the last merge
	while (errors == 0 && clusters > 1) {
		int ci = -1, cj = -1;
		for (int base = 0; base < changed && ci < 0; base += 64) {
			const int a = base + lane;
			const uint64_t mask = __ballot(mergeable(a < changed, a, changed));
			if (mask != 0) {
				ci = base + __ffsll(static_cast<unsigned long long>(mask)) - 1;
				cj = changed;
			}
		}
		if (ci < 0) {
			row = changed >= 0 ? changed : row;
			for (int r = row; r < clusters - 1 && ci < 0; ++r) {
				for (int base = r + 1; base < clusters && ci < 0; base += 64) {
					const int c = base + lane;
					const uint64_t mask = __ballot(mergeable(c < clusters, r, c));
					if (mask != 0) {
						ci = r;
						cj = base + __ffsll(static_cast<unsigned long long>(mask)) - 1;
					}
				}
			}
		}
		if (ci < 0) {
			break;
		}
		merge(ci, cj);
		changed = ci;
	}
	__threadfence(); // members written above are read by other lanes below
	__syncthreads();

	const RadarPostParams& p = it.p;
	const float lambda = 299792458.0f / p.frequency;
	const float lambdaSqrtDbsm = 10.0f * log10f(lambda * lambda);
	const float fourPiCubedDb = 10.0f * log10f(powf(4.0f * kPi, 3.0f));
	bool rcsNan = false;
	for (int c = lane; c < clusters && errors == 0; c += 64) {
		const float meanA = (b[kAMin][c] + b[kAMax][c]) / 2.0f;
		const float meanE = (b[kEMin][c] + b[kEMax][c]) / 2.0f;
		Cf au{0.0f, 0.0f}, ar{0.0f, 0.0f};
		float best = FLT_MAX;
		int centre = -1;
		for (int o = segHead[c]; o >= 0; o = segNext[o]) {
			const int end = offset[o] + size[o];
#pragma unroll 4
			for (int k = offset[o]; k < end; ++k) {
				const int pt = members[k];
				const float4 t = reinterpret_cast<const float4*>(it.terms)[pt];
				au = au + Cf{t.x, t.y};
				ar = ar + Cf{t.z, t.w};
				const float v = fabsf(azimuth[pt] - meanA) + fabsf(elevation[pt] - meanE);
				if (centre < 0) {
					centre = pt;
				}
				if (v < best) {
					best = v;
					centre = pt;
				}
			}
		}
		const float mu = hypotf(au.re, au.im), mr = hypotf(ar.re, ar.im);
		const float rcsDb = 10.0f * log10f(4.0f * kPi * (mu * mu + mr * mr));
		const float d = distance[centre];
		const float multiplier = fourPiCubedDb + 10.0f * log10f(powf(d, 4.0f));
		const float received = p.powerTransmitted + p.deviceGain + p.deviceGain + rcsDb + lambdaSqrtDbsm - multiplier;
		const float nz = p.noiseMean + p.noiseStDev * rngNormal(p.seed, p.run, p.tag, static_cast<uint32_t>(c));
		it.centres[c] = centre;
		it.rcs[c] = rcsDb;
		it.noise[c] = nz;
		it.power[c] = received + nz;
		it.snr[c] = it.power[c] - nz;
		rcsNan = rcsNan || isnan(rcsDb);
	}
	const bool anyNan = __ballot(rcsNan) != 0;
	if (lane == 0) {
		it.info[0] = errors == 0 ? clusters : 0;
		it.info[1] = errors | (anyNan ? kRadarErrRcsNan : 0);
	}
}

// out[f][k] = in[f][centre[k]] for every listed field.
__device__ void gatherOne(const RadarItem& it, int k)
{
	const FieldCopyList& list = *it.gather;
	const int src = it.centres[k];
	for (int f = 0; f < list.count; ++f) {
		const FieldCopy fc = list.f[f];
		const char* s = fc.src + static_cast<size_t>(src) * fc.size;
		char* d = fc.dst + static_cast<size_t>(k) * fc.size;
		if ((fc.size & 3) == 0) {
			for (int w = 0; w < fc.size / 4; ++w) {
				reinterpret_cast<int32_t*>(d)[w] = reinterpret_cast<const int32_t*>(s)[w];
			}
		} else {
			for (int w = 0; w < fc.size; ++w) {
				d[w] = s[w];
			}
		}
	}
}

__global__ void kRadarEnergy(RadarItem it)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < (it.count ? *it.count : it.capacity)) {
		energyOne(it, i);
	}
}

__global__ __launch_bounds__(64) void kRadarCluster(RadarItem it) { clusterOne(it); }

__global__ void kRadarGather(RadarItem it)
{
	const int k = blockIdx.x * kBlock + threadIdx.x;
	if (k < it.info[0]) {
		gatherOne(it, k);
	}
}

__device__ __forceinline__ int findRadarItem(const RadarItem* items, int count)
{
	return batchItemOf(items, count, static_cast<int>(blockIdx.x), [](const RadarItem& r) { return r.firstBlock; });
}

__global__ void kRadarEnergyBatch(const RadarItem* __restrict__ items, int count)
{
	const RadarItem& it = items[findRadarItem(items, count)];
	const int i = (static_cast<int>(blockIdx.x) - it.firstBlock) * kBlock + threadIdx.x;
	if (i < (it.count ? *it.count : it.capacity)) {
		energyOne(it, i);
	}
}

__global__ __launch_bounds__(64) void kRadarClusterBatch(const RadarItem* __restrict__ items) { clusterOne(items[blockIdx.x]); }

__global__ void kRadarGatherBatch(const RadarItem* __restrict__ items, int count)
{
	const RadarItem& it = items[findRadarItem(items, count)];
	const int k = (static_cast<int>(blockIdx.x) - it.firstBlock) * kBlock + threadIdx.x;
	if (k < it.info[0]) {
		gatherOne(it, k);
	}
}

} // namespace

int radarBlocks(int capacity) { return std::max(blocksFor(capacity), 1); }

void radarPostprocess(hipStream_t stream, const RadarItem& it)
{
	const int blocks = radarBlocks(it.capacity);
	hipLaunchKernelGGL(kRadarEnergy, dim3(blocks), dim3(kBlock), 0, stream, it);
	hipLaunchKernelGGL(kRadarCluster, dim3(1), dim3(64), 0, stream, it);
	hipLaunchKernelGGL(kRadarGather, dim3(blocks), dim3(kBlock), 0, stream, it);
}

void radarPostprocessBatch(hipStream_t stream, const RadarItem* items, int count, int totalBlocks)
{
	if (count <= 0) {
		return;
	}
	hipLaunchKernelGGL(kRadarEnergyBatch, dim3(totalBlocks), dim3(kBlock), 0, stream, items, count);
	hipLaunchKernelGGL(kRadarClusterBatch, dim3(count), dim3(64), 0, stream, items);
	hipLaunchKernelGGL(kRadarGatherBatch, dim3(totalBlocks), dim3(kBlock), 0, stream, items, count);
}

} // namespace gpu
} // namespace rgl
