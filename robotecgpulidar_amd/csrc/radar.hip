// Radar postprocessing with the reference's semantics (src/graph/RadarPostprocessPointsNode.cpp,
// src/gpu/nodeKernels.cu:82-101,209-320): per-point physical-optics scattering terms, greedy clustering of the
// detections by distance scope and separation thresholds, per-cluster RCS, received power, noise and SNR.
//
// The reference runs the clustering on the host after copying every input field back (a synchronisation inside the
// graph). Here it stays on the device: the clustering is sequential in point order (a point joins the first
// cluster, in creation order, that accepts it), so one wave assigns 64 points per step where the order provably
// does not change the answer and walks the others one at a time, testing 64 clusters per step; the merge phase
// tests 64 cluster pairs per step in the reference's pair order. Per-cluster sums then run with the wave spread over
// one cluster's members at a time.
#include <algorithm>
#include <cfloat>
#include <climits>

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
// Predicates of __builtin_amdgcn_fcmpf (LLVM's compare codes): ordered >=, ordered <=, unordered.
constexpr int kOGE = 3, kOLE = 5, kUNO = 8;

// Data-parallel-primitive controls (DPP, GFX9 encoding): shift right by 1/2/4/8 lanes within rows of 16, broadcast
// lane 15 of a row to the next row, lane 31 to rows 2 and 3, and shift the whole wave right by one lane.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// v from the lane the DPP control names, or `id` where there is none (or the row is masked off).
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ float dppOr(float id, float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(id), __float_as_int(v), kCtrl, kRowMask, 0xf, false));
}

// Inclusive scan of the wave's values, minimum or maximum (NaN-free values): within each row of 16, then the totals
// of rows 0 and 2 into rows 1 and 3, then that of rows 0..1 into rows 2 and 3.
template <bool kMax>
__device__ __forceinline__ float scanWave(float v)
{
	const float id = kMax ? -INFINITY : INFINITY;
	auto op = [](float a, float b) { return kMax ? (a < b ? b : a) : (b < a ? b : a); };
	v = op(v, dppOr<kRowShr1>(id, v));
	v = op(v, dppOr<kRowShr2>(id, v));
	v = op(v, dppOr<kRowShr4>(id, v));
	v = op(v, dppOr<kRowShr8>(id, v));
	v = op(v, dppOr<kRowBcast15, 0xa>(id, v));
	v = op(v, dppOr<kRowBcast31, 0xc>(id, v));
	return v;
}

// Bounds `nb` of a cluster with bounds `o` after adding point (d, a, r, e) (RadarCluster::addPoint,
// src/graph/RadarPostprocessPointsNode.cpp:274-291), or of a new cluster holding only that point when `fresh`
// (the constructor, :264-272): fmin2(v, v) = fmax2(v, v) = v, a NaN included. Radial-speed bounds holding a NaN take
// the point's speed.
__device__ __forceinline__ void addPoint(const float* o, float* nb, bool fresh, float d, float a, float r, float e)
{
	const float v[4] = {d, a, r, e}; // pairs (kDMin, kDMax), (kAMin, kAMax), (kRMin, kRMax), (kEMin, kEMax)
	for (int k = 0; k < 4; ++k) {
		const float lo = fresh ? v[k] : o[2 * k], hi = fresh ? v[k] : o[2 * k + 1];
		if (2 * k == kRMin) {
			nb[2 * k] = isnan(lo) ? v[k] : fmin2(lo, v[k]);
			nb[2 * k + 1] = isnan(hi) ? v[k] : fmax2(hi, v[k]);
		} else {
			nb[2 * k] = fmin2(lo, v[k]);
			nb[2 * k + 1] = fmax2(hi, v[k]);
		}
	}
}

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
// time with coalesced loads and their scopes found in parallel; the bounds and sizes of clusters 0..63 live in the
// registers of lane c, later clusters in LDS. A chunk step assigns a run of the 64 points at once, the walk takes
// the rest one by one (both described where they run). Each point records its cluster and its rank in it.
// Phase 2: the reference merges the first mergeable pair in (i, j) order and rescans from (0, 1) (:100-117,
// canMergeWith :350-377, takeIndicesFrom :379-393). Pairs of untouched clusters that were found unmergeable stay so,
// so after merging (ci, cj) the next first mergeable pair is either (a, ci) with a < ci or in row ci or later: the
// scan resumes there and finds the same pair. A cluster's member order is the reference's: the phase-1 clusters it
// absorbed form a chain, each contributing its points in index order.
// Phase 3: per cluster the scattered-field sums over its members (the wave across one cluster at a time), the
// directional centre (:379-398 of RadarCluster), then per cluster (one lane each) RCS, power, noise and SNR
// (:128-174).
__device__ void clusterOne(const RadarItem& it)
{
	__shared__ float b[kBounds][kRadarMaxClusters];
	__shared__ int32_t segHead[kRadarMaxClusters], segTail[kRadarMaxClusters]; // per current cluster
	__shared__ int32_t segNext[kRadarMaxClusters], size[kRadarMaxClusters];    // per phase-1 cluster
	__shared__ int32_t offset[kRadarMaxClusters];
	__shared__ float4 sums[kRadarMaxClusters];      // per final cluster: scattered-field sums (U re, im, R re, im)
	__shared__ int32_t centres[kRadarMaxClusters];  // per final cluster: its centre point
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
	uint64_t held = 0; // lanes holding a cluster: min(clusters, 64) low bits
	// -inf and +inf in registers the optimizer cannot see into, so that a median against them stays one instruction
	// rather than being turned back into a min/max with operand canonicalisation.
	float negInf = -INFINITY, posInf = INFINITY;
	asm volatile("" : "+v"(negInf), "+v"(posInf));
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
		// The ends of each point's acceptance windows, by the same additions the per-cluster test would make.
		const float dHi = pd + pdt, dLo = pd - pdt, aHi = pa + pat, aLo = pa - pat, rHi = pr + prt, rLo = pr - prt;
		int32_t myCluster = -1, myRank = 0; // lane j records point base + j
		// isCandidate for point j against clusters 0..63 (one per lane) as a lane mask: one compare per bound, the
		// masks combined on the scalar unit. Radial-speed bounds holding a NaN accept any speed.
		auto heldAccepting = [&](int j) -> uint64_t {
			const float sdHi = lanef(dHi, j), sdLo = lanef(dLo, j), saHi = lanef(aHi, j), saLo = lanef(aLo, j);
			const float srHi = lanef(rHi, j), srLo = lanef(rLo, j);
			const uint64_t dOk = __builtin_amdgcn_fcmpf(sdHi, rb[kDMin], kOGE) & __builtin_amdgcn_fcmpf(sdLo, rb[kDMax], kOLE);
			const uint64_t aOk = __builtin_amdgcn_fcmpf(saHi, rb[kAMin], kOGE) & __builtin_amdgcn_fcmpf(saLo, rb[kAMax], kOLE);
			const uint64_t rOk = __builtin_amdgcn_fcmpf(rb[kRMin], rb[kRMax], kUNO) |
			                     (__builtin_amdgcn_fcmpf(srHi, rb[kRMin], kOGE) & __builtin_amdgcn_fcmpf(srLo, rb[kRMax], kOLE));
			return dOk & aOk & rOk & held;
		};
		// Points outside every scope join no cluster; the others are walked in order, one per step. The inner loop
		// takes the common step (the point joins a cluster held in registers, whose lane grows its bounds) without
		// leaving it; a point that no such cluster accepts goes through the general step below.
		uint64_t pending = __builtin_amdgcn_ballot_w64(pi < n && ps >= 0);
		// Chunk step. Bounds only widen (minima fall, maxima rise), and a cluster's radial-speed bounds stay NaN until
		// a finite speed joins and stay finite after; so, as long as those bounds do not turn finite, the points a
		// cluster accepts only grow in number. Let f0(p) be the first cluster accepting point p at the start of the
		// step. If every earlier point q of the chunk joined f0(q), the walk's clusters before p are the starting
		// ones grown by those points, and p joins f0(p) unless a cluster below f0(p) that earlier points grew now
		// accepts it (clusters not grown still refuse it), f0(p) does not exist (p opens a cluster), or an earlier
		// point turned a grown cluster's radial-speed bounds finite. The step finds the first point where one of
		// these happens, from an inclusive scan of each grown cluster's members over the wave, and commits every
		// point before it. The walk then takes that point, and the step is tried again on the rest of the chunk
		// while a try commits at least kMinRun points. The step runs while every cluster is held in registers (at
		// most 64): over clusters held in LDS it is slower than the walk.
		constexpr int kMinRun = 16;
		auto below = [](int q) { return q >= 64 ? ~0ull : (1ull << q) - 1; };
		// The members' values, or the identity of each minimum (+inf) and maximum (-inf) for other lanes. A member's
		// distance and azimuth passed ordered compares, so they are not NaN, nor are radial speeds against finite
		// bounds; NaN elevations and speeds are left out, as std::min / std::max against a number leave them out.
		auto memberValues = [&](bool member, float* v) {
			const bool rNum = member && !isnan(pr), eNum = member && !isnan(pe);
			v[kDMin] = member ? pd : INFINITY;
			v[kDMax] = member ? pd : -INFINITY;
			v[kAMin] = member ? pa : INFINITY;
			v[kAMax] = member ? pa : -INFINITY;
			v[kRMin] = rNum ? pr : INFINITY;
			v[kRMax] = rNum ? pr : -INFINITY;
			v[kEMin] = eNum ? pe : INFINITY;
			v[kEMax] = eNum ? pe : -INFINITY;
		};
		// Bounds x grown by the values v. NaN radial-speed bounds take the speed among v when `rTook` says there is
		// one (there is at most one: a second would follow the point that turned the bounds finite); NaN elevation
		// bounds stay NaN, as in the walk.
		auto grow = [&](const float* x, const float* v, bool rTook, float* nb) {
			const bool rNan = __builtin_isunordered(x[kRMin], x[kRMax]);
			for (int q = 0; q < kBounds; q += 2) {
				nb[q] = fmin2(x[q], v[q]);
				nb[q + 1] = fmax2(x[q + 1], v[q + 1]);
			}
			nb[kRMin] = rNan ? (rTook ? v[kRMin] : x[kRMin]) : nb[kRMin];
			nb[kRMax] = rNan ? (rTook ? v[kRMax] : x[kRMax]) : nb[kRMax];
		};
		auto accepts = [&](const float* nb) {
			const bool rOk = __builtin_isunordered(nb[kRMin], nb[kRMax]) | ((rHi >= nb[kRMin]) & (rLo <= nb[kRMax]));
			return (dHi >= nb[kDMin]) & (dLo <= nb[kDMax]) & (aHi >= nb[kAMin]) & (aLo <= nb[kAMax]) & rOk;
		};
		auto chunkStep = [&](uint64_t pending) -> uint64_t {
			const bool live = (pending >> lane) & 1;
			// f0 cluster by cluster for all points at once (until every live point has one), or point by point
			// against all clusters at once, whichever takes fewer steps.
			int f0 = -1;
			if (clusters <= __builtin_popcountll(pending)) {
				for (int k = 0; k < clusters; ++k) {
					float x[kBounds];
					for (int q = 0; q < kBounds; ++q) {
						x[q] = lanef(rb[q], k);
					}
					f0 = ((f0 < 0) & accepts(x)) ? k : f0;
					if ((__builtin_amdgcn_ballot_w64(f0 < 0) & pending) == 0) {
						break;
					}
				}
			} else {
				for (uint64_t t = pending; t != 0; t &= t - 1) {
					const int j = __builtin_ctzll(t);
					const uint64_t hit = heldAccepting(j);
					f0 = lane == j ? (hit != 0 ? __builtin_ctzll(hit) : -1) : f0;
				}
			}
			const uint64_t none = __builtin_amdgcn_ballot_w64(f0 < 0) & pending;
			int stop = none != 0 ? __builtin_ctzll(none) : 64; // the first point the walk must take
			// Each grown cluster's bounds with all its live members are kept in its lane (tb, tsize) in case every
			// point commits; otherwise the commit below recomputes them up to `stop`. Ranks hold either way.
			float tb[kBounds];
			for (int q = 0; q < kBounds; ++q) {
				tb[q] = rb[q];
			}
			int tsize = rsize, trank = 0;
			for (uint64_t todo = pending & below(stop); todo != 0 && __builtin_ctzll(todo) < stop;) {
				const int k = lanei(f0, __builtin_ctzll(todo));
				const bool member = live && f0 == k;
				const uint64_t mk = __builtin_amdgcn_ballot_w64(member);
				todo &= ~mk;
				float x[kBounds], v[kBounds], nb[kBounds];
				for (int q = 0; q < kBounds; ++q) {
					x[q] = lanef(rb[q], k);
				}
				// A speed that turns NaN bounds finite ends the run after its point; no lane before it has one
				// among its earlier members, so the checks below keep the NaN bounds.
				const uint64_t finite = __builtin_amdgcn_ballot_w64(member && !isnan(pr));
				if (__builtin_isunordered(x[kRMin], x[kRMax])) {
					stop = finite != 0 ? min(stop, __builtin_ctzll(finite) + 1) : stop;
				}
				memberValues(member, v);
				for (int q = 0; q < kBounds; q += 2) { // inclusive scans over the wave
					v[q] = scanWave<false>(v[q]);
					v[q + 1] = scanWave<true>(v[q + 1]);
				}
				float total[kBounds];
				for (int q = 0; q < kBounds; ++q) {
					total[q] = lanef(v[q], 63);
				}
				grow(x, total, finite != 0, nb);
				const int size0 = lanei(rsize, k);
				if (lane == k) {
					for (int q = 0; q < kBounds; ++q) {
						tb[q] = nb[q];
					}
					tsize = size0 + __builtin_popcountll(mk);
				}
				const int before = static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mk >> 32),
				                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mk), 0u)));
				trank = member ? size0 + before : trank;
				// Would the cluster, grown by the members before a point whose f0 lies above it, take that point?
				if (__builtin_amdgcn_ballot_w64(live & (f0 > k)) != 0) {
					for (int q = 0; q < kBounds; q += 2) { // members before this lane
						v[q] = dppOr<kWaveShr1>(INFINITY, v[q]);
						v[q + 1] = dppOr<kWaveShr1>(-INFINITY, v[q + 1]);
					}
					grow(x, v, false, nb);
					const uint64_t taken = __builtin_amdgcn_ballot_w64(live & (f0 > k) & accepts(nb));
					stop = taken != 0 ? min(stop, __builtin_ctzll(taken)) : stop;
				}
			}
			if (stop == 64) { // every point commits, with the bounds kept above
				for (int q = 0; q < kBounds; ++q) {
					rb[q] = tb[q];
				}
				rsize = tsize;
				myCluster = live ? f0 : myCluster;
				myRank = live ? trank : myRank;
				return pending;
			}
			// Commit the points before `stop`: each grown cluster's bounds, size, and its members' ranks.
			const uint64_t done = pending & below(stop);
			const bool committed = (done >> lane) & 1;
			for (uint64_t todo = done; todo != 0;) {
				const int k = lanei(f0, __builtin_ctzll(todo));
				const bool member = committed && f0 == k;
				const uint64_t mk = __builtin_amdgcn_ballot_w64(member);
				todo &= ~mk;
				float x[kBounds], v[kBounds];
				for (int q = 0; q < kBounds; ++q) {
					x[q] = lanef(rb[q], k);
				}
				const bool rTook = __builtin_amdgcn_ballot_w64(member && !isnan(pr)) != 0;
				memberValues(member, v);
				for (int q = 0; q < kBounds; q += 2) { // the members' minima and maxima: the scans' last lane
					v[q] = lanef(scanWave<false>(v[q]), 63);
					v[q + 1] = lanef(scanWave<true>(v[q + 1]), 63);
				}
				float nb[kBounds];
				grow(x, v, rTook, nb);
				const int size0 = lanei(rsize, k);
				if (lane == k) {
					for (int q = 0; q < kBounds; ++q) {
						rb[q] = nb[q];
					}
					rsize = size0 + __builtin_popcountll(mk);
				}
				const int before = static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mk >> 32),
				                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mk), 0u)));
				myCluster = member ? k : myCluster;
				myRank = member ? size0 + before : myRank;
			}
			return done;
		};
		bool tryChunk = true;
		while (pending != 0) {
			tryChunk = tryChunk && clusters <= 64;
			if (tryChunk) {
				const uint64_t done = chunkStep(pending);
				pending &= ~done;
				tryChunk = __builtin_popcountll(done) >= kMinRun;
				if (pending == 0) {
					break;
				}
			}
			// The walk: one point while chunk steps are being tried, else the rest of the chunk.
			int j = __builtin_ctzll(pending);
			bool joined = false;
			for (uint64_t hit = heldAccepting(j); hit != 0;) {
				const int c = __builtin_ctzll(hit);
				const float d = lanef(pd, j), a = lanef(pa, j), r = lanef(pr, j), e = lanef(pe, j);
				const int rank = lanei(rsize, c);
				if (lane == c) {
					// Distance and azimuth passed ordered compares against these bounds, so none of the four is a
					// NaN and a median against -inf / +inf gives std::min / std::max (a tie of +0 and -0 may keep
					// either zero, which no comparison, difference or midpoint made from the bounds tells apart).
					rb[kDMin] = __builtin_amdgcn_fmed3f(rb[kDMin], d, negInf);
					rb[kDMax] = __builtin_amdgcn_fmed3f(rb[kDMax], d, posInf);
					rb[kAMin] = __builtin_amdgcn_fmed3f(rb[kAMin], a, negInf);
					rb[kAMax] = __builtin_amdgcn_fmed3f(rb[kAMax], a, posInf);
					// A cluster's two radial-speed bounds are NaN together or not at all, and a point with a NaN speed
					// passes the radial test only against NaN bounds: so "a NaN bound takes the speed, else std::min"
					// is "keep the bound if it is <= the speed, else take the speed" (>= for the upper bound).
					rb[kRMin] = rb[kRMin] <= r ? rb[kRMin] : r;
					rb[kRMax] = rb[kRMax] >= r ? rb[kRMax] : r;
					rb[kEMin] = fmin2(rb[kEMin], e);
					rb[kEMax] = fmax2(rb[kEMax], e);
					rsize = rank + 1;
				}
				myCluster = lane == j ? c : myCluster;
				myRank = lane == j ? rank : myRank;
				pending ^= 1ull << j;
				joined = true;
				if (pending == 0 || tryChunk) {
					break;
				}
				j = __builtin_ctzll(pending);
				hit = heldAccepting(j);
			}
			if (pending == 0) {
				break;
			}
			if (joined && tryChunk) {
				continue;
			}
			// General step for point j: the first accepting cluster among those held in LDS, else a new cluster.
			int found = -1;
			for (int cb = 64; cb < clusters && found < 0; cb += 64) {
				const int c = cb + lane;
				const int cc = c < clusters ? c : 0;
				const float sdHi = lanef(dHi, j), sdLo = lanef(dLo, j), saHi = lanef(aHi, j), saLo = lanef(aLo, j);
				const float srHi = lanef(rHi, j), srLo = lanef(rLo, j);
				const float rlo = b[kRMin][cc], rhi = b[kRMax][cc];
				const bool rOk = __builtin_isunordered(rlo, rhi) || (srHi >= rlo && srLo <= rhi);
				const bool cand = c < clusters && sdHi >= b[kDMin][cc] && sdLo <= b[kDMax][cc] && saHi >= b[kAMin][cc] &&
				                  saLo <= b[kAMax][cc] && rOk;
				const uint64_t mask = __builtin_amdgcn_ballot_w64(cand);
				if (mask != 0) {
					found = cb + __builtin_ctzll(mask);
				}
			}
			if (found < 0 && clusters == kRadarMaxClusters) {
				errors |= kRadarErrTooManyClusters;
				break;
			}
			const float d = lanef(pd, j), a = lanef(pa, j), r = lanef(pr, j), e = lanef(pe, j);
			const bool fresh = found < 0;
			const int c = fresh ? clusters : found;
			const int rank = fresh ? 0 : __builtin_amdgcn_readfirstlane(lane == 0 ? size[c] : 0);
			if (c < 64) { // a new cluster held in registers
				if (lane == c) {
					addPoint(rb, rb, true, d, a, r, e);
					rsize = 1;
				}
			} else if (lane == 0) {
				float lb[kBounds];
				for (int q = 0; q < kBounds; ++q) {
					lb[q] = b[q][c];
				}
				addPoint(lb, lb, fresh, d, a, r, e);
				for (int q = 0; q < kBounds; ++q) {
					b[q][c] = lb[q];
				}
				size[c] = rank + 1;
			}
			if (fresh) {
				if (lane == 0) {
					segHead[c] = segTail[c] = c;
					segNext[c] = -1;
				}
				++clusters;
				held = clusters >= 64 ? ~0ull : (1ull << clusters) - 1;
			}
			myCluster = lane == j ? c : myCluster;
			myRank = lane == j ? rank : myRank;
			pending &= pending - 1;
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

	// Cluster sums, one cluster at a time across the wave: lane l takes list positions l, l + 64, ... of the cluster's
	// members (four positions' loads issued together), and the 64 partial sums meet in a butterfly. The centre is the
	// first point in list order at the smallest |azimuth - mean| + |elevation - mean| (the first member when no
	// distance is below FLT_MAX): each lane keeps its own first minimum, the wave takes the least (distance,
	// position) pair. One lane per cluster walking its list would chain two dependent global loads per member.
	const float4* terms = reinterpret_cast<const float4*>(it.terms);
	for (int c = 0; c < clusters && errors == 0; ++c) {
		const float meanA = (b[kAMin][c] + b[kAMax][c]) / 2.0f;
		const float meanE = (b[kEMin][c] + b[kEMax][c]) / 2.0f;
		Cf au{0.0f, 0.0f}, ar{0.0f, 0.0f};
		float best = FLT_MAX;
		int bestPos = INT_MAX, bestPt = -1;
		int pos = 0; // list position of the segment's first member
		for (int o = segHead[c]; o >= 0; o = segNext[o]) {
			const int first = offset[o], count = size[o];
			for (int k0 = 0; k0 < count; k0 += 4 * 64) {
				int pt[4];
				float4 t[4];
				float v[4];
				for (int u = 0; u < 4; ++u) {
					const int k = k0 + u * 64 + lane;
					pt[u] = k < count ? members[first + k] : -1;
				}
				for (int u = 0; u < 4; ++u) {
					if (pt[u] >= 0) {
						t[u] = terms[pt[u]];
						v[u] = fabsf(azimuth[pt[u]] - meanA) + fabsf(elevation[pt[u]] - meanE);
					}
				}
				for (int u = 0; u < 4; ++u) {
					if (pt[u] >= 0) {
						au = au + Cf{t[u].x, t[u].y};
						ar = ar + Cf{t[u].z, t[u].w};
						if (v[u] < best) {
							best = v[u];
							bestPos = pos + k0 + u * 64 + lane;
							bestPt = pt[u];
						}
					}
				}
			}
			pos += count;
		}
		// Butterfly: both partners add the same two values, so every lane ends with the same sums.
		for (int m = 1; m < 64; m <<= 1) {
			au.re += __shfl_xor(au.re, m);
			au.im += __shfl_xor(au.im, m);
			ar.re += __shfl_xor(ar.re, m);
			ar.im += __shfl_xor(ar.im, m);
			const float ob = __shfl_xor(best, m);
			const int op = __shfl_xor(bestPos, m), opt = __shfl_xor(bestPt, m);
			if (ob < best || (ob == best && op < bestPos)) {
				best = ob;
				bestPos = op;
				bestPt = opt;
			}
		}
		if (lane == 0) {
			sums[c] = make_float4(au.re, au.im, ar.re, ar.im);
			centres[c] = bestPt >= 0 ? bestPt : members[offset[segHead[c]]];
		}
	}
	__syncthreads();

	const RadarPostParams& p = it.p;
	const float lambda = 299792458.0f / p.frequency;
	const float lambdaSqrtDbsm = 10.0f * log10f(lambda * lambda);
	const float fourPiCubedDb = 10.0f * log10f(powf(4.0f * kPi, 3.0f));
	bool rcsNan = false;
	for (int c = lane; c < clusters && errors == 0; c += 64) {
		const float4 s = sums[c];
		const Cf au{s.x, s.y}, ar{s.z, s.w};
		const int centre = centres[c];
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
