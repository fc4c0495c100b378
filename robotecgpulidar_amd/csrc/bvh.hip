// LBVH construction and refit on gfx950.
//
// Replaces the reference's OptiX GAS/IAS builds (src/scene/GASBuilder.hpp:31-100, src/scene/Scene.cpp:131-190),
// whose algorithms are closed source. The builder here is a Morton-code LBVH:
//   1. primitive boxes -> centroid bounds (ordered-int atomics),
//   2. 63-bit Morton keys (21 bits per axis),
//   3. stable LSD radix sort of (key, primitive) pairs, 8 bits per pass; each tile ranks its keys in LDS
//      with wave64 ballot matching,
//   4. Karras 2012 ("Maximizing parallelism in the construction of BVHs, octrees and k-d trees") topology,
//   5. bottom-up box fit with acquire/release counters; the same pass is the refit path for moved geometry.
// Node layout: BvhNode (two child boxes + two child references, 64 B).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <utility>

#include "device_types.hpp"
#include "kernels.hpp"

namespace rgl::gpu {

namespace {

constexpr int kBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems; // 4096 keys per tile
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ uint32_t floatToOrdered(float f)
{
	uint32_t b = __float_as_uint(f);
	return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float orderedToFloat(uint32_t o)
{
	uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
	return __uint_as_float(b);
}

__global__ void kInitBounds(uint32_t* bounds)
{
	if (threadIdx.x < 3) {
		bounds[threadIdx.x] = 0xffffffffu;     // min
		bounds[3 + threadIdx.x] = 0u;          // max
	}
}

__global__ void kCentroidBounds(int n, const Aabb* boxes, uint32_t* bounds)
{
	__shared__ float smin[3][kBlock];
	__shared__ float smax[3][kBlock];
	float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
	float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
		const Aabb b = boxes[i];
		for (int a = 0; a < 3; ++a) {
			const float c = 0.5f * (b.lo[a] + b.hi[a]);
			lo[a] = fminf(lo[a], c);
			hi[a] = fmaxf(hi[a], c);
		}
	}
	for (int a = 0; a < 3; ++a) {
		smin[a][threadIdx.x] = lo[a];
		smax[a][threadIdx.x] = hi[a];
	}
	__syncthreads();
	for (int s = kBlock / 2; s > 0; s >>= 1) {
		if (threadIdx.x < s) {
			for (int a = 0; a < 3; ++a) {
				smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
				smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
			}
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		for (int a = 0; a < 3; ++a) {
			atomicMin(&bounds[a], floatToOrdered(smin[a][0]));
			atomicMax(&bounds[3 + a], floatToOrdered(smax[a][0]));
		}
	}
}

__device__ __forceinline__ uint64_t expandBits21(uint32_t v)
{
	uint64_t x = v & 0x1fffffu;
	x = (x | (x << 32)) & 0x1f00000000ffffull;
	x = (x | (x << 16)) & 0x1f0000ff0000ffull;
	x = (x | (x << 8)) & 0x100f00f00f00f00full;
	x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
	x = (x | (x << 2)) & 0x1249249249249249ull;
	return x;
}

// Sort keys: 63-bit Morton code of the box centroid in the centroid bounds. With separateLarge, boxes whose diagonal
// is at least half the diagonal of the centroid bounds (ground planes, terrain, anything spanning the scene) get
// bit 63 set, so the first split of the tree separates them from the rest and their boxes do not inflate every
// node on the path of their centroid through the spatial subdivision of the other primitives.
__global__ void kMorton(int n, const Aabb* boxes, const uint32_t* bounds, uint64_t* keys, int32_t* vals, int separateLarge)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	float lo[3], ext[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = orderedToFloat(bounds[a]);
		ext[a] = orderedToFloat(bounds[3 + a]) - lo[a];
	}
	const Aabb b = boxes[i];
	uint32_t q[3];
	for (int a = 0; a < 3; ++a) {
		const float c = 0.5f * (b.lo[a] + b.hi[a]);
		float t = ext[a] > 0.0f ? (c - lo[a]) / ext[a] : 0.5f;
		t = fminf(fmaxf(t, 0.0f), 1.0f);
		q[a] = static_cast<uint32_t>(fminf(t * 2097152.0f, 2097151.0f));
	}
	uint64_t key = (expandBits21(q[0]) << 2) | (expandBits21(q[1]) << 1) | expandBits21(q[2]);
	if (separateLarge) {
		const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
		const float sceneDiag2 = ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2];
		if (dx * dx + dy * dy + dz * dz >= 0.25f * sceneDiag2) {
			key |= 1ull << 63;
		}
	}
	keys[i] = key;
	vals[i] = i;
}

// ---------------------------------------------------------------------------------------- radix sort
__global__ void kSortHistogram(int n, const uint64_t* keys, int shift, int numTiles, uint32_t* hist)
{
	__shared__ uint32_t sh[256];
	sh[threadIdx.x] = 0;
	__syncthreads();
	const int base = blockIdx.x * kSortTile;
	for (int r = 0; r < kSortItems; ++r) {
		const int i = base + r * kBlock + threadIdx.x;
		if (i < n) {
			atomicAdd(&sh[(keys[i] >> shift) & 0xffu], 1u);
		}
	}
	__syncthreads();
	hist[threadIdx.x * numTiles + blockIdx.x] = sh[threadIdx.x];
}

// Exclusive scan of the digit-major histogram in one block of 1024 threads.
__global__ void kSortScan(int count, uint32_t* hist)
{
	__shared__ uint32_t partial[1024];
	const int per = (count + 1023) / 1024;
	const int beg = threadIdx.x * per;
	const int end = min(beg + per, count);
	uint32_t sum = 0;
	for (int i = beg; i < end; ++i) {
		sum += hist[i];
	}
	partial[threadIdx.x] = sum;
	__syncthreads();
	for (int off = 1; off < 1024; off <<= 1) {
		uint32_t v = threadIdx.x >= off ? partial[threadIdx.x - off] : 0u;
		__syncthreads();
		partial[threadIdx.x] += v;
		__syncthreads();
	}
	uint32_t run = partial[threadIdx.x] - sum; // exclusive prefix of this chunk
	for (int i = beg; i < end; ++i) {
		const uint32_t v = hist[i];
		hist[i] = run;
		run += v;
	}
}

__global__ void kSortScatter(int n, const uint64_t* keysIn, const int32_t* valsIn, uint64_t* keysOut,
                             int32_t* valsOut, int shift, int numTiles, const uint32_t* hist)
{
	__shared__ uint32_t gOff[256];
	__shared__ uint32_t base[256];
	__shared__ uint32_t cnt[kWaves][256];
	const int tid = threadIdx.x;
	const int wave = tid >> 6;
	const int lane = tid & 63;
	gOff[tid] = hist[tid * numTiles + blockIdx.x];
	base[tid] = 0;
	const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	const int tileBase = blockIdx.x * kSortTile;
	for (int r = 0; r < kSortItems; ++r) {
		for (int w = 0; w < kWaves; ++w) {
			cnt[w][tid] = 0;
		}
		__syncthreads();
		const int i = tileBase + r * kBlock + tid;
		const bool valid = i < n;
		uint64_t key = valid ? keysIn[i] : 0ull;
		int32_t val = valid ? valsIn[i] : 0;
		const uint32_t digit = valid ? static_cast<uint32_t>((key >> shift) & 0xffu) : 0u;
		uint64_t peers = __ballot(valid);
		for (int b = 0; b < 8; ++b) {
			const bool bit = (digit >> b) & 1u;
			const uint64_t m = __ballot(bit);
			peers &= bit ? m : ~m;
		}
		const uint32_t rankInWave = static_cast<uint32_t>(__popcll(peers & lanesBelow));
		if (valid && rankInWave == 0) {
			cnt[wave][digit] = static_cast<uint32_t>(__popcll(peers));
		}
		__syncthreads();
		{
			uint32_t run = base[tid];
			for (int w = 0; w < kWaves; ++w) {
				const uint32_t c = cnt[w][tid];
				cnt[w][tid] = run;
				run += c;
			}
			base[tid] = run;
		}
		__syncthreads();
		if (valid) {
			const uint32_t pos = gOff[digit] + cnt[wave][digit] + rankInWave;
			keysOut[pos] = key;
			valsOut[pos] = val;
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------- topology
__device__ __forceinline__ int prefixDelta(int n, const uint64_t* keys, int i, int j)
{
	if (j < 0 || j >= n) {
		return -1;
	}
	const uint64_t a = keys[i], b = keys[j];
	if (a == b) {
		return 64 + __clz(static_cast<uint32_t>(i ^ j));
	}
	return __clzll(a ^ b);
}

// parentSlot[x] = (parent internal node << 1) | child slot; leaves at x in [0, n), internal node k at n + k.
__global__ void kKarras(int n, const uint64_t* keys, const int32_t* leafPrim, int leafRefIsPrim, BvhNode* nodes,
                        int32_t* parentSlot)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n - 1) {
		return;
	}
	const int d = (prefixDelta(n, keys, i, i + 1) - prefixDelta(n, keys, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = prefixDelta(n, keys, i, i - d);
	int lmax = 2;
	while (prefixDelta(n, keys, i, i + lmax * d) > dmin) {
		lmax <<= 1;
	}
	int l = 0;
	for (int t = lmax >> 1; t >= 1; t >>= 1) {
		if (prefixDelta(n, keys, i, i + (l + t) * d) > dmin) {
			l += t;
		}
	}
	const int j = i + l * d;
	const int dnode = prefixDelta(n, keys, i, j);
	int s = 0;
	int t = l;
	do {
		t = (t + 1) >> 1;
		if (prefixDelta(n, keys, i, i + (s + t) * d) > dnode) {
			s += t;
		}
	} while (t > 1);
	const int gamma = i + s * d + min(d, 0);
	const int lo = min(i, j), hi = max(i, j);
	int c0, c1;
	if (lo == gamma) {
		c0 = ~(leafRefIsPrim ? leafPrim[gamma] : gamma);
		parentSlot[gamma] = (i << 1) | 0;
	} else {
		c0 = gamma;
		parentSlot[n + gamma] = (i << 1) | 0;
	}
	if (hi == gamma + 1) {
		c1 = ~(leafRefIsPrim ? leafPrim[gamma + 1] : gamma + 1);
		parentSlot[gamma + 1] = (i << 1) | 1;
	} else {
		c1 = gamma + 1;
		parentSlot[n + gamma + 1] = (i << 1) | 1;
	}
	nodes[i].child[0] = c0;
	nodes[i].child[1] = c1;
	nodes[i].pad[0] = 0;
	nodes[i].pad[1] = 0;
	if (i == 0) {
		parentSlot[n] = -1;
	}
}

__global__ void kSingleLeafRoot(const Aabb* boxes, int leafRefIsPrim, BvhNode* nodes, int32_t* height)
{
	BvhNode r;
	for (int a = 0; a < 3; ++a) {
		r.f[a] = boxes[0].lo[a];
		r.f[3 + a] = boxes[0].hi[a];
		r.f[6 + a] = FLT_MAX;
		r.f[9 + a] = -FLT_MAX;
	}
	r.child[0] = ~0;
	r.child[1] = ~0;
	r.pad[0] = r.pad[1] = 0;
	nodes[0] = r;
	height[0] = 1;
	(void) leafRefIsPrim;
}

__global__ void kClearInts(int n, int32_t* p)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) {
		p[i] = 0;
	}
}

__device__ __forceinline__ void storeSlot(BvhNode* nodes, int slotCode, const float* box)
{
	float* f = nodes[slotCode >> 1].f + 6 * (slotCode & 1);
	for (int a = 0; a < 6; ++a) {
		f[a] = box[a];
	}
}

// Bottom-up fit. One thread per leaf; the second thread to reach an internal node unions its two child boxes
// and continues upward. Counters use acquire/release at agent scope so the sibling's box is visible.
// `height[k]` receives the height of internal node k (leaves have height 0); height[0] is the tree depth.
__global__ void kFit(int n, const Aabb* boxes, const int32_t* leafPrim, const int32_t* parentSlot, BvhNode* nodes,
                     int32_t* counters, int32_t* height)
{
	const int leaf = blockIdx.x * kBlock + threadIdx.x;
	if (leaf >= n) {
		return;
	}
	const Aabb b = boxes[leafPrim[leaf]];
	float box[6] = {b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]};
	int code = parentSlot[leaf];
	while (code >= 0) {
		storeSlot(nodes, code, box);
		const int p = code >> 1;
		const int prev = __hip_atomic_fetch_add(&counters[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		if (prev == 0) {
			return; // sibling not done yet; it will carry the union upward
		}
		// Both children are fitted: union them. Loads after the acquire see the sibling's stores.
		const float* f = nodes[p].f;
		for (int a = 0; a < 3; ++a) {
			box[a] = fminf(f[a], f[6 + a]);
			box[3 + a] = fmaxf(f[3 + a], f[9 + a]);
		}
		const int c0 = nodes[p].child[0], c1 = nodes[p].child[1];
		const int hc0 = c0 >= 0 ? height[c0] : 0;
		const int hc1 = c1 >= 0 ? height[c1] : 0;
		height[p] = 1 + max(hc0, hc1);
		code = parentSlot[n + p];
	}
}

__global__ void kTriangleBoxes(int n, const Vec3f* v, const Vec3i* idx, Aabb* out)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const Vec3i t = idx[i];
	const Vec3f a = v[t.x], b = v[t.y], c = v[t.z];
	Aabb r;
	r.lo[0] = fminf(a.x, fminf(b.x, c.x));
	r.lo[1] = fminf(a.y, fminf(b.y, c.y));
	r.lo[2] = fminf(a.z, fminf(b.z, c.z));
	r.hi[0] = fmaxf(a.x, fmaxf(b.x, c.x));
	r.hi[1] = fmaxf(a.y, fmaxf(b.y, c.y));
	r.hi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
	out[i] = r;
}

// Animated vertices: the displacement since the previous step, then the new positions.
__global__ void kUpdateVertices(int n, const Vec3f* fresh, Vec3f* vertices, Vec3f* displacement)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const Vec3f f = fresh[i];
	displacement[i] = f - vertices[i];
	vertices[i] = f;
}

// Linear blend skinning with up to four bones per vertex, summed in bone-slot order.
__global__ void kSkinVertices(int n, const Vec3f* rest, const BoneWeights* weights, const Mat3x4f* matrices,
                              Vec3f* vertices, Vec3f* displacement)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const Vec3f r = rest[i];
	const BoneWeights w = weights[i];
	Vec3f f = (matrices[w.bones[0]] * r) * w.weights[0];
	f = f + (matrices[w.bones[1]] * r) * w.weights[1];
	f = f + (matrices[w.bones[2]] * r) * w.weights[2];
	f = f + (matrices[w.bones[3]] * r) * w.weights[3];
	displacement[i] = f - vertices[i];
	vertices[i] = f;
}

__global__ void kGatherTris(int n, const int32_t* leafPrim, const Vec3f* v, const Vec3i* idx, LeafTri* out)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const int p = leafPrim[i];
	const Vec3i t = idx[p];
	const Vec3f a = v[t.x], b = v[t.y], c = v[t.z];
	LeafTri r;
	r.ax = a.x;
	r.ay = a.y;
	r.az = a.z;
	r.prim = p;
	r.bx = b.x;
	r.by = b.y;
	r.bz = b.z;
	r.pad0 = 0.0f;
	r.cx = c.x;
	r.cy = c.y;
	r.cz = c.z;
	r.pad1 = 0.0f;
	out[i] = r;
}

// World box of each instance: the 8 corners of its BLAS root box through objectToWorld, widened by a relative
// 2^-20 so that FP32 rounding of the corners never makes the box smaller than the true image.
__global__ void kInstanceBoxes(int n, const InstanceShade* shades, const InstanceRec* recs, Aabb* out)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const float* f = shades[i].blas2[0].f;
	float lo[3], hi[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = fminf(f[a], f[6 + a]);
		hi[a] = fmaxf(f[3 + a], f[9 + a]);
	}
	const Mat3x4f& m = shades[i].objectToWorld;
	Aabb r;
	for (int a = 0; a < 3; ++a) {
		r.lo[a] = FLT_MAX;
		r.hi[a] = -FLT_MAX;
	}
	for (int c = 0; c < 8; ++c) {
		const Vec3f p{(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
		const Vec3f w = m * p;
		for (int a = 0; a < 3; ++a) {
			r.lo[a] = fminf(r.lo[a], w[a]);
			r.hi[a] = fmaxf(r.hi[a], w[a]);
		}
	}
	for (int a = 0; a < 3; ++a) {
		const float pad = (fabsf(r.lo[a]) + fabsf(r.hi[a]) + (r.hi[a] - r.lo[a])) * 9.5367431640625e-7f + 1e-30f;
		r.lo[a] -= pad;
		r.hi[a] += pad;
	}
	out[i] = r;
}

// ------------------------------------------------------------------------------------------------ 4-wide nodes
// The 4-wide tree keeps every even-depth internal node of the binary LBVH; a kept node's children are the children
// of its two binary children (or those children themselves when they are leaves). The binary tree's depth is
// halved and the set of leaves under every node is unchanged.
constexpr int kScanBlock = 1024;

__global__ void kEvenDepthFlags(int inner, int n, const int32_t* parentSlot, int32_t* flags)
{
	const int k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= inner) {
		return;
	}
	int depth = 0;
	int code = n > 1 ? parentSlot[n + k] : -1;
	while (code >= 0) {
		++depth;
		code = parentSlot[n + (code >> 1)];
	}
	flags[k] = (depth & 1) == 0 ? 1 : 0;
}

// Exclusive scan, three passes: per-block scan (1024 elements), scan of the block sums, add block offsets.
__device__ int blockExclusiveScan(int v, int* total)
{
	__shared__ int waveSums[kScanBlock / 64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int x = v;
	for (int off = 1; off < 64; off <<= 1) {
		const int y = __shfl_up(x, off, 64);
		if (lane >= off) {
			x += y;
		}
	}
	if (lane == 63) {
		waveSums[wave] = x;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		int run = 0;
		for (int w = 0; w < kScanBlock / 64; ++w) {
			const int c = waveSums[w];
			waveSums[w] = run;
			run += c;
		}
		*total = run;
	}
	__syncthreads();
	return waveSums[wave] + x - v;
}

__global__ void __launch_bounds__(kScanBlock) kScanLocal(int n, const int32_t* in, int32_t* out, int32_t* blockSums)
{
	__shared__ int total;
	const int i = blockIdx.x * kScanBlock + threadIdx.x;
	const int v = i < n ? in[i] : 0;
	const int e = blockExclusiveScan(v, &total);
	if (i < n) {
		out[i] = e;
	}
	if (threadIdx.x == 0) {
		blockSums[blockIdx.x] = total;
	}
}

__global__ void __launch_bounds__(kScanBlock) kScanSums(int nb, int32_t* sums, int32_t* grandTotal)
{
	__shared__ int total;
	int carry = 0;
	for (int base = 0; base < nb; base += kScanBlock) {
		const int i = base + threadIdx.x;
		const int v = i < nb ? sums[i] : 0;
		const int e = blockExclusiveScan(v, &total);
		if (i < nb) {
			sums[i] = carry + e;
		}
		carry += total;
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		*grandTotal = carry;
	}
}

__global__ void kScanAdd(int n, int32_t* out, const int32_t* blockSums)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) {
		out[i] += blockSums[i / kScanBlock];
	}
}

__device__ __forceinline__ float pow2f(int e) { return __int_as_float((e + 127) << 23); }

// Smallest exponent e (scale 2^e) with lo + levels * 2^e >= hi in exact arithmetic (lo is an FP32 value; the sum is
// evaluated in FP64, where it is exact for the exponents that matter here).
__device__ int quantExponent(float lo, double hi, double levels)
{
	if (!(hi > lo)) {
		return -126;
	}
	int e = max(-126, min(127, ilogb((hi - lo) / levels)));
	while (e < 127 && static_cast<double>(lo) + levels * pow2f(e) < hi) {
		++e;
	}
	while (e > -126 && static_cast<double>(lo) + levels * pow2f(e - 1) >= hi) {
		--e;
	}
	return e;
}

// Quantized child bounds of one 4-wide node (Bvh4QNode with Q = uint8_t, Bvh4HNode with Q = uint16_t). Each decoded
// bound lies outside the child's FP32 bound by at least a margin of 2^-20 of the node's extent on that axis (checked
// in exact arithmetic), which covers the rounding of the traversal's folded decode (levels * 2^e < 2 * extent).
template<typename Node, typename Q>
__device__ void quantizeChildren(Node& o, const float (&clo)[4][3], const float (&chi)[4][3], int cnt)
{
	constexpr int kLevels = sizeof(Q) == 1 ? 255 : 65535;
	for (int a = 0; a < 3; ++a) {
		float lo0 = FLT_MAX, hi0 = -FLT_MAX;
		for (int c = 0; c < cnt; ++c) {
			lo0 = fminf(lo0, clo[c][a]);
			hi0 = fmaxf(hi0, chi[c][a]);
		}
		const double margin = ldexp(static_cast<double>(hi0) - lo0, -20);
		float lo = static_cast<float>(static_cast<double>(lo0) - margin);
		if (static_cast<double>(lo) > static_cast<double>(lo0) - margin) {
			lo = nextafterf(lo, -INFINITY);
		}
		const int e = quantExponent(lo, static_cast<double>(hi0) + margin, kLevels);
		const double sc = pow2f(e);
		o.origin[a] = lo;
		o.exp[a] = static_cast<uint8_t>(e + 127);
		for (int c = 0; c < 4; ++c) {
			if (c >= cnt) {
				o.qlo[a][c] = static_cast<Q>(kLevels);
				o.qhi[a][c] = 0;
				continue;
			}
			const double want = static_cast<double>(clo[c][a]) - margin, wantHi = static_cast<double>(chi[c][a]) + margin;
			int ql = static_cast<int>(floor((want - lo) / sc));
			ql = max(0, min(kLevels, ql));
			while (ql > 0 && static_cast<double>(lo) + ql * sc > want) {
				--ql;
			}
			int qh = static_cast<int>(ceil((wantHi - lo) / sc));
			qh = max(0, min(kLevels, qh));
			while (qh < kLevels && static_cast<double>(lo) + qh * sc < wantHi) {
				++qh;
			}
			o.qlo[a][c] = static_cast<Q>(ql);
			o.qhi[a][c] = static_cast<Q>(qh);
		}
	}
}

template<int kFmt>
__global__ void kCollapse4(int inner, const BvhNode* nodes, const int32_t* flags, const int32_t* rank, void* out)
{
	const int k = blockIdx.x * kBlock + threadIdx.x;
	if (k >= inner || flags[k] == 0) {
		return;
	}
	float clo[4][3], chi[4][3];
	int ref[4] = {0, 0, 0, 0};
	int cnt = 0;
	auto add = [&](const float* f, int r) {
		if (f[0] > f[3] || f[1] > f[4] || f[2] > f[5]) {
			return; // empty slot (inverted box)
		}
		for (int a = 0; a < 3; ++a) {
			clo[cnt][a] = f[a];
			chi[cnt][a] = f[3 + a];
		}
		ref[cnt] = r;
		++cnt;
	};
	const BvhNode b = nodes[k];
	for (int s = 0; s < 2; ++s) {
		const int c = b.child[s];
		if (c < 0) {
			add(b.f + 6 * s, c);
		} else {
			const BvhNode g = nodes[c];
			for (int t = 0; t < 2; ++t) {
				const int gc = g.child[t];
				add(g.f + 6 * t, gc < 0 ? gc : rank[gc]);
			}
		}
	}
	const int slot = rank[k];
	if constexpr (kFmt == kBvh4F) {
		Bvh4FNode o;
		for (int c = 0; c < 4; ++c) {
			for (int a = 0; a < 3; ++a) {
				o.lo[a][c] = c < cnt ? clo[c][a] : INFINITY;
				o.hi[a][c] = c < cnt ? chi[c][a] : INFINITY;
			}
			o.child[c] = c < cnt ? ref[c] : ref[0];
		}
		o.count = cnt;
		o.pad[0] = o.pad[1] = o.pad[2] = 0;
		static_cast<Bvh4FNode*>(out)[slot] = o;
		return;
	}
	auto finish = [&](auto& o) {
		for (int c = 0; c < 4; ++c) {
			o.child[c] = c < cnt ? ref[c] : ref[0];
		}
		o.count = static_cast<uint8_t>(cnt);
	};
	if constexpr (kFmt == kBvh4Q) {
		Bvh4QNode o;
		quantizeChildren<Bvh4QNode, uint8_t>(o, clo, chi, cnt);
		finish(o);
		o.pad[0] = o.pad[1] = 0;
		static_cast<Bvh4QNode*>(out)[slot] = o;
	} else {
		Bvh4HNode o;
		quantizeChildren<Bvh4HNode, uint16_t>(o, clo, chi, cnt);
		finish(o);
		static_cast<Bvh4HNode*>(out)[slot] = o;
	}
}

// ---------------------------------------------------------------------------------------- PLOC topology
// Agglomerative clustering (PLOC: Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
// Hierarchy Construction", IEEE TVCG 2018) over the Morton-sorted leaves, in one workgroup. Each round every
// cluster finds the neighbour within kPlocRadius positions whose union box has the smallest surface area; mutual
// nearest neighbours merge (the lower position creates the node and keeps the position), and the survivors are
// compacted in order. The result is a binary topology in the LBVH convention (root = node 0, child refs >= 0 are
// nodes, < 0 leaves; parentSlot[leaf] and parentSlot[n + node] = parent << 1 | slot); boxes and heights are filled by
// the same bottom-up fit as the LBVH. Surface-area-driven merging keeps scene-spanning boxes near the root and
// groups the rest by locality, which the Morton radix tree does only along its fixed splits.
constexpr int kPlocBlock = 1024;
constexpr int kPlocRadius = 16;
constexpr int kPlocMax = 16384;

__device__ __forceinline__ float halfArea(const Aabb& b)
{
	const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
	return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ Aabb unite(const Aabb& a, const Aabb& b)
{
	Aabb r;
	for (int k = 0; k < 3; ++k) {
		r.lo[k] = fminf(a.lo[k], b.lo[k]);
		r.hi[k] = fmaxf(a.hi[k], b.hi[k]);
	}
	return r;
}

// Exclusive scan of v over the workgroup's items [0, count) (each thread owns a contiguous chunk); returns the total.
__device__ int plocScan(int32_t* v, int count, int32_t* lds)
{
	const int per = (count + kPlocBlock - 1) / kPlocBlock;
	const int b = threadIdx.x * per, e = min(b + per, count);
	int sum = 0;
	for (int i = b; i < e; ++i) {
		sum += v[i];
	}
	// Inclusive scan of the per-thread sums: within each wave by shuffles, then over the 16 wave totals.
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int incl = sum;
	for (int off = 1; off < 64; off <<= 1) {
		const int y = __shfl_up(incl, off, 64);
		incl += lane >= off ? y : 0;
	}
	if (lane == 63) {
		lds[wave] = incl;
	}
	__syncthreads();
	if (threadIdx.x < 64) {
		const int w = threadIdx.x < kPlocBlock / 64 ? lds[threadIdx.x] : 0;
		int x = w;
		for (int off = 1; off < 64; off <<= 1) {
			const int y = __shfl_up(x, off, 64);
			x += threadIdx.x >= off ? y : 0;
		}
		if (threadIdx.x < kPlocBlock / 64) {
			lds[kPlocBlock / 64 + threadIdx.x] = x - w; // exclusive wave offsets
		}
		if (threadIdx.x == kPlocBlock / 64 - 1) {
			lds[2 * (kPlocBlock / 64)] = x; // total
		}
	}
	__syncthreads();
	const int total = lds[2 * (kPlocBlock / 64)];
	int run = lds[kPlocBlock / 64 + wave] + incl - sum;
	for (int i = b; i < e; ++i) {
		const int x = v[i];
		v[i] = run;
		run += x;
	}
	__syncthreads();
	return total;
}

// scratch: box[2][n], ref[2][n], nn[n], merge[n], keep[n]. Leaf clusters carry ref = ~leafIndex.
__global__ void __launch_bounds__(kPlocBlock) kPloc(int n, const Aabb* boxes, const int32_t* leafPrim, int leafRefIsPrim,
                                                    BvhNode* nodes, int32_t* parentSlot, Aabb* box0, Aabb* box1,
                                                    int32_t* ref0, int32_t* ref1, int32_t* nn, int32_t* merge,
                                                    int32_t* keep)
{
	__shared__ int32_t lds[kPlocBlock];
	for (int i = threadIdx.x; i < n; i += kPlocBlock) {
		box0[i] = boxes[leafPrim[i]];
		ref0[i] = ~i;
	}
	__syncthreads();
	Aabb* cb = box0;
	Aabb* nb = box1;
	int32_t* cr = ref0;
	int32_t* nr = ref1;
	int count = n;
	int next = n - 2; // index of the next node created (nodes are numbered downwards, so the root is node 0)
	while (count > 1) {
		for (int i = threadIdx.x; i < count; i += kPlocBlock) {
			const Aabb bi = cb[i];
			float best = INFINITY;
			int arg = -1;
			const int lo = max(0, i - kPlocRadius), hi = min(count - 1, i + kPlocRadius);
			for (int j = lo; j <= hi; ++j) {
				if (j == i) {
					continue;
				}
				float a = halfArea(unite(bi, cb[j]));
				a = isnan(a) ? INFINITY : a;
				if (a < best || arg < 0) { // ties keep the lower position, so some pair is always mutual
					best = a;
					arg = j;
				}
			}
			nn[i] = arg;
		}
		__syncthreads();
		for (int i = threadIdx.x; i < count; i += kPlocBlock) {
			const int j = nn[i];
			const bool mutual = nn[j] == i;
			merge[i] = (mutual && i < j) ? 1 : 0;
			keep[i] = (mutual && i > j) ? 0 : 1;
		}
		__syncthreads();
		const int merges = plocScan(merge, count, lds);
		const int kept = plocScan(keep, count, lds);
		for (int i = threadIdx.x; i < count; i += kPlocBlock) {
			const int j = nn[i];
			const bool mutual = nn[j] == i;
			if (mutual && i > j) {
				continue; // merged into j's node
			}
			const int pos = keep[i];
			if (mutual) {
				const int node = next - merge[i];
				const int c[2] = {cr[i], cr[j]};
				for (int slot = 0; slot < 2; ++slot) {
					const int code = (node << 1) | slot;
					if (c[slot] < 0) {
						const int leaf = ~c[slot];
						parentSlot[leaf] = code;
						nodes[node].child[slot] = ~(leafRefIsPrim ? leafPrim[leaf] : leaf);
					} else {
						parentSlot[n + c[slot]] = code;
						nodes[node].child[slot] = c[slot];
					}
				}
				nodes[node].pad[0] = 0;
				nodes[node].pad[1] = 0;
				nb[pos] = unite(cb[i], cb[j]);
				nr[pos] = node;
			} else {
				nb[pos] = cb[i];
				nr[pos] = cr[i];
			}
		}
		__syncthreads();
		next -= merges;
		count = kept;
		Aabb* tb = cb;
		cb = nb;
		nb = tb;
		int32_t* tr = cr;
		cr = nr;
		nr = tr;
	}
	if (threadIdx.x == 0) {
		parentSlot[n] = -1;
	}
}

inline int blocksFor(int n) { return (n + kBlock - 1) / kBlock; }

// Trace order of lidar rays: rays are binned into square tiles of their direction's (azimuth, elevation) in the
// sensor frame, tiles of kTraceTileCells x kTraceTileCells cells of the mean angular area per ray (so a tile holds
// about that many rays), listed elevation band by elevation band; a stable sort by tile keeps index order inside a
// tile. A channel-by-channel lidar pattern then becomes runs of consecutive rays of a few adjacent channels.
#ifndef RGL_TRACE_TILE_CELLS
#define RGL_TRACE_TILE_CELLS 6
#endif
constexpr float kTraceTileCells = RGL_TRACE_TILE_CELLS;

__device__ __forceinline__ void rayAngles(const Mat3x4f& r, float& az, float& el)
{
	const float dx = r.rc[0][2], dy = r.rc[1][2], dz = r.rc[2][2];
	az = atan2f(dy, dx);
	el = atan2f(dz, sqrtf(dx * dx + dy * dy));
}

// bounds: [0] min azimuth, [1] min elevation, [3] max azimuth, [4] max elevation (ordered-float encoding).
__global__ void kAngleBounds(int n, const Mat3x4f* rays, uint32_t* bounds)
{
	__shared__ float smin[2][kBlock], smax[2][kBlock];
	float lo[2] = {FLT_MAX, FLT_MAX}, hi[2] = {-FLT_MAX, -FLT_MAX};
	for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
		float a[2];
		rayAngles(rays[i], a[0], a[1]);
		for (int k = 0; k < 2; ++k) {
			lo[k] = fminf(lo[k], a[k]);
			hi[k] = fmaxf(hi[k], a[k]);
		}
	}
	for (int k = 0; k < 2; ++k) {
		smin[k][threadIdx.x] = lo[k];
		smax[k][threadIdx.x] = hi[k];
	}
	__syncthreads();
	for (int st = kBlock / 2; st > 0; st >>= 1) {
		if (threadIdx.x < st) {
			for (int k = 0; k < 2; ++k) {
				smin[k][threadIdx.x] = fminf(smin[k][threadIdx.x], smin[k][threadIdx.x + st]);
				smax[k][threadIdx.x] = fmaxf(smax[k][threadIdx.x], smax[k][threadIdx.x + st]);
			}
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		for (int k = 0; k < 2; ++k) {
			atomicMin(&bounds[k], floatToOrdered(smin[k][0]));
			atomicMax(&bounds[3 + k], floatToOrdered(smax[k][0]));
		}
	}
}

// Key = (elevation band, azimuth block) of the ray's tile; 0 for every ray when the directions span no area (one
// channel or one azimuth: index order is already coherent there).
__global__ void kTileKeys(int n, const Mat3x4f* rays, const uint32_t* bounds, uint64_t* keys, int32_t* vals)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const float az0 = orderedToFloat(bounds[0]), el0 = orderedToFloat(bounds[1]);
	const float azSpan = orderedToFloat(bounds[3]) - az0, elSpan = orderedToFloat(bounds[4]) - el0;
	uint64_t key = 0;
	const float cellArea = azSpan * elSpan / static_cast<float>(n);
	if (cellArea > 0.0f) {
		const float tile = kTraceTileCells * sqrtf(cellArea);
		float az, el;
		rayAngles(rays[i], az, el);
		const uint32_t band = static_cast<uint32_t>(fminf(floorf((el - el0) / tile), 1048575.0f));
		const uint32_t block = static_cast<uint32_t>(fminf(floorf((az - az0) / tile), 1048575.0f));
		key = (static_cast<uint64_t>(band) << 20) | block;
	}
	keys[i] = key;
	vals[i] = i;
}

void radixSortPairs(hipStream_t stream, int n, uint64_t* keys, uint64_t* keysAlt, int32_t* vals, int32_t* valsAlt,
                    uint32_t* hist)
{
	const int tiles = (n + kSortTile - 1) / kSortTile;
	uint64_t* kIn = keys;
	uint64_t* kOut = keysAlt;
	int32_t* vIn = vals;
	int32_t* vOut = valsAlt;
	for (int pass = 0; pass < 8; ++pass) {
		const int shift = pass * 8;
		hipLaunchKernelGGL(kSortHistogram, dim3(tiles), dim3(kBlock), 0, stream, n, kIn, shift, tiles, hist);
		hipLaunchKernelGGL(kSortScan, dim3(1), dim3(1024), 0, stream, 256 * tiles, hist);
		hipLaunchKernelGGL(kSortScatter, dim3(tiles), dim3(kBlock), 0, stream, n, kIn, vIn, kOut, vOut, shift, tiles,
		                   hist);
		std::swap(kIn, kOut);
		std::swap(vIn, vOut);
	}
	// 8 passes: the sorted pairs are back in (keys, vals).
}

} // namespace

size_t directionOrderScratchBytes(int n)
{
	const size_t tiles = (static_cast<size_t>(n) + kSortTile - 1) / kSortTile;
	return 2 * sizeof(uint64_t) * n + sizeof(int32_t) * n + sizeof(uint32_t) * 256 * tiles + 2048;
}

void directionOrder(hipStream_t stream, int n, const Mat3x4f* rays, int32_t* order, void* scratch)
{
	if (n <= 0) {
		return;
	}
	const int tiles = (n + kSortTile - 1) / kSortTile;
	char* s = static_cast<char*>(scratch);
	auto take = [&](size_t bytes) {
		char* p = s;
		s += (bytes + 255) & ~static_cast<size_t>(255);
		return p;
	};
	uint64_t* keys = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * n));
	uint64_t* keysAlt = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * n));
	int32_t* valsAlt = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
	uint32_t* hist = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 256 * tiles));
	uint32_t* bounds = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 8));
	hipLaunchKernelGGL(kInitBounds, dim3(1), dim3(64), 0, stream, bounds);
	hipLaunchKernelGGL(kAngleBounds, dim3(std::min(blocksFor(n), 1024)), dim3(kBlock), 0, stream, n, rays, bounds);
	hipLaunchKernelGGL(kTileKeys, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, rays, bounds, keys, order);
	radixSortPairs(stream, n, keys, keysAlt, order, valsAlt, hist);
}

size_t lbvhScratchBytes(int n)
{
	const size_t tiles = (static_cast<size_t>(n) + kSortTile - 1) / kSortTile;
	size_t bytes = 0;
	bytes += 2 * sizeof(uint64_t) * n;     // keys, keysAlt
	bytes += 2 * sizeof(int32_t) * n;      // vals, valsAlt
	bytes += sizeof(uint32_t) * 256 * tiles;
	bytes += sizeof(uint32_t) * 8;
	if (n <= kPlocMax) {
		bytes += (2 * sizeof(Aabb) + 5 * sizeof(int32_t)) * n; // PLOC clusters (two buffers), nn, merge, keep
	}
	bytes += 8192;                          // alignment slack
	return bytes;
}

void buildLbvh(hipStream_t stream, int n, const Aabb* boxes, bool leafRefIsPrim, BvhNode* nodes, int32_t* leafPrim,
               int32_t* parentSlot, int32_t* counters, int32_t* height, void* scratch)
{
	if (n <= 0) {
		return;
	}
	if (n == 1) {
		hipLaunchKernelGGL(kSingleLeafRoot, dim3(1), dim3(1), 0, stream, boxes, leafRefIsPrim ? 1 : 0, nodes, height);
		hipLaunchKernelGGL(kClearInts, dim3(1), dim3(kBlock), 0, stream, 1, leafPrim);
		return;
	}
	const int tiles = (n + kSortTile - 1) / kSortTile;
	char* s = static_cast<char*>(scratch);
	auto take = [&](size_t bytes) {
		char* p = s;
		s += (bytes + 255) & ~static_cast<size_t>(255);
		return p;
	};
	uint64_t* keys = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * n));
	uint64_t* keysAlt = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * n));
	int32_t* valsAlt = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
	uint32_t* hist = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 256 * tiles));
	uint32_t* bounds = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 8));

	hipLaunchKernelGGL(kInitBounds, dim3(1), dim3(64), 0, stream, bounds);
	const int rb = std::min(blocksFor(n), 1024);
	hipLaunchKernelGGL(kCentroidBounds, dim3(rb), dim3(kBlock), 0, stream, n, boxes, bounds);
	hipLaunchKernelGGL(kMorton, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, boxes, bounds, keys, leafPrim,
	                   leafRefIsPrim ? 1 : 0);

	radixSortPairs(stream, n, keys, keysAlt, leafPrim, valsAlt, hist);
	if (n <= kPlocMax) {
		// Small trees (every TLAS, most meshes): agglomerative topology, built in one workgroup.
		Aabb* b0 = reinterpret_cast<Aabb*>(take(sizeof(Aabb) * n));
		Aabb* b1 = reinterpret_cast<Aabb*>(take(sizeof(Aabb) * n));
		int32_t* r0 = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
		int32_t* r1 = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
		int32_t* nn = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
		int32_t* merge = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
		int32_t* keep = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
		hipLaunchKernelGGL(kPloc, dim3(1), dim3(kPlocBlock), 0, stream, n, boxes, leafPrim, leafRefIsPrim ? 1 : 0, nodes, parentSlot, b0,
		                   b1, r0, r1, nn, merge, keep);
	} else {
		hipLaunchKernelGGL(kKarras, dim3(blocksFor(n - 1)), dim3(kBlock), 0, stream, n, keys, leafPrim,
		                   leafRefIsPrim ? 1 : 0, nodes, parentSlot);
	}
	refitLbvh(stream, n, boxes, leafPrim, parentSlot, nodes, counters, height);
}

void refitLbvh(hipStream_t stream, int n, const Aabb* boxes, const int32_t* leafPrim, const int32_t* parentSlot,
               BvhNode* nodes, int32_t* counters, int32_t* height)
{
	if (n <= 1) {
		if (n == 1) {
			hipLaunchKernelGGL(kSingleLeafRoot, dim3(1), dim3(1), 0, stream, boxes, 0, nodes, height);
		}
		return;
	}
	hipLaunchKernelGGL(kClearInts, dim3(blocksFor(n - 1)), dim3(kBlock), 0, stream, n - 1, counters);
	hipLaunchKernelGGL(kFit, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, boxes, leafPrim, parentSlot, nodes,
	                   counters, height);
}

void triangleBoxes(hipStream_t stream, int n, const Vec3f* v, const Vec3i* idx, Aabb* out)
{
	if (n > 0) {
		hipLaunchKernelGGL(kTriangleBoxes, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, v, idx, out);
	}
}

void gatherTris(hipStream_t stream, int n, const int32_t* leafPrim, const Vec3f* v, const Vec3i* idx, LeafTri* out)
{
	if (n > 0) {
		hipLaunchKernelGGL(kGatherTris, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, leafPrim, v, idx, out);
	}
}

void updateVertices(hipStream_t stream, int n, const Vec3f* fresh, Vec3f* vertices, Vec3f* displacement)
{
	if (n > 0) {
		hipLaunchKernelGGL(kUpdateVertices, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, fresh, vertices, displacement);
	}
}

void skinVertices(hipStream_t stream, int n, const Vec3f* rest, const BoneWeights* weights, const Mat3x4f* matrices,
                  Vec3f* vertices, Vec3f* displacement)
{
	if (n > 0) {
		hipLaunchKernelGGL(kSkinVertices, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, rest, weights, matrices,
		                   vertices, displacement);
	}
}

void instanceBoxes(hipStream_t stream, int n, const InstanceShade* shades, const InstanceRec* recs, Aabb* out)
{
	if (n > 0) {
		hipLaunchKernelGGL(kInstanceBoxes, dim3(blocksFor(n)), dim3(kBlock), 0, stream, n, shades, recs, out);
	}
}

size_t bvh4ScratchBytes(int n)
{
	const size_t inner = static_cast<size_t>(std::max(n - 1, 1));
	const size_t blocks = (inner + kScanBlock - 1) / kScanBlock;
	return ((2 * inner * sizeof(int32_t) + 255) & ~static_cast<size_t>(255)) + (blocks + 64) * sizeof(int32_t);
}

void bvh4Ranks(hipStream_t stream, int n, const int32_t* parentSlot, void* scratch, int32_t* countDev)
{
	const int inner = std::max(n - 1, 1);
	int32_t* flags = static_cast<int32_t*>(scratch);
	int32_t* rank = flags + inner;
	int32_t* blockSums =
	    reinterpret_cast<int32_t*>(static_cast<char*>(scratch) + ((2 * inner * sizeof(int32_t) + 255) & ~static_cast<size_t>(255)));
	const int nb = (inner + kScanBlock - 1) / kScanBlock;
	hipLaunchKernelGGL(kEvenDepthFlags, dim3(blocksFor(inner)), dim3(kBlock), 0, stream, inner, n, parentSlot, flags);
	hipLaunchKernelGGL(kScanLocal, dim3(nb), dim3(kScanBlock), 0, stream, inner, flags, rank, blockSums);
	hipLaunchKernelGGL(kScanSums, dim3(1), dim3(kScanBlock), 0, stream, nb, blockSums, countDev);
	hipLaunchKernelGGL(kScanAdd, dim3(blocksFor(inner)), dim3(kBlock), 0, stream, inner, rank, blockSums);
}

void bvh4Collapse(hipStream_t stream, int n, const BvhNode* nodes, const void* scratch, int format, void* out)
{
	const int inner = std::max(n - 1, 1);
	const int32_t* flags = static_cast<const int32_t*>(scratch);
	const int32_t* rank = flags + inner;
	const dim3 grid(blocksFor(inner)), block(kBlock);
	switch (format) {
		case kBvh4Q: hipLaunchKernelGGL(kCollapse4<kBvh4Q>, grid, block, 0, stream, inner, nodes, flags, rank, out); break;
		case kBvh4H: hipLaunchKernelGGL(kCollapse4<kBvh4H>, grid, block, 0, stream, inner, nodes, flags, rank, out); break;
		default: hipLaunchKernelGGL(kCollapse4<kBvh4F>, grid, block, 0, stream, inner, nodes, flags, rank, out); break;
	}
}

} // namespace rgl::gpu
