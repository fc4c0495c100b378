// Point-cloud kernels: stream compaction, point transform, Gaussian noise, SoA<->AoS formatting, and their batched
// (many graphs per launch) forms.
//
// Compaction replaces thrust::inclusive_scan + one kApplyCompaction launch per field
// (src/gpu/nodeKernels.cu:58-68,424-436; src/graph/CompactByFieldPointsNode.cpp:33-99) with two launches in
// total: per-tile counts, then one pass that finds each tile's offset, ranks the kept elements inside the tile
// with wave ballots and scatters every requested field. Elements are kept in input order (stable), and the kept
// count stays on the device so nothing downstream waits for the host.
// Noise follows src/gpu/gaussianNoiseKernels.cu:20-66 with the seeded stream of noise_rng.hpp.
// Single-graph and batched kernels call the same element functions, so a graph's results do not depend on
// whether it ran alone or in a batch.

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "device_types.hpp"
#include "kernels.hpp"
#include "noise_rng.hpp"
#include "point_ops.hpp"
#include "runtime.hpp"

namespace rgl::gpu {

namespace {

constexpr int kBlock = 256;
// Points per thread of a compaction tile: 8 (tiles of 2048 points) by default; the batch autotuner also tries 4.
constexpr int kItems = 8;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ int liveCount(int capacity, const int32_t* count) { return count ? *count : capacity; }

// ------------------------------------------------------------------------------------------------ compaction
template<int kItems_ = kItems>
__device__ void tileCount(int capacity, const int32_t* countIn, const int32_t* flags, int tile, int32_t* tileCounts)
{
	__shared__ int32_t warpSums[kWaves];
	const int n = liveCount(capacity, countIn);
	const int base = tile * kBlock * kItems_;
	int c = 0;
	for (int k = 0; k < kItems_; ++k) {
		const int i = base + k * kBlock + threadIdx.x;
		c += (i < n && flags[i] != 0) ? 1 : 0;
	}
	for (int off = 32; off > 0; off >>= 1) {
		c += __shfl_down(c, off, 64);
	}
	if ((threadIdx.x & 63) == 0) {
		warpSums[threadIdx.x >> 6] = c;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		int s = 0;
		for (int w = 0; w < kWaves; ++w) {
			s += warpSums[w];
		}
		tileCounts[tile] = s;
	}
}

__device__ __forceinline__ void copyElem(const FieldCopy& f, int src, int dst)
{
	const char* s = f.src + static_cast<size_t>(src) * f.size;
	char* d = f.dst + static_cast<size_t>(dst) * f.size;
	switch (f.size) {
		case 1: *reinterpret_cast<uint8_t*>(d) = *reinterpret_cast<const uint8_t*>(s); break;
		case 2: *reinterpret_cast<uint16_t*>(d) = *reinterpret_cast<const uint16_t*>(s); break;
		case 4: *reinterpret_cast<uint32_t*>(d) = *reinterpret_cast<const uint32_t*>(s); break;
		case 8: *reinterpret_cast<uint64_t*>(d) = *reinterpret_cast<const uint64_t*>(s); break;
		case 12: {
			const uint32_t* s4 = reinterpret_cast<const uint32_t*>(s);
			uint32_t* d4 = reinterpret_cast<uint32_t*>(d);
			d4[0] = s4[0];
			d4[1] = s4[1];
			d4[2] = s4[2];
			break;
		}
		default:
			for (int b = 0; b < f.size; ++b) {
				d[b] = s[b];
			}
	}
}

// Copies the kept elements of one field of a tile: element base + k*kBlock goes to pos[k] when its bit is set.
template<typename T, int K>
__device__ __forceinline__ void scatterField(const FieldCopy& f, int base, const uint64_t* masks, const int* pos, int lane)
{
	const T* src = reinterpret_cast<const T*>(f.src);
	T* dst = reinterpret_cast<T*>(f.dst);
	T v[K];
#pragma unroll
	for (int k = 0; k < K; ++k) {
		if ((masks[k] >> lane) & 1ull) {
			v[k] = src[base + k * kBlock];
		}
	}
#pragma unroll
	for (int k = 0; k < K; ++k) {
		if ((masks[k] >> lane) & 1ull) {
			dst[pos[k]] = v[k];
		}
	}
}

// Three- and two-word groups of 12- and 8-byte records, which are only 4-byte aligned: the vector types would
// otherwise carry their natural 16- and 8-byte alignment, which the compiler may assume when it splits or merges
// the accesses.
typedef unsigned Word3 __attribute__((ext_vector_type(3), aligned(4)));
typedef unsigned Word2 __attribute__((ext_vector_type(2), aligned(4)));

// 4-byte accesses through global-address-space pointers: global (not flat) loads and stores count only on the vector
// memory counter, so reading the LDS-staged field list does not wait for the data loads in flight.
__device__ __forceinline__ uint32_t loadGlobalWord(const char* p)
{
	return *(const __attribute__((address_space(1))) uint32_t*)(p);
}
__device__ __forceinline__ void storeGlobalWord(char* p, uint32_t v)
{
	*(__attribute__((address_space(1))) uint32_t*)(p) = v;
}

// One 4-byte word of records of f.size bytes (f.src / f.dst point at the word inside the first record).
template<int K>
__device__ __forceinline__ void scatterStrided(const FieldCopy& f, int base, const uint64_t* masks, const int* pos, int lane)
{
	uint32_t v[K];
#pragma unroll
	for (int k = 0; k < K; ++k) {
		if ((masks[k] >> lane) & 1ull) {
			v[k] = *reinterpret_cast<const uint32_t*>(f.src + static_cast<size_t>(base + k * kBlock) * f.size);
		}
	}
#pragma unroll
	for (int k = 0; k < K; ++k) {
		if ((masks[k] >> lane) & 1ull) {
			*reinterpret_cast<uint32_t*>(f.dst + static_cast<size_t>(pos[k]) * f.size) = v[k];
		}
	}
}

// Offset of tile = exclusive scan of the tile counts (kTileScan). Element i = tile + k*256 + thread (k-major), so each of
// the kItems rounds reads and writes contiguous runs; ranks come from wave ballots and per-(round, wave) counts in
// LDS, which preserves input order.
template<int kItems_ = kItems>
__device__ void compactScatter(int capacity, const int32_t* countIn, const int32_t* flags, const FieldCopyList& fieldsIn,
                               const int32_t* tileCounts, int tile, int numTiles, int32_t* countOut)
{
	__shared__ int32_t waveBase[kItems_][kWaves];
	// The field list, staged in LDS: read from global memory inside the copy loops, every descriptor access would
	// wait for all outstanding data loads.
	__shared__ FieldCopy lf[kMaxCompactFields];
	__shared__ int32_t lfCount;
	if (threadIdx.x < kMaxCompactFields && static_cast<int>(threadIdx.x) < fieldsIn.count) {
		lf[threadIdx.x] = fieldsIn.f[threadIdx.x];
	}
	if (threadIdx.x == 0) {
		lfCount = fieldsIn.count;
	}
	const int n = liveCount(capacity, countIn);
	const int tileOffset = tileCounts[tile]; // exclusive offsets, written by the tile scan
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	const int base = tile * kBlock * kItems_ + threadIdx.x;
	uint64_t masks[kItems_];
#pragma unroll
	for (int k = 0; k < kItems_; ++k) {
		const int i = base + k * kBlock;
		masks[k] = __ballot(i < n && flags[i] != 0);
		if (lane == 0) {
			waveBase[k][wave] = static_cast<int32_t>(__popcll(masks[k]));
		}
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		int run = tileOffset;
		for (int k = 0; k < kItems_; ++k) {
			for (int w = 0; w < kWaves; ++w) {
				const int c = waveBase[k][w];
				waveBase[k][w] = run;
				run += c;
			}
		}
	}
	__syncthreads();
	int pos[kItems_];
#pragma unroll
	for (int k = 0; k < kItems_; ++k) {
		pos[k] = waveBase[k][wave] + static_cast<int>(__popcll(masks[k] & below));
	}
	// Fields whose size is a multiple of 4 B move as a sequence of streams of up to three words per element (a 12 B
	// vector is one stream of three-word accesses, a 48 B ray pose four), with the loads of the next stream issued
	// before the stores of the current one, so two streams are always in flight.
	const int fieldCount = lfCount;
	const FieldCopy* fields = lf;
	auto normalize = [&](int& f, int& w) {
		while (f < fieldCount && (fields[f].size % 4 != 0 || 4 * w >= fields[f].size)) {
			++f;
			w = 0;
		}
	};
	auto chunkOf = [&](int f, int w) { return min(3, fields[f].size / 4 - w); };
	auto advance = [&](int& f, int& w) {
		w += chunkOf(f, w);
		normalize(f, w);
	};
	// Loads are unconditional (clamped to the live range; unkept elements share cache lines with kept ones), so the
	// compiler knows how many are in flight and the stores of one stream wait only for that stream's loads.
	auto loadChunk = [&](int f, int w, Word3* v) {
		const FieldCopy& fc = fields[f];
		const int c = chunkOf(f, w);
#pragma unroll
		for (int k = 0; k < kItems_; ++k) {
			const char* p = fc.src + static_cast<size_t>(min(base + k * kBlock, n - 1)) * fc.size + 4 * w;
			if (c == 3) {
				v[k] = *(const __attribute__((address_space(1))) Word3*)(p);
			} else if (c == 2) {
				const Word2 x = *(const __attribute__((address_space(1))) Word2*)(p);
				v[k] = Word3{x.x, x.y, 0u};
			} else {
				v[k] = Word3{loadGlobalWord(p), 0u, 0u};
			}
		}
	};
	auto storeChunk = [&](int f, int w, const Word3* v) {
		const FieldCopy& fc = fields[f];
		const int c = chunkOf(f, w);
#pragma unroll
		for (int k = 0; k < kItems_; ++k) {
			if ((masks[k] >> lane) & 1ull) {
				char* p = fc.dst + static_cast<size_t>(pos[k]) * fc.size + 4 * w;
				if (c == 3) {
					*(__attribute__((address_space(1))) Word3*)(p) = v[k];
				} else if (c == 2) {
					*(__attribute__((address_space(1))) Word2*)(p) = Word2{v[k].x, v[k].y};
				} else {
					storeGlobalWord(p, v[k].x);
				}
			}
		}
	};
	int f = 0, w = 0;
	normalize(f, w);
	if (n <= 0) {
		f = fieldCount; // nothing live (no loads: the clamped index would be -1)
	}
	// Two register buffers used alternately (no copies between them, which would wait for the loads in flight).
	Word3 va[kItems_], vb[kItems_];
	if (f < fieldCount) {
		loadChunk(f, w, va);
	}
	while (f < fieldCount) {
		int f2 = f, w2 = w;
		advance(f2, w2);
		if (f2 < fieldCount) {
			loadChunk(f2, w2, vb);
		}
		storeChunk(f, w, va);
		if (f2 >= fieldCount) {
			break;
		}
		int f3 = f2, w3 = w2;
		advance(f3, w3);
		if (f3 < fieldCount) {
			loadChunk(f3, w3, va);
		}
		storeChunk(f2, w2, vb);
		f = f3;
		w = w3;
	}
	// 1- and 2-byte fields (and any other size) element by element.
	for (int g = 0; g < fieldCount; ++g) {
		const FieldCopy fc = fields[g];
		if (fc.size == 1) {
			scatterField<uint8_t, kItems_>(fc, base, masks, pos, lane);
		} else if (fc.size == 2) {
			scatterField<uint16_t, kItems_>(fc, base, masks, pos, lane);
		} else if (fc.size % 4 != 0) {
			for (int k = 0; k < kItems_; ++k) {
				if ((masks[k] >> lane) & 1ull) {
					copyElem(fc, base + k * kBlock, pos[k]);
				}
			}
		}
	}
}

// Exclusive scan of a graph's tile counts in place (tile counts become tile offsets) and its kept count: one block
// per graph, 256 tiles per round, shuffle scans within waves.
__device__ void tileScan(int32_t* counts, int numTiles, int32_t* countOut)
{
	__shared__ int32_t waveSums[kWaves];
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	int carry = 0;
	for (int base = 0; base < numTiles; base += kBlock) {
		const int i = base + threadIdx.x;
		const int v = i < numTiles ? counts[i] : 0;
		int x = v;
		for (int off = 1; off < 64; off <<= 1) {
			const int y = __shfl_up(x, off, 64);
			x += lane >= off ? y : 0;
		}
		if (lane == 63) {
			waveSums[wave] = x;
		}
		__syncthreads();
		int before = 0, total = 0;
		for (int w = 0; w < kWaves; ++w) {
			before += w < wave ? waveSums[w] : 0;
			total += waveSums[w];
		}
		if (i < numTiles) {
			counts[i] = carry + before + x - v;
		}
		carry += total;
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		*countOut = carry;
	}
}

__global__ void kTileScan(int32_t* counts, int numTiles, int32_t* countOut) { tileScan(counts, numTiles, countOut); }

__global__ void kTileCounts(int capacity, const int32_t* countIn, const int32_t* flags, int32_t* tileCounts)
{
	tileCount(capacity, countIn, flags, blockIdx.x, tileCounts);
}

__global__ void kCompactScatter(int capacity, const int32_t* countIn, const int32_t* flags, FieldCopyList fields,
                                const int32_t* tileCounts, int numTiles, int32_t* countOut)
{
	compactScatter(capacity, countIn, flags, fields, tileCounts, blockIdx.x, numTiles, countOut);
}

template<typename Item, typename Key>
__device__ __forceinline__ int findItem(const Item* items, int itemCount, Key key)
{
	return batchItemOf(items, itemCount, static_cast<int>(blockIdx.x), key);
}

// Tile counts of many graphs: one wave per tile of kBlock * kItems_ points, each lane summing 16-byte groups of four
// flags (tiles start at multiples of 1024 flags, so the groups are aligned; the group that crosses the live count is
// read flag by flag), then a wave reduction. No LDS or block barrier: with a block per tile the launch was bound by
// block scheduling rather than by its 4 bytes per point.
template<int kItems_>
__global__ void kTileCountsBatch(const CompactItem* __restrict__ items, int itemCount, int totalTiles)
{
	const int tile = static_cast<int>(blockIdx.x) * kWaves + static_cast<int>(threadIdx.x >> 6);
	if (tile >= totalTiles) {
		return; // the whole wave
	}
	const int lane = static_cast<int>(threadIdx.x & 63);
	const int i = batchItemOf(items, itemCount, tile, [](const CompactItem& c) { return c.firstTile; });
	const CompactItem& it = items[i];
	const int n = liveCount(it.capacity, it.countIn);
	const int local = tile - it.firstTile;
	constexpr int kTile = kBlock * kItems_;
	const int base = local * kTile;
	const bool aligned = (reinterpret_cast<uintptr_t>(it.flags) & 15u) == 0; // device buffers are; anything else reads flag by flag
	int c = 0;
#pragma unroll
	for (int g = 0; g < kTile / 256; ++g) {
		const int p = base + g * 256 + 4 * lane;
		if (aligned && p + 3 < n) {
			const int4 v = *reinterpret_cast<const int4*>(it.flags + p);
			c += (v.x != 0) + (v.y != 0) + (v.z != 0) + (v.w != 0);
		} else {
			for (int k = 0; k < 4; ++k) {
				c += (p + k < n && it.flags[p + k] != 0) ? 1 : 0;
			}
		}
	}
	for (int off = 32; off > 0; off >>= 1) {
		c += __shfl_down(c, off, 64);
	}
	if (lane == 0) {
		it.tileCounts[local] = c;
	}
}

__global__ void kTileScanBatch(const CompactItem* __restrict__ items)
{
	const CompactItem& it = items[blockIdx.x];
	tileScan(it.tileCounts, it.numTiles, it.countOut);
}

template<int kItems_>
__global__ void kCompactScatterBatch(const CompactItem* __restrict__ items, int itemCount)
{
	const int i = findItem(items, itemCount, [](const CompactItem& c) { return c.firstTile; });
	const CompactItem& it = items[i];
	compactScatter<kItems_>(it.capacity, it.countIn, it.flags, *it.fields, it.tileCounts,
	                        static_cast<int>(blockIdx.x) - it.firstTile, it.numTiles, it.countOut);
}

// ------------------------------------------------------------------------------------------------ element ops
__device__ __forceinline__ void opTransformPoints(int i, const Vec3f* in, Vec3f* out, const Mat3x4f& tf)
{
	out[i] = pointTransform(tf, in[i]);
}

__device__ __forceinline__ void opTransformRays(int i, const Mat3x4f* in, Mat3x4f* out, const Mat3x4f& tf)
{
	out[i] = tf * in[i];
}

// src/gpu/gaussianNoiseKernels.cu:20-29
__device__ __forceinline__ void opNoiseRays(int i, const Mat3x4f* in, Mat3x4f* out, float mean, float stDev, int axis,
                                            const Mat3x4f& lookAt, const Mat3x4f& lookAtInv, uint64_t seed, uint32_t run,
                                            uint32_t tag)
{
	const float err = mean + rngNormal(seed, run, tag, static_cast<uint32_t>(i)) * stDev;
	out[i] = lookAtInv * (Mat3x4f::rotationAxisRad(axis, err) * (lookAt * in[i]));
}

// src/gpu/gaussianNoiseKernels.cu:31-47
__device__ __forceinline__ void opNoiseHitpoint(int i, const Vec3f* inXyz, Vec3f* outXyz, float* outDistance, float mean,
                                                float stDev, int axis, const Mat3x4f& lookAt, const Mat3x4f& lookAtInv,
                                                uint64_t seed, uint32_t run, uint32_t tag)
{
	const float err = mean + rngNormal(seed, run, tag, static_cast<uint32_t>(i)) * stDev;
	const Vec3f q = Mat3x4f::rotationAxisRad(axis, err) * (lookAt * inXyz[i]);
	if (outDistance != nullptr) {
		outDistance[i] = length(q);
	}
	outXyz[i] = lookAtInv * q;
}

// src/gpu/gaussianNoiseKernels.cu:49-66
__device__ __forceinline__ void opNoiseDistance(int i, const Vec3f* inXyz, const float* inDistance, Vec3f* outXyz,
                                                float* outDistance, float mean, float stDevBase, float stDevRise,
                                                const Mat3x4f& lookAt, const Mat3x4f& lookAtInv, uint64_t seed,
                                                uint32_t run, uint32_t tag)
{
	Vec3f xyz;
	float distance;
	pointNoiseDistance(inXyz[i], inDistance[i], mean, stDevBase, stDevRise, lookAt, lookAtInv, seed, run, tag,
	                   static_cast<uint32_t>(i), xyz, distance);
	outXyz[i] = xyz;
	outDistance[i] = distance;
}

__device__ __forceinline__ void opSoaToAos(int i, int recordSize, const FieldCopyList& fields, char* out)
{
	char* rec = out + static_cast<size_t>(i) * recordSize;
	for (int f = 0; f < fields.count; ++f) {
		const FieldCopy& fc = fields.f[f];
		if (fc.src == nullptr) {
			for (int b = 0; b < fc.size; ++b) {
				rec[fc.offset + b] = 0; // padding
			}
			continue;
		}
		const char* s = fc.src + static_cast<size_t>(i) * fc.size;
		for (int b = 0; b < fc.size; ++b) {
			rec[fc.offset + b] = s[b];
		}
	}
}

// One element of an elementwise stage; shared by the single-graph and batched launches.
__device__ __forceinline__ void applyOp(const PointOpItem& it, int i)
{
	switch (it.op) {
		case kOpTransformPoints:
			opTransformPoints(i, static_cast<const Vec3f*>(it.in0), static_cast<Vec3f*>(it.out0), it.m0);
			break;
		case kOpNoiseHitpoint:
			opNoiseHitpoint(i, static_cast<const Vec3f*>(it.in0), static_cast<Vec3f*>(it.out0),
			                static_cast<float*>(it.out1), it.a, it.b, it.axis, it.m0, it.m1, it.seed, it.run, it.tag);
			break;
		case kOpNoiseDistance:
			opNoiseDistance(i, static_cast<const Vec3f*>(it.in0), static_cast<const float*>(it.in1),
			                static_cast<Vec3f*>(it.out0), static_cast<float*>(it.out1), it.a, it.b, it.c, it.m0, it.m1,
			                it.seed, it.run, it.tag);
			break;
		case kOpTransformRays:
			opTransformRays(i, static_cast<const Mat3x4f*>(it.in0), static_cast<Mat3x4f*>(it.out0), it.m0);
			break;
		case kOpNoiseRays:
			opNoiseRays(i, static_cast<const Mat3x4f*>(it.in0), static_cast<Mat3x4f*>(it.out0), it.a, it.b, it.axis,
			            it.m0, it.m1, it.seed, it.run, it.tag);
			break;
		case kOpFormat:
			opSoaToAos(i, it.recordSize, *it.format, static_cast<char*>(it.out0));
			break;
		case kOpToHost:
			for (int f = 0; f < it.format->count; ++f) {
				copyElem(it.format->f[f], i, i);
			}
			break;
		default:
			break;
	}
}

__global__ void kCopyFields(int capacity, const int32_t* count, FieldCopyList fields)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < liveCount(capacity, count)) {
		for (int f = 0; f < fields.count; ++f) {
			copyElem(fields.f[f], i, i);
		}
	}
}

__global__ void kPointOp(PointOpItem it)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < liveCount(it.capacity, it.count)) {
		applyOp(it, i);
	}
}

__global__ void kNoiseNormals(int n, float* out, uint64_t seed, uint32_t run, uint32_t tag)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) {
		out[i] = rngNormal(seed, run, tag, static_cast<uint32_t>(i));
	}
}

__global__ void kSoaToAos(int capacity, const int32_t* count, int recordSize, FieldCopyList fields, char* out)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < liveCount(capacity, count)) {
		opSoaToAos(i, recordSize, fields, out);
	}
}

__global__ void kAosToSoa(int n, int recordSize, const char* in, FieldCopyList fields)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const char* rec = in + static_cast<size_t>(i) * recordSize;
	for (int f = 0; f < fields.count; ++f) {
		const FieldCopy& fc = fields.f[f];
		if (fc.dst == nullptr) {
			continue;
		}
		char* d = fc.dst + static_cast<size_t>(i) * fc.size;
		for (int b = 0; b < fc.size; ++b) {
			d[b] = rec[fc.offset + b];
		}
	}
}

// Elementwise stages of many graphs in one launch.
__global__ void kPointOpsBatch(const PointOpItem* __restrict__ items, int itemCount)
{
	const int k = findItem(items, itemCount, [](const PointOpItem& p) { return p.firstBlock; });
	const PointOpItem& it = items[k];
	const int i = (static_cast<int>(blockIdx.x) - it.firstBlock) * kBlock + threadIdx.x;
	if (i < liveCount(it.capacity, it.count)) {
		applyOp(it, i);
	}
}

// dst[i] = *src[i]: collects the device-resident counts of many graphs for one copy to the host.
__global__ void kGatherInts(int n, const int32_t* const* src, int32_t* dst)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) {
		dst[i] = *src[i];
	}
}

// ------------------------------------------------------------------------------------------------ radar cells
__global__ void kRadarClear(int cells, unsigned long long* sums, uint32_t* counts)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < cells) {
		sums[i] = 0ull;
		counts[i] = 0u;
	}
}

// One thread per point: cell from the FP32 distance and the sensor-frame azimuth atan2(y, x); contribution
// d^2 * solid angle * cos(incidence), rounded to the fixed-point grid and added with a 64-bit integer atomic.
__global__ void kRadarBin(int capacity, const int32_t* count, const Vec3f* xyz, const float* distance,
                          const float* incident, const int32_t* isHit, Mat3x4f toSensor, RadarBinParams bp,
                          unsigned long long* sums, uint32_t* counts)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= liveCount(capacity, count) || isHit[i] == 0) {
		return;
	}
	const float d = distance[i];
	const Vec3f p = toSensor * xyz[i];
	const float az = atan2f(p.y, p.x);
	const float rf = floorf(d / bp.rangeStep);
	const float af = floorf((az - bp.azMin) / bp.azStep);
	if (!(rf >= 0.0f && rf < static_cast<float>(bp.rangeCells) && af >= 0.0f && af < static_cast<float>(bp.azCells))) {
		return;
	}
	const int cell = static_cast<int>(rf) * bp.azCells + static_cast<int>(af);
	const float sigma = d * d * bp.raySolidAngle * cosf(incident[i]);
	const unsigned long long q = static_cast<unsigned long long>(llrint(static_cast<double>(sigma) * kRcsFixedScale));
	atomicAdd(&sums[cell], q);
	atomicAdd(&counts[cell], 1u);
}

// Dense per-cell outputs and the keep flag of each cell (non-empty).
__global__ void kRadarCells(RadarBinParams bp, const unsigned long long* sums, const uint32_t* counts, Vec3f* xyz,
                            float* distance, float* azimuth, float* rcs, int32_t* isHit, int32_t* keep)
{
	const int c = blockIdx.x * kBlock + threadIdx.x;
	const int cells = bp.rangeCells * bp.azCells;
	if (c >= cells) {
		return;
	}
	const int r = c / bp.azCells, a = c % bp.azCells;
	const float dist = (static_cast<float>(r) + 0.5f) * bp.rangeStep;
	const float az = bp.azMin + (static_cast<float>(a) + 0.5f) * bp.azStep;
	xyz[c] = Vec3f{dist * cosf(az), dist * sinf(az), 0.0f};
	distance[c] = dist;
	azimuth[c] = az;
	const double sum = static_cast<double>(sums[c]) / kRcsFixedScale;
	rcs[c] = sum > 0.0 ? static_cast<float>(10.0 * log10(sum)) : -INFINITY;
	isHit[c] = 1;
	keep[c] = counts[c] > 0u ? 1 : 0;
}

__global__ void kFillInt(int n, int32_t* p, int32_t v)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) {
		p[i] = v;
	}
}

inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

} // namespace

int compactTiles(int capacity, int variant)
{
	const int tile = kBlock * (variant == 1 ? 4 : kItems);
	return capacity <= 0 ? 1 : (capacity + tile - 1) / tile;
}
int pointOpBlocks(int capacity) { return capacity <= 0 ? 1 : blocks(capacity); }

void compact(hipStream_t stream, int capacity, const int32_t* countIn, const int32_t* flags, const FieldCopyList& fields,
             int32_t* tileCounts, int32_t* countOut)
{
	if (capacity <= 0) {
		hipLaunchKernelGGL(kFillInt, dim3(1), dim3(kBlock), 0, stream, 1, countOut, 0);
		return;
	}
	const int tiles = compactTiles(capacity);
	hipLaunchKernelGGL(kTileCounts, dim3(tiles), dim3(kBlock), 0, stream, capacity, countIn, flags, tileCounts);
	hipLaunchKernelGGL(kTileScan, dim3(1), dim3(kBlock), 0, stream, tileCounts, tiles, countOut);
	hipLaunchKernelGGL(kCompactScatter, dim3(tiles), dim3(kBlock), 0, stream, capacity, countIn, flags, fields, tileCounts,
	                   tiles, countOut);
}

void compactBatch(hipStream_t stream, const CompactItem* items, int itemCount, int totalTiles, int variant)
{
	if (itemCount <= 0 || totalTiles <= 0) {
		return;
	}
	const int countBlocks = (totalTiles + kWaves - 1) / kWaves;
	if (variant == 1) {
		hipLaunchKernelGGL(kTileCountsBatch<4>, dim3(countBlocks), dim3(kBlock), 0, stream, items, itemCount, totalTiles);
		hipLaunchKernelGGL(kTileScanBatch, dim3(itemCount), dim3(kBlock), 0, stream, items);
		hipLaunchKernelGGL(kCompactScatterBatch<4>, dim3(totalTiles), dim3(kBlock), 0, stream, items, itemCount);
	} else {
		hipLaunchKernelGGL(kTileCountsBatch<kItems>, dim3(countBlocks), dim3(kBlock), 0, stream, items, itemCount, totalTiles);
		hipLaunchKernelGGL(kTileScanBatch, dim3(itemCount), dim3(kBlock), 0, stream, items);
		hipLaunchKernelGGL(kCompactScatterBatch<kItems>, dim3(totalTiles), dim3(kBlock), 0, stream, items, itemCount);
	}
}

void pointOpsBatch(hipStream_t stream, const PointOpItem* items, int itemCount, int totalBlocks)
{
	if (itemCount > 0 && totalBlocks > 0) {
		hipLaunchKernelGGL(kPointOpsBatch, dim3(totalBlocks), dim3(kBlock), 0, stream, items, itemCount);
	}
}

void pointOp(hipStream_t stream, const PointOpItem& it)
{
	if (it.op == kOpFormat || it.op == kOpToHost) {
		throw std::invalid_argument("pointOp: field-list stages take their list by value (formatSoaToAos, copyFields)");
	}
	if (it.capacity > 0) {
		hipLaunchKernelGGL(kPointOp, dim3(blocks(it.capacity)), dim3(kBlock), 0, stream, it);
	}
}

void copyFields(hipStream_t stream, int capacity, const int32_t* count, const FieldCopyList& fields)
{
	if (capacity > 0 && fields.count > 0) {
		hipLaunchKernelGGL(kCopyFields, dim3(blocks(capacity)), dim3(kBlock), 0, stream, capacity, count, fields);
	}
}

void gatherInts(hipStream_t stream, int n, const int32_t* const* src, int32_t* dst)
{
	if (n > 0) {
		hipLaunchKernelGGL(kGatherInts, dim3(blocks(n)), dim3(kBlock), 0, stream, n, src, dst);
	}
}

void noiseNormals(hipStream_t stream, int n, float* out, uint64_t seed, uint32_t run, uint32_t tag)
{
	if (n > 0) {
		hipLaunchKernelGGL(kNoiseNormals, dim3(blocks(n)), dim3(kBlock), 0, stream, n, out, seed, run, tag);
	}
}

void formatSoaToAos(hipStream_t stream, int capacity, const int32_t* count, int recordSize, const FieldCopyList& fields,
                    char* out)
{
	if (capacity > 0) {
		hipLaunchKernelGGL(kSoaToAos, dim3(blocks(capacity)), dim3(kBlock), 0, stream, capacity, count, recordSize, fields,
		                   out);
	}
}

void formatAosToSoa(hipStream_t stream, int n, int recordSize, const char* in, const FieldCopyList& fields)
{
	if (n > 0) {
		hipLaunchKernelGGL(kAosToSoa, dim3(blocks(n)), dim3(kBlock), 0, stream, n, recordSize, in, fields);
	}
}

void fillInt(hipStream_t stream, int n, int32_t* p, int32_t value)
{
	if (n > 0) {
		hipLaunchKernelGGL(kFillInt, dim3(blocks(n)), dim3(kBlock), 0, stream, n, p, value);
	}
}

void radarBin(hipStream_t stream, int capacity, const int32_t* count, const Vec3f* xyz, const float* distance,
              const float* incident, const int32_t* isHit, const Mat3x4f& toSensor, const RadarBinParams& bp,
              unsigned long long* sums, uint32_t* counts, Vec3f* cellXyz, float* cellDistance, float* cellAzimuth,
              float* cellRcs, int32_t* cellHit, int32_t* keep)
{
	const int cells = bp.rangeCells * bp.azCells;
	hipLaunchKernelGGL(kRadarClear, dim3(blocks(cells)), dim3(kBlock), 0, stream, cells, sums, counts);
	if (capacity > 0) {
		hipLaunchKernelGGL(kRadarBin, dim3(blocks(capacity)), dim3(kBlock), 0, stream, capacity, count, xyz, distance,
		                   incident, isHit, toSensor, bp, sums, counts);
	}
	hipLaunchKernelGGL(kRadarCells, dim3(blocks(cells)), dim3(kBlock), 0, stream, bp, sums, counts, cellXyz, cellDistance,
	                   cellAzimuth, cellRcs, cellHit, keep);
}

// ------------------------------------------------------------------------------------------------ no-alloc normals
// Staging array of noiseNormalsNoAlloc: a file-scope device array, one copy per device for the life of the process.
// Host code never passes its name to a kernel (on the host it is a stand-in, not a device address): the device
// address is looked up once per device with hipGetSymbolAddress and kept in scratchPtr.
__device__ __align__(256) float scratch[kNoAllocNormalsMax];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratchMutex;
float* scratchPtr[kMaxDevices] = {};

float* scratchOfCurrentDevice()
{
	int dev = 0;
	RGL_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices) {
		throw std::runtime_error("device index beyond the no-alloc staging table");
	}
	std::lock_guard<std::mutex> lock(scratchMutex);
	if (scratchPtr[dev] == nullptr) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || p == nullptr) {
			throw std::runtime_error(std::string("hipGetSymbolAddress of the no-alloc staging array failed: ") +
			                         hipGetErrorString(e));
		}
		scratchPtr[dev] = static_cast<float*>(p);
	}
	return scratchPtr[dev];
}
} // namespace

void noiseNormalsNoAlloc(int n, float* hostOut, uint64_t seed, uint32_t run, uint32_t tag)
{
	if (n < 0 || n > kNoAllocNormalsMax) {
		throw std::invalid_argument("RGL API Error: Invalid argument, condition unsatisfied: count <= " +
		                            std::to_string(kNoAllocNormalsMax) + " (the no-alloc staging array)");
	}
	if (n == 0) {
		return;
	}
	float* staging = scratchOfCurrentDevice();
	hipLaunchKernelGGL(kNoiseNormals, dim3(blocks(n)), dim3(kBlock), 0, nullptr, n, staging, seed, run, tag);
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipMemcpy(hostOut, staging, sizeof(float) * n, hipMemcpyDeviceToHost)); // null stream: after the kernel
}

} // namespace rgl::gpu
