// Point-cloud kernels: stream compaction, point transform, Gaussian noise, SoA<->AoS formatting.
//
// Compaction replaces thrust::inclusive_scan + one kApplyCompaction launch per field
// (src/gpu/nodeKernels.cu:58-68,424-436; src/graph/CompactByFieldPointsNode.cpp:33-99) with two launches in
// total: per-tile counts, then one pass that finds each tile's offset, ranks the kept elements inside the tile
// and scatters every requested field. Elements are kept in input order (stable), and the kept count stays on
// the device so nothing downstream waits for the host.
// Noise kernels follow src/gpu/gaussianNoiseKernels.cu:20-66 with the seeded stream of noise_rng.hpp.

#include <hip/hip_runtime.h>

#include "device_types.hpp"
#include "kernels.hpp"
#include "noise_rng.hpp"

namespace rgl::gpu {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems; // 2048 points per compaction tile

__device__ __forceinline__ int liveCount(int capacity, const int32_t* count) { return count ? *count : capacity; }

__global__ void kTileCounts(int capacity, const int32_t* countIn, const int32_t* flags, int32_t* tileCounts)
{
	__shared__ int32_t warpSums[kBlock / 64];
	const int n = liveCount(capacity, countIn);
	const int base = blockIdx.x * kTile;
	int c = 0;
	for (int k = 0; k < kItems; ++k) {
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
		for (int w = 0; w < kBlock / 64; ++w) {
			s += warpSums[w];
		}
		tileCounts[blockIdx.x] = s;
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

// Offset of this tile = sum of the counts of all previous tiles. Element i = tile + k*256 + thread (k-major), so
// each of the kItems rounds reads and writes contiguous runs; ranks come from wave ballots and per-(round, wave)
// counts in LDS, which preserves input order.
__global__ void kCompactScatter(int capacity, const int32_t* countIn, const int32_t* flags, FieldCopyList fields,
                                const int32_t* tileCounts, int numTiles, int32_t* countOut)
{
	__shared__ int32_t red[kBlock];
	__shared__ int32_t waveBase[kItems][kBlock / 64];
	const int n = liveCount(capacity, countIn);
	// 1. tile offset
	int s = 0;
	for (int t = threadIdx.x; t < static_cast<int>(blockIdx.x); t += kBlock) {
		s += tileCounts[t];
	}
	red[threadIdx.x] = s;
	__syncthreads();
	for (int off = kBlock / 2; off > 0; off >>= 1) {
		if (threadIdx.x < off) {
			red[threadIdx.x] += red[threadIdx.x + off];
		}
		__syncthreads();
	}
	const int tileOffset = red[0];
	if (blockIdx.x == numTiles - 1 && threadIdx.x == 0) {
		*countOut = tileOffset + tileCounts[blockIdx.x];
	}
	// 2. flags, ballots and per-(round, wave) counts
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	const int base = blockIdx.x * kTile + threadIdx.x;
	uint64_t masks[kItems];
	for (int k = 0; k < kItems; ++k) {
		const int i = base + k * kBlock;
		masks[k] = __ballot(i < n && flags[i] != 0);
		if (lane == 0) {
			waveBase[k][wave] = static_cast<int32_t>(__popcll(masks[k]));
		}
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		int run = tileOffset;
		for (int k = 0; k < kItems; ++k) {
			for (int w = 0; w < kBlock / 64; ++w) {
				const int c = waveBase[k][w];
				waveBase[k][w] = run;
				run += c;
			}
		}
	}
	__syncthreads();
	// 3. scatter every field
	for (int k = 0; k < kItems; ++k) {
		if ((masks[k] >> lane) & 1ull) {
			const int pos = waveBase[k][wave] + static_cast<int>(__popcll(masks[k] & below));
			for (int f = 0; f < fields.count; ++f) {
				copyElem(fields.f[f], base + k * kBlock, pos);
			}
		}
	}
}

__global__ void kTransformPoints(int capacity, const int32_t* count, const Vec3f* in, Vec3f* out, Mat3x4f tf)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < liveCount(capacity, count)) {
		out[i] = tf * in[i];
	}
}

// src/gpu/gaussianNoiseKernels.cu:20-29
__global__ void kNoiseAngularRays(int n, const Mat3x4f* in, Mat3x4f* out, float mean, float stDev, int axis,
                                  Mat3x4f lookAt, Mat3x4f lookAtInv, uint64_t seed, uint32_t run, uint32_t tag)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) {
		return;
	}
	const float err = mean + rngNormal(seed, run, tag, static_cast<uint32_t>(i)) * stDev;
	out[i] = lookAtInv * (Mat3x4f::rotationAxisRad(axis, err) * (lookAt * in[i]));
}

// src/gpu/gaussianNoiseKernels.cu:31-47
__global__ void kNoiseAngularHitpoint(int capacity, const int32_t* count, const Vec3f* inXyz, Vec3f* outXyz,
                                      float* outDistance, float mean, float stDev, int axis, Mat3x4f lookAt,
                                      Mat3x4f lookAtInv, uint64_t seed, uint32_t run, uint32_t tag)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= liveCount(capacity, count)) {
		return;
	}
	const float err = mean + rngNormal(seed, run, tag, static_cast<uint32_t>(i)) * stDev;
	const Vec3f q = Mat3x4f::rotationAxisRad(axis, err) * (lookAt * inXyz[i]);
	if (outDistance != nullptr) {
		outDistance[i] = length(q);
	}
	outXyz[i] = lookAtInv * q;
}

// src/gpu/gaussianNoiseKernels.cu:49-66
__global__ void kNoiseDistance(int capacity, const int32_t* count, const Vec3f* inXyz, const float* inDistance,
                               Vec3f* outXyz, float* outDistance, float mean, float stDevBase, float stDevRise,
                               Mat3x4f lookAt, Mat3x4f lookAtInv, uint64_t seed, uint32_t run, uint32_t tag)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= liveCount(capacity, count)) {
		return;
	}
	const float sigma = inDistance[i] * stDevRise + stDevBase;
	const float err = mean + rngNormal(seed, run, tag, static_cast<uint32_t>(i)) * sigma;
	const Vec3f pr = lookAt * inXyz[i];
	outXyz[i] = lookAtInv * (pr + normalized(pr) * err);
	outDistance[i] = inDistance[i] + err;
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
	if (i >= liveCount(capacity, count)) {
		return;
	}
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

__global__ void kFillInt(int n, int32_t* p, int32_t v)
{
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) {
		p[i] = v;
	}
}

inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

} // namespace

int compactTiles(int capacity) { return capacity <= 0 ? 1 : (capacity + kTile - 1) / kTile; }

void compact(hipStream_t stream, int capacity, const int32_t* countIn, const int32_t* flags, const FieldCopyList& fields,
             int32_t* tileCounts, int32_t* countOut)
{
	if (capacity <= 0) {
		hipLaunchKernelGGL(kFillInt, dim3(1), dim3(kBlock), 0, stream, 1, countOut, 0);
		return;
	}
	const int tiles = compactTiles(capacity);
	hipLaunchKernelGGL(kTileCounts, dim3(tiles), dim3(kBlock), 0, stream, capacity, countIn, flags, tileCounts);
	hipLaunchKernelGGL(kCompactScatter, dim3(tiles), dim3(kBlock), 0, stream, capacity, countIn, flags, fields, tileCounts,
	                   tiles, countOut);
}

void transformPoints(hipStream_t stream, int capacity, const int32_t* count, const Vec3f* in, Vec3f* out,
                     const Mat3x4f& tf)
{
	if (capacity > 0) {
		hipLaunchKernelGGL(kTransformPoints, dim3(blocks(capacity)), dim3(kBlock), 0, stream, capacity, count, in, out, tf);
	}
}

void noiseAngularRays(hipStream_t stream, int n, const Mat3x4f* in, Mat3x4f* out, float mean, float stDev, int axis,
                      const Mat3x4f& lookAt, uint64_t seed, uint32_t run, uint32_t tag)
{
	if (n > 0) {
		hipLaunchKernelGGL(kNoiseAngularRays, dim3(blocks(n)), dim3(kBlock), 0, stream, n, in, out, mean, stDev, axis, lookAt,
		                   lookAt.inverse(), seed, run, tag);
	}
}

void noiseAngularHitpoint(hipStream_t stream, int capacity, const int32_t* count, const Vec3f* inXyz, Vec3f* outXyz,
                          float* outDistance, float mean, float stDev, int axis, const Mat3x4f& lookAt, uint64_t seed,
                          uint32_t run, uint32_t tag)
{
	if (capacity > 0) {
		hipLaunchKernelGGL(kNoiseAngularHitpoint, dim3(blocks(capacity)), dim3(kBlock), 0, stream, capacity, count, inXyz,
		                   outXyz, outDistance, mean, stDev, axis, lookAt, lookAt.inverse(), seed, run, tag);
	}
}

void noiseDistance(hipStream_t stream, int capacity, const int32_t* count, const Vec3f* inXyz, const float* inDistance,
                   Vec3f* outXyz, float* outDistance, float mean, float stDevBase, float stDevRise, const Mat3x4f& lookAt,
                   uint64_t seed, uint32_t run, uint32_t tag)
{
	if (capacity > 0) {
		hipLaunchKernelGGL(kNoiseDistance, dim3(blocks(capacity)), dim3(kBlock), 0, stream, capacity, count, inXyz,
		                   inDistance, outXyz, outDistance, mean, stDevBase, stDevRise, lookAt, lookAt.inverse(), seed, run,
		                   tag);
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

} // namespace rgl::gpu
