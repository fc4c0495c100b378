// Elementwise point stages shared by their own launches (points.hip) and by the batched raytrace launch that folds
// them in (raytrace.hip). Every product and sum is rounded explicitly (no contraction), so a point gets the same
// bits from either kernel.
#pragma once

#include "math.hpp"
#include "noise_rng.hpp"

namespace rgl {

// src/gpu/nodeKernels.cu:51-56
__device__ __forceinline__ Vec3f pointTransform(const Mat3x4f& tf, Vec3f p) { return tf * p; }

// Distance noise of one point (src/gpu/gaussianNoiseKernels.cu:49-66): the point moves along its direction from the
// sensor by err = mean + N(0,1) * (distance * stDevRise + stDevBase), and the distance grows by err.
__device__ __forceinline__ void pointNoiseDistance(Vec3f xyz, float distance, float mean, float stDevBase,
                                                   float stDevRise, const Mat3x4f& lookAt, const Mat3x4f& lookAtInv,
                                                   uint64_t seed, uint32_t run, uint32_t tag, uint32_t index,
                                                   Vec3f& outXyz, float& outDistance)
{
	const float sigma = __fadd_rn(__fmul_rn(distance, stDevRise), stDevBase);
	const float err = __fadd_rn(mean, __fmul_rn(rngNormal(seed, run, tag, index), sigma));
	const Vec3f pr = lookAt * xyz;
	const float len = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(pr.x, pr.x), __fmul_rn(pr.y, pr.y)), __fmul_rn(pr.z, pr.z)));
	const Vec3f moved{__fadd_rn(pr.x, __fmul_rn(pr.x / len, err)), __fadd_rn(pr.y, __fmul_rn(pr.y / len, err)),
	                  __fadd_rn(pr.z, __fmul_rn(pr.z / len, err))};
	outXyz = lookAtInv * moved;
	outDistance = __fadd_rn(distance, err);
}

} // namespace rgl
