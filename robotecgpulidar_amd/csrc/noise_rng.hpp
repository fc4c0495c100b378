// Seeded, stateless Gaussian noise stream shared by host and device.
//
// The reference draws noise from cuRAND Philox states seeded with std::random_device
// (src/gpu/helpersKernels.cu:22-27, src/graph/NodesCore.hpp:536,563,590), which no run can reproduce.
// This library instead defines the stream as
//     normal(seed, run, tag, i) = BoxMuller(Philox4x32-10(key = seed, counter = {i, run, tag, 0}))
// with Philox4x32-10 from Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11), and
// Box-Muller written with the Cephes single-precision polynomials for log, sin and cos (S. L. Moshier,
// Cephes Math Library). Every float operation below is a single IEEE-rounded operation or an explicit fused
// multiply-add, so host and device produce bit-identical normals; this is what the CPU replay relies on.
#pragma once

#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RGL_RNG_HD __host__ __device__ __forceinline__
#else
#define RGL_RNG_HD inline
#endif

namespace rgl {

struct Philox4
{
	uint32_t v[4];
};

RGL_RNG_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32); }

// Philox4x32 with 10 rounds: round constants M0/M1 and Weyl key increments W0/W1 from the SC'11 paper.
RGL_RNG_HD Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1)
{
	const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
	const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
	for (int r = 0; r < 10; ++r) {
		if (r > 0) {
			k0 += W0;
			k1 += W1;
		}
		const uint32_t lo0 = M0 * ctr.v[0], hi0 = mulhi32(M0, ctr.v[0]);
		const uint32_t lo1 = M1 * ctr.v[2], hi1 = mulhi32(M1, ctr.v[2]);
		Philox4 n;
		n.v[0] = hi1 ^ ctr.v[1] ^ k0;
		n.v[1] = lo1;
		n.v[2] = hi0 ^ ctr.v[3] ^ k1;
		n.v[3] = lo0;
		ctr = n;
	}
	return ctr;
}

#if defined(__clang__)
#define RGL_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define RGL_NO_CONTRACT
#endif

RGL_RNG_HD float rngAsFloat(uint32_t u)
{
	union {
		uint32_t u;
		float f;
	} c{u};
	return c.f;
}
RGL_RNG_HD uint32_t rngAsUint(float f)
{
	union {
		float f;
		uint32_t u;
	} c{f};
	return c.u;
}

// Natural log for x in (0, 1], Cephes logf range reduction and polynomial (max error about 1 ulp).
RGL_RNG_HD float rngLog(float x)
{
	RGL_NO_CONTRACT
	uint32_t bits = rngAsUint(x);
	int e = static_cast<int>((bits >> 23) & 0xffu) - 126; // x = m * 2^e, m in [0.5, 1)
	float m = rngAsFloat((bits & 0x007fffffu) | 0x3f000000u);
	float f;
	if (m < 0.707106781186547524f) {
		e -= 1;
		f = (m + m) - 1.0f;
	} else {
		f = m - 1.0f;
	}
	const float z = f * f;
	float p = 7.0376836292e-2f;
	p = fmaf(p, f, -1.1514610310e-1f);
	p = fmaf(p, f, 1.1676998740e-1f);
	p = fmaf(p, f, -1.2420140846e-1f);
	p = fmaf(p, f, 1.4249322787e-1f);
	p = fmaf(p, f, -1.6668057665e-1f);
	p = fmaf(p, f, 2.0000714765e-1f);
	p = fmaf(p, f, -2.4999993993e-1f);
	p = fmaf(p, f, 3.3333331174e-1f);
	float y = (p * f) * z;
	const float fe = static_cast<float>(e);
	y = fmaf(fe, -2.12194440e-4f, y);
	y = fmaf(z, -0.5f, y);
	float r = f + y;
	r = fmaf(fe, 0.693359375f, r);
	return r;
}

// sin and cos of 2*pi*u for u in [0, 1): quadrant reduction then Cephes sinf/cosf polynomials on [-pi/4, pi/4].
RGL_RNG_HD void rngSinCos2Pi(float u, float* s, float* c)
{
	RGL_NO_CONTRACT
	const float v = u * 4.0f;        // exact
	const float q = rintf(v);        // nearest quadrant 0..4
	const float r = v - q;           // exact, in [-0.5, 0.5]
	const float t = r * 1.57079632679489661923f;
	const float z = t * t;
	float sp = -1.9515295891e-4f;
	sp = fmaf(sp, z, 8.3321608736e-3f);
	sp = fmaf(sp, z, -1.6666654611e-1f);
	const float st = fmaf(sp * z, t, t);
	float cp = 2.443315711809948e-5f;
	cp = fmaf(cp, z, -1.388731625493765e-3f);
	cp = fmaf(cp, z, 4.166664568298827e-2f);
	const float ct = fmaf(cp * z, z, fmaf(z, -0.5f, 1.0f));
	const int quadrant = static_cast<int>(q) & 3;
	switch (quadrant) {
		case 0: *s = st; *c = ct; break;
		case 1: *s = ct; *c = -st; break;
		case 2: *s = -st; *c = -ct; break;
		default: *s = -ct; *c = st; break;
	}
}

// One standard normal variate for element `index` of the stream (seed, run, tag).
RGL_RNG_HD float rngNormal(uint64_t seed, uint32_t run, uint32_t tag, uint32_t index)
{
	RGL_NO_CONTRACT
	Philox4 ctr{{index, run, tag, 0u}};
	Philox4 o = philox4x32_10(ctr, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
	const float u1 = static_cast<float>((o.v[0] >> 8) + 1u) * 5.9604644775390625e-8f; // (0, 1]
	const float u2 = static_cast<float>(o.v[1] >> 8) * 5.9604644775390625e-8f;        // [0, 1)
	const float l = rngLog(u1);
	const float rad = sqrtf(-2.0f * l);
	float s, c;
	rngSinCos2Pi(u2, &s, &c);
	return rad * c;
}

} // namespace rgl
