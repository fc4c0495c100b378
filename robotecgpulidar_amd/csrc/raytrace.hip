// Ray generation, two-level BVH traversal, closest-hit shading and beam reduction, fused in one kernel.
//
// Per-ray semantics follow the reference's OptiX programs and beam reduction:
//   raygen      src/gpu/optixPrograms.cu:46-87 (mask, distortion, ranges, 58-sample beams :75-86, :230-248)
//   closesthit  src/gpu/optixPrograms.cu:96-216 (FP64 distance :122-129, min range :132-136, normal :139-145,
//               incident angle :148-149, texture intensity :151-167, velocities :169-212)
//   miss        src/gpu/optixPrograms.cu:89-94
//   beam reduce src/gpu/nodeKernels.cu:334-422 with saveReturnAsHit / saveReturnAsNonHit :103-207
//   shared data src/gpu/optixPrograms.cu:295-326
// The reference stores 58 samples per beam in HBM and reduces them in a second kernel; here each thread keeps
// the five return trackers in registers and shades only the samples that are returned, so no sample buffer
// exists. Traversal replaces OptiX: closest-child-first stack traversal of a TLAS (instances) and per-mesh
// BLASes, conservative slab tests, and the watertight ray/triangle test of Woop, Benthin and Wald (JCGT 2013).
// Hits are accepted for t in (0, tmax]; equal distances are resolved by (instance index, triangle index) so the
// result does not depend on traversal order.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "device_types.hpp"
#include "kernels.hpp"
#include "point_ops.hpp"
#include "runtime.hpp"

namespace rgl::gpu {

namespace {

#ifndef RGL_RT_BLOCK
#define RGL_RT_BLOCK 128
#endif
constexpr int kRtBlock = RGL_RT_BLOCK; // threads per raytrace block (the LDS short stack is laid out per block)
// Occupancy target (waves per SIMD) of the raytrace kernels. More waves in flight hide the dependent node loads
// until register spills enter the traversal loop: on gfx950 the single-sample kernels fit 5 waves (95 VGPRs, no
// spills) and are slower at 6 (80 VGPRs, spills in the loop); the divergent-beam kernels reload a few beam values
// from scratch at 5 waves and still run ~7% faster than at the compiler's 4. 0 = compiler's choice.
#ifndef RGL_RT_WAVES
#define RGL_RT_WAVES 5
#endif
#if RGL_RT_WAVES > 0
#define RGL_RT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RGL_RT_WAVES, 8)))
#else
#define RGL_RT_OCCUPANCY
#endif
#ifndef RGL_LDS_STACK
#define RGL_LDS_STACK 24
#endif
constexpr int kLdsStack = RGL_LDS_STACK;
constexpr int kOvfStack = 104;
// Consecutive logical blocks per XCD turn of the raytrace launches (xcdLogicalBlock; <= 0: one range per XCD).
// Measured on one 262k-ray urban lidar: 8 -> 125 us, hardware order 128 us, 64 -> 149 us (a launch of about one
// occupancy generation then splits unevenly costly channels between XCDs); neutral on the 64-lidar batch.
#ifndef RGL_XCD_CHUNK
#define RGL_XCD_CHUNK 8
#endif
constexpr int kXcdChunk = RGL_XCD_CHUNK;
constexpr int kPop = INT_MIN;
constexpr int kBeamSamples = 58;  // 1 primary + 3 layers x 19 vertices (src/gpu/MultiReturn.hpp:19-21)
constexpr int kBeamLayers = 3;
constexpr int kBeamVertices = 19;
// Conservative slab test. The distance of the plane p on axis a is computed as one fma, fma(p, m, -c), with per-ray
// constants (SlabRay) instead of (p - o) * inv. With u = 2^-24, inv the hardware reciprocal (within 2u of 1 / dir),
// q = fl(o * m) and the slack S = 2^-22 = 4u:
//  * near planes use m = inv and c = fl(q + |q| S) >= o * inv, so the computed distance is at most
//    (p - o) * inv rounded once: within (1 + 2u)(1 + u) of the exact one, or <= 0 if that is negative;
//  * far planes use m = fl(inv * G), G = kSlabGrow = 1 + 16u, and c = fl(q - |q| S) <= o * m, so the computed
//    distance is at least the exact one times (1 - 2u)(1 + 16u)(1 - u)^2 > 1 + 8u when positive;
// which plane of an axis is near follows from the sign of inv. The slack moves each axis' interval by 2^-22 |o_a|
// in space at most, and the growth by 16u relative: every box the exact test enters ([max near, min far] meets
// [0, tBest]) is entered, since (1 + 2u)(1 + u) < 1 + 8u and the bound tBest is grown by G as well.
// The closest hit, and therefore every result, does not depend on which extra boxes are entered.
constexpr float kSlabGrow = 1.0f + 9.5367431640625e-7f;
constexpr float kSlabSlack = 2.384185791015625e-7f; // 2^-22

// Component selection by a one-hot vector e: e.x*v.x + e.y*v.y + e.z*v.z is exactly the selected component (the
// products are exact, the other terms are zeros). Selecting by a per-lane index otherwise compiles to branches or
// to an indexed scratch array.
struct OneHot3
{
	float x, y, z;
};
__device__ __forceinline__ OneHot3 oneHot(int k)
{
	return OneHot3{k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f};
}
__device__ __forceinline__ float pick(const Vec3f& v, const OneHot3& e) { return fmaf(e.x, v.x, fmaf(e.y, v.y, e.z * v.z)); }

// Ray of the traversal: origin, reciprocal direction for the slab tests, and the shear and axis permutation of
// the watertight triangle test (axes kx, ky, kz as one-hot selectors).
struct TravRay
{
	Vec3f o;
	Vec3f inv;
	float sx, sy, sz;
	OneHot3 ex, ey, ez;
};

// Per-axis reciprocal direction of the slab tests; a zero component is replaced by 1e-30. The hardware reciprocal
// (within one ulp) is used: the slab widening covers its error, and nothing but the slab tests reads it.
__device__ __forceinline__ Vec3f invDir(Vec3f d)
{
	const float tiny = 1e-30f;
	return Vec3f{__builtin_amdgcn_rcpf(d.x != 0.0f ? d.x : tiny), __builtin_amdgcn_rcpf(d.y != 0.0f ? d.y : tiny),
	             __builtin_amdgcn_rcpf(d.z != 0.0f ? d.z : tiny)};
}

__device__ __forceinline__ TravRay makeTravRay(Vec3f o, Vec3f d)
{
	TravRay r;
	r.o = o;
	r.inv = invDir(d);
	const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
	int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
	int kx = kz == 2 ? 0 : kz + 1;
	int ky = kx == 2 ? 0 : kx + 1;
	r.ez = oneHot(kz);
	const float dz = pick(d, r.ez);
	if (dz < 0.0f) {
		const int t = kx;
		kx = ky;
		ky = t;
	}
	r.ex = oneHot(kx);
	r.ey = oneHot(ky);
	r.sx = pick(d, r.ex) / dz;
	r.sy = pick(d, r.ey) / dz;
	r.sz = 1.0f / dz;
	return r;
}

// Object-space ray of an instance: the FP64 world origin through the FP64 world->object transform, rounded once, and
// the FP32 direction through the FP32 linear part.
__device__ __forceinline__ TravRay objectRay(const InstanceRec& ir, Vec3d o64, Vec3f d)
{
	return makeTravRay(round32(apply(ir.worldToObject64, o64)), mulDir(ir.worldToObject, d));
}

// FP64 world origin of the ray a thread traces, kept in LDS ([component][thread]) between ray setup, instance entries
// during traversal and shading, so it takes no registers in the traversal loop.
template<int kBlock>
__device__ __forceinline__ double* originSlot()
{
	__shared__ double slot[3 * kBlock];
	return slot + threadIdx.x;
}
template<int kBlock>
__device__ __forceinline__ void storeOrigin(Vec3d o)
{
	double* s = originSlot<kBlock>();
	s[0] = o.x;
	s[kBlock] = o.y;
	s[2 * kBlock] = o.z;
}
template<int kBlock>
__device__ __forceinline__ Vec3d loadOrigin()
{
	const double* s = originSlot<kBlock>();
	return Vec3d{s[0], s[kBlock], s[2 * kBlock]};
}

struct HitRec
{
	float t;
	int inst;
	int prim;
	int leaf;
	float u, v;
};

// Watertight test; on success writes t and the barycentric numerators: u = vNum / det is the weight of B and
// v = wNum / det the weight of C (divided only for the hits that are kept).
__device__ __forceinline__ bool intersectTri(const TravRay& r, const LeafTri& tri, float& tOut, float& vNum, float& wNum,
                                             float& detOut)
{
	const Vec3f A{tri.ax - r.o.x, tri.ay - r.o.y, tri.az - r.o.z};
	const Vec3f B{tri.bx - r.o.x, tri.by - r.o.y, tri.bz - r.o.z};
	const Vec3f C{tri.cx - r.o.x, tri.cy - r.o.y, tri.cz - r.o.z};
	const float Akz = pick(A, r.ez), Bkz = pick(B, r.ez), Ckz = pick(C, r.ez);
	const float Ax = pick(A, r.ex) - r.sx * Akz;
	const float Ay = pick(A, r.ey) - r.sy * Akz;
	const float Bx = pick(B, r.ex) - r.sx * Bkz;
	const float By = pick(B, r.ey) - r.sy * Bkz;
	const float Cx = pick(C, r.ex) - r.sx * Ckz;
	const float Cy = pick(C, r.ey) - r.sy * Ckz;
	float U = Cx * By - Cy * Bx;
	float V = Ax * Cy - Ay * Cx;
	float W = Bx * Ay - By * Ax;
	if (U == 0.0f || V == 0.0f || W == 0.0f) {
		const double CxBy = static_cast<double>(Cx) * By, CyBx = static_cast<double>(Cy) * Bx;
		const double AxCy = static_cast<double>(Ax) * Cy, AyCx = static_cast<double>(Ay) * Cx;
		const double BxAy = static_cast<double>(Bx) * Ay, ByAx = static_cast<double>(By) * Ax;
		U = static_cast<float>(CxBy - CyBx);
		V = static_cast<float>(AxCy - AyCx);
		W = static_cast<float>(BxAy - ByAx);
	}
	if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) {
		return false;
	}
	const float det = U + V + W;
	if (det == 0.0f) {
		return false;
	}
	const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
	const float T = U * Az + V * Bz + W * Cz;
	tOut = T / det;
	vNum = V;
	wNum = W;
	detOut = det;
	return true;
}

__device__ __forceinline__ bool better(float t, int inst, int prim, const HitRec& b)
{
	return t < b.t || (t == b.t && (inst < b.inst || (inst == b.inst && prim < b.prim)));
}

struct TravStats
{
	int nodes, tris, insts;
};

// Set of the 64 B lines a launch reads (footprint measurement): open addressing over mask + 1 keys, key = line + 1,
// 0 = empty. A line that finds no free slot within kLineProbes probes sets *full.
struct LineSet
{
	unsigned long long* keys;
	uint32_t mask;
	uint32_t* full;
};
constexpr int kLineProbes = 256;

__device__ __forceinline__ uint32_t lineHash(unsigned long long k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	return static_cast<uint32_t>(k);
}

__device__ void touchLines(const LineSet* s, const void* p, int bytes)
{
	const unsigned long long a = reinterpret_cast<unsigned long long>(p);
	for (unsigned long long line = a >> 6; line <= (a + bytes - 1) >> 6; ++line) {
		const unsigned long long key = line + 1;
		uint32_t h = lineHash(key) & s->mask;
		bool placed = false;
		for (int probe = 0; probe < kLineProbes && !placed; ++probe) {
			// Slots only change from empty to a key, so a plain read that sees the key is final (hot lines such as
			// the roots are then not contended by atomics).
			unsigned long long prev = __atomic_load_n(&s->keys[h], __ATOMIC_RELAXED);
			if (prev == 0ull) {
				prev = atomicCAS(&s->keys[h], 0ull, key);
			}
			placed = prev == 0ull || prev == key;
			h = (h + 1) & s->mask;
		}
		if (!placed) {
			atomicOr(s->full, 1u);
		}
	}
}

// Upper t bound of the slab tests while the closest hit so far is at tBest: widened like the far plane and kept
// finite, so that a box whose near distance is +inf (the empty-slot box of Bvh4FNode) is never entered.
__device__ __forceinline__ float slabBound(float tBest) { return fminf(tBest * kSlabGrow, FLT_MAX); }

// Slab-test constants of a ray at one level (world or object space): the distance of a plane p bounding axis a
// from below is fma(p, mlo.a, -clo.a), from above fma(p, mhi.a, -chi.a) (near or far constants by the sign of the
// reciprocal direction).
struct SlabRay
{
	Vec3f mlo, clo, mhi, chi;
};

__device__ __forceinline__ void slabAxis(float o, float inv, float& mlo, float& clo, float& mhi, float& chi)
{
	const float invF = inv * kSlabGrow;
	const float qN = o * inv, qF = o * invF;
	const float cN = fmaf(fabsf(qN), kSlabSlack, qN);
	const float cF = fmaf(-fabsf(qF), kSlabSlack, qF);
	const bool up = inv >= 0.0f; // the lower plane is the near one
	mlo = up ? inv : invF;
	clo = up ? cN : cF;
	mhi = up ? invF : inv;
	chi = up ? cF : cN;
}

__device__ __forceinline__ SlabRay slabRay(Vec3f o, Vec3f inv)
{
	SlabRay r;
	slabAxis(o.x, inv.x, r.mlo.x, r.clo.x, r.mhi.x, r.chi.x);
	slabAxis(o.y, inv.y, r.mlo.y, r.clo.y, r.mhi.y, r.chi.y);
	slabAxis(o.z, inv.z, r.mlo.z, r.clo.z, r.mhi.z, r.chi.z);
	return r;
}

// Conservative slab test of the box (lx, ly, lz) - (hx, hy, hz): entered if [max(tn, 0), min(tf, tBound)] is
// non-empty, tBound = slabBound(tBest); tNear = tn.
__device__ __forceinline__ void slab6(const SlabRay& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                      float tBound, bool& hit, float& tNear)
{
	const float tx0 = fmaf(lx, r.mlo.x, -r.clo.x), tx1 = fmaf(hx, r.mhi.x, -r.chi.x);
	const float ty0 = fmaf(ly, r.mlo.y, -r.clo.y), ty1 = fmaf(hy, r.mhi.y, -r.chi.y);
	const float tz0 = fmaf(lz, r.mlo.z, -r.clo.z), tz1 = fmaf(hz, r.mhi.z, -r.chi.z);
	const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
	const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
	hit = fmaxf(tn, 0.0f) <= fminf(tf, tBound);
	tNear = tn;
}

// The same test from the six plane distances (lower and upper plane per axis).
__device__ __forceinline__ void slabT(float tx0, float tx1, float ty0, float ty1, float tz0, float tz1, float tBound,
                                      bool& hit, float& tNear)
{
	const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
	const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
	hit = fmaxf(tn, 0.0f) <= fminf(tf, tBound);
	tNear = tn;
}

__device__ __forceinline__ void slab(const SlabRay& r, const float* f, float tBound, bool& hit, float& tNear)
{
	slab6(r, f[0], f[1], f[2], f[3], f[4], f[5], tBound, hit, tNear);
}

// Node and triangle records are read through global-address-space pointers, so the loads are global (not flat)
// loads and only wait on the vector memory counter.
typedef float RglF4 __attribute__((ext_vector_type(4)));
typedef int RglI4 __attribute__((ext_vector_type(4)));
typedef unsigned RglU4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) RglF4 GlobalFloat4;
template<typename T>
__device__ __forceinline__ T* globalPtr(const void* p)
{
	return (T*)(p);
}

// Records every active lane of the wave reads alike are fetched with scalar loads (through the scalar cache) instead:
// vector loads deliver each lane its own copy through the vector memory data path, which node and triangle fetches
// otherwise keep busy, while one scalar load serves the whole wave. constant-address-space pointers built from
// readfirstlane values compile to scalar loads; the records are not written while a launch reads them.
typedef const __attribute__((address_space(4))) RglF4 ConstFloat4;
__device__ __forceinline__ const void* firstLanePtr(const void* p)
{
	const uint64_t a = reinterpret_cast<uint64_t>(p);
	const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a));
	const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32));
	return reinterpret_cast<const void*>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ bool waveUniform(const void* p)
{
	return __ballot(p == firstLanePtr(p)) == __ballot(1);
}
template<typename T>
__device__ __forceinline__ T* scalarPtr(const void* uniformP)
{
	return (T*)(uniformP);
}

// A record of `kQuads` 16-byte quads at p (16-byte aligned), by scalar loads when every active lane reads the same
// one (about two thirds of the node steps of the 64-lidar frame), else by vector loads. (Serving the lanes of the first
// lane's record by scalar loads also in divergent steps, alone or one distinct record after another, measured 15-20%
// slower: the scalar load's latency then adds to the vector load's.)
template<int kQuads>
__device__ __forceinline__ void loadQuads(const void* p, RglF4 (&q)[kQuads])
{
	if (waveUniform(p)) {
		ConstFloat4* sp = scalarPtr<ConstFloat4>(firstLanePtr(p));
#pragma unroll
		for (int i = 0; i < kQuads; ++i) {
			q[i] = sp[i];
		}
		return;
	}
	GlobalFloat4* vp = globalPtr<GlobalFloat4>(p);
#pragma unroll
	for (int i = 0; i < kQuads; ++i) {
		q[i] = vp[i];
	}
}

__device__ __forceinline__ LeafTri loadLeafTri(const LeafTri* tris, int leaf)
{
	RglF4 q[3];
	loadQuads<3>(tris + leaf, q);
	const RglF4 a = q[0], b = q[1], c = q[2];
	LeafTri t;
	t.ax = a.x;
	t.ay = a.y;
	t.az = a.z;
	t.prim = __float_as_int(a.w);
	t.bx = b.x;
	t.by = b.y;
	t.bz = b.z;
	t.pad0 = 0.0f;
	t.cx = c.x;
	t.cy = c.y;
	t.cz = c.z;
	t.pad1 = 0.0f;
	return t;
}

__device__ __forceinline__ float ubyte(uint32_t w, int c) { return static_cast<float>((w >> (8 * c)) & 0xffu); }

// Closest hit of one ray against the two-level BVH; returns false if nothing is hit in (0, tmax].
// kFmt selects the node format of the TLAS and BLASes (BvhFormat). With kStats the canonical binary trees are
// traversed (sc.tlas2 and each instance's binary BLAS) and visits are counted. o is the FP32 rounding of the FP64
// origin the caller stored with storeOrigin<kBlock>; instance entries transform the FP64 one.
//
// One stack serves both levels. Entering an instance records the stack depth (instBase); when pops bring the depth
// back to it, the BLAS is finished and traversal continues in the TLAS. The ray of the node steps (origin and
// reciprocal direction of the current level) and the object-space ray record change only at the top of the node
// loop, on instance entry and on the return to the TLAS, so the pop loop carries no ray state.
// With kFoot every record read (nodes in format kFmt, leaf triangles, instance records) is added to *ls.
template<bool kStats, int kFmt, int kBlock = kRtBlock, bool kFoot = false>
__device__ bool traceClosest(const SceneView& sc, Vec3f o, Vec3f d, float tmax, HitRec& best, TravStats* st,
                             const LineSet* ls = nullptr)
{
	static_assert(!kStats || kFmt == kBvh2, "traversal statistics are defined on the binary trees");
	// Short stack in LDS, [depth][thread] so a wave's pushes hit distinct banks; deeper entries spill to scratch.
	__shared__ int32_t ldsStack[kLdsStack * kBlock];
	best.t = tmax;
	best.inst = INT_MAX;
	best.prim = INT_MAX;
	best.leaf = -1;
	best.u = best.v = 0.0f;
	if (sc.tlasNodes == nullptr || sc.instanceCount == 0 || !(tmax > 0.0f)) {
		return false;
	}
	const int tid = threadIdx.x;
	// The overflow part holds the same int32 entries as float bit patterns: with a different element type the
	// compiler cannot fold "LDS load or scratch load" into one load through a generic pointer.
	float ovf[kOvfStack];
	int sp = 0;
	bool overflow = false;
#define RGL_PUSH(v)                                                                                                    \
	do {                                                                                                               \
		if (sp < kLdsStack) {                                                                                          \
			ldsStack[sp * kBlock + tid] = (v);                                                                         \
			++sp;                                                                                                      \
		} else if (sp < kLdsStack + kOvfStack) {                                                                       \
			ovf[sp - kLdsStack] = __int_as_float(v);                                                                   \
			++sp;                                                                                                      \
		} else {                                                                                                       \
			overflow = true;                                                                                           \
		}                                                                                                              \
	} while (0)
#define RGL_POP() (--sp, sp < kLdsStack ? ldsStack[sp * kBlock + tid] : __float_as_int(ovf[sp - kLdsStack]))
#define RGL_CSWAP(a, b)                                                                                                \
	do {                                                                                                               \
		const bool sw = k##b < k##a;                                                                                   \
		const float ka_ = sw ? k##b : k##a, kb_ = sw ? k##a : k##b;                                                    \
		const int ra_ = sw ? r##b : r##a, rb_ = sw ? r##a : r##b;                                                      \
		k##a = ka_;                                                                                                    \
		k##b = kb_;                                                                                                    \
		r##a = ra_;                                                                                                    \
		r##b = rb_;                                                                                                    \
	} while (0)

	// Slab constants of the node steps at the current level (world in the TLAS, object space in a BLAS) and the
	// object-space ray record of the entered instance; both change only at the top of the node loop (instance
	// entry, and the return to the TLAS after the pops left the instance). The world constants are recomputed on
	// each return from the origin kept in LDS (o is its FP32 rounding) rather than held in registers.
	SlabRay sr = slabRay(o, invDir(d));
	TravRay obj;
	obj.o = Vec3f{0.0f, 0.0f, 0.0f};
	obj.inv = obj.o;
	obj.sx = obj.sy = obj.sz = 0.0f;
	obj.ex = obj.ey = obj.ez = oneHot(0);
	const void* tlas = kStats ? static_cast<const void*>(sc.tlas2) : sc.tlasNodes;
	const void* nodes = tlas;
	const LeafTri* tris = nullptr;
	int inst = -1;
	int instBase = -1;
	bool inBlas = false;
	bool found = false;
	float tBound = slabBound(best.t);
	// One node step at the current level: slab-test the children of `node`, push the far hit ones (far to near) and
	// return the nearest hit child, or kPop.
	auto stepNode = [&](int node) -> int {
		if (kStats) {
			st->nodes++;
		}
		if constexpr (kFoot) {
			const int nodeBytes = kFmt == kBvh4F   ? sizeof(Bvh4FNode)
			                      : kFmt == kBvh4Q ? sizeof(Bvh4QNode)
			                      : kFmt == kBvh4H ? sizeof(Bvh4HNode)
			                                       : sizeof(BvhNode);
			touchLines(ls, static_cast<const char*>(nodes) + static_cast<size_t>(node) * nodeBytes, nodeBytes);
		}
		int next;
		if constexpr (kFmt == kBvh2) {
			RglF4 q[4];
			loadQuads<4>(static_cast<const BvhNode*>(nodes) + node, q);
			const RglF4 q0 = q[0], q1 = q[1], q2 = q[2];
			const RglI4 q3 = __builtin_bit_cast(RglI4, q[3]);
			const float b0[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
			const float b1[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
			bool h0, h1;
			float n0, n1;
			slab(sr, b0, tBound, h0, n0);
			slab(sr, b1, tBound, h1, n1);
			if (h0 && h1) {
				const bool firstNear = n0 <= n1;
				RGL_PUSH(firstNear ? q3.y : q3.x);
				next = firstNear ? q3.x : q3.y;
			} else if (h0) {
				next = q3.x;
			} else if (h1) {
				next = q3.y;
			} else {
				next = kPop;
			}
		} else {
			// Four children: slab-test each, order the hit ones near to far, descend into the nearest.
			bool h0, h1, h2, h3;
			float k0, k1, k2, k3;
			int r0, r1, r2, r3;
			if constexpr (kFmt == kBvh4F) {
				// Empty slots hold boxes no ray enters, so the child count is not read.
				RglF4 q[7];
				loadQuads<7>(static_cast<const Bvh4FNode*>(nodes) + node, q);
				const RglF4 lx = q[0], ly = q[1], lz = q[2], hx = q[3], hy = q[4], hz = q[5];
				const RglI4 ch = __builtin_bit_cast(RglI4, q[6]);
				slab6(sr, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, tBound, h0, k0);
				slab6(sr, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, tBound, h1, k1);
				slab6(sr, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, tBound, h2, k2);
				slab6(sr, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, tBound, h3, k3);
				r0 = ch.x;
				r1 = ch.y;
				r2 = ch.z;
				r3 = ch.w;
			} else {
				// Quantized bounds (8-bit Bvh4QNode, 16-bit Bvh4HNode): the node origin, one power-of-two scale per axis
				// (exp byte = e + 127, so byte << 23 is the bit pattern of 2^e), the child count, then per axis and child
				// a lower and an upper level q.
				constexpr bool kWide16 = kFmt == kBvh4H;
				constexpr int kQ = kWide16 ? 5 : 4;
				RglF4 q[kQ];
				loadQuads<kQ>(static_cast<const char*>(nodes) + static_cast<size_t>(node) * (kWide16 ? sizeof(Bvh4HNode) : sizeof(Bvh4QNode)), q);
				const RglU4 u0 = __builtin_bit_cast(RglU4, q[0]), u1 = __builtin_bit_cast(RglU4, q[1]);
				const RglU4 u2 = __builtin_bit_cast(RglU4, q[2]), u3 = __builtin_bit_cast(RglU4, q[3]);
				const float ox = __uint_as_float(u0.x), oy = __uint_as_float(u0.y), oz = __uint_as_float(u0.z);
				const float sx = __uint_as_float((u0.w & 0xffu) << 23);
				const float sy = __uint_as_float(((u0.w >> 8) & 0xffu) << 23);
				const float sz = __uint_as_float(((u0.w >> 16) & 0xffu) << 23);
				const int cnt = static_cast<int>(u0.w >> 24);
				// Dequantization folded into the slab constants: the plane origin + q * 2^e at distance
				// fma(plane, m, -c) equals q * (2^e m) + (origin m - c), so per node and axis A = 2^e m (exact: a
				// power-of-two scaling) and B = fma(origin, m, -c) are formed once and each bound costs one conversion
				// and one fma. Against the plane's distance computed directly this adds an absolute error of at most
				// u |B| (u = 2^-24), and |B| <= |t| + L * 2^e |m| (L = 255 or 65535 levels): the |t| part stays inside
				// the widening of the plain test, and the rest is a shift of the plane by at most u * L * 2^e in space,
				// which the builder's outward margin of 2^-20 of the node extent (with L * 2^e < 2 * extent) exceeds.
				const float aLx = sx * sr.mlo.x, aHx = sx * sr.mhi.x, bLx = fmaf(ox, sr.mlo.x, -sr.clo.x);
				const float bHx = fmaf(ox, sr.mhi.x, -sr.chi.x);
				const float aLy = sy * sr.mlo.y, aHy = sy * sr.mhi.y, bLy = fmaf(oy, sr.mlo.y, -sr.clo.y);
				const float bHy = fmaf(oy, sr.mhi.y, -sr.chi.y);
				const float aLz = sz * sr.mlo.z, aHz = sz * sr.mhi.z, bLz = fmaf(oz, sr.mlo.z, -sr.clo.z);
				const float bHz = fmaf(oz, sr.mhi.z, -sr.chi.z);
				if constexpr (kWide16) {
					// Words: u1 = lo x (children 0-1, 2-3), lo y; u2 = lo z, hi x; u3 = hi y, hi z.
					auto lv = [](uint32_t w0, uint32_t w1, int c) {
						const uint32_t w = c < 2 ? w0 : w1;
						return static_cast<float>((c & 1) ? (w >> 16) : (w & 0xffffu));
					};
#define RGL_HBOX(c, h, k)                                                                                              \
	slabT(fmaf(lv(u1.x, u1.y, c), aLx, bLx), fmaf(lv(u2.z, u2.w, c), aHx, bHx), fmaf(lv(u1.z, u1.w, c), aLy, bLy),       \
	      fmaf(lv(u3.x, u3.y, c), aHy, bHy), fmaf(lv(u2.x, u2.y, c), aLz, bLz), fmaf(lv(u3.z, u3.w, c), aHz, bHz), tBound, \
	      h, k)
					RGL_HBOX(0, h0, k0);
					RGL_HBOX(1, h1, k1);
					RGL_HBOX(2, h2, k2);
					RGL_HBOX(3, h3, k3);
#undef RGL_HBOX
					const RglI4 ch = __builtin_bit_cast(RglI4, q[kQ - 1]);
					r0 = ch.x;
					r1 = ch.y;
					r2 = ch.z;
					r3 = ch.w;
				} else {
#define RGL_QBOX(c, h, k)                                                                                              \
	slabT(fmaf(ubyte(u1.x, c), aLx, bLx), fmaf(ubyte(u1.w, c), aHx, bHx), fmaf(ubyte(u1.y, c), aLy, bLy),               \
	      fmaf(ubyte(u2.x, c), aHy, bHy), fmaf(ubyte(u1.z, c), aLz, bLz), fmaf(ubyte(u2.y, c), aHz, bHz), tBound, h, k)
					RGL_QBOX(0, h0, k0);
					RGL_QBOX(1, h1, k1);
					RGL_QBOX(2, h2, k2);
					RGL_QBOX(3, h3, k3);
#undef RGL_QBOX
					r0 = static_cast<int>(u2.z);
					r1 = static_cast<int>(u2.w);
					r2 = static_cast<int>(u3.x);
					r3 = static_cast<int>(u3.y);
				}
				// Empty slots decode to inverted boxes, which the slab test does not reject.
				h1 = h1 && cnt > 1;
				h2 = h2 && cnt > 2;
				h3 = h3 && cnt > 3;
			}
			const float kMiss = INFINITY;
			k0 = h0 ? k0 : kMiss;
			k1 = h1 ? k1 : kMiss;
			k2 = h2 ? k2 : kMiss;
			k3 = h3 ? k3 : kMiss;
			RGL_CSWAP(0, 1);
			RGL_CSWAP(2, 3);
			RGL_CSWAP(0, 2);
			RGL_CSWAP(1, 3);
			RGL_CSWAP(1, 2);
			if (__all(sp <= kLdsStack - 3)) {
				// Every lane has room for three LDS entries: write all three slots unconditionally and advance the
				// stack pointer past the hit ones only (children are sorted, so misses come last and are
				// overwritten or left above the top).
				int t = sp;
				ldsStack[t * kBlock + tid] = r3;
				t += k3 != kMiss ? 1 : 0;
				ldsStack[t * kBlock + tid] = r2;
				t += k2 != kMiss ? 1 : 0;
				ldsStack[t * kBlock + tid] = r1;
				t += k1 != kMiss ? 1 : 0;
				sp = t;
			} else {
				if (k3 != kMiss) {
					RGL_PUSH(r3);
				}
				if (k2 != kMiss) {
					RGL_PUSH(r2);
				}
				if (k1 != kMiss) {
					RGL_PUSH(r1);
				}
			}
			next = k0 != kMiss ? r0 : kPop;
		}
		return next;
	};
	// Triangle leaf test of the entered instance; keeps the closest hit (ties: lower instance, then triangle).
	auto testLeaf = [&](int leaf) {
		if (kStats) {
			st->tris++;
		}
		if constexpr (kFoot) {
			touchLines(ls, tris + leaf, sizeof(LeafTri));
		}
		const LeafTri lt = loadLeafTri(tris, leaf);
		float t, vNum, wNum, det;
		if (intersectTri(obj, lt, t, vNum, wNum, det) && t > 0.0f && t <= tmax && better(t, inst, lt.prim, best)) {
			best.t = t;
			best.inst = inst;
			best.prim = lt.prim;
			best.leaf = leaf;
			best.u = vNum / det;
			best.v = wNum / det;
			tBound = slabBound(t);
			found = true;
		}
	};
	// Entering the instance of TLAS leaf `leafRef`: its BLAS with the object-space ray (FP64 origin mapped once).
	auto enterInstance = [&](int leafRef) {
		inst = ~leafRef;
		instBase = sp;
		static_assert(sizeof(InstanceRec) == 10 * sizeof(RglF4), "instance record read as 10 quads");
		InstanceRec ir;
		{
			RglF4 q[10];
			loadQuads<10>(sc.instances + inst, q);
			__builtin_memcpy(&ir, q, sizeof(ir));
		}
		if constexpr (kFoot) {
			touchLines(ls, sc.instances + inst, sizeof(InstanceRec));
		}
		obj = objectRay(ir, loadOrigin<kBlock>(), d);
		sr = slabRay(obj.o, obj.inv);
		nodes = kStats ? static_cast<const void*>(sc.shades[inst].blas2) : ir.blasNodes;
		tris = ir.blasTris;
		inBlas = true;
		if (kStats) {
			st->insts++;
		}
	};
	auto returnToTlas = [&]() {
		// Opaque copies keep the recomputation here: hoisted out of the loop, its twelve results would take
		// registers (or scratch) for the whole traversal.
		Vec3d o64 = loadOrigin<kBlock>();
		Vec3f dd = d;
		asm volatile("" : "+v"(o64.x), "+v"(o64.y), "+v"(o64.z), "+v"(dd.x), "+v"(dd.y), "+v"(dd.z));
		sr = slabRay(round32(o64), invDir(dd));
		nodes = tlas;
		inBlas = false;
	};
	auto popEntry = [&]() -> int {
		if (__all(sp <= kLdsStack)) {
			--sp;
			return ldsStack[sp * kBlock + tid];
		}
		return RGL_POP();
	};
	// Closest child first; a triangle leaf is tested as soon as it is reached. (Postponing leaves until most lanes of
	// the wave hold one, with or without speculative node steps meanwhile, measured 20-30% slower on the 64-lidar
	// batch: the extra loop control costs more than the fuller triangle tests save.)
	int node = 0;
	while (true) {
		if (node < 0) {
			enterInstance(node);
			node = 0;
		} else if (inBlas && inst < 0) {
			returnToTlas();
		}
		int next = stepNode(node);
		// Pops and triangle leaves until the next inner node or instance leaf (one exit, so the loop-carried
		// values need no copies at divergent exits).
		bool done = false;
		bool tri;
		do {
			if (next == kPop) {
				if (sp == instBase) {
					inst = -1; // the entered instance's BLAS is finished
					instBase = -1;
				}
				done = sp == 0 || overflow;
				if (!done) {
					next = popEntry();
				}
			}
			tri = !done && next < 0 && inst >= 0;
			if (tri) {
				testLeaf(~next);
				next = kPop;
			}
		} while (tri);
		if (done) {
			break;
		}
		node = next;
	}
	if (overflow) {
		*sc.overflow = 1;
	}
	return found;
#undef RGL_PUSH
#undef RGL_POP
#undef RGL_CSWAP
}

// Beam sample transform (src/gpu/optixPrograms.cu:230-248): rotation about X by the vertical angle times
// rotation about Y by the horizontal angle, on a ring shrinking with the layer index.
__device__ Mat3x4f beamSampleTransform(float hHalf, float vHalf, int layer, int vertex)
{
	const float shrink = 1.0f - static_cast<float>(layer) / kBeamLayers;
	const float hCur = hHalf * shrink;
	const float vCur = vHalf * shrink;
	const float step = 2.0f * static_cast<float>(M_PI) / kBeamVertices;
	const float hAngle = hCur * cosf(static_cast<float>(vertex) * step);
	const float vAngle = vCur * sinf(static_cast<float>(vertex) * step);
	return Mat3x4f::rotationRad(vAngle, 0.0f, 0.0f) * Mat3x4f::rotationRad(0.0f, hAngle, 0.0f);
}

// World-frame pose of ray i of the rays chain in FP64; a transform the chain deferred to this kernel is composed here
// in FP64 (exact products of the FP32 operands, fixed rounding order), so every kernel gets the same bits.
__device__ __forceinline__ Mat3x4d rayWorldPose64(const RaytraceParams& p, int i)
{
	return p.rayPreTransformed ? widen(p.rayPreTransform) * widen(p.rays[i]) : widen(p.rays[i]);
}
__device__ __forceinline__ Mat3x4f rayWorldPose(const RaytraceParams& p, int i) { return round32(rayWorldPose64(p, i)); }

// Ray origin = M*(0,0,0) and direction = M*(0,0,1) - M*(0,0,0), i.e. the matrix's fourth and third columns. (The
// reference subtracts in FP32, src/gpu/optixPrograms.cu:222-228, which rounds the direction to the ulp of the
// translation: ~4e-6 for a sensor 40 m from the origin, millimetres at grazing incidence. The exact column is used.)
__device__ __forceinline__ void rayOriginDir(const Mat3x4f& m, Vec3f& o, Vec3f& d)
{
	o = Vec3f{m.rc[0][3], m.rc[1][3], m.rc[2][3]};
	d = Vec3f{m.rc[0][2], m.rc[1][2], m.rc[2][2]};
}
__device__ __forceinline__ Vec3d originOf(const Mat3x4d& m) { return Vec3d{m.rc[0][3], m.rc[1][3], m.rc[2][3]}; }
__device__ __forceinline__ Vec3f directionOf(const Mat3x4d& m)
{
	return Vec3f{static_cast<float>(m.rc[0][2]), static_cast<float>(m.rc[1][2]), static_cast<float>(m.rc[2][2])};
}

struct Shade
{
	int isHit;
	float distance;
	float intensity;
	Vec3f normal;
	float incident;
	Vec3f absVel, relVel;
	float radial;
	int entityId;
	float retro;
};

// Closest-hit program for one sample (src/gpu/optixPrograms.cu:96-216). `full` also computes the fields that the
// beam reduction does not need (normal, velocities).
// so64: the sample's FP64 world origin.
__device__ Shade shadeSample(const SceneView& sc, const RaytraceParams& p, float minRange, Vec3d so64, Vec3f sd,
                             const HitRec& h, bool full)
{
	Shade s;
	const InstanceShade& is = sc.shades[h.inst];
	const Vec3i tri = is.indices[h.prim];
	const Vec3f A = is.vertices[tri.x], B = is.vertices[tri.y], C = is.vertices[tri.z];
	// Barycentric point A + u(B-A) + v(C-A): exact when A, B, C share a coordinate (axis-aligned faces). It goes to
	// world space and to the distance in FP64 (src/gpu/optixPrograms.cu:122-129 takes the length in FP64).
	const Vec3f hitObject = A + h.u * (B - A) + h.v * (C - A);
	const Vec3d hitWorld64 = apply(is.objectToWorld, Vec3d{hitObject.x, hitObject.y, hitObject.z});
	const Vec3f hitWorld = round32(hitWorld64);
	const double dx = hitWorld64.x - so64.x;
	const double dy = hitWorld64.y - so64.y;
	const double dz = hitWorld64.z - so64.z;
	const double distance = sqrt(dx * dx + dy * dy + dz * dz);
	if (distance < minRange) {
		s.isHit = 0;
		s.distance = p.nearNonHitDistance;
		s.intensity = 0.0f;
		return s;
	}
	s.isHit = 1;
	s.distance = static_cast<float>(distance);
	// World normal from the world-space edges (src/gpu/optixPrograms.cu:139-145), formed as the entity's linear part
	// applied to the object-space edges: the same vectors without the translation, so large world coordinates do not
	// cancel in the differences.
	const Vec3f n = normalized(cross(mulDir(is.objectToWorld, B - A), mulDir(is.objectToWorld, C - A)));
	const float cosI = fabsf(dot(n, sd));
	s.normal = n;
	s.incident = acosf(cosI);
	float intensity = p.defaultIntensity;
	if (is.uvs != nullptr && is.texels != nullptr) {
		const Vec2f ua = is.uvs[tri.x], ub = is.uvs[tri.y], uc = is.uvs[tri.z];
		const float w = 1.0f - h.u - h.v;
		const float tu = w * ua.x + h.u * ub.x + h.v * uc.x;
		const float tv = w * ua.y + h.u * ub.y + h.v * uc.y;
		// Point sampling, normalized coordinates, wrap addressing on both axes (src/scene/Texture.cpp:52-56).
		const float fu = tu - floorf(tu), fv = tv - floorf(tv);
		int ix = static_cast<int>(floorf(fu * is.texWidth));
		int iy = static_cast<int>(floorf(fv * is.texHeight));
		ix = min(max(ix, 0), is.texWidth - 1);
		iy = min(max(iy, 0), is.texHeight - 1);
		intensity = static_cast<float>(is.texels[static_cast<size_t>(iy) * is.texWidth + ix]);
	}
	s.intensity = intensity * cosI;
	s.entityId = is.entityId;
	s.retro = is.laserRetro;
	const float qnan = __int_as_float(0x7fc00000);
	s.absVel = Vec3f{qnan, qnan, qnan};
	s.relVel = Vec3f{qnan, qnan, qnan};
	s.radial = qnan;
	if (full && p.velocityRequested && p.sceneDeltaTime > 0.0f) {
		Vec3f disp{0.0f, 0.0f, 0.0f};
		if (is.hasPrev) {
			disp = hitWorld - is.prevObjectToWorld * hitObject;
		}
		if (is.displacement != nullptr) {
			// Vertex animation: the vertex displacements rotated to world space, interpolated at the hit.
			const Mat3x4f rot = is.objectToWorld.rotation();
			const Vec3f dA = rot * is.displacement[tri.x];
			const Vec3f dB = rot * is.displacement[tri.y];
			const Vec3f dC = rot * is.displacement[tri.z];
			disp = disp + ((1.0f - h.u - h.v) * dA + h.u * dB + h.v * dC);
		}
		const Vec3f absVel = disp / p.sceneDeltaTime;
		const Mat3x4f rotInv = p.rayOriginToWorld.rotation().inverse();
		const Vec3f absInSensor = rotInv * absVel;
		const Vec3f relLin = absInSensor - p.sensorLinearVelocity;
		const Vec3f hitRays = round32(apply(p.worldToRayOrigin64, hitWorld64));
		const Vec3f relAng = Vec3f{0.0f, 0.0f, 0.0f} - cross(p.sensorAngularVelocity, hitRays);
		s.absVel = absVel;
		s.relVel = relLin + relAng;
		s.radial = dot(normalized(hitRays), s.relVel);
	}
	return s;
}

// o, d: origin and direction of the beam's world-frame ray (before distortion).
__device__ void writeReturnHit(const RaytraceOutputs& out, Vec3f o, Vec3f d, int rp, uint8_t type, const Shade& s)
{
	if (out.xyz) {
		out.xyz[rp] = o + normalized(d) * s.distance;
	}
	if (out.isHit) out.isHit[rp] = 1;
	if (out.returnType) out.returnType[rp] = type;
	if (out.distance) out.distance[rp] = s.distance;
	if (out.intensityF32) out.intensityF32[rp] = s.intensity;
	if (out.intensityU8) {
		out.intensityU8[rp] = s.intensity < 255.0f ? static_cast<uint8_t>(roundf(s.intensity)) : static_cast<uint8_t>(255);
	}
	if (out.entityId) out.entityId[rp] = s.entityId;
	if (out.absVelocity) out.absVelocity[rp] = s.absVel;
	if (out.relVelocity) out.relVelocity[rp] = s.relVel;
	if (out.radialSpeed) out.radialSpeed[rp] = s.radial;
	if (out.normal) out.normal[rp] = s.normal;
	if (out.incidentAngle) out.incidentAngle[rp] = s.incident;
	if (out.laserRetro) out.laserRetro[rp] = s.retro;
}

__device__ void writeReturnNonHit(const RaytraceOutputs& out, Vec3f o, Vec3f d, int rp, uint8_t type, float d0)
{
	const float qnan = __int_as_float(0x7fc00000);
	if (out.xyz) {
		Vec3f disp = normalized(d) * d0;
		disp.x = isnan(disp.x) ? 0.0f : disp.x;
		disp.y = isnan(disp.y) ? 0.0f : disp.y;
		disp.z = isnan(disp.z) ? 0.0f : disp.z;
		out.xyz[rp] = o + disp;
	}
	if (out.isHit) out.isHit[rp] = 0;
	if (out.returnType) out.returnType[rp] = type;
	if (out.distance) out.distance[rp] = d0;
	if (out.intensityF32) out.intensityF32[rp] = 0.0f;
	if (out.intensityU8) out.intensityU8[rp] = 0;
	if (out.entityId) out.entityId[rp] = 0;
	if (out.absVelocity) out.absVelocity[rp] = Vec3f{qnan, qnan, qnan};
	if (out.relVelocity) out.relVelocity[rp] = Vec3f{qnan, qnan, qnan};
	if (out.radialSpeed) out.radialSpeed[rp] = 0.0f;
	if (out.normal) out.normal[rp] = Vec3f{qnan, qnan, qnan};
	if (out.incidentAngle) out.incidentAngle[rp] = qnan;
	if (out.laserRetro) out.laserRetro[rp] = 0.0f;
}

__device__ __forceinline__ uint8_t returnTypeOf(int mode, int r) { return static_cast<uint8_t>((mode >> (8 * r)) & 0xff); }

// Beam-shared fields (src/gpu/optixPrograms.cu:295-326) of every return point of the beam.
__device__ void writeBeamShared(const RaytraceOutputs& out, const RaytraceParams& p, int beam, const Mat3x4f& rayWorld)
{
	if (!(out.rayIdx || out.azimuth || out.elevation || out.timeStampF64 || out.timeStampU32 || out.ringId)) {
		return;
	}
	const int R = p.returnCount;
	const Mat3x4f rayLocal = p.worldToRayOrigin * rayWorld;
	for (int r = 0; r < R; ++r) {
		const int rp = beam * R + r;
		if (out.rayIdx) out.rayIdx[rp] = static_cast<uint32_t>(beam);
		if (out.azimuth) out.azimuth[rp] = rayLocal.toRotationYOrderZXYLeftHandRad();
		if (out.elevation) out.elevation[rp] = rayLocal.toRotationXOrderZXYLeftHandRad();
		if (out.timeStampF64) {
			out.timeStampF64[rp] = p.doDistortion ? static_cast<double>(p.timeOffsetsMs[beam] * 1e-3f) : 0.0;
		}
		if (out.timeStampU32) {
			out.timeStampU32[rp] = p.doDistortion ? static_cast<uint32_t>(p.timeOffsetsMs[beam] * 1e6f) : 0u;
		}
		if (out.ringId && p.ringIds) out.ringId[rp] = static_cast<uint16_t>(p.ringIds[beam % p.ringIdsCount]);
	}
}

// The beam's ray after velocity distortion (src/gpu/optixPrograms.cu:61-70): the sensor-frame ray moved by the
// sensor's motion over the ray's time offset, composed in FP64 (the motion itself, TRS, is formed in FP32).
__device__ __forceinline__ Mat3x4d distortedRay64(const RaytraceParams& p, int beam, const Mat3x4d& rayWorld)
{
	if (!p.doDistortion) {
		return rayWorld;
	}
	const Mat3x4d rayLocal = p.worldToRayOrigin64 * rayWorld;
	const float toDeg = static_cast<float>(180.0 / M_PI);
	const float off = p.timeOffsetsMs[beam];
	const Vec3f t = (off * p.sensorLinearVelocity) * 0.001f;
	const Vec3f rdeg = (off * (p.sensorAngularVelocity * toDeg)) * 0.001f;
	return widen(p.rayOriginToWorld) * (widen(TRS(t, rdeg, Vec3f{1.0f, 1.0f, 1.0f})) * rayLocal);
}

// One beam traced with a single sample (no divergence): the beam reduction degenerates to "every return is the
// primary sample" (src/gpu/nodeKernels.cu:345-354,399-421 with one sample), so no trackers are kept.
template<int kFmt, int kBlock = kRtBlock>
__device__ __forceinline__ void raytraceBeam(const SceneView& sc, const RaytraceParams& p, const RaytraceOutputs& out,
                                             const int beam)
{
	const Mat3x4d rayWorld64 = rayWorldPose64(p, beam);
	const int R = p.returnCount;
	writeBeamShared(out, p, beam, round32(rayWorld64));
	const Vec2f range = p.rangesCount == 1 ? p.ranges[0] : p.ranges[beam];
	if (p.mask != nullptr && p.mask[beam] == 0) {
		const Vec3f wo = round32(originOf(rayWorld64)), wd = directionOf(rayWorld64);
		for (int r = 0; r < R; ++r) {
			writeReturnNonHit(out, wo, wd, beam * R + r, returnTypeOf(p.returnMode, r), p.farNonHitDistance);
		}
		return;
	}
	const Mat3x4d ray64 = distortedRay64(p, beam, rayWorld64);
	const Vec3f so = round32(originOf(ray64)), sd = directionOf(ray64);
	storeOrigin<kBlock>(originOf(ray64));
	HitRec h;
	float d0 = p.farNonHitDistance;
	const bool hit = traceClosest<false, kFmt, kBlock>(sc, so, sd, range.y, h, nullptr);
	// Points are placed along the undistorted world ray; without distortion that is the traced ray, otherwise it is
	// recomputed here rather than kept in registers through the traversal.
	Vec3f wo = so, wd = sd;
	if (p.doDistortion) {
		rayOriginDir(rayWorldPose(p, beam), wo, wd);
	}
	if (hit) {
		const Shade sh = shadeSample(sc, p, range.x, loadOrigin<kBlock>(), sd, h, p.velocityRequested != 0);
		if (sh.isHit) {
			for (int r = 0; r < R; ++r) {
				const uint8_t type = returnTypeOf(p.returnMode, r);
				if (type >= 1 && type <= 5) {
					writeReturnHit(out, wo, wd, beam * R + r, type, sh);
				}
			}
			return;
		}
		d0 = sh.distance;
	}
	for (int r = 0; r < R; ++r) {
		writeReturnNonHit(out, wo, wd, beam * R + r, returnTypeOf(p.returnMode, r), d0);
	}
}

// One divergent beam per wave: lane s < 58 traces sample s (the reference's 58-sample beam,
// src/gpu/optixPrograms.cu:75-86,230-248), so the samples of a beam, which are nearly parallel, traverse together.
// The beam reduction of src/gpu/nodeKernels.cu:334-422 (first hit seeds every tracker; then strict comparisons in
// sample order) runs on every lane over the samples' (hit, distance, intensity) read across lanes, which keeps it
// wave-uniform and exact; the lane holding each selected sample shades it in full and writes the return.
template<int kFmt>
__device__ __forceinline__ void raytraceBeamWave(const SceneView& sc, const RaytraceParams& p, const RaytraceOutputs& out,
                                                 const int beam)
{
	const int lane = threadIdx.x & 63;
	const Mat3x4d rayWorld64 = rayWorldPose64(p, beam);
	const Mat3x4f rayWorld = round32(rayWorld64);
	const int R = p.returnCount;
	if (lane == 0) {
		writeBeamShared(out, p, beam, rayWorld);
	}
	Vec3f wo, wd;
	rayOriginDir(rayWorld, wo, wd);
	const Vec2f range = p.rangesCount == 1 ? p.ranges[0] : p.ranges[beam];
	if (p.mask != nullptr && p.mask[beam] == 0) {
		if (lane == 0) {
			for (int r = 0; r < R; ++r) {
				writeReturnNonHit(out, wo, wd, beam * R + r, returnTypeOf(p.returnMode, r), p.farNonHitDistance);
			}
		}
		return;
	}
	const Mat3x4d ray64 = distortedRay64(p, beam, rayWorld64);
	const Mat3x4f ray = round32(ray64);
	storeOrigin<kRtBlock>(originOf(ray64)); // every sample of the beam starts at the beam's origin
	Vec3f so{0.0f, 0.0f, 0.0f}, sd{0.0f, 0.0f, 1.0f};
	HitRec h;
	h.t = 0.0f;
	h.inst = h.prim = h.leaf = -1;
	h.u = h.v = 0.0f;
	int isHit = 0;
	float dist = p.farNonHitDistance, inten = 0.0f;
	if (lane < kBeamSamples) {
		Mat3x4f sr = ray;
		if (lane > 0) {
			const int layer = (lane - 1) / kBeamVertices;
			const int vertex = (lane - 1) % kBeamVertices;
			sr = ray * beamSampleTransform(p.hHalfDivergence, p.vHalfDivergence, layer, vertex);
		}
		rayOriginDir(sr, so, sd);
		if (traceClosest<false, kFmt>(sc, so, sd, range.y, h, nullptr)) {
			const Shade sh = shadeSample(sc, p, range.x, loadOrigin<kRtBlock>(), sd, h, false);
			isHit = sh.isHit;
			dist = sh.distance;
			inten = sh.intensity;
		}
	}
	// Trackers hold (sample, key): distance for FIRST / SECOND / LAST, intensity for STRONGEST / SECOND_STRONGEST.
	// The samples that hit are visited in sample order through the ballot of their lanes; the sample index is
	// wave-uniform, so each value is read from its lane with v_readlane into a scalar register (a cross-lane shuffle
	// would be an LDS permute and a wait per value in this sequential loop).
	int fS = -1, sS = -1, lS = -1, stS = -1, ssS = -1;
	float fD = 0.0f, sD = 0.0f, lD = 0.0f, stI = 0.0f, ssI = 0.0f;
	const float d0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dist), 0));
	for (uint64_t pending = __ballot(isHit != 0); pending != 0; pending &= pending - 1) {
		const int k = static_cast<int>(__builtin_ctzll(pending));
		const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dist), k));
		const float I = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(inten), k));
		if (fS < 0) {
			fS = sS = lS = stS = ssS = k;
			fD = sD = lD = d;
			stI = ssI = I;
			continue;
		}
		if (d < fD) {
			sS = fS;
			sD = fD;
			fS = k;
			fD = d;
		} else if (fD < d && d < sD) {
			sS = k;
			sD = d;
		}
		if (d > lD) {
			lS = k;
			lD = d;
		}
		if (I > stI) {
			ssS = stS;
			ssI = stI;
			stS = k;
			stI = I;
		} else if (stI > I && I > ssI) {
			ssS = k;
			ssI = I;
		}
	}
	const bool needVel = p.velocityRequested != 0;
	for (int r = 0; r < R; ++r) {
		const uint8_t type = returnTypeOf(p.returnMode, r);
		const int rp = beam * R + r;
		if (fS < 0) {
			if (lane == 0) {
				writeReturnNonHit(out, wo, wd, rp, type, d0);
			}
			continue;
		}
		if (type < 1 || type > 5) { // RGL_RETURN_TYPE_UNKNOWN or invalid: nothing is written, as in the reference
			continue;
		}
		const int sel = type == 3 ? sS : (type == 2 ? lS : (type == 1 ? stS : (type == 5 ? ssS : fS)));
		if (lane == sel) {
			const Shade sh = shadeSample(sc, p, range.x, loadOrigin<kRtBlock>(), sd, h, needVel);
			writeReturnHit(out, wo, wd, rp, type, sh);
		}
	}
}

constexpr int kBeamsPerMultiBlock = kRtBlock / 64; // divergent beams: one wave each

// Traces the beam(s) of one thread (single-sample: thread t of the launch traces beam order[t]) and returns the beam
// index of a single-sample thread, or -1.
template<int kFmt, bool kMulti, int kBlock = kRtBlock>
__device__ __forceinline__ int raytraceOne(const SceneView& sc, const RaytraceParams& p, const RaytraceOutputs& out,
                                           int localBlock)
{
	if constexpr (kMulti) {
		const int beam = localBlock * kBeamsPerMultiBlock + static_cast<int>(threadIdx.x >> 6);
		if (beam < p.rayCount) {
			raytraceBeamWave<kFmt>(sc, p, out, beam);
		}
		return -1;
	} else {
		const int t = localBlock * kBlock + threadIdx.x;
		if (t >= p.rayCount) {
			return -1;
		}
		const int beam = p.order != nullptr ? p.order[t] : t;
		raytraceBeam<kFmt, kBlock>(sc, p, out, beam);
		return beam;
	}
}

// The point stages folded into a batched launch, for the return points of one beam. The thread reads back the XYZ
// and DISTANCE it has just written (its own stores) and carries them through the stages in registers.
__device__ __forceinline__ void applyFusedStages(const RaytraceBatchItem& it, int beam)
{
	const int R = it.p.returnCount;
	for (int r = 0; r < R; ++r) {
		const int i = beam * R + r;
		Vec3f xyz = it.out.xyz[i];
		float distance = it.out.distance != nullptr ? it.out.distance[i] : 0.0f;
		for (int k = 0; k < it.fusedCount; ++k) {
			const PointOpItem& op = it.fused[k];
			if (op.op == kOpTransformPoints) {
				xyz = pointTransform(op.m0, xyz);
				static_cast<Vec3f*>(op.out0)[i] = xyz;
			} else {
				pointNoiseDistance(xyz, distance, op.a, op.b, op.c, op.m0, op.m1, op.seed, op.run, op.tag,
				                   static_cast<uint32_t>(i), xyz, distance);
				static_cast<Vec3f*>(op.out0)[i] = xyz;
				static_cast<float*>(op.out1)[i] = distance;
			}
		}
	}
}

// Many sensors (possibly on different scenes) in one launch: block b belongs to the last item whose firstBlock <= b.
// The item index is wave-uniform, so its descriptor is read with scalar loads.
template<int kFmt, bool kMulti, int kBlock = kRtBlock>
__device__ __forceinline__ void raytraceBatchOne(const RaytraceBatchItem* items, int itemCount, int block)
{
	const int item = batchItemOf(items, itemCount, block, [](const RaytraceBatchItem& r) { return r.firstBlock; });
	const RaytraceBatchItem& it = items[item];
	const int localBlock = block - it.firstBlock;
	const int beam = raytraceOne<kFmt, kMulti, kBlock>(it.sc, it.p, it.out, localBlock);
	if constexpr (!kMulti) {
		if (it.fusedCount > 0 && beam >= 0) {
			applyFusedStages(it, beam);
		}
	}
}

// Single-sample beams (no divergence), one per thread; divergent beams, one per wave.
template<int kFmt>
__global__ void __launch_bounds__(kRtBlock) RGL_RT_OCCUPANCY kRaytrace(SceneView sc, RaytraceParams p, RaytraceOutputs out)
{
	raytraceOne<kFmt, false>(sc, p, out, xcdLogicalBlock(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), kXcdChunk));
}

template<int kFmt>
__global__ void __launch_bounds__(kRtBlock) RGL_RT_OCCUPANCY kRaytraceMulti(SceneView sc, RaytraceParams p, RaytraceOutputs out)
{
	raytraceOne<kFmt, true>(sc, p, out, xcdLogicalBlock(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), kXcdChunk));
}

template<int kFmt, int kBlock>
__global__ void __launch_bounds__(kBlock) RGL_RT_OCCUPANCY kRaytraceBatch(const RaytraceBatchItem* __restrict__ items, int itemCount)
{
	raytraceBatchOne<kFmt, false, kBlock>(
	    items, itemCount, xcdLogicalBlock(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), kXcdChunk));
}

// Persistent form (launch variant 2, only when forced): a grid of about one occupancy generation of 64-ray waves; block
// b traces the run of RGL_PERSISTENT_RUN logical blocks at run b first, then claims runs gridDim.x, gridDim.x + 1, ...
// from a counter that the launch's descriptor upload zeroes, until the claims pass the last block, so a wave moves on
// without a block launch. The claim for the next run is issued before the current one is traced, so its latency
// overlaps the ray setup. Every logical block is traced once (runs below gridDim.x statically, the rest by the claim
// that returns them), and every wave exits once its claim is past totalBlocks. On the 64-lidar batch it measured
// slower than hardware dispatch of 64-thread blocks (0.98 ms): 1.92 ms with one block per claim (the claims of all
// 8 XCDs serialize on the one counter) and 1.37 / 1.65 / 2.15 ms with runs of 4 / 8 / 16 (coarser claims leave a
// longer tail), so the autotuner does not consider it.
#ifndef RGL_PERSISTENT_RUN
#define RGL_PERSISTENT_RUN 4
#endif
template<int kFmt>
__global__ void __launch_bounds__(64) RGL_RT_OCCUPANCY kRaytraceBatchPersistent(const RaytraceBatchItem* __restrict__ items,
                                                                                int itemCount, int totalBlocks,
                                                                                uint32_t* claims)
{
	constexpr int kRun = RGL_PERSISTENT_RUN; // logical blocks per claim
	const int grid = static_cast<int>(gridDim.x);
	int block = xcdLogicalBlock(static_cast<int>(blockIdx.x), grid, kXcdChunk) * kRun;
	while (block < totalBlocks) {
		uint32_t claim = 0;
		if (threadIdx.x == 0) {
			claim = atomicAdd(claims, 1u);
		}
		const int end = min(block + kRun, totalBlocks);
		for (int b = block; b < end; ++b) {
			raytraceBatchOne<kFmt, false, 64>(items, itemCount, b);
		}
		block = (grid + static_cast<int>(__builtin_amdgcn_readfirstlane(claim))) * kRun;
	}
}

template<int kFmt>
__global__ void __launch_bounds__(kRtBlock) RGL_RT_OCCUPANCY kRaytraceBatchMulti(const RaytraceBatchItem* __restrict__ items, int itemCount)
{
	raytraceBatchOne<kFmt, true>(
	    items, itemCount, xcdLogicalBlock(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), kXcdChunk));
}

__global__ void __launch_bounds__(kRtBlock) kTraversalStats(SceneView sc, RaytraceParams p, int32_t* nodesVisited,
                                                             int32_t* trisTested, int32_t* instancesEntered)
{
	const int beam = blockIdx.x * kRtBlock + threadIdx.x;
	if (beam >= p.rayCount) {
		return;
	}
	const Vec2f range = p.rangesCount == 1 ? p.ranges[0] : p.ranges[beam];
	const Mat3x4d world = rayWorldPose64(p, beam);
	const Vec3f so = round32(originOf(world)), sd = directionOf(world);
	storeOrigin<kRtBlock>(originOf(world));
	HitRec h;
	TravStats st{0, 0, 0};
	const bool masked = p.mask != nullptr && p.mask[beam] == 0;
	if (!masked) {
		traceClosest<true, kBvh2>(sc, so, sd, range.y, h, &st);
	}
	nodesVisited[beam] = st.nodes;
	trisTested[beam] = st.tris;
	instancesEntered[beam] = st.insts;
}

// Footprint of a raytrace launch: every record its rays read is added to *ls, including the shading records of the
// hits (instance shading record, index triple, vertices, texture coordinates and texel). Divergent beams trace their
// 58 samples in one thread.
template<int kFmt>
__global__ void __launch_bounds__(kRtBlock) kTraversalFootprint(SceneView sc, RaytraceParams p, LineSet ls)
{
	const int beam = blockIdx.x * kRtBlock + threadIdx.x;
	if (beam >= p.rayCount || (p.mask != nullptr && p.mask[beam] == 0)) {
		return;
	}
	const Vec2f range = p.rangesCount == 1 ? p.ranges[0] : p.ranges[beam];
	const Mat3x4d ray64 = distortedRay64(p, beam, rayWorldPose64(p, beam));
	const Mat3x4f ray = round32(ray64);
	storeOrigin<kRtBlock>(originOf(ray64));
	const int samples = p.hHalfDivergence > 0.0f && p.vHalfDivergence > 0.0f ? kBeamSamples : 1;
	for (int s = 0; s < samples; ++s) {
		Mat3x4f sr = ray;
		if (s > 0) {
			sr = ray * beamSampleTransform(p.hHalfDivergence, p.vHalfDivergence, (s - 1) / kBeamVertices,
			                               (s - 1) % kBeamVertices);
		}
		Vec3f so, sd;
		rayOriginDir(sr, so, sd);
		HitRec h;
		if (!traceClosest<false, kFmt, kRtBlock, true>(sc, so, sd, range.y, h, nullptr, &ls)) {
			continue;
		}
		const InstanceShade& is = sc.shades[h.inst];
		touchLines(&ls, &is, sizeof(InstanceShade));
		touchLines(&ls, is.indices + h.prim, sizeof(Vec3i));
		const Vec3i tri = is.indices[h.prim];
		const int v[3] = {tri.x, tri.y, tri.z};
		for (int k = 0; k < 3; ++k) {
			touchLines(&ls, is.vertices + v[k], sizeof(Vec3f));
			if (is.uvs != nullptr && is.texels != nullptr) {
				touchLines(&ls, is.uvs + v[k], sizeof(Vec2f));
			}
		}
	}
}

__global__ void kCountKeys(const unsigned long long* keys, size_t n, unsigned long long* out)
{
	unsigned long long c = 0;
	for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x) {
		c += keys[i] != 0ull ? 1 : 0;
	}
	for (int off = 32; off > 0; off >>= 1) {
		c += __shfl_down(c, off, 64);
	}
	if ((threadIdx.x & 63) == 0) {
		atomicAdd(out, c);
	}
}

__global__ void kBruteForce(SceneView sc, const int32_t* triCounts, RaytraceParams p, float* outT, int32_t* outInst,
                            int32_t* outPrim)
{
	const int beam = blockIdx.x * blockDim.x + threadIdx.x;
	if (beam >= p.rayCount) {
		return;
	}
	const Vec2f range = p.rangesCount == 1 ? p.ranges[0] : p.ranges[beam];
	const Mat3x4d world = rayWorldPose64(p, beam);
	const Vec3d o64 = originOf(world);
	const Vec3f d = directionOf(world);
	HitRec best;
	best.t = range.y;
	best.inst = INT_MAX;
	best.prim = INT_MAX;
	bool found = false;
	for (int i = 0; i < sc.instanceCount; ++i) {
		const InstanceRec& ir = sc.instances[i];
		const TravRay r = objectRay(ir, o64, d);
		for (int k = 0; k < triCounts[i]; ++k) {
			const LeafTri tri = ir.blasTris[k];
			float t, vNum, wNum, det;
			if (intersectTri(r, tri, t, vNum, wNum, det) && t > 0.0f && t <= range.y && better(t, i, tri.prim, best)) {
				best.t = t;
				best.inst = i;
				best.prim = tri.prim;
				found = true;
			}
		}
	}
	outT[beam] = found ? best.t : -1.0f;
	outInst[beam] = found ? best.inst : -1;
	outPrim[beam] = found ? best.prim : -1;
}

__global__ void kTransformRays(int n, const Mat3x4f* in, Mat3x4f* out, Mat3x4f tf)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		out[i] = round32(widen(tf) * widen(in[i])); // the same FP64 composition as a deferred transform
	}
}

} // namespace

namespace {
template<int kFmt>
void launchRaytrace(hipStream_t stream, int blocks, bool multi, const SceneView& scene, const RaytraceParams& params,
                    const RaytraceOutputs& out)
{
	if (multi) {
		hipLaunchKernelGGL(kRaytraceMulti<kFmt>, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, out);
	} else {
		hipLaunchKernelGGL(kRaytrace<kFmt>, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, out);
	}
}

// Blocks of the persistent launch: about one occupancy generation (resident 64-thread blocks per CU times CUs).
template<int kFmt>
int persistentGrid(int totalBlocks)
{
	static int resident = 0;
	if (resident == 0) {
		int device = 0, perCu = 0;
		hipDeviceProp_t prop;
		RGL_HIP(hipGetDevice(&device));
		RGL_HIP(hipGetDeviceProperties(&prop, device));
		RGL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kRaytraceBatchPersistent<kFmt>, 64, 0));
		resident = std::max(1, perCu) * prop.multiProcessorCount;
	}
	return std::min((totalBlocks + RGL_PERSISTENT_RUN - 1) / RGL_PERSISTENT_RUN, resident);
}

template<int kFmt>
void launchRaytraceBatch(hipStream_t stream, int blocks, bool multi, int variant, const RaytraceBatchItem* items,
                         int itemCount, uint32_t* claims)
{
	if (multi) {
		hipLaunchKernelGGL(kRaytraceBatchMulti<kFmt>, dim3(blocks), dim3(kRtBlock), 0, stream, items, itemCount);
	} else if (variant == 2) {
		hipLaunchKernelGGL(kRaytraceBatchPersistent<kFmt>, dim3(persistentGrid<kFmt>(blocks)), dim3(64), 0, stream, items,
		                   itemCount, blocks, claims);
	} else if (variant == 1) {
		hipLaunchKernelGGL((kRaytraceBatch<kFmt, 64>), dim3(blocks), dim3(64), 0, stream, items, itemCount);
	} else {
		hipLaunchKernelGGL((kRaytraceBatch<kFmt, kRtBlock>), dim3(blocks), dim3(kRtBlock), 0, stream, items, itemCount);
	}
}
} // namespace

bool beamIsMulti(const RaytraceParams& p) { return p.hHalfDivergence > 0.0f && p.vHalfDivergence > 0.0f; }

void raytrace(hipStream_t stream, const SceneView& scene, const RaytraceParams& params, const RaytraceOutputs& out)
{
	if (params.rayCount <= 0) {
		return;
	}
	const bool multi = beamIsMulti(params);
	const int blocks = raytraceBlocks(params.rayCount, multi, 0);
	switch (scene.format) {
		case kBvh4F: launchRaytrace<kBvh4F>(stream, blocks, multi, scene, params, out); break;
		case kBvh4Q: launchRaytrace<kBvh4Q>(stream, blocks, multi, scene, params, out); break;
		case kBvh4H: launchRaytrace<kBvh4H>(stream, blocks, multi, scene, params, out); break;
		default: launchRaytrace<kBvh2>(stream, blocks, multi, scene, params, out); break;
	}
}

int raytraceBlocks(int rayCount, bool multi, int variant)
{
	const int per = multi ? kBeamsPerMultiBlock : (variant >= 1 ? 64 : kRtBlock);
	return (rayCount + per - 1) / per;
}

void raytraceBatch(hipStream_t stream, const RaytraceBatchItem* items, int itemCount, int totalBlocks, int format,
                   bool multi, int variant, uint32_t* claims)
{
	if (itemCount <= 0 || totalBlocks <= 0) {
		return;
	}
	if (variant == 2 && !multi && claims == nullptr) {
		throw std::invalid_argument("raytraceBatch: the persistent variant needs a zeroed claim counter");
	}
	switch (format) {
		case kBvh4F: launchRaytraceBatch<kBvh4F>(stream, totalBlocks, multi, variant, items, itemCount, claims); break;
		case kBvh4Q: launchRaytraceBatch<kBvh4Q>(stream, totalBlocks, multi, variant, items, itemCount, claims); break;
		case kBvh4H: launchRaytraceBatch<kBvh4H>(stream, totalBlocks, multi, variant, items, itemCount, claims); break;
		default: launchRaytraceBatch<kBvh2>(stream, totalBlocks, multi, variant, items, itemCount, claims); break;
	}
}

void traversalStats(hipStream_t stream, const SceneView& scene, const RaytraceParams& params, int32_t* nodesVisited,
                    int32_t* trisTested, int32_t* instancesEntered)
{
	if (params.rayCount <= 0) {
		return;
	}
	const int blocks = (params.rayCount + kRtBlock - 1) / kRtBlock;
	hipLaunchKernelGGL(kTraversalStats, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, nodesVisited, trisTested,
	                   instancesEntered);
}

void traversalFootprint(hipStream_t stream, const SceneView& scene, const RaytraceParams& params,
                        unsigned long long* keys, uint32_t mask, uint32_t* full)
{
	if (params.rayCount <= 0 || scene.tlasNodes == nullptr) {
		return;
	}
	const LineSet ls{keys, mask, full};
	const int blocks = (params.rayCount + kRtBlock - 1) / kRtBlock;
	switch (scene.format) {
		case kBvh4F: hipLaunchKernelGGL(kTraversalFootprint<kBvh4F>, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, ls); break;
		case kBvh4Q: hipLaunchKernelGGL(kTraversalFootprint<kBvh4Q>, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, ls); break;
		case kBvh4H: hipLaunchKernelGGL(kTraversalFootprint<kBvh4H>, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, ls); break;
		default: hipLaunchKernelGGL(kTraversalFootprint<kBvh2>, dim3(blocks), dim3(kRtBlock), 0, stream, scene, params, ls); break;
	}
}

void countNonzeroKeys(hipStream_t stream, const unsigned long long* keys, size_t n, unsigned long long* out)
{
	hipLaunchKernelGGL(kCountKeys, dim3(1024), dim3(256), 0, stream, keys, n, out);
}

void bruteForceClosest(hipStream_t stream, const SceneView& scene, const int32_t* triCounts, const RaytraceParams& params,
                       float* outT, int32_t* outInstance, int32_t* outPrim)
{
	if (params.rayCount <= 0) {
		return;
	}
	hipLaunchKernelGGL(kBruteForce, dim3((params.rayCount + 127) / 128), dim3(128), 0, stream, scene, triCounts, params,
	                   outT, outInstance, outPrim);
}

void transformRays(hipStream_t stream, int n, const Mat3x4f* in, Mat3x4f* out, const Mat3x4f& tf)
{
	if (n > 0) {
		hipLaunchKernelGGL(kTransformRays, dim3((n + 255) / 256), dim3(256), 0, stream, n, in, out, tf);
	}
}

int maxTraversalStack() { return kLdsStack + kOvfStack; }

} // namespace rgl::gpu
