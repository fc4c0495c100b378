// FP64 CPU oracle of RGL's raytrace-and-point-cloud semantics -- TEST INFRASTRUCTURE ONLY.
//
// Restates, in double precision and with its own data structures, what the reference computes per ray:
//   raygen        src/gpu/optixPrograms.cu:46-87, :222-248 (mask, distortion, ranges, 58-sample beams)
//   closest hit   src/gpu/optixPrograms.cu:96-216 (distance, min range, normal, incident angle, intensity,
//                 entity id, laser retro, velocities)
//   miss          src/gpu/optixPrograms.cu:89-94
//   beam reduce   src/gpu/nodeKernels.cu:103-207, :334-422
//   shared data   src/gpu/optixPrograms.cu:295-326
//   return modes  src/returnModeUtils.h:22-30, include/rgl/api/core.h:453-504
//   transforms    src/math/Mat3x4f.hpp:38-267
// Closest hit is exact FP64 ray/triangle intersection (edge functions in a sheared frame, Woop et al. 2013)
// over a top-down median-split BVH built here, or over all triangles (bruteForce). Hits count for t in (0, tmax];
// equal t resolves to the lower (instance, triangle) index. Per beam it also reports whether the result is
// ambiguous at relative tolerance eps: a candidate within eps of a triangle edge, a hit of another instance within
// eps of the closest distance, or a distance within eps of a range limit. FP32 and FP64 may legitimately disagree
// only on such beams.
// The noise stream (Philox4x32-10 from the SC'11 paper, Box-Muller with Cephes single-precision log/sin/cos) is
// written again here, independently of the library's header, and is pinned by published Philox known-answer
// vectors and by the statistics of the normals.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// Deliberate mistake planted by orc_plant_mistake (0 = none); only tests set it.
int gMistake = 0;

struct V3
{
	double x, y, z;
	double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
	double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a)
{
	const double l = len(a);
	return {a.x / l, a.y / l, a.z / l};
}

struct M34
{
	double m[3][4];
	static M34 identity()
	{
		M34 r{};
		r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0;
		return r;
	}
	static M34 fromFloat(const float* f)
	{
		M34 r;
		for (int i = 0; i < 12; ++i) {
			r.m[i / 4][i % 4] = f[i];
		}
		return r;
	}
	V3 point(V3 p) const
	{
		return {m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3], m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
		        m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]};
	}
	V3 dir(V3 p) const
	{
		return {m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z, m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z,
		        m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z};
	}
	M34 operator*(const M34& b) const
	{
		M34 r;
		for (int i = 0; i < 3; ++i) {
			for (int j = 0; j < 4; ++j) {
				r.m[i][j] = m[i][0] * b.m[0][j] + m[i][1] * b.m[1][j] + m[i][2] * b.m[2][j] + (j == 3 ? m[i][3] : 0.0);
			}
		}
		return r;
	}
	// Affine inverse; zero matrix when singular (src/math/Mat3x4f.hpp:188-245).
	M34 inverse() const
	{
		const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0], h = m[2][1],
		             k = m[2][2];
		const double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
		M34 r{};
		if (det == 0.0) {
			return r;
		}
		const double i = 1.0 / det;
		r.m[0][0] = (e * k - f * h) * i;
		r.m[0][1] = (c * h - b * k) * i;
		r.m[0][2] = (b * f - c * e) * i;
		r.m[1][0] = (f * g - d * k) * i;
		r.m[1][1] = (a * k - c * g) * i;
		r.m[1][2] = (c * d - a * f) * i;
		r.m[2][0] = (d * h - e * g) * i;
		r.m[2][1] = (b * g - a * h) * i;
		r.m[2][2] = (a * e - b * d) * i;
		for (int row = 0; row < 3; ++row) {
			r.m[row][3] = -(r.m[row][0] * m[0][3] + r.m[row][1] * m[1][3] + r.m[row][2] * m[2][3]);
		}
		return r;
	}
	M34 rotationOnly() const
	{
		M34 r = *this;
		r.m[0][3] = r.m[1][3] = r.m[2][3] = 0.0;
		return r;
	}
};

// Euler composition of src/math/Mat3x4f.hpp:38-60 (Rz(z) * Ry(y) * Rx(x)).
M34 rotationRad(double x, double y, double z)
{
	const double cr = std::cos(x), sr = std::sin(x), cp = std::cos(y), sp = std::sin(y), cy = std::cos(z), sy = std::sin(z);
	M34 r = M34::identity();
	r.m[0][0] = cp * cy;
	r.m[0][1] = sr * sp * cy - cr * sy;
	r.m[0][2] = sr * sy + cr * sp * cy;
	r.m[1][0] = cp * sy;
	r.m[1][1] = cr * cy + sr * sp * sy;
	r.m[1][2] = cr * sp * sy - sr * cy;
	r.m[2][0] = -sp;
	r.m[2][1] = sr * cp;
	r.m[2][2] = cr * cp;
	return r;
}

M34 trsDeg(V3 t, V3 rdeg)
{
	const double k = M_PI / 180.0;
	M34 r = rotationRad(rdeg.x * k, rdeg.y * k, rdeg.z * k);
	r.m[0][3] = t.x;
	r.m[1][3] = t.y;
	r.m[2][3] = t.z;
	return r;
}

// Ray origin M*(0,0,0) and direction M*(0,0,1) - M*(0,0,0) (src/gpu/optixPrograms.cu:222-228), evaluated exactly:
// the direction is the matrix's third column (the reference's float subtraction rounds it to the ulp of the
// translation, which the library does not reproduce).
void rayOf(const float* m, V3& o, V3& d)
{
	o = {m[3], m[7], m[11]};
	d = {m[2], m[6], m[10]};
}
void rayOf(const M34& m, V3& o, V3& d)
{
	o = {m.m[0][3], m.m[1][3], m.m[2][3]};
	d = {m.m[0][2], m.m[1][2], m.m[2][2]};
}

struct Box
{
	V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
	void grow(V3 p)
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], p[a]);
			hi[a] = std::max(hi[a], p[a]);
		}
	}
	void grow(const Box& b)
	{
		grow(b.lo);
		grow(b.hi);
	}
	// Slab test over [0, tmax]; exact enough in FP64 and widened by a relative 1e-9.
	bool hit(V3 o, V3 inv, double tmax) const
	{
		double t0 = 0.0, t1 = tmax * (1.0 + 1e-9);
		for (int a = 0; a < 3; ++a) {
			double ta = (lo[a] - o[a]) * inv[a], tb = (hi[a] - o[a]) * inv[a];
			if (std::isnan(ta)) ta = -INFINITY;
			if (std::isnan(tb)) tb = INFINITY;
			if (ta > tb) std::swap(ta, tb);
			t0 = std::max(t0, ta * (1.0 - 1e-9));
			t1 = std::min(t1, tb * (1.0 + 1e-9));
		}
		return t0 <= t1;
	}
};

// Top-down median-split BVH over primitive boxes, at most 4 primitives per leaf.
struct Bvh
{
	struct Node
	{
		Box box;
		int left = -1, right = -1, first = 0, count = 0;
	};
	std::vector<Node> nodes;
	std::vector<int> prims;

	void build(const std::vector<Box>& boxes)
	{
		prims.resize(boxes.size());
		for (size_t i = 0; i < boxes.size(); ++i) {
			prims[i] = static_cast<int>(i);
		}
		nodes.clear();
		nodes.reserve(2 * boxes.size() + 1);
		if (!boxes.empty()) {
			rec(boxes, 0, static_cast<int>(boxes.size()));
		}
	}
	int rec(const std::vector<Box>& boxes, int first, int count)
	{
		const int id = static_cast<int>(nodes.size());
		nodes.emplace_back();
		Box b, cb;
		for (int i = first; i < first + count; ++i) {
			b.grow(boxes[prims[i]]);
			cb.grow((boxes[prims[i]].lo + boxes[prims[i]].hi) * 0.5);
		}
		nodes[id].box = b;
		if (count <= 4) {
			nodes[id].first = first;
			nodes[id].count = count;
			return id;
		}
		int axis = 0;
		for (int a = 1; a < 3; ++a) {
			if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
		}
		const int mid = first + count / 2;
		std::nth_element(prims.begin() + first, prims.begin() + mid, prims.begin() + first + count, [&](int x, int y) {
			const double cx = boxes[x].lo[axis] + boxes[x].hi[axis], cy = boxes[y].lo[axis] + boxes[y].hi[axis];
			return cx < cy || (cx == cy && x < y);
		});
		const int l = rec(boxes, first, mid - first);
		const int r = rec(boxes, mid, first + count - mid);
		nodes[id].left = l;
		nodes[id].right = r;
		return id;
	}
	template<typename LeafFn>
	void traverse(V3 o, V3 d, const double& tLimit, LeafFn&& leaf) const
	{
		if (nodes.empty()) return;
		const V3 inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
		int stack[128];
		int sp = 0;
		stack[sp++] = 0;
		while (sp) {
			const Node& n = nodes[stack[--sp]];
			if (!n.box.hit(o, inv, tLimit)) continue;
			if (n.left < 0) {
				for (int i = n.first; i < n.first + n.count; ++i) leaf(prims[i]);
			} else {
				stack[sp++] = n.left;
				stack[sp++] = n.right;
			}
		}
	}
};

struct Tri
{
	V3 a, b, c;
};

struct MeshData
{
	std::vector<Tri> tris;
	std::vector<int> idx; // 3 per triangle
	Bvh bvh;
	Box box;
};

struct Inst
{
	int mesh;
	M34 toWorld, toObject, prev;
	bool hasPrev;
	const float* disp; // object-space vertex displacement since the previous frame, or nullptr
	int id;
	float retro;
	int texture;
	Box worldBox;
};

struct Candidate
{
	double t;
	int inst, prim;
	double u, v; // weights of B and C
	double margin;
};

// Exact FP64 ray/triangle test in the sheared frame (Woop, Benthin, Wald 2013); no face culling.
bool intersect(V3 o, V3 d, const Tri& tri, Candidate& c)
{
	int kz = 0;
	if (std::fabs(d.y) > std::fabs(d[kz])) kz = 1;
	if (std::fabs(d.z) > std::fabs(d[kz])) kz = 2;
	int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
	if (d[kz] < 0) std::swap(kx, ky);
	const double sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1.0 / d[kz];
	const V3 A = tri.a - o, B = tri.b - o, C = tri.c - o;
	const double ax = A[kx] - sx * A[kz], ay = A[ky] - sy * A[kz];
	const double bx = B[kx] - sx * B[kz], by = B[ky] - sy * B[kz];
	const double cx = C[kx] - sx * C[kz], cy = C[ky] - sy * C[kz];
	const double U = cx * by - cy * bx, V = ax * cy - ay * cx, W = bx * ay - by * ax;
	const double det = U + V + W;
	if (det == 0.0) return false;
	const double u = U / det, v = V / det, w = W / det;
	const double T = U * (sz * A[kz]) + V * (sz * B[kz]) + W * (sz * C[kz]);
	c.t = T / det;
	c.u = v;
	c.v = w;
	c.margin = std::min(u, std::min(v, w));
	return true;
}

struct World
{
	std::vector<MeshData> meshes;
	std::vector<Inst> insts;
	Bvh tlas;
	const OrcScene* sc;
};

struct Hit
{
	bool found = false;
	Candidate c{};
};

bool better(const Candidate& a, const Candidate& b)
{
	return a.t < b.t || (a.t == b.t && (a.inst < b.inst || (a.inst == b.inst && a.prim < b.prim)));
}

// Closest hit over (0, tmax]; optionally the ambiguity diagnostic at relative tolerance eps.
Hit closest(const World& w, V3 o, V3 d, double tmax, bool brute, double eps, bool* ambiguous, bool* instanceTie)
{
	Hit best;
	best.c.t = tmax;
	best.c.inst = std::numeric_limits<int>::max();
	best.c.prim = std::numeric_limits<int>::max();
	auto visitInstance = [&](int ii, double limit, auto&& onCand) {
		const Inst& in = w.insts[ii];
		const V3 oo = in.toObject.point(o), dd = in.toObject.dir(d);
		const MeshData& m = w.meshes[in.mesh];
		auto tri = [&](int p) {
			Candidate c;
			if (!intersect(oo, dd, m.tris[p], c)) return;
			c.inst = ii;
			c.prim = p;
			onCand(c);
		};
		if (brute) {
			for (int p = 0; p < static_cast<int>(m.tris.size()); ++p) tri(p);
		} else {
			m.bvh.traverse(oo, dd, limit, tri);
		}
	};
	auto pass = [&](double& limit, auto&& onCand) {
		if (brute) {
			for (int i = 0; i < static_cast<int>(w.insts.size()); ++i) visitInstance(i, limit, onCand);
		} else {
			w.tlas.traverse(o, d, limit, [&](int i) { visitInstance(i, limit, onCand); });
		}
	};
	double limit = tmax;
	pass(limit, [&](const Candidate& c) {
		if (c.margin >= 0.0 && c.t > 0.0 && c.t <= tmax && better(c, best.c)) {
			best.c = c;
			best.found = true;
			limit = c.t;
		}
	});
	if (ambiguous != nullptr && eps > 0.0) {
		bool amb = false;
		double lim2 = (best.found ? best.c.t : tmax) * (1.0 + eps);
		const double ref = best.found ? best.c.t : tmax;
		pass(lim2, [&](const Candidate& c) {
			if (c.margin < -eps || !(c.t > 0.0) || c.t > lim2) return;
			if (c.margin < eps) amb = true; // grazing: within eps of an edge
			if (best.found && c.inst != best.c.inst && c.margin >= 0.0 && c.t >= best.c.t * (1.0 - eps)) {
				amb = true;
				if (instanceTie != nullptr) *instanceTie = true;
			}
			if (!best.found && c.margin >= 0.0) amb = true; // hit just beyond tmax
		});
		if (best.found && best.c.t >= tmax * (1.0 - eps)) amb = true;
		(void) ref;
		*ambiguous = *ambiguous || amb;
	}
	return best;
}

struct Shade
{
	bool isHit;
	float distance;
	float intensity;
	double base; // intensity before the incidence factor (texel or default intensity)
	V3 normal;
	float incident;
	V3 absVel, relVel;
	float radial;
	int entityId;
	float retro;
};

Shade shade(const World& w, const OrcRequest& q, const M34& rayToWorld, double minRange, V3 o, V3 d, const Hit& h,
            bool* ambiguous, double eps, bool* texelEdge)
{
	Shade s{};
	const Inst& in = w.insts[h.c.inst];
	const MeshData& m = w.meshes[in.mesh];
	const Tri& t = m.tris[h.c.prim];
	const V3 hitWorld = o + d * h.c.t;
	const double distance = len(hitWorld - o);
	if (ambiguous && eps > 0.0 && std::fabs(distance - minRange) <= eps * std::max(minRange, 1e-9)) {
		*ambiguous = true;
	}
	if (distance < minRange) {
		s.isHit = false;
		s.distance = q.nearDistance;
		s.intensity = 0.0f;
		return s;
	}
	s.isHit = true;
	s.distance = static_cast<float>(distance);
	const V3 wA = in.toWorld.point(t.a), wB = in.toWorld.point(t.b), wC = in.toWorld.point(t.c);
	const V3 n = unit(cross(wB - wA, wC - wA));
	const double cosI = std::fabs(dot(gMistake == 6 ? cross(wB - wA, wC - wA) : n, d));
	s.normal = n;
	s.incident = static_cast<float>(std::acos(std::min(cosI, 1.0)));
	double intensity = q.defaultIntensity;
	const OrcScene& sc = *w.sc;
	if (in.texture >= 0 && sc.uvs != nullptr && sc.meshHasUV[in.mesh]) {
		const int base = sc.meshVertOffset[in.mesh];
		const int* ix = &m.idx[3 * h.c.prim];
		const float* uvA = sc.uvs + 2 * (base + ix[0]);
		const float* uvB = sc.uvs + 2 * (base + ix[1]);
		const float* uvC = sc.uvs + 2 * (base + ix[2]);
		const double wA0 = 1.0 - h.c.u - h.c.v;
		const double tu = wA0 * uvA[0] + h.c.u * uvB[0] + h.c.v * uvC[0];
		const double tv = wA0 * uvA[1] + h.c.u * uvB[1] + h.c.v * uvC[1];
		const double fu = tu - std::floor(tu), fv = tv - std::floor(tv);
		const int W = sc.texWidth[in.texture], H = sc.texHeight[in.texture];
		// Point sampling jumps at texel edges: a coordinate within FP32's reach of one may pick the neighbour.
		if (texelEdge != nullptr && eps > 0.0) {
			const double gu = fu * W, gv = fv * H;
			if (std::fabs(gu - std::round(gu)) <= eps * W * std::max(1.0, std::fabs(tu)) ||
			    std::fabs(gv - std::round(gv)) <= eps * H * std::max(1.0, std::fabs(tv))) {
				*texelEdge = true;
			}
		}
		int x = std::min(std::max(static_cast<int>(std::floor(fu * W)), 0), W - 1);
		int y = std::min(std::max(static_cast<int>(std::floor(fv * H)), 0), H - 1);
		if (gMistake == 4) {
			std::swap(x, y);
		}
		intensity = sc.texels[sc.texOffset[in.texture] + static_cast<int64_t>(y) * W + x];
	}
	s.base = intensity;
	s.intensity = static_cast<float>(intensity * cosI);
	s.entityId = in.id;
	s.retro = in.retro;
	s.absVel = s.relVel = {NAN, NAN, NAN};
	s.radial = NAN;
	if (q.sceneDeltaTime > 0.0f) {
		V3 disp{0, 0, 0};
		if (in.hasPrev) {
			const V3 hitObject = in.toObject.point(hitWorld);
			disp = hitWorld - in.prev.point(hitObject);
		}
		if (in.disp != nullptr) {
			// src/gpu/optixPrograms.cu:186-196: vertex displacements in world orientation, interpolated at the hit.
			const int* ix = &m.idx[3 * h.c.prim];
			V3 dv[3];
			for (int k = 0; k < 3; ++k) {
				const float* p = in.disp + 3 * ix[k];
				dv[k] = in.toWorld.rotationOnly().dir(V3{p[0], p[1], p[2]});
			}
			disp = disp + dv[0] * (1.0 - h.c.u - h.c.v) + dv[1] * h.c.u + dv[2] * h.c.v;
		}
		const V3 absVel = disp * (1.0 / q.sceneDeltaTime);
		const V3 absInSensor = rayToWorld.rotationOnly().inverse().dir(absVel);
		const V3 lin{q.linVel[0], q.linVel[1], q.linVel[2]};
		const V3 ang{q.angVel[0], q.angVel[1], q.angVel[2]};
		const V3 hitRays = rayToWorld.inverse().point(hitWorld);
		const V3 rel = (absInSensor - lin) - cross(ang, hitRays);
		s.absVel = absVel;
		s.relVel = rel;
		s.radial = static_cast<float>(dot(unit(hitRays), rel));
	}
	return s;
}

void put3(float* dst, int i, V3 v)
{
	dst[3 * i] = static_cast<float>(v.x);
	dst[3 * i + 1] = static_cast<float>(v.y);
	dst[3 * i + 2] = static_cast<float>(v.z);
}

// Euler angles of the Unity-style ZXY left-handed convention (src/math/Mat3x4f.hpp:146-175).
float azimuthOf(const M34& r)
{
	const double m20 = gMistake == 5 ? r.m[0][2] : r.m[2][0];
	return (r.m[2][1] < 1 && r.m[2][1] > -1) ? static_cast<float>(std::atan2(-m20, r.m[2][2])) : 0.0f;
}
float elevationOf(const M34& r)
{
	if (r.m[1][2] < 1) {
		return r.m[1][2] > -1 ? static_cast<float>(std::asin(r.m[1][2])) : static_cast<float>(-M_PI / 2);
	}
	return static_cast<float>(M_PI / 2);
}

void beam(const World& w, const OrcRequest& q, OrcOutputs& out, int b)
{
	const int R = static_cast<int>(static_cast<uint32_t>(q.returnMode) >> 24);
	const M34 toWorld = M34::fromFloat(q.rayOriginToWorld);
	const M34 toRays = toWorld.inverse();
	const float* rayF = q.raysF64 != nullptr ? nullptr : q.rays + 12 * b;
	M34 rayWorld;
	if (q.raysF64 != nullptr) {
		for (int i = 0; i < 12; ++i) {
			rayWorld.m[i / 4][i % 4] = q.raysF64[12 * b + i];
		}
	} else {
		rayWorld = M34::fromFloat(rayF);
	}
	// Origin and direction of the beam's world ray: the float rounding of the reference for float rays, the exact
	// third column for FP64 rays.
	auto beamRay = [&](V3& o, V3& d) {
		if (rayF != nullptr) {
			rayOf(rayF, o, d);
		} else {
			rayOf(rayWorld, o, d);
		}
	};
	const M34 rayLocal = toRays * rayWorld;
	const double eps = q.ambiguityEps;
	bool ambiguous = false;
	for (int r = 0; r < R; ++r) {
		const int rp = b * R + r;
		if (out.rayIdx) out.rayIdx[rp] = static_cast<uint32_t>(b);
		if (out.azimuth) out.azimuth[rp] = azimuthOf(rayLocal);
		if (out.elevation) out.elevation[rp] = elevationOf(rayLocal);
		// Time stamps are computed in float by the reference (src/gpu/optixPrograms.cu:313-321).
		if (out.timeStampF64) out.timeStampF64[rp] = q.distortion ? static_cast<double>(q.timeOffsetsMs[b] * 1e-3f) : 0.0;
		if (out.timeStampU32) out.timeStampU32[rp] = q.distortion ? static_cast<uint32_t>(q.timeOffsetsMs[b] * 1e6f) : 0u;
		if (out.ringId && q.ringIds) out.ringId[rp] = static_cast<uint16_t>(q.ringIds[b % q.nRingIds]);
	}
	const float* range = q.ranges + 2 * (q.nRanges == 1 ? 0 : b);
	const double minRange = range[0], maxRange = range[1];

	V3 o0, d0;
	beamRay(o0, d0);
	auto writeNonHit = [&](int rp, uint8_t type, float dist) {
		if (out.xyz) {
			V3 disp = unit(d0) * dist;
			for (int a = 0; a < 3; ++a) if (std::isnan(disp[a])) disp[a] = 0.0;
			put3(out.xyz, rp, o0 + disp);
		}
		if (out.isHit) out.isHit[rp] = 0;
		if (out.returnType) out.returnType[rp] = type;
		if (out.distance) out.distance[rp] = dist;
		if (out.intensityF32) out.intensityF32[rp] = 0.0f;
		if (out.intensityU8) out.intensityU8[rp] = 0;
		if (out.entityId) out.entityId[rp] = 0;
		if (out.absVelocity) put3(out.absVelocity, rp, {NAN, NAN, NAN});
		if (out.relVelocity) put3(out.relVelocity, rp, {NAN, NAN, NAN});
		if (out.radialSpeed) out.radialSpeed[rp] = 0.0f;
		if (out.normal) put3(out.normal, rp, {NAN, NAN, NAN});
		if (out.incidentAngle) out.incidentAngle[rp] = NAN;
		if (out.laserRetro) out.laserRetro[rp] = 0.0f;
	};
	auto typeOf = [&](int r) { return static_cast<uint8_t>((q.returnMode >> (8 * r)) & 0xff); };
	if (out.primaryT) out.primaryT[b] = -1.0;
	if (out.primaryInst) out.primaryInst[b] = -1;
	if (out.primaryPrim) out.primaryPrim[b] = -1;
	if (q.mask != nullptr && q.mask[b] == 0) {
		for (int r = 0; r < R; ++r) writeNonHit(b * R + r, typeOf(r), q.farDistance);
		if (out.ambiguous) out.ambiguous[b] = 0;
		return;
	}
	M34 ray = rayWorld;
	if (q.distortion) {
		const double off = gMistake == 3 ? -q.timeOffsetsMs[b] : q.timeOffsetsMs[b];
		const V3 t{off * q.linVel[0] * 1e-3, off * q.linVel[1] * 1e-3, off * q.linVel[2] * 1e-3};
		const double toDeg = 180.0 / M_PI;
		const V3 rd{off * q.angVel[0] * toDeg * 1e-3, off * q.angVel[1] * toDeg * 1e-3, off * q.angVel[2] * toDeg * 1e-3};
		ray = toWorld * (trsDeg(t, rd) * rayLocal);
	}
	const double hHalf = q.hDivergence / 2.0, vHalf = q.vDivergence / 2.0;
	const bool divergent = hHalf > 0.0 && vHalf > 0.0;
	const int samples = divergent ? 58 : 1;
	struct Track
	{
		int sample = -1;
		float distance = 0, intensity = 0;
		Shade sh{};
	};
	Track first, second, last, strongest, secondStrongest;
	float sample0Distance = q.farDistance;
	// Every (first, second) distance pair and (strongest, second strongest) intensity pair the reduction can reach
	// when each comparison whose operands are within FP32's reach of each other may go either way (at most kStates
	// each; `overflow` if more). The returned SECOND and SECOND_STRONGEST lie within the range of the second
	// elements.
	constexpr size_t kStates = 512;
	std::vector<std::pair<double, double>> dStates, iStates;
	bool statesOverflow = false;
	// Beam reduction compares samples strictly and in sample order (nodeKernels.cu:369-397): a comparison whose
	// operands are within FP32's reach of each other may go the other way on the GPU and change the selection.
	// Distances carry ~eps relative error; intensities texel*cos(incidence) carry ~eps relative plus about
	// 4e-6 of the texel value from the FP32 normal (dominant at grazing incidence).
	bool distTie = false, intTie = false, texelEdge = false, instanceTie = false;
	auto nearD = [&](double a, double b) { return std::fabs(a - b) <= eps * std::max(std::fabs(a), std::fabs(b)); };
	auto nearI = [&](double a, double b, double baseA, double baseB) {
		return std::fabs(a - b) <= eps * std::max(std::fabs(a), std::fabs(b)) + 4e-6 * std::max(baseA, baseB);
	};
	for (int s = 0; s < samples; ++s) {
		M34 sr = ray;
		if (s > 0) {
			const int layer = (s - 1) / 19, vertex = (s - 1) % 19;
			const double shrink = 1.0 - layer / 3.0;
			const double step = 2.0 * M_PI / 19.0;
			const double hA = hHalf * shrink * std::cos(vertex * step), vA = vHalf * shrink * std::sin(vertex * step);
			sr = ray * (rotationRad(vA, 0, 0) * rotationRad(0, hA, 0));
		}
		V3 so, sd;
		if (s == 0 && !q.distortion) {
			beamRay(so, sd);
		} else {
			rayOf(sr, so, sd);
		}
		const Hit h = closest(w, so, sd, maxRange, q.bruteForce != 0, eps, out.ambiguous ? &ambiguous : nullptr,
		                      out.ambiguous ? &instanceTie : nullptr);
		Track c;
		c.sample = s;
		bool isHit = false;
		if (h.found) {
			c.sh = shade(w, q, toWorld, minRange, so, sd, h, out.ambiguous ? &ambiguous : nullptr, eps,
			             out.ambiguous ? &texelEdge : nullptr);
			isHit = c.sh.isHit;
			c.distance = c.sh.distance;
			c.intensity = c.sh.intensity;
			if (s == 0) {
				if (out.primaryT) out.primaryT[b] = h.c.t;
				if (out.primaryInst) out.primaryInst[b] = h.c.inst;
				if (out.primaryPrim) out.primaryPrim[b] = h.c.prim;
			}
		} else {
			c.distance = q.farDistance;
		}
		if (s == 0) sample0Distance = c.distance;
		if (!isHit) continue;
		if (samples > 1 && eps > 0.0 && first.sample >= 0) {
			const double dd = c.distance, ii = c.intensity, bb = c.sh.base;
			if (nearD(dd, first.distance) || nearD(dd, second.distance) || nearD(dd, last.distance)) distTie = true;
			if (nearI(ii, strongest.intensity, bb, strongest.sh.base) ||
			    nearI(ii, secondStrongest.intensity, bb, secondStrongest.sh.base)) {
				intTie = true;
			}
		}
		if (samples > 1 && out.secondRange != nullptr) {
			// lt(a, b): the outcomes a < b can take on the GPU (bit 0: false possible, bit 1: true possible).
			auto outcomes = [&](double a, double b, bool nearAB) {
				const bool exact = a < b;
				return nearAB ? 3 : (exact ? 2 : 1);
			};
			auto advance = [&](std::vector<std::pair<double, double>>& st, double x, double base, bool byIntensity) {
				if (st.empty()) {
					st.push_back({x, x});
					return;
				}
				std::vector<std::pair<double, double>> next;
				auto add = [&](double f, double g) {
					for (auto& e : next) {
						if (e.first == f && e.second == g) {
							return;
						}
					}
					next.push_back({f, g});
				};
				for (const auto& [f, g] : st) {
					// Distances: (x < f) -> (x, f); else if (f < x && x < g) -> (f, x). Intensities mirror it
					// with > (the strongest tracker).
					auto near2 = [&](double a, double b2) {
						return byIntensity ? nearI(a, b2, 255.0, 255.0) : nearD(a, b2);
					};
					const int o1 = byIntensity ? outcomes(f, x, near2(f, x)) : outcomes(x, f, near2(x, f));
					if (o1 & 2) {
						add(x, f);
					}
					if (o1 & 1) {
						const int o2 = byIntensity ? outcomes(x, f, near2(x, f)) : outcomes(f, x, near2(f, x));
						const int o3 = byIntensity ? outcomes(g, x, near2(g, x)) : outcomes(x, g, near2(x, g));
						if ((o2 & 2) && (o3 & 2)) {
							add(f, x);
						}
						if ((o2 & 1) || (o3 & 1)) {
							add(f, g);
						}
					}
				}
				if (next.size() > kStates) {
					statesOverflow = true;
					next.resize(kStates);
				}
				st.swap(next);
			};
			advance(dStates, c.distance, 0.0, false);
			advance(iStates, c.intensity, c.sh.base, true);
		}
		if (first.sample < 0) {
			first = second = last = strongest = secondStrongest = c;
			if (gMistake == 1) {
				second.distance = INFINITY;
				secondStrongest.intensity = -INFINITY;
			}
			continue;
		}
		const float dd = c.distance;
		if (dd < first.distance) {
			second = first;
			first = c;
		} else if (first.distance < dd && dd < second.distance) {
			second = c;
		}
		if (gMistake == 2 ? dd < last.distance : dd > last.distance) last = c;
		const float I = c.intensity;
		if (I > strongest.intensity) {
			secondStrongest = strongest;
			strongest = c;
		} else if (strongest.intensity > I && I > secondStrongest.intensity) {
			secondStrongest = c;
		}
	}
	if (out.ambiguous) {
		out.ambiguous[b] =
		    static_cast<int8_t>((ambiguous ? 1 : 0) | (distTie ? 2 : 0) | (intTie ? 4 : 0) | (texelEdge ? 8 : 0) |
		                        (instanceTie ? 16 : 0));
	}
	if (out.secondRange != nullptr) {
		double dLo = INFINITY, dHi = -INFINITY, iLo = INFINITY, iHi = -INFINITY;
		for (const auto& e : dStates) {
			dLo = std::min(dLo, e.second);
			dHi = std::max(dHi, e.second);
		}
		for (const auto& e : iStates) {
			iLo = std::min(iLo, e.second);
			iHi = std::max(iHi, e.second);
		}
		if (statesOverflow || samples == 1) {
			dLo = iLo = -INFINITY;
			dHi = iHi = INFINITY;
		}
		out.secondRange[4 * b + 0] = static_cast<float>(dLo);
		out.secondRange[4 * b + 1] = static_cast<float>(dHi);
		out.secondRange[4 * b + 2] = static_cast<float>(iLo);
		out.secondRange[4 * b + 3] = static_cast<float>(iHi);
	}
	for (int r = 0; r < R; ++r) {
		const int rp = b * R + r;
		const uint8_t type = typeOf(r);
		if (first.sample < 0) {
			writeNonHit(rp, type, sample0Distance);
			continue;
		}
		const Track* t = nullptr;
		switch (type) {
			case 4: t = &first; break;
			case 3: t = &second; break;
			case 2: t = &last; break;
			case 1: t = &strongest; break;
			case 5: t = &secondStrongest; break;
			default: break;
		}
		if (!t) continue;
		const Shade& s = t->sh;
		if (out.xyz) put3(out.xyz, rp, o0 + unit(d0) * static_cast<double>(s.distance));
		if (out.isHit) out.isHit[rp] = 1;
		if (out.returnType) out.returnType[rp] = type;
		if (out.distance) out.distance[rp] = s.distance;
		if (out.intensityF32) out.intensityF32[rp] = s.intensity;
		if (out.intensityU8) out.intensityU8[rp] = s.intensity < 255.0f ? static_cast<uint8_t>(std::round(s.intensity)) : 255;
		if (out.entityId) out.entityId[rp] = s.entityId;
		if (out.absVelocity) put3(out.absVelocity, rp, s.absVel);
		if (out.relVelocity) put3(out.relVelocity, rp, s.relVel);
		if (out.radialSpeed) out.radialSpeed[rp] = s.radial;
		if (out.normal) put3(out.normal, rp, s.normal);
		if (out.incidentAngle) out.incidentAngle[rp] = s.incident;
		if (out.laserRetro) out.laserRetro[rp] = s.retro;
	}
}

World makeWorld(const OrcScene* sc, bool brute)
{
	World w;
	w.sc = sc;
	w.meshes.resize(sc->nMeshes);
	for (int m = 0; m < sc->nMeshes; ++m) {
		MeshData& md = w.meshes[m];
		const int v0 = sc->meshVertOffset[m];
		const int t0 = sc->meshTriOffset[m], t1 = sc->meshTriOffset[m + 1];
		std::vector<Box> boxes;
		for (int t = t0; t < t1; ++t) {
			Tri tri;
			V3* vs[3] = {&tri.a, &tri.b, &tri.c};
			Box b;
			for (int k = 0; k < 3; ++k) {
				const int vi = v0 + sc->indices[3 * t + k];
				md.idx.push_back(sc->indices[3 * t + k]);
				*vs[k] = {sc->vertices[3 * vi], sc->vertices[3 * vi + 1], sc->vertices[3 * vi + 2]};
				b.grow(*vs[k]);
			}
			md.tris.push_back(tri);
			boxes.push_back(b);
			md.box.grow(b);
		}
		if (!brute) md.bvh.build(boxes);
	}
	std::vector<Box> instBoxes;
	for (int i = 0; i < sc->nInstances; ++i) {
		Inst in;
		in.mesh = sc->instMesh[i];
		in.toWorld = M34::fromFloat(sc->instTransform + 12 * i);
		in.toObject = in.toWorld.inverse();
		in.hasPrev = sc->instHasPrev != nullptr && sc->instHasPrev[i] != 0;
		in.prev = in.hasPrev ? M34::fromFloat(sc->instPrevTransform + 12 * i) : M34::identity();
		in.disp = (sc->instDispOffset != nullptr && sc->instDispOffset[i] >= 0) ? sc->displacements + sc->instDispOffset[i]
		                                                                          : nullptr;
		in.id = sc->instId[i];
		in.retro = sc->instRetro ? sc->instRetro[i] : 0.0f;
		in.texture = sc->instTexture ? sc->instTexture[i] : -1;
		const Box& mb = w.meshes[in.mesh].box;
		for (int c = 0; c < 8; ++c) {
			in.worldBox.grow(in.toWorld.point({(c & 1) ? mb.hi.x : mb.lo.x, (c & 2) ? mb.hi.y : mb.lo.y, (c & 4) ? mb.hi.z : mb.lo.z}));
		}
		instBoxes.push_back(in.worldBox);
		w.insts.push_back(in);
	}
	if (!brute) w.tlas.build(instBoxes);
	return w;
}

} // namespace

static int traceWorld(const World& w, const OrcRequest* q, OrcOutputs* out);

extern "C" int orc_raytrace(const OrcScene* scene, const OrcRequest* q, OrcOutputs* out)
{
	if (scene == nullptr || q == nullptr || out == nullptr || q->nRays < 0) {
		return 1;
	}
	const World w = makeWorld(scene, q->bruteForce != 0);
	return traceWorld(w, q, out);
}

// Persistent world: build the BVHs once, trace many requests (the scene arrays must outlive the world).
extern "C" void* orc_world_create(const OrcScene* scene, int bruteForce)
{
	return new World(makeWorld(scene, bruteForce != 0));
}

extern "C" int orc_world_raytrace(void* world, const OrcRequest* q, OrcOutputs* out)
{
	if (world == nullptr || q == nullptr || out == nullptr || q->nRays < 0) {
		return 1;
	}
	return traceWorld(*static_cast<World*>(world), q, out);
}

extern "C" void orc_world_destroy(void* world) { delete static_cast<World*>(world); }

extern "C" void orc_plant_mistake(int which) { gMistake = which; }

static int traceWorld(const World& w, const OrcRequest* q, OrcOutputs* out)
{
	const int threads = std::max(1, q->threads);
	std::atomic<int> next{0};
	auto worker = [&]() {
		for (;;) {
			const int beg = next.fetch_add(64);
			if (beg >= q->nRays) return;
			const int end = std::min(beg + 64, q->nRays);
			for (int b = beg; b < end; ++b) beam(w, *q, *out, b);
		}
	};
	std::vector<std::thread> pool;
	for (int i = 1; i < threads; ++i) pool.emplace_back(worker);
	worker();
	for (auto& t : pool) t.join();
	return 0;
}

// ============================================================================================ noise stream
extern "C" void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for (int round = 0; round < 10; ++round) {
		const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
		const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
		const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
		const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
		c0 = n0;
		c1 = static_cast<uint32_t>(p1);
		c2 = n2;
		c3 = static_cast<uint32_t>(p0);
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

// Cephes logf (S. L. Moshier): x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log(1+f) by a degree-9 polynomial.
extern "C" float orc_logf(float x)
{
	int e = 0;
	float m = std::frexp(x, &e); // m in [0.5, 1)
	float f;
	if (m < 0.707106781186547524f) {
		e -= 1;
		f = m + m - 1.0f;
	} else {
		f = m - 1.0f;
	}
	const float z = f * f;
	static const float c[9] = {7.0376836292e-2f, -1.1514610310e-1f, 1.1676998740e-1f, -1.2420140846e-1f, 1.4249322787e-1f,
	                           -1.6668057665e-1f, 2.0000714765e-1f,  -2.4999993993e-1f, 3.3333331174e-1f};
	float p = c[0];
	for (int i = 1; i < 9; ++i) p = std::fma(p, f, c[i]);
	float y = (p * f) * z;
	const float fe = static_cast<float>(e);
	y = std::fma(fe, -2.12194440e-4f, y);
	y = std::fma(z, -0.5f, y);
	float r = f + y;
	return std::fma(fe, 0.693359375f, r);
}

// sin and cos of 2 pi u: nearest quarter turn, then Cephes sinf/cosf polynomials on [-pi/4, pi/4].
extern "C" void orc_sincos2pi(float u, float* s, float* c)
{
	const float v = u * 4.0f;
	const float q = std::nearbyint(v);
	const float t = (v - q) * 1.57079632679489661923f;
	const float z = t * t;
	float sp = std::fma(std::fma(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
	const float st = std::fma(sp * z, t, t);
	float cp = std::fma(std::fma(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
	const float ct = std::fma(cp * z, z, std::fma(z, -0.5f, 1.0f));
	switch (static_cast<int>(q) & 3) {
		case 0: *s = st; *c = ct; break;
		case 1: *s = ct; *c = -st; break;
		case 2: *s = -st; *c = -ct; break;
		default: *s = -ct; *c = st; break;
	}
}

extern "C" float orc_normal_at(uint64_t seed, uint32_t run, uint32_t tag, uint32_t index)
{
	const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
	const uint32_t ctr[4] = {index, run, tag, 0u};
	uint32_t o[4];
	orc_philox4x32_10(ctr, key, o);
	const float u1 = static_cast<float>((o[0] >> 8) + 1u) * 5.9604644775390625e-8f;
	const float u2 = static_cast<float>(o[1] >> 8) * 5.9604644775390625e-8f;
	const float rad = std::sqrt(-2.0f * orc_logf(u1));
	float s, c;
	orc_sincos2pi(u2, &s, &c);
	return rad * c;
}

extern "C" void orc_normals(uint64_t seed, uint32_t run, uint32_t tag, float* out, int count)
{
	for (int i = 0; i < count; ++i) {
		out[i] = orc_normal_at(seed, run, tag, static_cast<uint32_t>(i));
	}
}
