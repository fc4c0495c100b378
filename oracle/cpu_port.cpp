// FP32 CPU port of the library's lidar graph -- TEST INFRASTRUCTURE ONLY (the timed CPU baseline of bench.py).
//
// The graph rays -> pose -> range -> raytrace -> points transform -> distance noise -> compact, computed the way the
// HIP library computes it, on the host's cores:
//   * acceleration structure: the canonical LBVH, built here on the CPU (63-bit Morton codes of box centroids, 21 bits
//     per axis, in the centroid bounds; radix tree over the sorted codes, split at the highest differing bit;
//     bottom-up boxes), one binary BLAS per mesh over its triangles and a TLAS over the instances' world boxes;
//   * traversal: two levels, closest child first, FP32 slab tests widened by 2^-20 relative, FP32 watertight triangle
//     test (Woop, Benthin, Wald, JCGT 2013) with the same ties rule (lower instance, then lower triangle index);
//   * shading: FP64 hit point and distance, normal from the transformed object-space edges, entity id
//     (src/gpu/optixPrograms.cu:96-149), miss -> non-hit (src/gpu/nodeKernels.cu:153-207);
//   * point stages: FP32 points transform (src/gpu/nodeKernels.cu:51-56) and distance noise from the library's seeded
//     normal stream (src/gpu/gaussianNoiseKernels.cu:49-66), then stable compaction of the hits
//     (src/graph/CompactByFieldPointsNode.cpp:33-105).
// Work is split into 64-ray tiles over all sensors of a frame and taken by `threads` std::threads.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct F3
{
	float x, y, z;
};
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
inline F3 unit(F3 a)
{
	const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
	return {a.x / l, a.y / l, a.z / l};
}

struct M34
{
	float m[12];
	F3 point(F3 p) const
	{
		return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
		        m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
	}
	F3 dir(F3 p) const
	{
		return {m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z, m[8] * p.x + m[9] * p.y + m[10] * p.z};
	}
};

// FP64 inverse of an affine 3x4 matrix (world -> object), kept in FP64 for origins and rounded for directions.
void inverse64(const float* a, double* r)
{
	const double c00 = double(a[5]) * a[10] - double(a[6]) * a[9];
	const double c01 = double(a[6]) * a[8] - double(a[4]) * a[10];
	const double c02 = double(a[4]) * a[9] - double(a[5]) * a[8];
	const double det = double(a[0]) * c00 + double(a[1]) * c01 + double(a[2]) * c02;
	std::memset(r, 0, 12 * sizeof(double));
	if (det == 0.0) {
		return;
	}
	const double id = 1.0 / det;
	r[0] = c00 * id;
	r[1] = (double(a[2]) * a[9] - double(a[1]) * a[10]) * id;
	r[2] = (double(a[1]) * a[6] - double(a[2]) * a[5]) * id;
	r[4] = c01 * id;
	r[5] = (double(a[0]) * a[10] - double(a[2]) * a[8]) * id;
	r[6] = (double(a[2]) * a[4] - double(a[0]) * a[6]) * id;
	r[8] = c02 * id;
	r[9] = (double(a[1]) * a[8] - double(a[0]) * a[9]) * id;
	r[10] = (double(a[0]) * a[5] - double(a[1]) * a[4]) * id;
	for (int y = 0; y < 3; ++y) {
		r[4 * y + 3] = -(r[4 * y] * a[3] + r[4 * y + 1] * a[7] + r[4 * y + 2] * a[11]);
	}
}

struct Box
{
	float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	void grow(F3 p)
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], comp(p, a));
			hi[a] = std::max(hi[a], comp(p, a));
		}
	}
	void grow(const Box& b)
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], b.lo[a]);
			hi[a] = std::max(hi[a], b.hi[a]);
		}
	}
};

// Binary node: both child boxes and both child references (>= 0 internal node, < 0 ~leaf index).
struct Node
{
	Box child[2];
	int ref[2];
};

inline uint64_t spread21(uint64_t v)
{
	v &= 0x1fffff;
	v = (v | v << 32) & 0x1f00000000ffffull;
	v = (v | v << 16) & 0x1f0000ff0000ffull;
	v = (v | v << 8) & 0x100f00f00f00f00full;
	v = (v | v << 4) & 0x10c30c30c30c30c3ull;
	v = (v | v << 2) & 0x1249249249249249ull;
	return v;
}

struct Bvh
{
	std::vector<Node> nodes;
	std::vector<int> leafPrim; // leaf order -> primitive
	std::vector<uint64_t> keys;
	const std::vector<Box>* boxes = nullptr;

	void build(const std::vector<Box>& b)
	{
		boxes = &b;
		const int n = static_cast<int>(b.size());
		nodes.clear();
		leafPrim.resize(n);
		if (n == 0) {
			return;
		}
		Box cb;
		for (const Box& x : b) {
			cb.grow(F3{0.5f * (x.lo[0] + x.hi[0]), 0.5f * (x.lo[1] + x.hi[1]), 0.5f * (x.lo[2] + x.hi[2])});
		}
		std::vector<std::pair<uint64_t, int>> order(n);
		for (int i = 0; i < n; ++i) {
			uint64_t code = 0;
			for (int a = 0; a < 3; ++a) {
				const float c = 0.5f * (b[i].lo[a] + b[i].hi[a]);
				const float ext = cb.hi[a] - cb.lo[a];
				const float q = ext > 0.0f ? (c - cb.lo[a]) / ext : 0.0f;
				const uint64_t k = static_cast<uint64_t>(std::min(std::max(q, 0.0f), 1.0f) * 2097151.0f);
				code |= spread21(k) << (2 - a);
			}
			order[i] = {code, i};
		}
		std::sort(order.begin(), order.end());
		keys.resize(n);
		for (int i = 0; i < n; ++i) {
			keys[i] = order[i].first;
			leafPrim[i] = order[i].second;
		}
		nodes.reserve(std::max(n - 1, 1));
		if (n == 1) {
			Node root;
			root.child[0] = b[leafPrim[0]];
			root.ref[0] = ~0;
			root.ref[1] = ~0;
			for (int a = 0; a < 3; ++a) { // a point box at +inf: no ray with a finite bound enters it
				root.child[1].lo[a] = root.child[1].hi[a] = INFINITY;
			}
			nodes.push_back(root);
			return;
		}
		Box rootBox;
		split(0, n - 1, rootBox);
	}

	// Radix-tree node over sorted leaves [first, last]; returns its reference and its box.
	int split(int first, int last, Box& out)
	{
		if (first == last) {
			out = (*boxes)[leafPrim[first]];
			return ~first;
		}
		const int id = static_cast<int>(nodes.size());
		nodes.emplace_back();
		int mid;
		const uint64_t a = keys[first], z = keys[last];
		if (a == z) {
			mid = (first + last) >> 1;
		} else {
			const int common = __builtin_clzll(a ^ z);
			mid = first;
			int step = last - first;
			do { // last index whose key shares more than `common` leading bits with the first key
				step = (step + 1) >> 1;
				const int probe = mid + step;
				if (probe < last && __builtin_clzll(a ^ keys[probe]) > common) {
					mid = probe;
				}
			} while (step > 1);
		}
		Box l, r;
		const int lr = split(first, mid, l);
		const int rr = split(mid + 1, last, r);
		nodes[id].child[0] = l;
		nodes[id].child[1] = r;
		nodes[id].ref[0] = lr;
		nodes[id].ref[1] = rr;
		out = l;
		out.grow(r);
		return id;
	}
};

struct Mesh
{
	std::vector<F3> v;
	std::vector<int> idx;
	std::vector<F3> leafA, leafB, leafC; // triangles in leaf order
	std::vector<int> leafTri;
	std::vector<Box> triBoxes;
	Bvh bvh;
};

struct Inst
{
	int mesh;
	M34 toWorld;     // as given
	double toObject64[12];
	M34 toObject;    // FP32 rounding (directions)
	int id;
};

struct Scene
{
	std::vector<Mesh> meshes;
	std::vector<Inst> insts;
	std::vector<Box> instBoxes;
	Bvh tlas;
};

constexpr float kSlabGrow = 1.0f + 9.5367431640625e-7f;

struct Ray
{
	F3 o, inv;
	int kx, ky, kz;
	float sx, sy, sz;
};

Ray makeRay(F3 o, F3 d)
{
	Ray r;
	r.o = o;
	r.inv = {1.0f / (d.x != 0.0f ? d.x : 1e-30f), 1.0f / (d.y != 0.0f ? d.y : 1e-30f), 1.0f / (d.z != 0.0f ? d.z : 1e-30f)};
	const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
	r.kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
	r.kx = r.kz == 2 ? 0 : r.kz + 1;
	r.ky = r.kx == 2 ? 0 : r.kx + 1;
	const float dz = comp(d, r.kz);
	if (dz < 0.0f) {
		std::swap(r.kx, r.ky);
	}
	r.sx = comp(d, r.kx) / dz;
	r.sy = comp(d, r.ky) / dz;
	r.sz = 1.0f / dz;
	return r;
}

inline bool slab(const Ray& r, const Box& b, float tBound, float& tNear)
{
	const float tx0 = (b.lo[0] - r.o.x) * r.inv.x, tx1 = (b.hi[0] - r.o.x) * r.inv.x;
	const float ty0 = (b.lo[1] - r.o.y) * r.inv.y, ty1 = (b.hi[1] - r.o.y) * r.inv.y;
	const float tz0 = (b.lo[2] - r.o.z) * r.inv.z, tz1 = (b.hi[2] - r.o.z) * r.inv.z;
	const float tn = std::max(std::max(std::min(tx0, tx1), std::min(ty0, ty1)), std::min(tz0, tz1));
	const float tf = std::min(std::min(std::max(tx0, tx1), std::max(ty0, ty1)), std::max(tz0, tz1));
	tNear = tn;
	return std::max(tn, 0.0f) <= std::min(tf * kSlabGrow, tBound);
}

// Watertight test; on success t and the barycentric weights of B (u) and C (v).
inline bool intersect(const Ray& r, F3 a, F3 b, F3 c, float& t, float& u, float& v)
{
	const F3 A = a - r.o, B = b - r.o, C = c - r.o;
	const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
	const float Ax = comp(A, r.kx) - r.sx * Akz, Ay = comp(A, r.ky) - r.sy * Akz;
	const float Bx = comp(B, r.kx) - r.sx * Bkz, By = comp(B, r.ky) - r.sy * Bkz;
	const float Cx = comp(C, r.kx) - r.sx * Ckz, Cy = comp(C, r.ky) - r.sy * Ckz;
	float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
	if (U == 0.0f || V == 0.0f || W == 0.0f) {
		U = static_cast<float>(double(Cx) * By - double(Cy) * Bx);
		V = static_cast<float>(double(Ax) * Cy - double(Ay) * Cx);
		W = static_cast<float>(double(Bx) * Ay - double(By) * Ax);
	}
	if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) {
		return false;
	}
	const float det = U + V + W;
	if (det == 0.0f) {
		return false;
	}
	const float T = U * (r.sz * Akz) + V * (r.sz * Bkz) + W * (r.sz * Ckz);
	t = T / det;
	u = V / det;
	v = W / det;
	return true;
}

struct Hit
{
	float t = 0.0f, u = 0.0f, v = 0.0f;
	int inst = INT_MAX, prim = INT_MAX;
};

inline bool better(float t, int inst, int prim, const Hit& h)
{
	return t < h.t || (t == h.t && (inst < h.inst || (inst == h.inst && prim < h.prim)));
}

template<typename Leaf>
void traverse(const Bvh& bvh, const Ray& r, const float& tBest, Leaf&& leaf)
{
	if (bvh.nodes.empty()) {
		return;
	}
	int stack[192];
	int sp = 0;
	int node = 0;
	for (;;) {
		const Node& n = bvh.nodes[node];
		const float bound = std::min(tBest * kSlabGrow, FLT_MAX);
		float t0, t1;
		const bool h0 = slab(r, n.child[0], bound, t0), h1 = slab(r, n.child[1], bound, t1);
		int next = INT_MIN;
		auto take = [&](int ref) {
			if (ref < 0) {
				leaf(~ref);
				return false;
			}
			return true;
		};
		if (h0 && h1) {
			const int nearRef = t0 <= t1 ? n.ref[0] : n.ref[1], farRef = t0 <= t1 ? n.ref[1] : n.ref[0];
			stack[sp++] = farRef; // a leaf is tested when popped, after the near subtree
			if (take(nearRef)) {
				next = nearRef;
			}
		} else if (h0) {
			if (take(n.ref[0])) next = n.ref[0];
		} else if (h1) {
			if (take(n.ref[1])) next = n.ref[1];
		}
		while (next == INT_MIN) {
			if (sp == 0) {
				return;
			}
			const int ref = stack[--sp];
			if (take(ref)) {
				next = ref;
			}
		}
		node = next;
	}
}

bool closest(const Scene& sc, F3 o32, const double* o64, F3 d, float tmax, Hit& best)
{
	best = Hit{};
	best.t = tmax;
	bool found = false;
	const Ray world = makeRay(o32, d);
	traverse(sc.tlas, world, best.t, [&](int leaf) {
		const int ii = sc.tlas.leafPrim[leaf];
		const Inst& in = sc.insts[ii];
		const double* w = in.toObject64;
		const F3 oo{static_cast<float>(w[0] * o64[0] + w[1] * o64[1] + w[2] * o64[2] + w[3]),
		            static_cast<float>(w[4] * o64[0] + w[5] * o64[1] + w[6] * o64[2] + w[7]),
		            static_cast<float>(w[8] * o64[0] + w[9] * o64[1] + w[10] * o64[2] + w[11])};
		const Ray obj = makeRay(oo, in.toObject.dir(d));
		const Mesh& m = sc.meshes[in.mesh];
		traverse(m.bvh, obj, best.t, [&](int tl) {
			float t, u, v;
			const int prim = m.leafTri[tl];
			if (intersect(obj, m.leafA[tl], m.leafB[tl], m.leafC[tl], t, u, v) && t > 0.0f && t <= tmax &&
			    better(t, ii, prim, best)) {
				best.t = t;
				best.u = u;
				best.v = v;
				best.inst = ii;
				best.prim = prim;
				found = true;
			}
		});
	});
	return found;
}

void compose64(const float* a, const float* b, double* r)
{
	for (int y = 0; y < 3; ++y) {
		for (int x = 0; x < 4; ++x) {
			r[4 * y + x] = double(a[4 * y]) * b[x] + double(a[4 * y + 1]) * b[4 + x] + double(a[4 * y + 2]) * b[8 + x] +
			               (x == 3 ? double(a[4 * y + 3]) : 0.0);
		}
	}
}

// One beam of one sensor: uncompacted outputs at index i.
void beam(const Scene& sc, const CpuFrame& f, const CpuOut& o, int i)
{
	double w[12];
	compose64(f.pose, f.rays + 12 * i, w);
	const double o64[3] = {w[3], w[7], w[11]};
	const F3 org{static_cast<float>(w[3]), static_cast<float>(w[7]), static_cast<float>(w[11])};
	const F3 dir{static_cast<float>(w[2]), static_cast<float>(w[6]), static_cast<float>(w[10])};
	Hit h;
	bool hit = closest(sc, org, o64, dir, f.range[1], h);
	float distance = INFINITY;
	F3 normal{NAN, NAN, NAN};
	int entity = 0;
	if (hit) {
		const Inst& in = sc.insts[h.inst];
		const Mesh& m = sc.meshes[in.mesh];
		const F3 A = m.v[m.idx[3 * h.prim]], B = m.v[m.idx[3 * h.prim + 1]], C = m.v[m.idx[3 * h.prim + 2]];
		const F3 p = A + (B - A) * h.u + (C - A) * h.v;
		const float* t = in.toWorld.m;
		const double hx = double(t[0]) * p.x + double(t[1]) * p.y + double(t[2]) * p.z + t[3] - o64[0];
		const double hy = double(t[4]) * p.x + double(t[5]) * p.y + double(t[6]) * p.z + t[7] - o64[1];
		const double hz = double(t[8]) * p.x + double(t[9]) * p.y + double(t[10]) * p.z + t[11] - o64[2];
		const double d = std::sqrt(hx * hx + hy * hy + hz * hz);
		if (d < f.range[0]) {
			hit = false; // a too-near hit occludes (src/gpu/optixPrograms.cu:132-136)
		} else {
			distance = static_cast<float>(d);
			normal = unit(cross(in.toWorld.dir(B - A), in.toWorld.dir(C - A)));
			entity = in.id;
		}
	}
	const F3 xyz = org + unit(dir) * (hit ? distance : 0.0f); // non-hits keep the origin (NaN components -> 0)
	// points transform, then distance noise in the look-at frame of the noise node
	const M34& T = *reinterpret_cast<const M34*>(f.toSensor);
	F3 q = T.point(xyz);
	float dist = distance;
	{
		const float sigma = dist * f.noiseRise + f.noiseBase;
		const float err = f.noiseMean + orc_normal_at(f.seed, f.run, f.tag, static_cast<uint32_t>(i)) * sigma;
		const M34& L = *reinterpret_cast<const M34*>(f.lookAt);
		const M34& Li = *reinterpret_cast<const M34*>(f.lookAtInv);
		const F3 pr = L.point(q);
		const float len = std::sqrt(pr.x * pr.x + pr.y * pr.y + pr.z * pr.z);
		q = Li.point(F3{pr.x + pr.x / len * err, pr.y + pr.y / len * err, pr.z + pr.z / len * err});
		dist = dist + err;
	}
	o.fullXyz[3 * i] = q.x;
	o.fullXyz[3 * i + 1] = q.y;
	o.fullXyz[3 * i + 2] = q.z;
	o.fullDistance[i] = dist;
	o.fullEntityId[i] = entity;
	o.fullNormal[3 * i] = normal.x;
	o.fullNormal[3 * i + 1] = normal.y;
	o.fullNormal[3 * i + 2] = normal.z;
	o.fullIsHit[i] = hit ? 1 : 0;
}

} // namespace

extern "C" void* cpu_scene_create(const OrcScene* s)
{
	auto* sc = new Scene();
	sc->meshes.resize(s->nMeshes);
	for (int m = 0; m < s->nMeshes; ++m) {
		Mesh& me = sc->meshes[m];
		const int v0 = s->meshVertOffset[m], v1 = s->meshVertOffset[m + 1];
		const int t0 = s->meshTriOffset[m], t1 = s->meshTriOffset[m + 1];
		for (int v = v0; v < v1; ++v) {
			me.v.push_back(F3{s->vertices[3 * v], s->vertices[3 * v + 1], s->vertices[3 * v + 2]});
		}
		for (int t = t0; t < t1; ++t) {
			Box b;
			for (int k = 0; k < 3; ++k) {
				me.idx.push_back(s->indices[3 * t + k]);
				b.grow(me.v[s->indices[3 * t + k]]);
			}
			me.triBoxes.push_back(b);
		}
		me.bvh.build(me.triBoxes);
		for (int leaf = 0; leaf < static_cast<int>(me.bvh.leafPrim.size()); ++leaf) {
			const int t = me.bvh.leafPrim[leaf];
			me.leafA.push_back(me.v[me.idx[3 * t]]);
			me.leafB.push_back(me.v[me.idx[3 * t + 1]]);
			me.leafC.push_back(me.v[me.idx[3 * t + 2]]);
			me.leafTri.push_back(t);
		}
	}
	for (int i = 0; i < s->nInstances; ++i) {
		Inst in;
		in.mesh = s->instMesh[i];
		std::memcpy(in.toWorld.m, s->instTransform + 12 * i, 12 * sizeof(float));
		inverse64(in.toWorld.m, in.toObject64);
		for (int k = 0; k < 12; ++k) {
			in.toObject.m[k] = static_cast<float>(in.toObject64[k]);
		}
		in.id = s->instId[i];
		Box mb;
		for (const Box& b : sc->meshes[in.mesh].triBoxes) {
			mb.grow(b);
		}
		Box wb;
		for (int c = 0; c < 8; ++c) {
			wb.grow(in.toWorld.point(F3{(c & 1) ? mb.hi[0] : mb.lo[0], (c & 2) ? mb.hi[1] : mb.lo[1], (c & 4) ? mb.hi[2] : mb.lo[2]}));
		}
		for (int a = 0; a < 3; ++a) { // the same outward padding as the library's instance boxes
			const float pad = (std::fabs(wb.lo[a]) + std::fabs(wb.hi[a]) + (wb.hi[a] - wb.lo[a])) * 9.5367431640625e-7f + 1e-30f;
			wb.lo[a] -= pad;
			wb.hi[a] += pad;
		}
		sc->instBoxes.push_back(wb);
		sc->insts.push_back(in);
	}
	sc->tlas.build(sc->instBoxes);
	return sc;
}

extern "C" void cpu_scene_destroy(void* scene) { delete static_cast<Scene*>(scene); }

extern "C" int cpu_run(void* const* scenes, const CpuFrame* frames, CpuOut* outs, int count, int threads)
{
	struct Tile
	{
		int sensor, begin, end;
	};
	std::vector<Tile> tiles;
	for (int k = 0; k < count; ++k) {
		for (int b = 0; b < frames[k].nRays; b += 64) {
			tiles.push_back({k, b, std::min(b + 64, frames[k].nRays)});
		}
	}
	std::vector<int> tileHits(tiles.size());
	std::atomic<size_t> next{0};
	auto trace = [&]() {
		for (size_t t; (t = next.fetch_add(1)) < tiles.size();) {
			const Tile& tl = tiles[t];
			int hits = 0;
			for (int i = tl.begin; i < tl.end; ++i) {
				beam(*static_cast<const Scene*>(scenes[tl.sensor]), frames[tl.sensor], outs[tl.sensor], i);
				hits += outs[tl.sensor].fullIsHit[i];
			}
			tileHits[t] = hits;
		}
	};
	std::vector<std::thread> pool;
	for (int i = 1; i < threads; ++i) {
		pool.emplace_back(trace);
	}
	trace();
	for (auto& th : pool) {
		th.join();
	}
	// Stable compaction: tile offsets per sensor, then every tile copies its kept points.
	std::vector<int> offset(tiles.size());
	for (int k = 0; k < count; ++k) {
		outs[k].kept = 0;
	}
	for (size_t t = 0; t < tiles.size(); ++t) {
		offset[t] = outs[tiles[t].sensor].kept;
		outs[tiles[t].sensor].kept += tileHits[t];
	}
	next = 0;
	auto scatter = [&]() {
		for (size_t t; (t = next.fetch_add(1)) < tiles.size();) {
			const Tile& tl = tiles[t];
			const CpuOut& o = outs[tl.sensor];
			int at = offset[t];
			for (int i = tl.begin; i < tl.end; ++i) {
				if (!o.fullIsHit[i]) {
					continue;
				}
				std::memcpy(o.xyz + 3 * at, o.fullXyz + 3 * i, 12);
				o.distance[at] = o.fullDistance[i];
				o.entityId[at] = o.fullEntityId[i];
				std::memcpy(o.normal + 3 * at, o.fullNormal + 3 * i, 12);
				++at;
			}
		}
	};
	pool.clear();
	for (int i = 1; i < threads; ++i) {
		pool.emplace_back(scatter);
	}
	scatter();
	for (auto& th : pool) {
		th.join();
	}
	return 0;
}
