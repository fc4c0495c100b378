// Data layouts that live in HBM and are shared between the host runtime and the HIP kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include "math.hpp"

namespace rgl {

// ------------------------------------------------------------------------------------------------ BVH
// Binary BVH node, 64 B: the two child boxes stored in the parent (so one node fetch decides both children)
// and the two child references.
//   f[0..5]   child 0 box (lo.x lo.y lo.z hi.x hi.y hi.z)
//   f[6..11]  child 1 box
//   i[12..13] child references: >= 0 internal node index, < 0 leaf (~leafIndex)
//   i[14..15] unused (keeps the node at 64 B so it is fetched as four 16 B loads)
// An empty child slot has an inverted box (lo = +inf, hi = -inf), which no slab test accepts.
struct alignas(16) BvhNode
{
	float f[12];
	int32_t child[2];
	int32_t pad[2];
};
static_assert(sizeof(BvhNode) == 64, "BVH node must be 64 B");

// Node formats traversal can use. Every format is derived from the same binary LBVH, so the set of triangles a ray
// can reach, and therefore the closest hit, does not depend on the format.
enum BvhFormat : int32_t
{
	kBvh2 = 0,   // BvhNode: binary, FP32 child boxes, 64 B
	kBvh4F = 1,  // Bvh4FNode: 4-wide, FP32 child boxes, 128 B
	kBvh4Q = 2,  // Bvh4QNode: 4-wide, child boxes quantized to 8 bits relative to the node box, 64 B
	kBvh4H = 3,  // Bvh4HNode: 4-wide, child boxes quantized to 16 bits relative to the node box, 80 B
};
constexpr int kBvhFormats = 4;

// 4-wide node with FP32 child boxes stored per axis ([axis][child]) so the four slab tests load contiguous lanes.
// Children occupy slots [0, count); child references follow the BvhNode convention. An empty slot holds the point
// box lo = hi = (+inf, +inf, +inf), which no ray with a finite t bound can hit, so traversal needs no count test.
struct alignas(16) Bvh4FNode
{
	float lo[3][4];
	float hi[3][4];
	int32_t child[4];
	int32_t count;
	int32_t pad[3];
};
static_assert(sizeof(Bvh4FNode) == 128, "BVH4 FP32 node must be 128 B");

// 4-wide node with quantized child boxes. A child bound on axis a stands for the plane origin[a] + q * 2^(exp[a] - 127)
// (exact real value); the builder picks q so that this plane lies outside the child's FP32 bound by a margin of 2^-20
// of the node's extent on that axis (lower bounds rounded down, upper bounds rounded up, checked in exact arithmetic).
// Traversal evaluates the plane distances with the decode folded into its slab constants, whose extra rounding the
// margin covers, so traversal with quantized boxes culls nothing the exact boxes would keep.
struct alignas(16) Bvh4QNode
{
	float origin[3];
	uint8_t exp[3];
	uint8_t count;
	uint8_t qlo[3][4]; // [axis][child]
	uint8_t qhi[3][4];
	int32_t child[4];
	int32_t pad[2];
};
static_assert(sizeof(Bvh4QNode) == 64, "BVH4 quantized node must be 64 B");

// The same with 16-bit bounds (q in [0, 65535]): boxes within 2^-16 of the node extent of the exact ones.
struct alignas(16) Bvh4HNode
{
	float origin[3];
	uint8_t exp[3];
	uint8_t count;
	uint16_t qlo[3][4]; // [axis][child]
	uint16_t qhi[3][4];
	int32_t child[4];
};
static_assert(sizeof(Bvh4HNode) == 80, "BVH4 16-bit node must be 80 B");

// Leaf triangle in BVH leaf order, 48 B: object-space vertices A, B, C. w of the first vec4 holds the
// triangle's index in the mesh index buffer (bit pattern of an int32), which the shading stage uses for UVs
// and which breaks distance ties deterministically.
struct alignas(16) LeafTri
{
	float ax, ay, az;
	int32_t prim;
	float bx, by, bz, pad0;
	float cx, cy, cz, pad1;
};
static_assert(sizeof(LeafTri) == 48, "leaf triangle must be 48 B");

// Instance record read by traversal when it enters an instance (160 B).
struct alignas(16) InstanceRec
{
	Mat3x4f worldToObject;      // FP32 rounding of worldToObject64 (its linear part maps ray directions)
	const void* blasNodes;      // BLAS of the instance's mesh, in the scene's node format
	const LeafTri* blasTris;
	Mat3x4d worldToObject64;    // FP64 inverse of the entity transform (maps the FP64 ray origin)
};
static_assert(sizeof(InstanceRec) == 160, "instance record must be 160 B");

// Per-instance data used only when shading a hit (closest-hit stage).
struct alignas(16) InstanceShade
{
	Mat3x4f objectToWorld;      // entity transform as the user gave it (FP32)
	Mat3x4f prevObjectToWorld;  // previous-frame transform for velocities
	const Vec3f* vertices;      // mesh vertices (object space)
	const Vec3i* indices;       // mesh index buffer
	const Vec2f* uvs;           // nullptr if the mesh has no texture coordinates
	const uint8_t* texels;      // nullptr if the entity has no intensity texture
	const BvhNode* blas2;       // binary BLAS of the mesh (its root holds the mesh's object-space box)
	const Vec3f* displacement;  // animated entity: object-space vertex displacement since the previous frame, or nullptr
	int32_t texWidth;
	int32_t texHeight;
	int32_t entityId;
	float laserRetro;
	int32_t hasPrev;
	int32_t pad[3];
};
static_assert(sizeof(InstanceShade) == 176, "instance shading record must be 176 B");

// Skinning weights of one vertex (the layout of rgl_bone_weights_t, include/rgl/api/core.h:160-171).
struct BoneWeights
{
	float weights[4];
	int32_t bones[4];
};
static_assert(sizeof(BoneWeights) == 32, "bone weights record must be 32 B");

// Device view of a scene for one launch.
struct SceneView
{
	const void* tlasNodes;          // nullptr if the scene is empty; node format `format`
	const BvhNode* tlas2;           // the binary TLAS (canonical tree for traversal statistics)
	const InstanceRec* instances;
	const InstanceShade* shades;
	int32_t instanceCount;
	int32_t format;                 // BvhFormat of the TLAS and of every BLAS
	int32_t* overflow;              // set to 1 if a traversal stack would overflow (the ray is then cut short)
};

// ------------------------------------------------------------------------------------------------ raytrace
// Output field pointers of a raytrace launch (nullptr = field not requested). Layout: return point
// `beam * returnCount + r`.
struct RaytraceOutputs
{
	Vec3f* xyz;
	int32_t* isHit;
	uint32_t* rayIdx;
	uint16_t* ringId;
	uint8_t* returnType;
	float* distance;
	float* intensityF32;
	uint8_t* intensityU8;
	float* laserRetro;
	double* timeStampF64;
	uint32_t* timeStampU32;
	int32_t* entityId;
	Vec3f* absVelocity;
	Vec3f* relVelocity;
	float* radialSpeed;
	float* azimuth;
	float* elevation;
	Vec3f* normal;
	float* incidentAngle;
};


struct RaytraceParams
{
	const Mat3x4f* rays;          // rays of the rays-node chain (world frame unless rayPreTransformed)
	const int32_t* order;         // single-sample launches: thread t traces beam order[t] (nullptr: beam t)
	int32_t rayCount;
	int32_t rayPreTransformed;    // 1: the world-frame ray is rayPreTransform * rays[i], composed in FP64
	Mat3x4f rayPreTransform;
	Mat3x4f rayOriginToWorld;     // cumulative ray transform of the rays chain
	Mat3x4f worldToRayOrigin;     // its inverse, computed once on the host (FP64, rounded)
	Mat3x4d worldToRayOrigin64;   // the same inverse in FP64 (velocity distortion)
	const Vec2f* ranges;          // 1 or rayCount entries
	int32_t rangesCount;
	const int32_t* ringIds;
	int32_t ringIdsCount;
	const float* timeOffsetsMs;
	const int8_t* mask;
	int32_t doDistortion;
	Vec3f sensorLinearVelocity;
	Vec3f sensorAngularVelocity;
	float nearNonHitDistance;
	float farNonHitDistance;
	float defaultIntensity;
	float sceneDeltaTime;         // seconds, 0 if unknown
	float hHalfDivergence;        // radians
	float vHalfDivergence;
	int32_t returnMode;
	int32_t returnCount;
	int32_t velocityRequested;
};

#if defined(__HIPCC__)
// Item of a batched launch that block `block` belongs to: the last item whose first block (key) is <= block; items
// are sorted by first block and the first one starts at 0. Each lane of a wave tests one item and a ballot counts the
// matches, so the lookup costs one round of loads per 64 items rather than a chain of dependent loads (which
// dominated short blocks). The result is wave-uniform. Every lane of the wave must be active.
template<typename Item, typename Key>
__device__ __forceinline__ int batchItemOf(const Item* items, int itemCount, int block, Key key)
{
	const int lane = static_cast<int>(threadIdx.x & 63);
	int count = 0;
	for (int base = 0; base < itemCount; base += 64) {
		const int k = base + lane;
		const bool le = k < itemCount && key(items[k]) <= block;
		count += __popcll(__ballot(le));
	}
	return __builtin_amdgcn_readfirstlane(count - 1);
}

// Workgroups of a launch are dealt to the 8 XCDs of the device (each with its own L2) in turn: hardware block b
// runs on XCD b % 8. The logical block of b gives XCD x the chunks x, x + 8, x + 16, ... of `chunk` consecutive
// logical blocks, so the rays one L2 serves at a time are neighbours (they read the same node and triangle records)
// while the XCDs still share the work evenly; chunk <= 0 spreads each XCD's share in one contiguous range. Blocks
// past the last whole round keep their index, so the mapping is a permutation of [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcdLogicalBlock(int b, int n, int chunk)
{
	if (chunk <= 0) {
		chunk = max(1, n / kXcds);
	}
	const int round = kXcds * chunk;
	const int whole = n / round * round;
	if (b >= whole) {
		return b;
	}
	const int x = b % kXcds, j = b / kXcds;
	return ((j / chunk) * kXcds + x) * chunk + j % chunk;
}
#endif

// ------------------------------------------------------------------------------------------------ points
constexpr int kMaxFormatFields = 32;
constexpr int kMaxCompactFields = 32;

struct FieldCopy
{
	const char* src;
	char* dst;
	int32_t size;    // bytes per element
	int32_t offset;  // byte offset inside an AoS record (format / from-array only)
};

struct FieldCopyList
{
	FieldCopy f[kMaxCompactFields];
	int32_t count;
};

// ------------------------------------------------------------------------------------------------ radar cells
// Range x azimuth binning of radar hits (build-defined). Cell (r, a) covers distances [r*rangeStep, (r+1)*rangeStep)
// and sensor-frame azimuths [azMin + a*azStep, azMin + (a+1)*azStep).
struct RadarBinParams
{
	float rangeStep;
	float azStep;
	float azMin;
	float raySolidAngle; // steradians per ray
	int32_t rangeCells;
	int32_t azCells;
};
// Radar postprocessing with the reference's semantics (src/graph/RadarPostprocessPointsNode.cpp): distance scopes
// with their separation thresholds (the layout of rgl_radar_scope_t), energy and link-budget parameters, and the
// node's noise stream.
struct RadarScope
{
	float begin, end, distanceThreshold, radialSpeedThreshold, azimuthThreshold;
};
constexpr int kMaxRadarScopes = 16;
constexpr int kRadarMaxClusters = 1024; // clusters held in LDS by the clustering wave
struct RadarPostParams
{
	RadarScope scopes[kMaxRadarScopes];
	int32_t scopeCount;
	float azimuthStep, elevationStep, frequency;
	float powerTransmitted, deviceGain, noiseMean, noiseStDev;
	Mat3x4f lookAt; // point frame -> sensor frame
	uint64_t seed;
	uint32_t run, tag;
};
// One radar postprocess node's launch description (alone or inside a batched launch).
struct RadarItem
{
	RadarPostParams p;
	const Mat3x4f* rayPose;
	const float* distance;
	const float* azimuth;
	const float* elevation;
	const float* radialSpeed;
	const Vec3f* normal;
	const int32_t* count; // live input count on the device, nullptr = capacity
	float* terms;         // 4 floats per point
	int32_t* work;        // 3 ints per point
	int32_t* centres;
	float* power;
	float* rcs;
	float* noise;
	float* snr;
	int32_t* info;                // {cluster count, error bits}
	const FieldCopyList* gather;  // centre-point fields to gather (device memory)
	int32_t capacity;
	int32_t firstBlock;           // energy / gather launches: blocks [firstBlock, firstBlock + blocks(capacity))
};
// Error bits reported by the clustering kernel.
constexpr int32_t kRadarErrRcsNan = 1;
constexpr int32_t kRadarErrTooManyClusters = 2;

// Fixed-point scale of the RCS sums: contributions are rounded to multiples of 2^-24 m^2 and summed as 64-bit
// integers, so a cell's sum does not depend on the order in which hits are added.
constexpr double kRcsFixedScale = 16777216.0;

// ------------------------------------------------------------------------------------------------ batches
// Elementwise point / ray stage of one graph inside a batched launch.
enum PointOpKind : int32_t
{
	kOpTransformPoints = 0, // out0 = m0 * in0                                   (nodeKernels.cu:51-56)
	kOpNoiseHitpoint = 1,   // angular hit-point noise                           (gaussianNoiseKernels.cu:31-47)
	kOpNoiseDistance = 2,   // distance noise                                    (gaussianNoiseKernels.cu:49-66)
	kOpTransformRays = 3,   // out0 = m0 * in0 for rays                          (nodeKernels.cu:45-49)
	kOpNoiseRays = 4,       // angular ray noise                                 (gaussianNoiseKernels.cu:20-29)
	kOpFormat = 5,          // SoA -> AoS                                         (nodeKernels.cu:25-33)
	kOpToHost = 6,          // copy fields (SoA) into host-mapped pinned memory  (YieldPointsNode.cpp:31-35)
};

struct PointOpItem
{
	int32_t op;
	int32_t capacity;
	const int32_t* count; // live count on the device, nullptr = capacity
	const void* in0;
	const void* in1;
	void* out0;
	void* out1;
	Mat3x4f m0; // transform, or look-at transform for noise
	Mat3x4f m1; // inverse look-at transform for noise
	float a, b, c; // mean, st dev (base), st dev rise
	int32_t axis;
	uint64_t seed;
	uint32_t run, tag;
	const FieldCopyList* format; // kOpFormat / kOpToHost field list (device memory)
	int32_t recordSize;
	int32_t firstBlock;
};

// One sensor of a batched raytrace launch; blocks [firstBlock, next item's firstBlock) trace its beams.
// fused[0..fusedCount) are elementwise stages of the same graph folded into the launch (points transform, distance
// noise; single-sample beams only): after a beam's return points are written, stage 0 reads their XYZ and DISTANCE
// back and every later stage takes the previous one's, each writing its own outputs as its own launch would.
struct RaytraceBatchItem
{
	SceneView sc;
	RaytraceParams p;
	RaytraceOutputs out;
	int32_t firstBlock;
	int32_t fusedCount;
	PointOpItem fused[2];
};

// Compaction of one graph inside a batched launch; tiles [firstTile, firstTile + numTiles).
struct CompactItem
{
	int32_t capacity;
	const int32_t* countIn;
	const int32_t* flags;
	const FieldCopyList* fields; // device memory
	int32_t* tileCounts;
	int32_t* countOut;
	int32_t numTiles;
	int32_t firstTile;
};

} // namespace rgl
