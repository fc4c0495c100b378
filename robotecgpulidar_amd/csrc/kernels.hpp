// Host-callable launch wrappers of the HIP kernels. Every wrapper only enqueues work on `stream`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_types.hpp"

namespace rgl {

struct Aabb
{
	float lo[3];
	float hi[3];
};

namespace gpu {

// ---------------------------------------------------------------- BVH (bvh.hip)
size_t lbvhScratchBytes(int n);
// Builds an LBVH over n primitive boxes. nodes: max(n-1, 1) entries; leafPrim: n (primitive of each leaf in
// leaf order); parentSlot: 2n entries; counters: max(n-1,1); height: max(n-1,1) (height[0] = tree depth).
// When leafRefIsPrim, leaf references in nodes hold the primitive index (TLAS); otherwise the leaf index
// (BLAS, whose triangles are gathered in leaf order). Trees of up to 16384 leaves take their topology from
// agglomerative clustering over the Morton order instead of the radix tree.
void buildLbvh(hipStream_t stream, int n, const Aabb* boxes, bool leafRefIsPrim, BvhNode* nodes, int32_t* leafPrim,
               int32_t* parentSlot, int32_t* counters, int32_t* height, void* scratch);
// Refits the boxes of an existing LBVH topology to new primitive boxes.
void refitLbvh(hipStream_t stream, int n, const Aabb* boxes, const int32_t* leafPrim, const int32_t* parentSlot,
               BvhNode* nodes, int32_t* counters, int32_t* height);
// 4-wide nodes from a binary LBVH over n leaves (built by buildLbvh). bvh4Ranks marks the kept (even-depth) internal
// nodes in `scratch` (bvh4ScratchBytes(n)) and writes their number to *countDev; bvh4Collapse then writes that many
// nodes of `format` (kBvh4F, kBvh4Q or kBvh4H) to `out`. Leaf references keep the binary tree's convention.
// Trace order of n rays: a permutation of [0, n) that lists rays tile by tile of their directions' (azimuth,
// elevation), so that the rays one wave traces point into a compact patch of directions and still form runs of
// consecutive indices (`scratch`: directionOrderScratchBytes(n)).
size_t directionOrderScratchBytes(int n);
void directionOrder(hipStream_t stream, int n, const Mat3x4f* rays, int32_t* order, void* scratch);
size_t bvh4ScratchBytes(int n);
void bvh4Ranks(hipStream_t stream, int n, const int32_t* parentSlot, void* scratch, int32_t* countDev);
void bvh4Collapse(hipStream_t stream, int n, const BvhNode* nodes, const void* scratch, int format, void* out);
void triangleBoxes(hipStream_t stream, int n, const Vec3f* v, const Vec3i* idx, Aabb* out);
void gatherTris(hipStream_t stream, int n, const int32_t* leafPrim, const Vec3f* v, const Vec3i* idx, LeafTri* out);
// Animation vertex updates (src/gpu/sceneKernels.cu:18-49). External: displacement = fresh - vertices, vertices =
// fresh. Skinning: fresh = sum_k w_k (matrices[bone_k] * rest), then the same update.
void updateVertices(hipStream_t stream, int n, const Vec3f* fresh, Vec3f* vertices, Vec3f* displacement);
void skinVertices(hipStream_t stream, int n, const Vec3f* rest, const BoneWeights* weights, const Mat3x4f* matrices,
                  Vec3f* vertices, Vec3f* displacement);
void instanceBoxes(hipStream_t stream, int n, const InstanceShade* shades, const InstanceRec* recs, Aabb* out);

// ---------------------------------------------------------------- raytrace (raytrace.hip)
void raytrace(hipStream_t stream, const SceneView& scene, const RaytraceParams& params, const RaytraceOutputs& out);
// Per-ray traversal statistics (internal nodes visited, triangles tested, instances entered) on the same BVH and
// with the same traversal order as raytrace(); used for algorithmic-byte accounting.
void traversalStats(hipStream_t stream, const SceneView& scene, const RaytraceParams& params, int32_t* nodesVisited,
                    int32_t* trisTested, int32_t* instancesEntered);
// Footprint of raytrace(): inserts every 64 B line its rays read (BVH nodes in the scene's format, leaf triangles,
// instance records, and the shading records, index triples, vertices and texture coordinates of the hits) into the
// open-addressing set keys[0..mask] (key = line + 1, 0 = empty); *full is set if the set ran out of room.
void traversalFootprint(hipStream_t stream, const SceneView& scene, const RaytraceParams& params,
                        unsigned long long* keys, uint32_t mask, uint32_t* full);
// *out += number of nonzero keys[0..n) (out must be zeroed by the caller).
void countNonzeroKeys(hipStream_t stream, const unsigned long long* keys, size_t n, unsigned long long* out);
// Brute-force closest hit over every triangle of every instance (testing aid; same intersection routine).
void bruteForceClosest(hipStream_t stream, const SceneView& scene, const int32_t* triCounts, const RaytraceParams& params,
                       float* outT, int32_t* outInstance, int32_t* outPrim);
void transformRays(hipStream_t stream, int n, const Mat3x4f* in, Mat3x4f* out, const Mat3x4f& tf);
// Launch variants of the batched launches: raytrace single-sample blocks of 128 (0) or 64 (1) threads, chosen by the
// batch autotuner, or (2, only when forced) a persistent grid of 64-thread blocks that claim runs of 64-ray blocks;
// compaction tiles of 8 (0) or 4 (1) points per thread.
constexpr int kRaytraceVariants = 3;
constexpr int kRaytraceTunedVariants = 2;
constexpr int kCompactVariants = 2;
// Blocks of a raytrace launch over rayCount beams: one beam per thread, or one wave per divergent beam.
int raytraceBlocks(int rayCount, bool multi, int variant);
// True if the beams of `p` are traced with the 58-sample divergence model.
bool beamIsMulti(const RaytraceParams& p);
// Batched launches: `items` live in device memory; every item's scene uses node format `format`; with `multi` every
// item has divergent beams (blocks from raytraceBlocks(n, true)), otherwise none has. The persistent variant (2) takes
// `claims`, a device counter that must be zero when the launch starts.
void raytraceBatch(hipStream_t stream, const RaytraceBatchItem* items, int itemCount, int totalBlocks, int format,
                   bool multi, int variant, uint32_t* claims = nullptr);
int pointOpBlocks(int capacity);
void pointOpsBatch(hipStream_t stream, const PointOpItem* items, int itemCount, int totalBlocks);
void compactBatch(hipStream_t stream, const CompactItem* items, int itemCount, int totalTiles, int variant);

// ---------------------------------------------------------------- point-cloud kernels (points.hip)
// Point clouds are processed over `capacity` elements; `count` (device, may be nullptr) holds the live count
// when it is only known on the device (after compaction).
int compactTiles(int capacity, int variant = 0);
void compact(hipStream_t stream, int capacity, const int32_t* countIn, const int32_t* flags, const FieldCopyList& fields,
             int32_t* tileCounts, int32_t* countOut);
// One elementwise stage (any PointOpKind but kOpFormat) over it.capacity elements (it.count live on the device).
void pointOp(hipStream_t stream, const PointOpItem& it);
// Copies element i of every field (f.src -> f.dst, f.size bytes) for i below the live count.
void copyFields(hipStream_t stream, int capacity, const int32_t* count, const FieldCopyList& fields);
// dst[i] = *src[i] for i < n (`src` lives in device memory).
void gatherInts(hipStream_t stream, int n, const int32_t* const* src, int32_t* dst);
void noiseNormals(hipStream_t stream, int n, float* out, uint64_t seed, uint32_t run, uint32_t tag);
// The same normals into host memory through a fixed per-device staging array (no device allocation); n <=
// kNoAllocNormalsMax. Synchronous; calls that overlap on one device share the array.
constexpr int kNoAllocNormalsMax = 1 << 20;
void noiseNormalsNoAlloc(int n, float* hostOut, uint64_t seed, uint32_t run, uint32_t tag);
// SoA -> AoS (format) and AoS -> SoA (from-array) over `capacity` records of `recordSize` bytes.
void formatSoaToAos(hipStream_t stream, int capacity, const int32_t* count, int recordSize, const FieldCopyList& fields,
                    char* out);
void formatAosToSoa(hipStream_t stream, int n, int recordSize, const char* in, const FieldCopyList& fields);
void fillInt(hipStream_t stream, int n, int32_t* p, int32_t value);
// Radar range x azimuth binning: dense per-cell outputs (rangeCells * azCells entries) and keep flags; the caller
// compacts the non-empty cells. sums / counts are scratch of one entry per cell.
void radarBin(hipStream_t stream, int capacity, const int32_t* count, const Vec3f* xyz, const float* distance,
              const float* incident, const int32_t* isHit, const Mat3x4f& toSensor, const RadarBinParams& bp,
              unsigned long long* sums, uint32_t* counts, Vec3f* cellXyz, float* cellDistance, float* cellAzimuth,
              float* cellRcs, int32_t* cellHit, int32_t* keep);

// Radar postprocessing (src/gpu/nodeKernels.cu:82-101,209-320; src/graph/RadarPostprocessPointsNode.cpp:74-174):
// per-point scattered-field terms, then one wave per node clusters the points in order, merges clusters, and writes
// per cluster its centre point, RCS, noise, power and SNR (info = {cluster count, error bits}); then the centre
// points' fields are gathered. `it.gather` must point to device memory. The batched form runs several nodes in
// one set of launches (items in device memory, energy and gather blocks laid out by firstBlock).
int radarBlocks(int capacity);
void radarPostprocess(hipStream_t stream, const RadarItem& it);
void radarPostprocessBatch(hipStream_t stream, const RadarItem* items, int count, int totalBlocks);

} // namespace gpu
} // namespace rgl
