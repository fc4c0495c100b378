// Scene model: meshes (with their BLAS), textures, entities (instances) and the default scene (TLAS).
//
// Mirrors the reference's object model (src/scene/Mesh.cpp:20-41, src/scene/Entity.cpp:26-114,
// src/scene/Texture.cpp:17-76, src/scene/Scene.cpp:41-233). Acceleration structures are this library's own
// LBVHs: one BLAS per mesh, shared by every entity of that mesh, and a TLAS over entity world boxes that is
// rebuilt (or refitted, when only transforms moved) on the scene stream when the scene changed.
#pragma once

#include <map>
#include <memory>
#include <optional>
#include <vector>

#include "device_types.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rgl {

class Scene;

// Bottom-level acceleration structure over one triangle set: a mesh's own vertices, or the vertices of an animated
// entity (which the reference gives its own refittable GAS, src/scene/Scene.cpp:195-233).
struct Blas
{
	// Full LBVH build over triangles `idx[0..triCount)` of `v`; waits for it to finish.
	void build(hipStream_t stream, int triCount, const Vec3f* v, const Vec3i* idx, bool refittable);
	// Same topology, new vertex positions: refits boxes bottom-up and regathers leaf triangles, and re-collapses the
	// 4-wide trees already built; asynchronous on `stream`.
	void refit(hipStream_t stream, const Vec3f* v, const Vec3i* idx);
	// Nodes in `format` (4-wide trees are collapsed from the binary one on first use; waits for that).
	const void* nodesFor(hipStream_t stream, int format);
	int64_t bytes() const;
	void release();

	bool built = false;
	bool keepScratch = false; // refittable: keeps the 4-wide collapse ranks for refits
	int depth = 0;
	int triCount = 0;
	DeviceBuffer nodes, tris, leafPrim, parentSlot, counters, height, boxes, scratch;
	DeviceBuffer wide[kBvhFormats]; // 4-wide BLAS per format (index kBvh4F, kBvh4Q)
	DeviceBuffer wideRanks;   // 4-wide collapse ranks (topology only)
	int32_t wideCount = 0;    // 4-wide node count (topology only)
	bool wideBuilt[kBvhFormats] = {};
};

struct Mesh
{
	Mesh(const Vec3f* vertices, int vertexCount, const Vec3i* indices, int indexCount);
	void setTexCoords(const Vec2f* uvs, int count);
	// Skinning data (src/scene/Mesh.cpp:43-68): one weight record per vertex; rest poses per bone.
	void setBoneWeights(const BoneWeights* weights, int count);
	void setRestposes(const Mat3x4f* restposes, int count);
	// Builds the BLAS on `stream` if needed (first use); waits for it to finish.
	void ensureBlas(hipStream_t stream);
	// BLAS nodes in `format` (built on first use from the binary BLAS; waits for it to finish).
	const void* nodesFor(hipStream_t stream, int format);
	// Device bytes of the acceleration structures (built, or estimated for `format` when not built yet).
	int64_t accelBytes(int format) const;
	// Frees the acceleration structures (rebuilt on next use). The caller makes every scene using the mesh stale.
	void releaseAccel();

	int vertexCount = 0;
	int triCount = 0;
	bool hasUVs = false;
	DeviceBuffer vertices, indices, uvs;
	Blas blas;
	DeviceBuffer boneWeights; // BoneWeights per vertex
	bool hasBoneWeights = false;
	int maxBoneIndex = -1;
	std::vector<Mat3x4f> restposes;
};

// Estimated device bytes of a BLAS over `tris` triangles in `format`.
int64_t estimatedBlasBytes(int64_t tris, int format);

struct Texture
{
	Texture(const void* texels, int width, int height);
	int width = 0, height = 0;
	DeviceBuffer texels;
};

struct Time
{
	uint64_t ns = 0;
};

// Vertices of an animated entity (src/scene/animator/ExternalAnimator.cpp, SkeletonAnimator.cpp): the current
// vertices, their displacement since the previous animation step, and the entity's own BLAS over them.
struct Animation
{
	enum Kind
	{
		External,
		Skeleton,
	};
	Kind kind;
	DeviceBuffer vertices, displacement, staging, matrices;
	Blas blas;
	bool blasStale = true;
};

struct Entity
{
	Entity(std::shared_ptr<Mesh> m, Scene* owner) : mesh(std::move(m)), scene(owner) {}
	void setTransform(const Mat3x4f& tf);
	std::optional<Mat3x4f> previousFrameTransform() const;
	// Replaces the entity's vertices with `vertices` (object space, one per mesh vertex).
	void applyExternalAnimation(const Vec3f* vertices, int count);
	// Skins the mesh's rest vertices with `pose[b] * restpose[b]` per bone (src/gpu/sceneKernels.cu:18-37).
	void setPoseAndAnimate(const Mat3x4f* pose, int bones);
	// Per-vertex displacement of the last animation step, if that step happened at the scene's previous time
	// (src/scene/Entity.cpp:103-117); nullptr otherwise.
	const Vec3f* displacementSincePrevFrame() const;
	// Vertices and BLAS traced for this entity: the animated ones if animated, else the mesh's.
	const Vec3f* tracedVertices() const;
	Blas& tracedBlas(hipStream_t stream);

	std::shared_ptr<Mesh> mesh;
	Scene* scene;
	std::shared_ptr<Texture> texture;
	Mat3x4f transform = Mat3x4f::identity();
	std::optional<uint64_t> transformTime;       // scene time when the transform was set
	Mat3x4f formerTransform = Mat3x4f::identity();
	std::optional<uint64_t> formerTransformTime;
	bool formerTransformValid = false;
	int32_t id = RGL_DEFAULT_ENTITY_ID;
	float laserRetro = 0.0f;
	uint64_t order = 0;                          // creation order: instance index order in the TLAS
	std::unique_ptr<Animation> animation;
	std::optional<uint64_t> animationTime, formerAnimationTime;

private:
	Animation& animationOf(Animation::Kind kind);
	void animated();
};

// How the TLAS follows entity motion when no entity was added or removed.
enum class TlasUpdate
{
	Rebuild,
	Refit,
};

class Scene
{
public:
	Scene();
	~Scene();
	Scene(const Scene&) = delete;
	Scene& operator=(const Scene&) = delete;
	// The implicit default scene (rgl_scene_t == NULL).
	static Scene& instance();
	static std::shared_ptr<Scene> defaultShared();
	// Marks every live scene stale (e.g. a mesh shared by several scenes changed).
	static void markAllDirty(bool topologyChanged);
	// Node format used by traversal in every scene (BvhFormat); changing it marks every scene stale.
	static int bvhFormat();
	static void setBvhFormat(int format);

	void addEntity(const std::shared_ptr<Entity>& e);
	void removeEntity(const std::shared_ptr<Entity>& e);
	void clear();
	void markDirty(bool topologyChanged);

	void setTime(uint64_t ns)
	{
		prevTime = time;
		time = ns;
	}
	std::optional<uint64_t> getTime() const { return time; }
	std::optional<uint64_t> getPrevTime() const { return prevTime; }
	// Seconds since the previous scene time, 0 if unknown.
	float deltaTimeSeconds() const;

	// Builds whatever is stale on the scene stream and makes `graphStream` wait for it.
	SceneView prepare(hipStream_t graphStream);
	// Records that work enqueued on `s` so far reads the scene's device data; the next rebuild overwrites that data
	// only after this work.
	void noteUse(hipStream_t s);
	// Same, with an event the caller has recorded on `s` after that work (a batch's completion event), so no event of
	// the scene's own is recorded.
	void noteUse(hipStream_t s, std::shared_ptr<Event> done);
	// Traversal-stack overflow flag written by the raytrace kernel (never expected to be set).
	int32_t* overflowFlag();
	size_t entityCount() const { return entities.size(); }
	const std::vector<std::shared_ptr<Entity>>& getEntities() const { return entities; }
	// Distinct meshes of the scene's entities.
	std::vector<Mesh*> meshes() const;
	// Device bytes of the scene's acceleration structures (BLASes of its meshes and animated entities plus TLAS),
	// built or estimated.
	int64_t accelBytes(int format) const;
	// Frees the BLASes of the scene's animated entities (rebuilt on next use).
	void releaseEntityAccel();
	std::vector<int32_t> instanceTriCounts() const;

	TlasUpdate tlasUpdate = TlasUpdate::Rebuild;
	int lastTlasDepth = 0;

private:
	void rebuild(hipStream_t s);

	std::vector<std::shared_ptr<Entity>> entities;
	std::optional<uint64_t> time, prevTime;
	bool dirty = true;
	bool topologyDirty = true;
	int maxBlasDepth = 0;
	std::unique_ptr<Stream> stream;
	std::unique_ptr<Event> ready;
	// Two sets of pinned staging arrays for the instance uploads, used in turn; `staged[k]` records the end of the
	// upload from set k, so a rebuild waits on the host only while the one before the last still reads its set.
	PinnedBuffer hostInst[2], hostShade[2];
	std::unique_ptr<Event> staged[2];
	int stagingSet = 0;
	DeviceBuffer dInst, dShade, dBoxes, tlasNodes, tlasLeafPrim, tlasParent, tlasCounters, tlasHeight, tlasScratch;
	DeviceBuffer dOverflow;
	DeviceBuffer tlasWide, tlasWideScratch, tlasWideCount; // 4-wide TLAS (when the format is 4-wide)
	std::map<hipStream_t, std::shared_ptr<Event>> uses;
	uint64_t readyGeneration = 0;
	hipStream_t waitedStream = nullptr;
	uint64_t waitedGeneration = 0;
	SceneView view{};
	uint64_t nextOrder = 0;
	friend struct Entity;
};

} // namespace rgl
