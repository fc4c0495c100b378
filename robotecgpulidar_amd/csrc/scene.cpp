#include "scene.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <set>

namespace rgl {

// ------------------------------------------------------------------------------------------------ Mesh
Mesh::Mesh(const Vec3f* v, int vc, const Vec3i* idx, int ic) : vertexCount(vc), triCount(ic)
{
	for (int i = 0; i < ic; ++i) {
		const Vec3i t = idx[i];
		if (t.x < 0 || t.y < 0 || t.z < 0 || t.x >= vc || t.y >= vc || t.z >= vc) {
			throw std::invalid_argument("mesh index out of range of the vertex array");
		}
	}
	RGL_HIP(hipMemcpy(vertices.ensure(sizeof(Vec3f) * vc), v, sizeof(Vec3f) * vc, hipMemcpyHostToDevice));
	RGL_HIP(hipMemcpy(indices.ensure(sizeof(Vec3i) * ic), idx, sizeof(Vec3i) * ic, hipMemcpyHostToDevice));
}

void Mesh::setTexCoords(const Vec2f* uv, int count)
{
	if (count != vertexCount) {
		throw std::invalid_argument("Cannot set texture coordinates because vertex count do not match with texture "
		                            "coordinates count: vertexCount=" +
		                            std::to_string(vertexCount) + ", textureCoords=" + std::to_string(count));
	}
	RGL_HIP(hipMemcpy(uvs.ensure(sizeof(Vec2f) * count), uv, sizeof(Vec2f) * count, hipMemcpyHostToDevice));
	hasUVs = true;
}

void Mesh::setBoneWeights(const BoneWeights* w, int count)
{
	if (count != vertexCount) {
		throw std::invalid_argument("Cannot set bone weights because vertex count do not match with bone weights count: "
		                            "vertexCount=" +
		                            std::to_string(vertexCount) + ", boneWeightsCount=" + std::to_string(count));
	}
	int maxBone = -1;
	for (int i = 0; i < count; ++i) {
		for (int k = 0; k < 4; ++k) {
			if (w[i].bones[k] < 0) {
				throw std::invalid_argument("bone index must not be negative (vertex " + std::to_string(i) + ")");
			}
			maxBone = std::max(maxBone, w[i].bones[k]);
		}
	}
	RGL_HIP(hipMemcpy(boneWeights.ensure(sizeof(BoneWeights) * count), w, sizeof(BoneWeights) * count,
	                  hipMemcpyHostToDevice));
	hasBoneWeights = true;
	maxBoneIndex = maxBone;
}

void Mesh::setRestposes(const Mat3x4f* r, int count) { restposes.assign(r, r + count); }

void Mesh::ensureBlas(hipStream_t s)
{
	if (!blas.built) {
		blas.build(s, triCount, vertices.as<Vec3f>(), indices.as<Vec3i>(), false);
	}
}

const void* Mesh::nodesFor(hipStream_t s, int format)
{
	ensureBlas(s);
	return blas.nodesFor(s, format);
}

int64_t Mesh::accelBytes(int format) const { return blas.built ? blas.bytes() : estimatedBlasBytes(triCount, format); }

void Mesh::releaseAccel() { blas.release(); }

// ------------------------------------------------------------------------------------------------ Blas
namespace {
size_t nodeBytes(int format)
{
	switch (format) {
		case kBvh4F: return sizeof(Bvh4FNode);
		case kBvh4Q: return sizeof(Bvh4QNode);
		case kBvh4H: return sizeof(Bvh4HNode);
		default: return sizeof(BvhNode);
	}
}
} // namespace

int64_t estimatedBlasBytes(int64_t t, int format)
{
	// Binary nodes 64 B, leaf triangles 48 B, leaf order 4 B, parent links 8 B, fit counters and heights 8 B,
	// primitive boxes 24 B per triangle; at most one 4-wide node per binary internal node (an upper bound: about half
	// of them are kept in practice).
	const int64_t wideBytes = format == kBvh2 ? 0 : static_cast<int64_t>(nodeBytes(format)) * t;
	return t * (64 + 48 + 4 + 8 + 8 + 24) + wideBytes;
}

void Blas::build(hipStream_t s, int n, const Vec3f* v, const Vec3i* idx, bool refittable)
{
	release();
	triCount = n;
	keepScratch = refittable;
	const int inner = std::max(n - 1, 1);
	boxes.ensure(sizeof(Aabb) * n);
	nodes.ensure(sizeof(BvhNode) * inner);
	tris.ensure(sizeof(LeafTri) * n);
	leafPrim.ensure(sizeof(int32_t) * n);
	parentSlot.ensure(sizeof(int32_t) * 2 * n);
	counters.ensure(sizeof(int32_t) * inner);
	height.ensure(sizeof(int32_t) * inner);
	scratch.ensure(gpu::lbvhScratchBytes(n));
	gpu::triangleBoxes(s, n, v, idx, boxes.as<Aabb>());
	gpu::buildLbvh(s, n, boxes.as<Aabb>(), false, nodes.as<BvhNode>(), leafPrim.as<int32_t>(), parentSlot.as<int32_t>(),
	               counters.as<int32_t>(), height.as<int32_t>(), scratch.get());
	gpu::gatherTris(s, n, leafPrim.as<int32_t>(), v, idx, tris.as<LeafTri>());
	int32_t d = 0;
	RGL_HIP(hipMemcpyAsync(&d, height.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
	RGL_HIP(hipStreamSynchronize(s));
	RGL_HIP(hipGetLastError());
	depth = d;
	scratch.release();
	built = true;
}

void Blas::refit(hipStream_t s, const Vec3f* v, const Vec3i* idx)
{
	const int n = triCount;
	gpu::triangleBoxes(s, n, v, idx, boxes.as<Aabb>());
	gpu::refitLbvh(s, n, boxes.as<Aabb>(), leafPrim.as<int32_t>(), parentSlot.as<int32_t>(), nodes.as<BvhNode>(),
	               counters.as<int32_t>(), height.as<int32_t>());
	gpu::gatherTris(s, n, leafPrim.as<int32_t>(), v, idx, tris.as<LeafTri>());
	for (int f = kBvh4F; f < kBvhFormats; ++f) {
		if (wideBuilt[f]) {
			gpu::bvh4Collapse(s, n, nodes.as<BvhNode>(), wideRanks.get(), f, wide[f].get());
		}
	}
	RGL_HIP(hipGetLastError());
}

const void* Blas::nodesFor(hipStream_t s, int format)
{
	if (format == kBvh2) {
		return nodes.get();
	}
	if (format != kBvh4F && format != kBvh4Q && format != kBvh4H) {
		throw std::invalid_argument("unknown BVH node format " + std::to_string(format));
	}
	if (!wideBuilt[format]) {
		const int n = triCount;
		if (wideRanks.get() == nullptr) {
			DeviceBuffer countDev;
			wideRanks.ensure(gpu::bvh4ScratchBytes(n));
			countDev.ensure(sizeof(int32_t));
			gpu::bvh4Ranks(s, n, parentSlot.as<int32_t>(), wideRanks.get(), countDev.as<int32_t>());
			RGL_HIP(hipMemcpyAsync(&wideCount, countDev.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
			RGL_HIP(hipStreamSynchronize(s));
		}
		const size_t m = static_cast<size_t>(wideCount);
		gpu::bvh4Collapse(s, n, nodes.as<BvhNode>(), wideRanks.get(), format, wide[format].ensure(nodeBytes(format) * m));
		RGL_HIP(hipGetLastError());
		RGL_HIP(hipStreamSynchronize(s));
		wideBuilt[format] = true;
		if (!keepScratch) {
			wideRanks.release();
		}
	}
	return wide[format].get();
}

int64_t Blas::bytes() const
{
	int64_t b = 0;
	for (const DeviceBuffer* d : {&nodes, &tris, &leafPrim, &parentSlot, &counters, &height, &boxes, &scratch, &wideRanks}) {
		b += static_cast<int64_t>(d->capacity());
	}
	for (const DeviceBuffer& w : wide) {
		b += static_cast<int64_t>(w.capacity());
	}
	return b;
}

void Blas::release()
{
	for (DeviceBuffer* d : {&nodes, &tris, &leafPrim, &parentSlot, &counters, &height, &boxes, &scratch, &wideRanks}) {
		d->release();
	}
	for (int f = 0; f < kBvhFormats; ++f) {
		wide[f].release();
		wideBuilt[f] = false;
	}
	built = false;
	depth = 0;
	wideCount = 0;
}

// ------------------------------------------------------------------------------------------------ Texture
Texture::Texture(const void* texels_, int w, int h) : width(w), height(h)
{
	const size_t bytes = static_cast<size_t>(w) * h;
	RGL_HIP(hipMemcpy(texels.ensure(bytes), texels_, bytes, hipMemcpyHostToDevice));
}

// ------------------------------------------------------------------------------------------------ Entity
void Entity::setTransform(const Mat3x4f& tf)
{
	formerTransform = transform;
	formerTransformTime = transformTime;
	formerTransformValid = true;
	transform = tf;
	transformTime = scene->getTime();
	scene->markDirty(false);
}

// src/scene/Entity.cpp:66-86
std::optional<Mat3x4f> Entity::previousFrameTransform() const
{
	const bool setOnlyOnce = transformTime.has_value() && !formerTransformTime.has_value();
	if (setOnlyOnce) {
		return transform;
	}
	const bool setInPrecedingFrame = formerTransformTime.has_value() &&
	                                 formerTransformTime == scene->getPrevTime();
	if (!setInPrecedingFrame) {
		return std::nullopt;
	}
	return formerTransform;
}

Animation& Entity::animationOf(Animation::Kind kind)
{
	// A different kind of animation starts again from the mesh's own vertices (src/scene/Entity.cpp:88-101).
	if (!animation || animation->kind != kind) {
		auto a = std::make_unique<Animation>();
		a->kind = kind;
		const size_t bytes = sizeof(Vec3f) * static_cast<size_t>(mesh->vertexCount);
		RGL_HIP(hipMemcpy(a->vertices.ensure(bytes), mesh->vertices.get(), bytes, hipMemcpyDeviceToDevice));
		RGL_HIP(hipMemset(a->displacement.ensure(bytes), 0, bytes));
		animation = std::move(a);
	}
	return *animation;
}

void Entity::animated()
{
	formerAnimationTime = animationTime;
	animationTime = scene->getTime();
	animation->blasStale = true;
	scene->markDirty(false);
}

// src/scene/animator/ExternalAnimator.cpp:23-40
void Entity::applyExternalAnimation(const Vec3f* v, int count)
{
	if (count != mesh->vertexCount) {
		throw std::invalid_argument("Invalid argument: cannot apply external animation because vertex counts do not match: "
		                            "old=" +
		                            std::to_string(mesh->vertexCount) + ", new=" + std::to_string(count));
	}
	Animation& a = animationOf(Animation::External);
	const size_t bytes = sizeof(Vec3f) * static_cast<size_t>(count);
	RGL_HIP(hipMemcpy(a.staging.ensure(bytes), v, bytes, hipMemcpyHostToDevice));
	gpu::updateVertices(nullptr, count, a.staging.as<Vec3f>(), a.vertices.as<Vec3f>(), a.displacement.as<Vec3f>());
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipStreamSynchronize(nullptr));
	animated();
}

// src/scene/animator/SkeletonAnimator.cpp:19-58
void Entity::setPoseAndAnimate(const Mat3x4f* pose, int bones)
{
	if (!mesh->hasBoneWeights) {
		throw std::invalid_argument("Cannot create SkeletonAnimator because mesh has no bone weights defined.");
	}
	if (mesh->restposes.empty()) {
		throw std::invalid_argument("Cannot create SkeletonAnimator because mesh has no restposes defined.");
	}
	if (static_cast<size_t>(bones) != mesh->restposes.size()) {
		throw std::invalid_argument("Cannot perform skeleton animation because bones count do not match restposes count: "
		                            "bones=" +
		                            std::to_string(bones) + ", restposes=" + std::to_string(mesh->restposes.size()));
	}
	if (mesh->maxBoneIndex >= bones) {
		throw std::invalid_argument("bone weights refer to bone " + std::to_string(mesh->maxBoneIndex) + " of " +
		                            std::to_string(bones));
	}
	Animation& a = animationOf(Animation::Skeleton);
	std::vector<Mat3x4f> m(static_cast<size_t>(bones));
	for (int b = 0; b < bones; ++b) {
		m[b] = pose[b] * mesh->restposes[b];
	}
	const size_t mb = sizeof(Mat3x4f) * m.size();
	RGL_HIP(hipMemcpy(a.matrices.ensure(mb), m.data(), mb, hipMemcpyHostToDevice));
	gpu::skinVertices(nullptr, mesh->vertexCount, mesh->vertices.as<Vec3f>(), mesh->boneWeights.as<BoneWeights>(),
	                  a.matrices.as<Mat3x4f>(), a.vertices.as<Vec3f>(), a.displacement.as<Vec3f>());
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipStreamSynchronize(nullptr));
	animated();
}

const Vec3f* Entity::displacementSincePrevFrame() const
{
	if (!animation || !formerAnimationTime.has_value() || formerAnimationTime != scene->getPrevTime()) {
		return nullptr;
	}
	return animation->displacement.as<Vec3f>();
}

const Vec3f* Entity::tracedVertices() const
{
	return animation ? animation->vertices.as<Vec3f>() : mesh->vertices.as<Vec3f>();
}

Blas& Entity::tracedBlas(hipStream_t s)
{
	if (!animation) {
		mesh->ensureBlas(s);
		return mesh->blas;
	}
	Animation& a = *animation;
	if (!a.blas.built) {
		a.blas.build(s, mesh->triCount, a.vertices.as<Vec3f>(), mesh->indices.as<Vec3i>(), true);
	} else if (a.blasStale) {
		a.blas.refit(s, a.vertices.as<Vec3f>(), mesh->indices.as<Vec3i>());
	}
	a.blasStale = false;
	return a.blas;
}

// ------------------------------------------------------------------------------------------------ Scene
namespace {
std::set<Scene*>& liveScenes()
{
	static std::set<Scene*> s;
	return s;
}
} // namespace

Scene& Scene::instance()
{
	static Scene scene;
	return scene;
}

std::shared_ptr<Scene> Scene::defaultShared()
{
	static std::shared_ptr<Scene> p(&instance(), [](Scene*) {});
	return p;
}

namespace {
int& formatSetting()
{
	static int f = kBvh4F; // fastest on gfx950 for the lidar workloads measured (latency-bound traversal)
	return f;
}
} // namespace

int Scene::bvhFormat() { return formatSetting(); }

void Scene::setBvhFormat(int format)
{
	if (format < 0 || format >= kBvhFormats) {
		throw std::invalid_argument("unknown BVH node format " + std::to_string(format));
	}
	if (format != formatSetting()) {
		formatSetting() = format;
		markAllDirty(true);
	}
}

void Scene::markAllDirty(bool topologyChanged)
{
	for (Scene* s : liveScenes()) {
		s->markDirty(topologyChanged);
	}
}

Scene::Scene() { liveScenes().insert(this); }

Scene::~Scene() { liveScenes().erase(this); }

void Scene::addEntity(const std::shared_ptr<Entity>& e)
{
	e->order = nextOrder++;
	entities.push_back(e);
	markDirty(true);
}

void Scene::removeEntity(const std::shared_ptr<Entity>& e)
{
	entities.erase(std::remove(entities.begin(), entities.end(), e), entities.end());
	markDirty(true);
}

void Scene::clear()
{
	entities.clear();
	time.reset();
	prevTime.reset();
	markDirty(true);
}

void Scene::markDirty(bool topologyChanged)
{
	dirty = true;
	topologyDirty = topologyDirty || topologyChanged;
}

float Scene::deltaTimeSeconds() const
{
	if (!prevTime.has_value() || !time.has_value()) {
		return 0.0f;
	}
	const double dt = (static_cast<double>(*time) - static_cast<double>(*prevTime)) * 1e-9;
	return static_cast<float>(dt);
}

void Scene::noteUse(hipStream_t s)
{
	auto& ev = uses[s];
	if (!ev || ev.use_count() > 1) { // never re-record an event that is shared with its recorder
		ev = std::make_shared<Event>();
	}
	ev->record(s);
}

void Scene::noteUse(hipStream_t s, std::shared_ptr<Event> done)
{
	uses[s] = std::move(done);
}

int32_t* Scene::overflowFlag()
{
	if (dOverflow.get() == nullptr) {
		dOverflow.ensure(sizeof(int32_t));
		RGL_HIP(hipMemset(dOverflow.get(), 0, sizeof(int32_t)));
	}
	return dOverflow.as<int32_t>();
}

std::vector<Mesh*> Scene::meshes() const
{
	std::vector<Mesh*> out;
	for (auto& e : entities) {
		if (std::find(out.begin(), out.end(), e->mesh.get()) == out.end()) {
			out.push_back(e->mesh.get());
		}
	}
	return out;
}

int64_t Scene::accelBytes(int format) const
{
	int64_t b = 0;
	for (Mesh* m : meshes()) {
		b += m->accelBytes(format);
	}
	for (auto& e : entities) {
		if (e->animation) {
			const Blas& a = e->animation->blas;
			b += a.built ? a.bytes() : estimatedBlasBytes(e->mesh->triCount, format);
		}
	}
	const int64_t n = static_cast<int64_t>(entities.size());
	return b + n * (64 + 160 + 24 + 64 + 4 + 8 + 8) + n * 128;
}

void Scene::releaseEntityAccel()
{
	for (auto& e : entities) {
		if (e->animation) {
			e->animation->blas.release();
		}
	}
	markDirty(false);
}

std::vector<int32_t> Scene::instanceTriCounts() const
{
	std::vector<int32_t> c;
	for (auto& e : entities) {
		c.push_back(e->mesh->triCount);
	}
	return c;
}

void Scene::rebuild(hipStream_t s)
{
	// Raytrace launches still in flight read the device arrays rewritten below: the scene stream waits for them.
	for (auto& [user, ev] : uses) {
		RGL_HIP(hipStreamWaitEvent(s, ev->get(), 0));
	}
	const int n = static_cast<int>(entities.size());
	const int fmt = bvhFormat();
	maxBlasDepth = 0;
	for (auto& e : entities) {
		maxBlasDepth = std::max(maxBlasDepth, e->tracedBlas(s).depth);
	}
	if (n == 0) {
		view = SceneView{};
		view.format = fmt;
		view.overflow = overflowFlag();
		return;
	}
	const int set = stagingSet;
	stagingSet ^= 1;
	if (!staged[set]) {
		staged[set] = std::make_unique<Event>();
	} else if (staged[set]->wasRecorded()) {
		staged[set]->sync(); // the upload before the last one still reads this set
	}
	auto* inst = static_cast<InstanceRec*>(hostInst[set].ensure(sizeof(InstanceRec) * n));
	auto* shade = static_cast<InstanceShade*>(hostShade[set].ensure(sizeof(InstanceShade) * n));
	for (int i = 0; i < n; ++i) {
		const Entity& e = *entities[i];
		Blas& blas = entities[i]->tracedBlas(s);
		inst[i].worldToObject64 = inverse64(widen(e.transform));
		inst[i].worldToObject = round32(inst[i].worldToObject64);
		inst[i].blasNodes = blas.nodesFor(s, fmt);
		inst[i].blasTris = blas.tris.as<LeafTri>();
		InstanceShade& sh = shade[i];
		std::memset(&sh, 0, sizeof(sh));
		sh.objectToWorld = e.transform;
		const auto prev = e.previousFrameTransform();
		sh.prevObjectToWorld = prev.value_or(Mat3x4f::identity());
		sh.hasPrev = prev.has_value() ? 1 : 0;
		sh.vertices = e.tracedVertices();
		sh.displacement = e.displacementSincePrevFrame();
		sh.indices = e.mesh->indices.as<Vec3i>();
		sh.uvs = e.mesh->hasUVs ? e.mesh->uvs.as<Vec2f>() : nullptr;
		sh.texels = e.texture ? e.texture->texels.as<uint8_t>() : nullptr;
		sh.blas2 = blas.nodes.as<BvhNode>();
		sh.texWidth = e.texture ? e.texture->width : 0;
		sh.texHeight = e.texture ? e.texture->height : 0;
		sh.entityId = e.id;
		sh.laserRetro = e.laserRetro;
	}
	const bool refit = !topologyDirty && tlasUpdate == TlasUpdate::Refit && tlasNodes.get() != nullptr;
	RGL_HIP(hipMemcpyAsync(dInst.ensure(sizeof(InstanceRec) * n), inst, sizeof(InstanceRec) * n, hipMemcpyHostToDevice, s));
	RGL_HIP(hipMemcpyAsync(dShade.ensure(sizeof(InstanceShade) * n), shade, sizeof(InstanceShade) * n,
	                       hipMemcpyHostToDevice, s));
	staged[set]->record(s);
	const int inner = std::max(n - 1, 1);
	dBoxes.ensure(sizeof(Aabb) * n);
	gpu::instanceBoxes(s, n, dShade.as<InstanceShade>(), dInst.as<InstanceRec>(), dBoxes.as<Aabb>());
	if (refit) {
		gpu::refitLbvh(s, n, dBoxes.as<Aabb>(), tlasLeafPrim.as<int32_t>(), tlasParent.as<int32_t>(),
		               tlasNodes.as<BvhNode>(), tlasCounters.as<int32_t>(), tlasHeight.as<int32_t>());
	} else {
		tlasNodes.ensure(sizeof(BvhNode) * inner);
		tlasLeafPrim.ensure(sizeof(int32_t) * n);
		tlasParent.ensure(sizeof(int32_t) * 2 * n);
		tlasCounters.ensure(sizeof(int32_t) * inner);
		tlasHeight.ensure(sizeof(int32_t) * inner);
		tlasScratch.ensure(gpu::lbvhScratchBytes(n));
		gpu::buildLbvh(s, n, dBoxes.as<Aabb>(), true, tlasNodes.as<BvhNode>(), tlasLeafPrim.as<int32_t>(),
		               tlasParent.as<int32_t>(), tlasCounters.as<int32_t>(), tlasHeight.as<int32_t>(), tlasScratch.get());
	}
	view.tlasNodes = tlasNodes.get();
	view.tlas2 = tlasNodes.as<BvhNode>();
	if (fmt != kBvh2) {
		// Upper bound on the 4-wide node count: the binary tree's internal node count.
		tlasWideScratch.ensure(gpu::bvh4ScratchBytes(n));
		gpu::bvh4Ranks(s, n, tlasParent.as<int32_t>(), tlasWideScratch.get(), tlasWideCount.ensureT<int32_t>(1));
		void* wideNodes = tlasWide.ensure(nodeBytes(fmt) * inner);
		gpu::bvh4Collapse(s, n, tlasNodes.as<BvhNode>(), tlasWideScratch.get(), fmt, wideNodes);
		view.tlasNodes = wideNodes;
	}
	RGL_HIP(hipGetLastError());
	view.instances = dInst.as<InstanceRec>();
	view.shades = dShade.as<InstanceShade>();
	view.instanceCount = n;
	view.format = fmt;
	view.overflow = overflowFlag();
}

SceneView Scene::prepare(hipStream_t graphStream)
{
	if (!stream) {
		stream = std::make_unique<Stream>();
		ready = std::make_unique<Event>();
	}
	if (dirty) {
		rebuild(stream->get());
		ready->record(stream->get());
		++readyGeneration;
		dirty = false;
		topologyDirty = false;
	}
	// A stream that already waited for this build needs no second wait.
	if (ready->wasRecorded() && !(waitedStream == graphStream && waitedGeneration == readyGeneration)) {
		RGL_HIP(hipStreamWaitEvent(graphStream, ready->get(), 0));
		waitedStream = graphStream;
		waitedGeneration = readyGeneration;
	}
	if (view.overflow == nullptr) {
		view.overflow = overflowFlag();
	}
	return view;
}

} // namespace rgl
