#include "graph.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <functional>

#include "kernels.hpp"

namespace rgl {

// ------------------------------------------------------------------------------------------------ registry
namespace {
std::map<const void*, Node::Ptr>& nodeMap()
{
	static std::map<const void*, Node::Ptr> m;
	return m;
}
} // namespace

void registerNode(const Node::Ptr& n) { nodeMap()[static_cast<const void*>(n.get())] = n; }

Node::Ptr lookupNode(const void* handle)
{
	auto it = nodeMap().find(handle);
	if (it == nodeMap().end()) {
		char buf[64];
		std::snprintf(buf, sizeof(buf), "%p", handle);
		throw InvalidAPIObject(std::string("RGL API Error: Object does not exist: Node ") + buf);
	}
	return it->second;
}

bool nodeAlive(const void* handle) { return nodeMap().count(handle) != 0; }

void releaseNode(const Node* n) { nodeMap().erase(static_cast<const void*>(n)); }

std::vector<Node::Ptr> allNodes()
{
	std::vector<Node::Ptr> v;
	for (auto& [k, n] : nodeMap()) {
		v.push_back(n);
	}
	return v;
}

// ------------------------------------------------------------------------------------------------ Node
// src/graph/Node.cpp:23-60
void Node::addChild(const Ptr& child)
{
	if (child == nullptr) {
		throw InvalidPipeline("attempted to set an empty child for " + getName());
	}
	if (std::find(outputs.begin(), outputs.end(), child) != outputs.end()) {
		throw InvalidPipeline("attempted to add child " + child->getName() + " to parent " + getName() + " twice");
	}
	if (hasCtx()) {
		getCtx()->detachAndDestroy();
	}
	if (child->hasCtx()) {
		child->getCtx()->detachAndDestroy();
	}
	outputs.push_back(child);
	child->inputs.push_back(shared_from_this());
	++structureEpoch();
	std::stable_sort(outputs.begin(), outputs.end(), [](const Ptr& a, const Ptr& b) { return a->priority > b->priority; });
	const int32_t maxChild = outputs.front()->priority;
	if (maxChild > priority) {
		setPriority(maxChild);
	}
	dirty = true;
	child->dirty = true;
}

// src/graph/Node.cpp:62-100
void Node::removeChild(const Ptr& child)
{
	if (child == nullptr) {
		throw InvalidPipeline("attempted to remove an empty child for " + getName());
	}
	auto it = std::find(outputs.begin(), outputs.end(), child);
	if (it == outputs.end()) {
		throw InvalidPipeline("attempted to remove child " + child->getName() + " from " + getName() +
		                      ", but it was not found");
	}
	auto self = shared_from_this();
	auto pit = std::find(child->inputs.begin(), child->inputs.end(), self);
	if (pit == child->inputs.end()) {
		throw InvalidPipeline("attempted to remove parent " + getName() + " from " + child->getName() +
		                      ", but it was not found");
	}
	if (hasCtx()) {
		getCtx()->detachAndDestroy();
	}
	outputs.erase(it);
	child->inputs.erase(pit);
	dirty = true;
	child->dirty = true;
	++structureEpoch();
}

// src/graph/Node.cpp:186-214
void Node::setPriority(int32_t requested)
{
	if (requested == priority) {
		return;
	}
	if (!outputs.empty() && outputs.front()->priority > requested) {
		throw InvalidPipeline("cannot set priority of node " + getName() + " to " + std::to_string(requested) +
		                      ", because it's children " + outputs.front()->getName() + " has higher priority " +
		                      std::to_string(outputs.front()->priority));
	}
	priority = requested;
	for (auto& in : inputs) {
		std::stable_sort(in->outputs.begin(), in->outputs.end(),
		                 [](const Ptr& a, const Ptr& b) { return a->priority > b->priority; });
		if (in->priority < requested) {
			in->setPriority(requested);
		}
	}
	if (hasCtx()) {
		ctx->clearExecutionOrder();
	}
}

std::set<Node::Ptr> Node::connectedComponent()
{
	if (hasCtx()) {
		return ctx->getNodes();
	}
	std::set<Ptr> visited;
	std::function<void(const Ptr&)> dfs = [&](const Ptr& cur) {
		visited.insert(cur);
		for (auto& o : cur->outputs) {
			if (!visited.count(o)) {
				dfs(o);
			}
		}
		for (auto& i : cur->inputs) {
			if (!visited.count(i)) {
				dfs(i);
			}
		}
	};
	dfs(shared_from_this());
	return visited;
}

std::set<Node::Ptr> Node::disconnectConnectedNodes()
{
	auto nodes = connectedComponent();
	for (auto& n : nodes) {
		n->inputs.clear();
		n->outputs.clear();
		n->setCtx(nullptr);
	}
	return nodes;
}

void Node::setCtx(std::shared_ptr<GraphRunCtx> c)
{
	ctx = std::move(c);
	dirty = true;
}

hipStream_t Node::stream() const
{
	if (!ctx) {
		throw std::logic_error(getName() + ": no graph context");
	}
	return ctx->stream();
}

void Node::validate()
{
	if (isValid()) {
		return;
	}
	if (!hasCtx()) {
		throw std::logic_error(getName() + ": attempted to call validate() despite !hasGraphRunCtx()");
	}
	validateImpl();
	dirty = false;
	++structureEpoch();
}

void Node::enqueue()
{
	if (!isValid()) {
		throw std::logic_error(getName() + ": attempted to call enqueueExec() despite !isValid()");
	}
	enqueueImpl();
	if (!done || done.use_count() > 1) {
		done = std::make_shared<Event>();
	}
	done->record(stream());
	ready = nullptr;
	counted = nullptr;
	ranOnce = true;
}

void Node::pointOp(PointOpItem&, FieldCopyList&)
{
	throw std::logic_error(getName() + " has no elementwise stage");
}

void Node::markEnqueued(const std::shared_ptr<Event>& ev, const std::shared_ptr<Event>& readyAt,
                        const std::shared_ptr<Event>& countedAt)
{
	done = ev;
	ready = readyAt;
	counted = countedAt;
	ranOnce = true;
}

void Node::waitForResultsIfRun()
{
	if (ranOnce && done) {
		done->sync();
	}
}

void Node::waitForResults()
{
	if (!isValid() || !ranOnce) {
		throw InvalidPipeline("Cannot get results from " + getName() + "; it hasn't been run yet, or the run has failed");
	}
	if (ready) {
		ready->sync();
	} else if (done) {
		done->sync();
	}
}

void Node::waitForCounts()
{
	if (counted) {
		if (!isValid() || !ranOnce) {
			throw InvalidPipeline("Cannot get results from " + getName() + "; it hasn't been run yet, or the run has failed");
		}
		counted->sync();
		return;
	}
	waitForResults();
}

// ------------------------------------------------------------------------------------------------ GraphRunCtx
std::vector<std::shared_ptr<GraphRunCtx>>& GraphRunCtx::instances()
{
	static std::vector<std::shared_ptr<GraphRunCtx>> v;
	return v;
}

std::shared_ptr<GraphRunCtx> GraphRunCtx::createAndAttach(const Node::Ptr& node)
{
	auto ctx = std::shared_ptr<GraphRunCtx>(new GraphRunCtx());
	ctx->nodes = node->connectedComponent();
	for (auto& n : ctx->nodes) {
		if (n->hasCtx()) {
			throw std::logic_error("attempted to replace existing graphRunCtx in node " + n->getName());
		}
		n->setCtx(ctx);
	}
	instances().push_back(ctx);
	return ctx;
}

// src/graph/GraphRunCtx.cpp:107-142: reversed post-order of a DFS from the entry nodes in ascending priority.
std::vector<Node::Ptr> GraphRunCtx::findExecutionOrder(std::set<Node::Ptr> nodes)
{
	std::vector<Node::Ptr> reverseOrder;
	std::function<void(const Node::Ptr&)> dfs = [&](const Node::Ptr& cur) {
		nodes.erase(cur);
		const auto& outs = cur->getOutputs();
		for (auto it = outs.rbegin(); it != outs.rend(); ++it) {
			if (nodes.count(*it)) {
				dfs(*it);
			}
		}
		reverseOrder.push_back(cur);
	};
	std::vector<Node::Ptr> entries;
	for (auto& n : nodes) {
		if (n->getInputs().empty()) {
			entries.push_back(n);
		}
	}
	std::stable_sort(entries.begin(), entries.end(),
	                 [](const Node::Ptr& a, const Node::Ptr& b) { return a->getPriority() < b->getPriority(); });
	for (auto& e : entries) {
		if (nodes.count(e)) {
			dfs(e);
		}
	}
	if (!nodes.empty()) {
		throw std::logic_error("bug: detected unreachable nodes");
	}
	return {reverseOrder.rbegin(), reverseOrder.rend()};
}

const std::vector<Node::Ptr>& GraphRunCtx::validatedOrder()
{
	if (order.empty()) {
		order = findExecutionOrder(nodes);
	}
	for (auto& n : order) {
		n->validate();
	}
	return order;
}

void GraphRunCtx::executeAsync()
{
	validatedOrder();
	if (batchDone) {
		// The previous batch wrote this graph's buffers on the batch stream.
		RGL_HIP(hipStreamWaitEvent(stream_->get(), batchDone->get(), 0));
		batchDone.reset();
	}
	running = true;
	lastRunSolo = true;
	for (auto& n : order) {
		n->enqueue();
	}
}

void GraphRunCtx::beginBatch(hipStream_t batchStream, Event& scratch)
{
	if (running && lastRunSolo) {
		scratch.record(stream_->get());
		RGL_HIP(hipStreamWaitEvent(batchStream, scratch.get(), 0));
	}
	if (batchDone && batchDoneStream != batchStream) {
		// The previous batch ran on another batch stream.
		RGL_HIP(hipStreamWaitEvent(batchStream, batchDone->get(), 0));
	}
	streamOverride = batchStream;
	running = true;
	lastRunSolo = false;
}

void GraphRunCtx::endBatch(hipStream_t batchStream, std::shared_ptr<Event> completion)
{
	streamOverride = nullptr;
	if (completion) {
		batchDone = std::move(completion);
		batchDoneStream = batchStream;
	}
}

bool GraphRunCtx::inFlight() const
{
	if (!batchDone || !batchDone->wasRecorded()) {
		return false;
	}
	const hipError_t e = hipEventQuery(batchDone->get());
	if (e == hipErrorNotReady) {
		return true;
	}
	RGL_HIP(e);
	return false;
}

void GraphRunCtx::synchronize()
{
	if (running) {
		if (batchDone) {
			batchDone->sync();
		}
		stream_->sync();
		running = false;
	}
}

void GraphRunCtx::markNodesDirty()
{
	for (auto& n : nodes) {
		n->markDirty();
	}
}

void GraphRunCtx::detachAndDestroy()
{
	synchronize();
	auto self = shared_from_this();
	for (auto& n : nodes) {
		n->setCtx(nullptr);
	}
	auto& v = instances();
	v.erase(std::remove(v.begin(), v.end(), self), v.end());
}

void GraphRunCtx::synchronizeAll()
{
	for (auto& c : instances()) {
		c->synchronize();
	}
}

// ------------------------------------------------------------------------------------------------ interfaces
void IRaysNodeSingleInput::validateImpl() { input = exactlyOneInputOfType<IRaysNode>(); }

void IPointsNodeSingleInput::validateImpl()
{
	input = exactlyOneInputOfType<IPointsNode>();
	for (auto f : requiredFields()) {
		if (!isPadding(f) && !input->hasField(f)) {
			throw InvalidPipeline(getName() + " requires " + fieldName(f) + " to be present");
		}
	}
}

static void requireNoInputs(const Node& n, const std::vector<Node::Ptr>& inputs)
{
	if (!inputs.empty()) {
		throw InvalidPipeline("inputs for node " + n.getName() + " are not allowed");
	}
}

// ------------------------------------------------------------------------------------------------ rays nodes
// Rays are traced tile by tile of their directions once there are enough of them to fill several waves: lidar
// patterns list rays row by row (one elevation, many azimuths), and a wave of 64 consecutive rays then spans a thin
// strip of directions; the rays of a compact patch (a few channels x a few azimuths) descend the same BVH nodes
// longer. The order is computed once per update.
void FromMat3x4fRaysNode::setParameters(const Mat3x4f* r, int n)
{
	constexpr int kOrderMinRays = 4096;
	RGL_HIP(hipMemcpy(buf.ensure(sizeof(Mat3x4f) * n), r, sizeof(Mat3x4f) * n, hipMemcpyHostToDevice));
	count = n;
	ordered = n >= kOrderMinRays;
	if (ordered) {
		// One stream and one scratch buffer serve every rays node (API calls come from one thread); the scratch only
		// grows, so repeated updates allocate nothing. Both live until the process ends (never destroyed during
		// static destruction, after which the HIP runtime may be gone).
		static Stream* s = new Stream();
		static DeviceBuffer* scratch = new DeviceBuffer();
		gpu::directionOrder(s->get(), n, buf.as<Mat3x4f>(), order.ensureT<int32_t>(n),
		                    scratch->ensure(gpu::directionOrderScratchBytes(n)));
		RGL_HIP(hipGetLastError());
		s->sync();
	} else {
		order.release();
	}
}
void FromMat3x4fRaysNode::validateImpl() { requireNoInputs(*this, inputs); }

void SetRingIdsRaysNode::setParameters(const int32_t* ids, int n)
{
	RGL_HIP(hipMemcpy(buf.ensure(sizeof(int32_t) * n), ids, sizeof(int32_t) * n, hipMemcpyHostToDevice));
	count = n;
}
void SetRingIdsRaysNode::validateImpl()
{
	IRaysNodeSingleInput::validateImpl();
	if (input->rayCount() % count != 0) {
		throw InvalidPipeline("ring ids doesn't match number of rays. RayCount(" + std::to_string(input->rayCount()) +
		                      ") mod RingIdsCount(" + std::to_string(count) + ") should be zero");
	}
}

// src/graph/SetRangeRaysNode.cpp:17-47
void SetRangeRaysNode::setParameters(const Vec2f* r, int n)
{
	for (int i = 0; i < n; ++i) {
		if (std::isnan(r[i].x) || std::isnan(r[i].y)) {
			throw std::invalid_argument("range must be a number");
		}
		if (r[i].x > r[i].y) {
			throw std::invalid_argument("minimum range must be lower or equal to maximum range");
		}
		if (r[i].x < 0.0f) {
			throw std::invalid_argument("minimum range must not be negative");
		}
	}
	RGL_HIP(hipMemcpy(buf.ensure(sizeof(Vec2f) * n), r, sizeof(Vec2f) * n, hipMemcpyHostToDevice));
	count = n;
}
void SetRangeRaysNode::validateImpl()
{
	IRaysNodeSingleInput::validateImpl();
	if (!(count == 1 || count == input->rayCount())) {
		throw InvalidPipeline("ranges doesn't match number of rays. RangesCount(" + std::to_string(count) +
		                      ") should be one or equal to RayCount(" + std::to_string(input->rayCount()) + ")");
	}
}

void SetTimeOffsetsRaysNode::setParameters(const float* o, int n)
{
	RGL_HIP(hipMemcpy(buf.ensure(sizeof(float) * n), o, sizeof(float) * n, hipMemcpyHostToDevice));
	count = n;
}
void SetTimeOffsetsRaysNode::validateImpl()
{
	IRaysNodeSingleInput::validateImpl();
	if (input->rayCount() != count) {
		throw InvalidPipeline("time offsets don't match number of rays. RayCount(" + std::to_string(input->rayCount()) +
		                      ") and TimeOffsetsCount(" + std::to_string(count) + ") should be equal");
	}
}

void TransformRaysNode::pointOp(PointOpItem& it, FieldCopyList&)
{
	const int n = rayCount();
	it.op = kOpTransformRays;
	it.capacity = n;
	it.in0 = input->rays();
	it.out0 = out.ensureT<Mat3x4f>(std::max(n, 1));
	it.m0 = transform;
}
// Depends on the graph's structure and on validated field sets only, so it is cached per structure epoch (batch
// planning asks for it several times per node and frame).
bool TransformRaysNode::deferToRaytrace() const
{
	if (deferEpoch != structureEpoch()) {
		deferCached = computeDeferToRaytrace();
		deferEpoch = structureEpoch();
	}
	return deferCached;
}

bool TransformRaysNode::computeDeferToRaytrace() const
{
	if (outputs.empty() || input->deferredRays().hasPre) {
		return false;
	}
	std::function<bool(const Node&)> accepts = [&](const Node& n) -> bool {
		if (dynamic_cast<const SetRangeRaysNode*>(&n) != nullptr || dynamic_cast<const SetRingIdsRaysNode*>(&n) != nullptr ||
		    dynamic_cast<const SetTimeOffsetsRaysNode*>(&n) != nullptr) {
			const auto& outs = n.getOutputs();
			return !outs.empty() && std::all_of(outs.begin(), outs.end(), [&](const Node::Ptr& o) { return accepts(*o); });
		}
		const auto* rt = dynamic_cast<const RaytraceNode*>(&n);
		return rt != nullptr && !rt->hasField(RGL_FIELD_RAY_POSE_MAT3x4_F32);
	};
	return std::all_of(outputs.begin(), outputs.end(), [&](const Node::Ptr& o) { return accepts(*o); });
}

IRaysNode::DeferredRays TransformRaysNode::deferredRays() const
{
	if (deferToRaytrace()) {
		return {input->rays(), transform, true};
	}
	return {out.as<Mat3x4f>(), Mat3x4f::identity(), false};
}

void TransformRaysNode::enqueueImpl()
{
	if (deferToRaytrace()) {
		return;
	}
	PointOpItem it{};
	FieldCopyList unused;
	pointOp(it, unused);
	gpu::pointOp(stream(), it);
}

// Seeds: a process-wide sequence (base from RGL_NOISE_SEED, default fixed) so runs are reproducible.
void NoiseStream::init(uint32_t tagValue)
{
	static std::atomic<uint64_t> counter{0};
	static const uint64_t base = []() {
		const char* env = std::getenv("RGL_NOISE_SEED");
		return env ? std::strtoull(env, nullptr, 0) : 0x52474C2D4D493335ull;
	}();
	uint64_t z = base + 0x9E3779B97F4A7C15ull * (counter.fetch_add(1) + 1);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	seed = z ^ (z >> 31);
	run = 0;
	tag = tagValue;
}

void GaussianNoiseAngularRaysNode::setParameters(float m, float s, int a)
{
	mean = m;
	stDev = s;
	axis = a;
}
void GaussianNoiseAngularRaysNode::pointOp(PointOpItem& it, FieldCopyList&)
{
	const int n = rayCount();
	it.op = kOpNoiseRays;
	it.capacity = n;
	it.in0 = input->rays();
	it.out0 = out.ensureT<Mat3x4f>(std::max(n, 1));
	it.m0 = input->cumulativeRayTransform().inverse();
	it.m1 = it.m0.inverse();
	it.a = mean;
	it.b = stDev;
	it.axis = axis;
	it.seed = rng.seed;
	it.run = rng.run++;
	it.tag = rng.tag;
}
void GaussianNoiseAngularRaysNode::enqueueImpl()
{
	PointOpItem it{};
	FieldCopyList unused;
	pointOp(it, unused);
	gpu::pointOp(stream(), it);
}

// ------------------------------------------------------------------------------------------------ raytrace
// src/graph/RaytraceNode.cpp:166-191: XYZ plus every field required by point nodes reachable through inputs and
// outputs.
std::set<rgl_field_t> RaytraceNode::findFieldsToCompute()
{
	std::set<rgl_field_t> out{RGL_FIELD_XYZ_VEC3_F32};
	std::function<void(const Node::Ptr&, bool)> dfs = [&](const Node::Ptr& cur, bool viaInputs) {
		const auto& next = viaInputs ? cur->getInputs() : cur->getOutputs();
		for (auto& n : next) {
			if (auto pn = std::dynamic_pointer_cast<IPointsNode>(n)) {
				for (auto f : pn->requiredFields()) {
					if (!isPadding(f)) {
						out.insert(f);
					}
				}
				dfs(n, viaInputs);
			}
		}
	};
	dfs(shared_from_this(), true);
	dfs(shared_from_this(), false);
	return out;
}

void RaytraceNode::validateImpl()
{
	fields = findFieldsToCompute();
	raysNode = exactlyOneInputOfType<IRaysNode>();
	if (fields.count(RGL_FIELD_RING_ID_U16) && raysNode->ringIds() == nullptr) {
		throw InvalidPipeline("requested for field RING_ID_U16, but RaytraceNode cannot get ring ids");
	}
	if (fields.count(RGL_FIELD_TIME_STAMP_F64) && !scene->getTime().has_value()) {
		throw InvalidPipeline("requested for field TIME_STAMP_F64, but RaytraceNode cannot get time from scene");
	}
	if (distortion && raysNode->timeOffsets() == nullptr) {
		throw InvalidPipeline("requested for raytrace with velocity distortion, but RaytraceNode cannot get time offsets");
	}
	if (hasMask && maskCount != raysNode->rayCount()) {
		throw InvalidPipeline("Mask size does not match the number of rays");
	}
	for (auto it = data.begin(); it != data.end();) {
		it = fields.count(it->first) ? std::next(it) : data.erase(it);
	}
}

void RaytraceNode::setMask(const int8_t* m, int n)
{
	RGL_HIP(hipMemcpy(maskBuf.ensure(n), m, n, hipMemcpyHostToDevice));
	maskCount = n;
	hasMask = true;
}

RaytraceParams RaytraceNode::buildParams(const SceneView&) const
{
	RaytraceParams p{};
	const IRaysNode::DeferredRays dr = raysNode->deferredRays();
	p.rays = dr.rays;
	p.order = raysNode->traceOrder();
	p.rayPreTransformed = dr.hasPre ? 1 : 0;
	p.rayPreTransform = dr.pre;
	p.rayCount = raysNode->rayCount();
	p.rayOriginToWorld = raysNode->cumulativeRayTransform();
	p.worldToRayOrigin64 = inverse64(widen(p.rayOriginToWorld));
	p.worldToRayOrigin = round32(p.worldToRayOrigin64);
	if (raysNode->ranges() != nullptr) {
		p.ranges = raysNode->ranges();
		p.rangesCount = raysNode->rangesCount();
	} else {
		p.ranges = defaultRange.as<Vec2f>();
		p.rangesCount = 1;
	}
	p.ringIds = raysNode->ringIds();
	p.ringIdsCount = raysNode->ringIdsCount();
	p.timeOffsetsMs = raysNode->timeOffsets();
	p.mask = hasMask ? maskBuf.as<int8_t>() : nullptr;
	p.doDistortion = distortion ? 1 : 0;
	p.sensorLinearVelocity = linVel;
	p.sensorAngularVelocity = angVel;
	p.nearNonHitDistance = nearDist;
	p.farNonHitDistance = farDist;
	p.defaultIntensity = defaultIntensity;
	p.sceneDeltaTime = scene->deltaTimeSeconds();
	p.hHalfDivergence = hHalf;
	p.vHalfDivergence = vHalf;
	p.returnMode = static_cast<int32_t>(mode);
	p.returnCount = returnCount();
	p.velocityRequested = (fields.count(RGL_FIELD_ABSOLUTE_VELOCITY_VEC3_F32) ||
	                       fields.count(RGL_FIELD_RELATIVE_VELOCITY_VEC3_F32) || fields.count(RGL_FIELD_RADIAL_SPEED_F32))
	                          ? 1
	                          : 0;
	return p;
}

RaytraceBatchItem RaytraceNode::batchItem(hipStream_t s)
{
	if (defaultRange.get() == nullptr) {
		const Vec2f r{0.0f, FLT_MAX};
		RGL_HIP(hipMemcpy(defaultRange.ensure(sizeof(Vec2f)), &r, sizeof(Vec2f), hipMemcpyHostToDevice));
	}
	const int cap = capacity();
	for (auto f : fields) {
		if (f == RGL_FIELD_RAY_POSE_MAT3x4_F32) {
			continue;
		}
		auto& b = data[f];
		if (!b) {
			b = std::make_unique<DeviceBuffer>();
		}
		b->ensure(fieldSizeChecked(f) * std::max(cap, 1));
	}
	RaytraceBatchItem it{};
	it.sc = scene->prepare(s);
	it.p = buildParams(it.sc);
	auto ptr = [&](rgl_field_t f) -> void* { return data.count(f) ? data[f]->get() : nullptr; };
	RaytraceOutputs& o = it.out;
	o.xyz = static_cast<Vec3f*>(ptr(RGL_FIELD_XYZ_VEC3_F32));
	o.isHit = static_cast<int32_t*>(ptr(RGL_FIELD_IS_HIT_I32));
	o.rayIdx = static_cast<uint32_t*>(ptr(RGL_FIELD_RAY_IDX_U32));
	o.ringId = static_cast<uint16_t*>(ptr(RGL_FIELD_RING_ID_U16));
	o.returnType = static_cast<uint8_t*>(ptr(RGL_FIELD_RETURN_TYPE_U8));
	o.distance = static_cast<float*>(ptr(RGL_FIELD_DISTANCE_F32));
	o.intensityF32 = static_cast<float*>(ptr(RGL_FIELD_INTENSITY_F32));
	o.intensityU8 = static_cast<uint8_t*>(ptr(RGL_FIELD_INTENSITY_U8));
	o.laserRetro = static_cast<float*>(ptr(RGL_FIELD_LASER_RETRO_F32));
	o.timeStampF64 = static_cast<double*>(ptr(RGL_FIELD_TIME_STAMP_F64));
	o.timeStampU32 = static_cast<uint32_t*>(ptr(RGL_FIELD_TIME_STAMP_U32));
	o.entityId = static_cast<int32_t*>(ptr(RGL_FIELD_ENTITY_ID_I32));
	o.absVelocity = static_cast<Vec3f*>(ptr(RGL_FIELD_ABSOLUTE_VELOCITY_VEC3_F32));
	o.relVelocity = static_cast<Vec3f*>(ptr(RGL_FIELD_RELATIVE_VELOCITY_VEC3_F32));
	o.radialSpeed = static_cast<float*>(ptr(RGL_FIELD_RADIAL_SPEED_F32));
	o.azimuth = static_cast<float*>(ptr(RGL_FIELD_AZIMUTH_F32));
	o.elevation = static_cast<float*>(ptr(RGL_FIELD_ELEVATION_F32));
	o.normal = static_cast<Vec3f*>(ptr(RGL_FIELD_NORMAL_VEC3_F32));
	o.incidentAngle = static_cast<float*>(ptr(RGL_FIELD_INCIDENT_ANGLE_F32));
	overflowSlots.reset();
	return it;
}

void RaytraceNode::enqueueImpl()
{
	RaytraceBatchItem it = batchItem(stream());
	it.p.order = nullptr; // a single sensor's launch spans about one occupancy generation: index order (see batch.cpp)
	if (timing) {
		if (timingUsed == timingEvents.size()) {
			hipEvent_t a, b;
			RGL_HIP(hipEventCreate(&a));
			RGL_HIP(hipEventCreate(&b));
			timingEvents.emplace_back(a, b);
		}
		RGL_HIP(hipEventRecord(timingEvents[timingUsed].first, stream()));
	}
	gpu::raytrace(stream(), it.sc, it.p, it.out);
	RGL_HIP(hipGetLastError());
	if (timing) {
		RGL_HIP(hipEventRecord(timingEvents[timingUsed].second, stream()));
		++timingUsed;
	}
	scene->noteUse(stream());
	RGL_HIP(hipMemcpyAsync(overflowHost.ensure(sizeof(int32_t)), it.sc.overflow, sizeof(int32_t), hipMemcpyDeviceToHost,
	                       stream()));
}

void RaytraceNode::setTiming(bool enable)
{
	timing = enable;
	timingUsed = 0;
}

void RaytraceNode::getTiming(double* totalMs, int32_t* launches)
{
	double total = 0.0;
	for (size_t i = 0; i < timingUsed; ++i) {
		RGL_HIP(hipEventSynchronize(timingEvents[i].second));
		float ms = 0.0f;
		RGL_HIP(hipEventElapsedTime(&ms, timingEvents[i].first, timingEvents[i].second));
		total += ms;
	}
	*totalMs = total;
	*launches = static_cast<int32_t>(timingUsed);
}

RaytraceNode::~RaytraceNode()
{
	for (auto& [a, b] : timingEvents) {
		(void) hipEventDestroy(a);
		(void) hipEventDestroy(b);
	}
}

bool RaytraceNode::overflowed()
{
	const auto* flag = overflowSlots ? static_cast<const int32_t*>(overflowSlots->get()) + overflowSlot
	                                 : static_cast<const int32_t*>(overflowHost.get());
	return flag != nullptr && *flag != 0;
}

const void* RaytraceNode::fieldData(rgl_field_t f)
{
	if (f == RGL_FIELD_RAY_POSE_MAT3x4_F32) {
		return raysNode->rays();
	}
	auto it = data.find(f);
	if (it == data.end()) {
		throw InvalidPipeline("RaytraceNode does not provide field " + fieldName(f));
	}
	return it->second->get();
}

void RaytraceNode::traversalStats(int32_t* nodes, int32_t* tris, int32_t* insts)
{
	const SceneView view = scene->prepare(stream());
	gpu::traversalStats(stream(), view, buildParams(view), nodes, tris, insts);
	RGL_HIP(hipGetLastError());
}

void RaytraceNode::footprint(unsigned long long* keys, uint32_t mask, uint32_t* full)
{
	const SceneView view = scene->prepare(stream());
	gpu::traversalFootprint(stream(), view, buildParams(view), keys, mask, full);
	RGL_HIP(hipGetLastError());
}

void RaytraceNode::bruteForce(float* t, int32_t* inst, int32_t* prim)
{
	const SceneView view = scene->prepare(stream());
	const auto counts = scene->instanceTriCounts();
	DeviceBuffer dc;
	RGL_HIP(hipMemcpy(dc.ensure(sizeof(int32_t) * std::max<size_t>(counts.size(), 1)), counts.data(),
	                  sizeof(int32_t) * counts.size(), hipMemcpyHostToDevice));
	gpu::bruteForceClosest(stream(), view, dc.as<int32_t>(), buildParams(view), t, inst, prim);
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipStreamSynchronize(stream())); // the kernel reads `dc`, released on return
	RGL_HIP(hipStreamSynchronize(stream()));
}

// ------------------------------------------------------------------------------------------------ point nodes
void TransformPointsNode::pointOp(PointOpItem& it, FieldCopyList&)
{
	const int cap = capacity();
	it.op = kOpTransformPoints;
	it.capacity = cap;
	it.count = countDevice();
	it.in0 = input->fieldData(RGL_FIELD_XYZ_VEC3_F32);
	it.out0 = out.ensureT<Vec3f>(std::max(cap, 1));
	it.m0 = transform;
}
void TransformPointsNode::enqueueImpl()
{
	PointOpItem it{};
	FieldCopyList unused;
	pointOp(it, unused);
	gpu::pointOp(stream(), it);
}
const void* TransformPointsNode::fieldData(rgl_field_t f)
{
	return f == RGL_FIELD_XYZ_VEC3_F32 ? out.get() : input->fieldData(f);
}

void GaussianNoiseAngularHitpointNode::setParameters(float m, float s, int a)
{
	mean = m;
	stDev = s;
	axis = a;
}
void GaussianNoiseAngularHitpointNode::validateImpl()
{
	IPointsNodeSingleInput::validateImpl();
	withDistance = input->hasField(RGL_FIELD_DISTANCE_F32);
}
void GaussianNoiseAngularHitpointNode::pointOp(PointOpItem& it, FieldCopyList&)
{
	const int cap = capacity();
	it.op = kOpNoiseHitpoint;
	it.capacity = cap;
	it.count = countDevice();
	it.in0 = input->fieldData(RGL_FIELD_XYZ_VEC3_F32);
	it.out0 = outXyz.ensureT<Vec3f>(std::max(cap, 1));
	it.out1 = withDistance ? outDistance.ensureT<float>(std::max(cap, 1)) : nullptr;
	it.m0 = input->lookAtOriginTransform();
	it.m1 = it.m0.inverse();
	it.a = mean;
	it.b = stDev;
	it.axis = axis;
	it.seed = rng.seed;
	it.run = rng.run++;
	it.tag = rng.tag;
}
void GaussianNoiseAngularHitpointNode::enqueueImpl()
{
	PointOpItem it{};
	FieldCopyList unused;
	pointOp(it, unused);
	gpu::pointOp(stream(), it);
}
const void* GaussianNoiseAngularHitpointNode::fieldData(rgl_field_t f)
{
	if (f == RGL_FIELD_XYZ_VEC3_F32) {
		return outXyz.get();
	}
	if (f == RGL_FIELD_DISTANCE_F32 && withDistance) {
		return outDistance.get();
	}
	return input->fieldData(f);
}

void GaussianNoiseDistanceNode::setParameters(float m, float b, float r)
{
	mean = m;
	base = b;
	rise = r;
}
void GaussianNoiseDistanceNode::pointOp(PointOpItem& it, FieldCopyList&)
{
	const int cap = capacity();
	it.op = kOpNoiseDistance;
	it.capacity = cap;
	it.count = countDevice();
	it.in0 = input->fieldData(RGL_FIELD_XYZ_VEC3_F32);
	it.in1 = input->fieldData(RGL_FIELD_DISTANCE_F32);
	it.out0 = outXyz.ensureT<Vec3f>(std::max(cap, 1));
	it.out1 = outDistance.ensureT<float>(std::max(cap, 1));
	it.m0 = input->lookAtOriginTransform();
	it.m1 = it.m0.inverse();
	it.a = mean;
	it.b = base;
	it.c = rise;
	it.seed = rng.seed;
	it.run = rng.run++;
	it.tag = rng.tag;
}
void GaussianNoiseDistanceNode::enqueueImpl()
{
	PointOpItem it{};
	FieldCopyList unused;
	pointOp(it, unused);
	gpu::pointOp(stream(), it);
}
const void* GaussianNoiseDistanceNode::fieldData(rgl_field_t f)
{
	if (f == RGL_FIELD_XYZ_VEC3_F32) {
		return outXyz.get();
	}
	if (f == RGL_FIELD_DISTANCE_F32) {
		return outDistance.get();
	}
	return input->fieldData(f);
}

// src/graph/CompactByFieldPointsNode.cpp:33-105; every field of the input is compacted in one pass.
CompactItem CompactByFieldPointsNode::compactItem(FieldCopyList& list, int variant)
{
	const int cap = capacity();
	list = FieldCopyList{};
	std::set<rgl_field_t> keep;
	for (auto f : input->fieldList()) {
		if (isPadding(f) || f == RGL_FIELD_DYNAMIC_FORMAT) {
			continue;
		}
		keep.insert(f);
		auto& b = data[f];
		if (!b) {
			b = std::make_unique<DeviceBuffer>();
		}
		const size_t sz = fieldSizeChecked(f);
		b->ensure(sz * std::max(cap, 1));
		if (list.count == kMaxCompactFields) {
			throw std::logic_error("too many fields to compact");
		}
		list.f[list.count++] = FieldCopy{static_cast<const char*>(input->fieldData(f)), static_cast<char*>(b->get()),
		                                 static_cast<int32_t>(sz), 0};
	}
	for (auto it = data.begin(); it != data.end();) {
		it = keep.count(it->first) ? std::next(it) : data.erase(it);
	}
	CompactItem it{};
	it.capacity = cap;
	it.countIn = input->countDevice();
	it.flags = cap > 0 ? static_cast<const int32_t*>(input->fieldData(by)) : nullptr;
	it.numTiles = gpu::compactTiles(cap, variant);
	it.tileCounts = tileCounts.ensureT<int32_t>(it.numTiles);
	it.countOut = count.ensureT<int32_t>(1);
	countSlots.reset();
	return it;
}

void CompactByFieldPointsNode::enqueueImpl()
{
	FieldCopyList list;
	const CompactItem it = compactItem(list);
	gpu::compact(stream(), it.capacity, it.countIn, it.flags, list, it.tileCounts, it.countOut);
	RGL_HIP(hipMemcpyAsync(countHost.ensure(sizeof(int32_t)), it.countOut, sizeof(int32_t), hipMemcpyDeviceToHost,
	                       stream()));
}

int CompactByFieldPointsNode::hostCount()
{
	waitForResults();
	if (countSlots) {
		return static_cast<const int32_t*>(countSlots->get())[countSlot];
	}
	return *static_cast<int32_t*>(countHost.get());
}

const void* CompactByFieldPointsNode::fieldData(rgl_field_t f)
{
	auto it = data.find(f);
	if (it == data.end()) {
		throw InvalidPipeline("CompactByFieldPointsNode: input does not provide field " + fieldName(f));
	}
	return it->second->get();
}

void YieldPointsNode::setParameters(const std::vector<rgl_field_t>& f)
{
	if (std::find(f.begin(), f.end(), RGL_FIELD_DYNAMIC_FORMAT) != f.end()) {
		throw std::invalid_argument("cannot yield field 'RGL_FIELD_DYNAMIC_FORMAT'");
	}
	fields = f;
}
bool YieldPointsNode::hasField(rgl_field_t f) const { return std::find(fields.begin(), fields.end(), f) != fields.end(); }

void YieldPointsNode::setPrefetch(bool on)
{
	prefetch = on;
	prefetched.clear();
}

void YieldPointsNode::pointOp(PointOpItem& it, FieldCopyList& layout)
{
	const int cap = capacity();
	layout = FieldCopyList{};
	prefetched.clear();
	for (auto f : fields) {
		if (isPadding(f) || prefetched.count(f)) {
			continue;
		}
		const size_t sz = fieldSizeChecked(f);
		MappedHostBuffer& hb = host[f];
		if (hb.capacity() < sz * std::max(cap, 1) && hb.capacity() > 0) {
			waitForResultsIfRun(); // the previous run may still write into the buffer being replaced
		}
		hb.ensure(sz * std::max(cap, 1));
		if (layout.count == kMaxCompactFields) {
			throw std::logic_error("too many fields to prefetch");
		}
		layout.f[layout.count++] = FieldCopy{static_cast<const char*>(input->fieldData(f)), static_cast<char*>(hb.device()),
		                                     static_cast<int32_t>(sz), 0};
		prefetched.insert(f);
	}
	it.op = kOpToHost;
	it.capacity = cap;
	it.count = countDevice();
}

void YieldPointsNode::enqueueImpl()
{
	if (!prefetch) {
		return;
	}
	PointOpItem it{};
	FieldCopyList layout;
	pointOp(it, layout);
	gpu::copyFields(stream(), it.capacity, it.count, layout);
}

const void* YieldPointsNode::hostData(rgl_field_t f) const
{
	if (!prefetch || !prefetched.count(f)) {
		return nullptr;
	}
	return host.at(f).host();
}

void FormatPointsNode::setParameters(const std::vector<rgl_field_t>& f)
{
	if (std::find(f.begin(), f.end(), RGL_FIELD_DYNAMIC_FORMAT) != f.end()) {
		throw std::invalid_argument("cannot format field 'RGL_FIELD_DYNAMIC_FORMAT'");
	}
	for (auto x : f) {
		fieldSizeChecked(x);
	}
	fields = f;
}
bool FormatPointsNode::hasField(rgl_field_t f) const
{
	return f == RGL_FIELD_DYNAMIC_FORMAT || std::find(fields.begin(), fields.end(), f) != fields.end();
}
size_t FormatPointsNode::fieldPointSize(rgl_field_t f) const
{
	return f == RGL_FIELD_DYNAMIC_FORMAT ? pointSize(fields) : fieldSizeChecked(f);
}
FieldCopyList FormatPointsNode::makeLayout()
{
	FieldCopyList list{};
	size_t off = 0;
	for (auto f : fields) {
		if (list.count == kMaxFormatFields) {
			throw std::logic_error("too many fields to format");
		}
		const size_t sz = fieldSizeChecked(f);
		const char* src = isPadding(f) ? nullptr : static_cast<const char*>(input->fieldData(f));
		list.f[list.count++] = FieldCopy{src, nullptr, static_cast<int32_t>(sz), static_cast<int32_t>(off)};
		off += sz;
	}
	return list;
}
void FormatPointsNode::pointOp(PointOpItem& it, FieldCopyList& layout)
{
	const int cap = capacity();
	const size_t rec = pointSize(fields);
	layout = makeLayout();
	it.op = kOpFormat;
	it.capacity = cap;
	it.count = countDevice();
	it.out0 = out.ensureT<char>(std::max<size_t>(rec * cap, 1));
	it.recordSize = static_cast<int32_t>(rec);
}
void FormatPointsNode::enqueueImpl()
{
	PointOpItem it{};
	FieldCopyList layout;
	pointOp(it, layout);
	gpu::formatSoaToAos(stream(), it.capacity, it.count, it.recordSize, layout, static_cast<char*>(it.out0));
}
const void* FormatPointsNode::fieldData(rgl_field_t f)
{
	return f == RGL_FIELD_DYNAMIC_FORMAT ? out.get() : input->fieldData(f);
}

// ------------------------------------------------------------------------------------------------ radar cells
namespace {
const rgl_field_t kRadarFields[] = {RGL_FIELD_XYZ_VEC3_F32, RGL_FIELD_DISTANCE_F32, RGL_FIELD_AZIMUTH_F32,
                                    RGL_FIELD_RCS_F32, RGL_FIELD_IS_HIT_I32};
}

void RadarBinPointsNode::setParameters(float rangeStep, float rangeMax, float azStep, float azMin, float azMax,
                                       float raySolidAngle)
{
	if (!(rangeStep > 0.0f) || !(azStep > 0.0f) || !(rangeMax > 0.0f) || !(azMax > azMin) || !(raySolidAngle > 0.0f)) {
		throw std::invalid_argument("radar binning needs positive steps, range and solid angle and azimuth_max > "
		                            "azimuth_min");
	}
	const double rc = std::ceil(static_cast<double>(rangeMax) / rangeStep);
	const double ac = std::ceil((static_cast<double>(azMax) - azMin) / azStep);
	if (rc * ac > 64.0 * 1024 * 1024) {
		throw std::invalid_argument("radar binning grid exceeds 64M cells");
	}
	bp = RadarBinParams{rangeStep, azStep, azMin, raySolidAngle, static_cast<int32_t>(rc), static_cast<int32_t>(ac)};
}

bool RadarBinPointsNode::hasField(rgl_field_t f) const
{
	return std::find(std::begin(kRadarFields), std::end(kRadarFields), f) != std::end(kRadarFields);
}

std::vector<rgl_field_t> RadarBinPointsNode::fieldList() const
{
	return {std::begin(kRadarFields), std::end(kRadarFields)};
}

void RadarBinPointsNode::enqueueImpl()
{
	const int cells = capacity();
	FieldCopyList list{};
	for (rgl_field_t f : kRadarFields) {
		const size_t sz = fieldSizeChecked(f);
		list.f[list.count++] = FieldCopy{static_cast<const char*>(dense[f].ensure(sz * cells)),
		                                 static_cast<char*>(out[f].ensure(sz * cells)), static_cast<int32_t>(sz), 0};
	}
	const int cap = input->capacity();
	gpu::radarBin(stream(), cap, input->countDevice(),
	              static_cast<const Vec3f*>(input->fieldData(RGL_FIELD_XYZ_VEC3_F32)),
	              static_cast<const float*>(input->fieldData(RGL_FIELD_DISTANCE_F32)),
	              static_cast<const float*>(input->fieldData(RGL_FIELD_INCIDENT_ANGLE_F32)),
	              static_cast<const int32_t*>(input->fieldData(RGL_FIELD_IS_HIT_I32)), input->lookAtOriginTransform(), bp,
	              sums.ensureT<unsigned long long>(cells), counts.ensureT<uint32_t>(cells),
	              dense[RGL_FIELD_XYZ_VEC3_F32].as<Vec3f>(), dense[RGL_FIELD_DISTANCE_F32].as<float>(),
	              dense[RGL_FIELD_AZIMUTH_F32].as<float>(), dense[RGL_FIELD_RCS_F32].as<float>(),
	              dense[RGL_FIELD_IS_HIT_I32].as<int32_t>(), keep.ensureT<int32_t>(cells));
	gpu::compact(stream(), cells, nullptr, keep.as<int32_t>(), list, tileCounts.ensureT<int32_t>(gpu::compactTiles(cells)),
	             count.ensureT<int32_t>(1));
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipMemcpyAsync(countHost.ensure(sizeof(int32_t)), count.get(), sizeof(int32_t), hipMemcpyDeviceToHost,
	                       stream()));
}

int RadarBinPointsNode::hostCount()
{
	waitForResults();
	return *static_cast<int32_t*>(countHost.get());
}

const void* RadarBinPointsNode::fieldData(rgl_field_t f)
{
	auto it = out.find(f);
	if (it == out.end() || !hasField(f)) {
		throw InvalidPipeline("RadarBinPointsNode does not provide field " + fieldName(f));
	}
	return it->second.get();
}

// ------------------------------------------------------------------------------------------------ radar postprocess
namespace {
const rgl_field_t kRadarPostFields[] = {RGL_FIELD_POWER_F32, RGL_FIELD_RCS_F32, RGL_FIELD_NOISE_F32, RGL_FIELD_SNR_F32};
bool isRadarPostField(rgl_field_t f)
{
	return std::find(std::begin(kRadarPostFields), std::end(kRadarPostFields), f) != std::end(kRadarPostFields);
}
} // namespace

void RadarPostprocessPointsNode::setParameters(const std::vector<RadarScope>& scopes, float azimuthStep,
                                               float elevationStep, float frequency, float powerTransmitted,
                                               float deviceGain, float noiseMean, float noiseStDev)
{
	if (scopes.size() > static_cast<size_t>(kMaxRadarScopes)) {
		throw std::invalid_argument("at most " + std::to_string(kMaxRadarScopes) + " radar scopes are supported");
	}
	RadarPostParams p{};
	for (size_t i = 0; i < scopes.size(); ++i) {
		p.scopes[i] = scopes[i];
	}
	p.scopeCount = static_cast<int32_t>(scopes.size());
	p.azimuthStep = azimuthStep;
	p.elevationStep = elevationStep;
	p.frequency = frequency;
	p.powerTransmitted = powerTransmitted;
	p.deviceGain = deviceGain;
	p.noiseMean = noiseMean;
	p.noiseStDev = noiseStDev;
	params = p;
}

// src/graph/RadarPostprocessPointsNode.cpp:37-49
void RadarPostprocessPointsNode::validateImpl()
{
	IPointsNodeSingleInput::validateImpl();
	if (!input->isDense()) {
		throw InvalidPipeline("RadarComputeEnergyPointsNode requires dense input");
	}
}

bool RadarPostprocessPointsNode::hasField(rgl_field_t f) const { return isRadarPostField(f) || input->hasField(f); }

std::vector<rgl_field_t> RadarPostprocessPointsNode::fieldList() const
{
	std::vector<rgl_field_t> out = input->fieldList();
	for (rgl_field_t f : kRadarPostFields) {
		if (std::find(out.begin(), out.end(), f) == out.end()) {
			out.push_back(f);
		}
	}
	return out;
}

RadarItem RadarPostprocessPointsNode::radarItem(FieldCopyList& list)
{
	const int cap = input->capacity();
	const size_t n = static_cast<size_t>(std::max(cap, 1));
	RadarItem it{};
	it.p = params;
	it.p.lookAt = input->lookAtOriginTransform();
	it.p.seed = rng.seed;
	it.p.run = rng.run++;
	it.p.tag = rng.tag;
	it.rayPose = static_cast<const Mat3x4f*>(input->fieldData(RGL_FIELD_RAY_POSE_MAT3x4_F32));
	it.distance = static_cast<const float*>(input->fieldData(RGL_FIELD_DISTANCE_F32));
	it.azimuth = static_cast<const float*>(input->fieldData(RGL_FIELD_AZIMUTH_F32));
	it.elevation = static_cast<const float*>(input->fieldData(RGL_FIELD_ELEVATION_F32));
	it.radialSpeed = static_cast<const float*>(input->fieldData(RGL_FIELD_RADIAL_SPEED_F32));
	it.normal = static_cast<const Vec3f*>(input->fieldData(RGL_FIELD_NORMAL_VEC3_F32));
	it.count = input->countDevice();
	it.terms = static_cast<float*>(terms.ensure(16 * n));
	it.work = work.ensureT<int32_t>(3 * n);
	it.centres = centres.ensureT<int32_t>(n);
	it.power = power.ensureT<float>(n);
	it.rcs = rcs.ensureT<float>(n);
	it.noise = noise.ensureT<float>(n);
	it.snr = snr.ensureT<float>(n);
	it.info = info.ensureT<int32_t>(2);
	it.capacity = cap;
	// The centre point's fields of every cluster (src/graph/RadarPostprocessPointsNode.cpp:215-257 gathers them on
	// request; here every input field is gathered with the clusters).
	list = FieldCopyList{};
	for (rgl_field_t f : input->fieldList()) {
		if (isRadarPostField(f) || isPadding(f)) {
			continue;
		}
		if (list.count == kMaxCompactFields) {
			throw InvalidPipeline("RadarPostprocessPointsNode: too many input fields");
		}
		const size_t sz = fieldSizeChecked(f);
		list.f[list.count++] = FieldCopy{static_cast<const char*>(input->fieldData(f)),
		                                 static_cast<char*>(gathered[f].ensure(sz * n)), static_cast<int32_t>(sz), 0};
	}
	infoSlots.reset();
	return it;
}

void RadarPostprocessPointsNode::enqueueImpl()
{
	FieldCopyList list{};
	RadarItem it = radarItem(list);
	// From pageable memory: the copy has read `list` when the call returns.
	RGL_HIP(hipMemcpyAsync(gatherList.ensure(sizeof(FieldCopyList)), &list, sizeof(FieldCopyList),
	                       hipMemcpyHostToDevice, stream()));
	it.gather = gatherList.as<FieldCopyList>();
	gpu::radarPostprocess(stream(), it);
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipMemcpyAsync(infoHost.ensure(2 * sizeof(int32_t)), info.get(), 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
	                       stream()));
}

int RadarPostprocessPointsNode::hostCount()
{
	waitForResults();
	const auto* h = infoSlots ? static_cast<const int32_t*>(infoSlots->get()) + infoSlot
	                          : static_cast<const int32_t*>(infoHost.get());
	if (h[1] & kRadarErrRcsNan) {
		throw InvalidPipeline("RCS is NaN");
	}
	if (h[1] & kRadarErrTooManyClusters) {
		throw InvalidPipeline("radar postprocessing supports at most " + std::to_string(kRadarMaxClusters) +
		                      " clusters per frame");
	}
	return h[0];
}

const void* RadarPostprocessPointsNode::fieldData(rgl_field_t f)
{
	switch (f) {
		case RGL_FIELD_POWER_F32: return power.get();
		case RGL_FIELD_RCS_F32: return rcs.get();
		case RGL_FIELD_NOISE_F32: return noise.get();
		case RGL_FIELD_SNR_F32: return snr.get();
		default: break;
	}
	auto it = gathered.find(f);
	if (it == gathered.end()) {
		throw InvalidPipeline("RadarPostprocessPointsNode does not provide field " + fieldName(f));
	}
	return it->second.get();
}

// src/graph/FromArrayPointsNode.cpp:18-46
void FromArrayPointsNode::setParameters(const void* points, int n, const std::vector<rgl_field_t>& f)
{
	if (std::find(f.begin(), f.end(), RGL_FIELD_DYNAMIC_FORMAT) != f.end()) {
		throw std::invalid_argument("cannot create point cloud from field 'RGL_FIELD_DYNAMIC_FORMAT'");
	}
	data.clear();
	width = n;
	const size_t rec = pointSize(f);
	FieldCopyList list{};
	size_t off = 0;
	for (auto x : f) {
		const size_t sz = fieldSizeChecked(x);
		if (!isPadding(x) && !data.count(x)) {
			auto b = std::make_unique<DeviceBuffer>();
			b->ensure(sz * std::max(n, 1));
			if (list.count == kMaxFormatFields) {
				throw std::invalid_argument("too many fields");
			}
			list.f[list.count++] = FieldCopy{nullptr, static_cast<char*>(b->get()), static_cast<int32_t>(sz),
			                                 static_cast<int32_t>(off)};
			data[x] = std::move(b);
		}
		off += sz;
	}
	DeviceBuffer staging;
	RGL_HIP(hipMemcpy(staging.ensure(rec * n), points, rec * n, hipMemcpyHostToDevice));
	gpu::formatAosToSoa(nullptr, n, static_cast<int>(rec), staging.as<char>(), list);
	RGL_HIP(hipGetLastError());
	RGL_HIP(hipDeviceSynchronize());
}
void FromArrayPointsNode::validateImpl() { requireNoInputs(*this, inputs); }
std::vector<rgl_field_t> FromArrayPointsNode::fieldList() const
{
	std::vector<rgl_field_t> v;
	for (auto& [f, b] : data) {
		v.push_back(f);
	}
	return v;
}
const void* FromArrayPointsNode::fieldData(rgl_field_t f)
{
	auto it = data.find(f);
	if (it == data.end()) {
		throw InvalidPipeline("FromArrayPointsNode does not provide field " + fieldName(f));
	}
	return it->second->get();
}

} // namespace rgl
