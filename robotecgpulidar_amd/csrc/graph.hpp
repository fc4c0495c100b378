// Graph model: nodes, the run context of a connected graph, and the node kinds on the raytrace-and-point-cloud
// path.
//
// Structure follows the reference (src/graph/Node.{hpp,cpp}, src/graph/Interfaces.hpp:24-145,
// src/graph/GraphRunCtx.cpp:23-211, src/graph/NodesCore.hpp): create-or-update nodes, a DAG executed in the
// reference's priority-DFS order, one stream per graph, and node interfaces for rays and point clouds.
// Differences that matter for MI355X: nodes are enqueued from the calling thread (HIP launches are already
// asynchronous, so no thread is spawned per run), and point counts that are only known on the device after
// compaction stay on the device; downstream kernels read them there and the host reads them only when a result
// is requested.
#pragma once

#include <map>
#include <memory>
#include <optional>
#include <set>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "runtime.hpp"
#include "scene.hpp"

namespace rgl {

struct GraphRunCtx;

// How a node takes part in a batched run.
enum class BatchStage
{
	Passive,  // nothing to launch (configuration or view nodes)
	Raytrace, // one item of the batched raytrace launch
	RadarPost, // one item of the batched radar postprocess launches
	PointOp,  // one item of the batched elementwise launch
	Compact,  // one item of the batched compaction launches
	Other,    // enqueued on its own, on the batch stream
};

struct Node : std::enable_shared_from_this<Node>
{
	using Ptr = std::shared_ptr<Node>;
	virtual ~Node() = default;

	void addChild(const Ptr& child);
	void removeChild(const Ptr& child);
	void setPriority(int32_t p);
	int32_t getPriority() const { return priority; }

	const std::vector<Ptr>& getInputs() const { return inputs; }
	const std::vector<Ptr>& getOutputs() const { return outputs; }
	std::set<Ptr> connectedComponent();
	std::set<Ptr> disconnectConnectedNodes();

	bool hasCtx() const { return ctx != nullptr; }
	std::shared_ptr<GraphRunCtx> getCtx() const { return ctx; }
	void setCtx(std::shared_ptr<GraphRunCtx> c);

	void validate();
	void enqueue();
	// Marks the node's work as enqueued by a batch: `ev` records the completion of the whole batch (every launch that
	// may still read or write the node's buffers), `ready` (recorded no later than `ev`, or null for `ev`) the point
	// from which the node's results are complete, `counted` (no later than `ready`, or null for `ready`) the point
	// from which the host copies of the batch's point counts and overflow flags are.
	void markEnqueued(const std::shared_ptr<Event>& ev, const std::shared_ptr<Event>& ready = nullptr,
	                  const std::shared_ptr<Event>& counted = nullptr);
	// Waits until this node's results of the last run are complete; throws if the node never ran.
	void waitForResults();
	// Waits until the host copies of the last run's point counts and overflow flags are complete (the results
	// themselves may still be in flight); throws if the node never ran.
	void waitForCounts();
	// Waits for all of the last run's device work if there was one (before the node's buffers are replaced).
	void waitForResultsIfRun();
	void markDirty()
	{
		dirty = true;
		++structureEpoch();
	}
	bool isValid() const { return !dirty; }
	// Advances whenever any node is marked dirty, linked, unlinked or validated: a value derived from graph structure
	// or validated node state may be cached for as long as the epoch is unchanged.
	static uint64_t& structureEpoch()
	{
		static uint64_t epoch = 0;
		return epoch;
	}

	virtual std::string getName() const = 0;
	hipStream_t stream() const;

	virtual BatchStage batchStage() const { return BatchStage::Other; }
	// Elementwise stages: fills the launch description of this run (firstBlock unset; `layout` is the field layout
	// of a format stage, which `it.format` must point to in device memory).
	virtual void pointOp(PointOpItem& it, FieldCopyList& layout);

	template<typename T>
	std::shared_ptr<T> exactlyOneInputOfType() const;

protected:
	virtual void validateImpl() = 0;
	virtual void enqueueImpl() = 0;

	std::vector<Ptr> inputs;
	std::vector<Ptr> outputs; // sorted by priority, descending
	int32_t priority = 0;
	bool dirty = true;
	bool ranOnce = false;
	std::shared_ptr<GraphRunCtx> ctx;
	std::shared_ptr<Event> done;  // shared by every node of a batch
	std::shared_ptr<Event> ready;   // results complete (batched runs; null = `done`)
	std::shared_ptr<Event> counted; // counts and flags on the host (batched runs; null = `ready`)
	int32_t batchLevel = 0;      // longest path from an entry node (batched runs)
	bool batchFused = false;     // computed inside its raytrace node's batched launch this batch (no launch of its own)
	friend struct GraphRunCtx;
	friend struct GraphBatch;
};

// ------------------------------------------------------------------------------------------------ rays
struct IRaysNode : virtual Node
{
	static constexpr const char* kTypeName = "IRaysNode";
	virtual int rayCount() const = 0;
	virtual const Mat3x4f* rays() const = 0;
	virtual const int32_t* ringIds() const { return nullptr; }
	virtual int ringIdsCount() const { return 0; }
	virtual const Vec2f* ranges() const { return nullptr; }
	virtual int rangesCount() const { return 0; }
	virtual const float* timeOffsets() const { return nullptr; }
	virtual int timeOffsetsCount() const { return 0; }
	virtual Mat3x4f cumulativeRayTransform() const { return Mat3x4f::identity(); }
	// Order in which a raytrace launch traces the rays (a permutation of [0, rayCount) grouping rays of close
	// directions), or nullptr for index order. Results do not depend on it: every ray writes its own index.
	virtual const int32_t* traceOrder() const { return nullptr; }
	// The rays as a raytrace launch consumes them: `rays`, with `pre` still to be applied when `hasPre` (a transform
	// the chain left to the raytrace kernel).
	struct DeferredRays
	{
		const Mat3x4f* rays;
		Mat3x4f pre;
		bool hasPre;
	};
	virtual DeferredRays deferredRays() const { return {rays(), Mat3x4f::identity(), false}; }
};

struct IRaysNodeSingleInput : IRaysNode
{
	void validateImpl() override;
	int rayCount() const override { return input->rayCount(); }
	const Mat3x4f* rays() const override { return input->rays(); }
	const int32_t* ringIds() const override { return input->ringIds(); }
	int ringIdsCount() const override { return input->ringIdsCount(); }
	const Vec2f* ranges() const override { return input->ranges(); }
	int rangesCount() const override { return input->rangesCount(); }
	const float* timeOffsets() const override { return input->timeOffsets(); }
	int timeOffsetsCount() const override { return input->timeOffsetsCount(); }
	Mat3x4f cumulativeRayTransform() const override { return input->cumulativeRayTransform(); }
	const int32_t* traceOrder() const override { return input->traceOrder(); }

protected:
	std::shared_ptr<IRaysNode> input;
};

// ------------------------------------------------------------------------------------------------ points
struct IPointsNode : virtual Node
{
	static constexpr const char* kTypeName = "IPointsNode";
	virtual std::vector<rgl_field_t> requiredFields() const { return {}; }
	virtual bool isDense() const = 0;
	virtual bool hasField(rgl_field_t f) const = 0;
	virtual std::vector<rgl_field_t> fieldList() const = 0;
	// Upper bound of the point count, known on the host at enqueue time.
	virtual int capacity() const = 0;
	// Device-resident live point count, or nullptr when the count equals capacity().
	virtual const int32_t* countDevice() const { return nullptr; }
	// Live point count on the host; the caller has waited for this node's results.
	virtual int hostCount() { return capacity(); }
	virtual rgl_return_mode_t returnMode() const = 0;
	virtual int returnCount() const = 0;
	virtual Mat3x4f lookAtOriginTransform() const { return Mat3x4f::identity(); }
	// Device pointer of a field (SoA), valid for the live count.
	virtual const void* fieldData(rgl_field_t f) = 0;
	virtual size_t fieldPointSize(rgl_field_t f) const { return fieldSizeChecked(f); }
};

struct IPointsNodeSingleInput : IPointsNode
{
	void validateImpl() override;
	bool isDense() const override { return input->isDense(); }
	bool hasField(rgl_field_t f) const override { return input->hasField(f); }
	std::vector<rgl_field_t> fieldList() const override { return input->fieldList(); }
	int capacity() const override { return input->capacity(); }
	const int32_t* countDevice() const override { return input->countDevice(); }
	int hostCount() override { return input->hostCount(); }
	rgl_return_mode_t returnMode() const override { return input->returnMode(); }
	int returnCount() const override { return input->returnCount(); }
	Mat3x4f lookAtOriginTransform() const override { return input->lookAtOriginTransform(); }
	const void* fieldData(rgl_field_t f) override { return input->fieldData(f); }

protected:
	std::shared_ptr<IPointsNode> input;
};

// ------------------------------------------------------------------------------------------------ run context
struct GraphRunCtx : std::enable_shared_from_this<GraphRunCtx>
{
	static std::shared_ptr<GraphRunCtx> createAndAttach(const Node::Ptr& node);
	static void synchronizeAll();
	static std::vector<std::shared_ptr<GraphRunCtx>>& instances();

	void executeAsync();
	void synchronize();
	void detachAndDestroy();
	void markNodesDirty();
	void clearExecutionOrder() { order.clear(); }
	hipStream_t stream() const { return streamOverride ? streamOverride : stream_->get(); }
	hipStream_t ownStream() const { return stream_->get(); }
	const std::set<Node::Ptr>& getNodes() const { return nodes; }
	// Validated execution order (computed on first use).
	const std::vector<Node::Ptr>& validatedOrder();
	// Batched runs: orders `batchStream` after this graph's earlier work on its own stream and makes every node
	// enqueue on `batchStream` until endBatch(), which hands the batch's completion event to later runs.
	void beginBatch(hipStream_t batchStream, Event& scratch);
	void endBatch(hipStream_t batchStream, std::shared_ptr<Event> completion);
	// The graph's last batch is still executing.
	bool inFlight() const;

	static std::vector<Node::Ptr> findExecutionOrder(std::set<Node::Ptr> nodes);

private:
	std::set<Node::Ptr> nodes;
	std::vector<Node::Ptr> order;
	std::unique_ptr<Stream> stream_ = std::make_unique<Stream>();
	hipStream_t streamOverride = nullptr;
	bool running = false;
	bool lastRunSolo = false;
	std::shared_ptr<Event> batchDone; // completion of the last batch this graph ran in
	hipStream_t batchDoneStream = nullptr;
};

struct RaytraceNode;

// Runs several independent graphs as one batch: stages of the same kind at the same depth of every graph execute in
// one launch (raytrace, elementwise point/ray stages, compaction), on one stream.
struct GraphBatch
{
	// The batch behind rgl_graph_run_batch. Servers own further instances (one stream each) to keep several
	// batches in flight.
	static GraphBatch& instance();
	GraphBatch();
	GraphBatch(const GraphBatch&) = delete;
	GraphBatch& operator=(const GraphBatch&) = delete;
	~GraphBatch();
	// Runs the graphs that contain `entryNodes` (any node of each graph; duplicates are allowed).
	void run(const std::vector<Node::Ptr>& entryNodes);
	void synchronize();
	// True while the last run's work is still executing on the device.
	bool busy() const;
	// HIP-event timing of the batched raytrace launches while enabled.
	void setTiming(bool enable);
	void getTiming(double* totalMs, int32_t* launches);
	// Autotuning of the raytrace block size and the compaction tile: while a family is untuned, successive batches
	// cycle through its variants, timed with HIP events (only launches of at least 100k elements count); once every
	// variant has three timings the one with the lowest time per element is kept. -1 = still tuning.
	void setAutotune(bool enable);
	// Disables autotuning and uses the given variants.
	void forceVariants(int raytraceVariant, int compactVariant);
	// Capture of repeated launch sequences into HIP graphs (off by default: a batch is a handful of launches, and on the
	// 64-lidar frame replay measured no faster than launching them one by one); counters of captures and replays.
	void setGraphs(bool enable);
	int64_t graphCaptures = 0, graphReplays = 0;
	int64_t fusedStages = 0; // point stages folded into batched raytrace launches
	void autotuneChoice(int32_t* raytraceVariant, int32_t* compactVariant) const;

private:
	struct Plan;
	struct Tuner
	{
		int variants;
		int chosen = -1;
		double ns[4] = {0, 0, 0, 0};
		double units[4] = {0, 0, 0, 0};
		int samples[4] = {0, 0, 0, 0};
	};
	struct PendingTiming
	{
		hipEvent_t a, b;
		Tuner* tuner;
		int variant;
		double units;
	};
	int pick(const Tuner& t) const;
	void fusePointStages(RaytraceNode* rt, RaytraceBatchItem& it);
	void harvest(bool wait);
	void timeStart(hipStream_t s, PendingTiming& p);
	std::shared_ptr<PinnedBuffer> hostSlots(size_t bytes);
	std::shared_ptr<Event> completionEvent();

	std::unique_ptr<Plan> plan;
	std::shared_ptr<Event> lastCompletion;
	std::unique_ptr<Stream> stream_;
	std::unique_ptr<Event> uploaded, scratchEvent;
	PinnedBuffer staging;
	DeviceBuffer items;
	DeviceBuffer slotsDev;
	std::vector<std::shared_ptr<PinnedBuffer>> slotPool; // host copies of kept counts and overflow flags
	std::vector<std::shared_ptr<Event>> eventPool;
	bool timing = false;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> timingEvents;
	size_t timingUsed = 0;
	bool autotune = true;
	int forcedRt = 0, forcedCp = 0;
	Tuner rtTune{2}, cpTune{2}; // tuned raytrace and compaction launch variants (gpu::kRaytraceTunedVariants, kCompactVariants)
	std::vector<PendingTiming> pending;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> tuneEvents; // free pairs
	// Captured kernel sequences, keyed by their launch signature; a signature is captured the second time it is
	// seen among the last kRecentSigs batches, and at most kCapturedGraphs graphs are kept (least recently used go).
	struct CapturedGraph
	{
		std::vector<int64_t> sig;
		hipGraph_t graph = nullptr;
		hipGraphExec_t exec = nullptr;
		uint64_t lastUse = 0;
		void release();
	};
	static constexpr size_t kCapturedGraphs = 4, kRecentSigs = 8;
	bool graphs = false;
	std::vector<CapturedGraph> captured;
	std::vector<std::vector<int64_t>> recentSigs;
	uint64_t graphClock = 0;
};

// ------------------------------------------------------------------------------------------------ node kinds
struct FromMat3x4fRaysNode : IRaysNode
{
	void setParameters(const Mat3x4f* r, int count);
	std::string getName() const override { return "FromMat3x4fRaysNode"; }
	BatchStage batchStage() const override { return BatchStage::Passive; }
	int rayCount() const override { return count; }
	const Mat3x4f* rays() const override { return buf.as<Mat3x4f>(); }
	const int32_t* traceOrder() const override { return ordered ? order.as<int32_t>() : nullptr; }

protected:
	void validateImpl() override;
	void enqueueImpl() override {}

private:
	DeviceBuffer buf;
	DeviceBuffer order;
	bool ordered = false;
	int count = 0;
};

struct SetRingIdsRaysNode : IRaysNodeSingleInput
{
	void setParameters(const int32_t* ids, int count);
	std::string getName() const override { return "SetRingIdsRaysNode"; }
	DeferredRays deferredRays() const override { return input->deferredRays(); }
	BatchStage batchStage() const override { return BatchStage::Passive; }
	const int32_t* ringIds() const override { return buf.as<int32_t>(); }
	int ringIdsCount() const override { return count; }

protected:
	void validateImpl() override;
	void enqueueImpl() override {}

private:
	DeviceBuffer buf;
	int count = 0;
};

struct SetRangeRaysNode : IRaysNodeSingleInput
{
	void setParameters(const Vec2f* r, int count);
	std::string getName() const override { return "SetRangeRaysNode"; }
	DeferredRays deferredRays() const override { return input->deferredRays(); }
	BatchStage batchStage() const override { return BatchStage::Passive; }
	const Vec2f* ranges() const override { return buf.as<Vec2f>(); }
	int rangesCount() const override { return count; }

protected:
	void validateImpl() override;
	void enqueueImpl() override {}

private:
	DeviceBuffer buf;
	int count = 0;
};

struct SetTimeOffsetsRaysNode : IRaysNodeSingleInput
{
	void setParameters(const float* offsets, int count);
	std::string getName() const override { return "SetTimeOffsetsRaysNode"; }
	DeferredRays deferredRays() const override { return input->deferredRays(); }
	BatchStage batchStage() const override { return BatchStage::Passive; }
	const float* timeOffsets() const override { return buf.as<float>(); }
	int timeOffsetsCount() const override { return count; }

protected:
	void validateImpl() override;
	void enqueueImpl() override {}

private:
	DeviceBuffer buf;
	int count = 0;
};

// When every consumer of the transformed rays is a raytrace node that does not output ray poses (possibly behind
// range / ring id / time offset nodes), the transform is not materialized: the raytrace kernel applies it to each
// ray it loads (the device matrix product rounds identically in every kernel, so the poses are bit-identical),
// which saves a pass over the rays.
struct TransformRaysNode : IRaysNodeSingleInput
{
	void setParameters(const Mat3x4f& tf) { transform = tf; }
	std::string getName() const override { return "TransformRaysNode"; }
	const Mat3x4f* rays() const override { return out.as<Mat3x4f>(); }
	DeferredRays deferredRays() const override;
	Mat3x4f cumulativeRayTransform() const override { return transform * input->cumulativeRayTransform(); }
	BatchStage batchStage() const override { return deferToRaytrace() ? BatchStage::Passive : BatchStage::PointOp; }
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;
	bool deferToRaytrace() const;

protected:
	void enqueueImpl() override;

private:
	bool computeDeferToRaytrace() const;
	Mat3x4f transform = Mat3x4f::identity();
	DeviceBuffer out;
	mutable uint64_t deferEpoch = ~uint64_t(0);
	mutable bool deferCached = false;
};

// Seeded noise: stream identity of one noise node.
struct NoiseStream
{
	uint64_t seed = 0;
	uint32_t run = 0;
	uint32_t tag = 0;
	void init(uint32_t tagValue);
};

struct GaussianNoiseAngularRaysNode : IRaysNodeSingleInput
{
	GaussianNoiseAngularRaysNode() { rng.init(1); }
	void setParameters(float mean, float stDev, int axis);
	std::string getName() const override { return "GaussianNoiseAngularRaysNode"; }
	const Mat3x4f* rays() const override { return out.as<Mat3x4f>(); }
	BatchStage batchStage() const override { return BatchStage::PointOp; }
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;
	NoiseStream rng;

protected:
	void enqueueImpl() override;

private:
	float mean = 0, stDev = 0;
	int axis = 3;
	DeviceBuffer out;
};

struct RaytraceNode : IPointsNode
{
	void setParameters(std::shared_ptr<Scene> s) { scene = std::move(s); }
	Scene& getScene() const { return *scene; }
	std::string getName() const override { return "RaytraceNode"; }
	BatchStage batchStage() const override { return BatchStage::Raytrace; }
	bool isDense() const override { return false; }
	bool hasField(rgl_field_t f) const override { return fields.count(f) != 0; }
	std::vector<rgl_field_t> fieldList() const override { return {fields.begin(), fields.end()}; }
	int capacity() const override { return raysNode ? returnCount() * raysNode->rayCount() : 0; }
	rgl_return_mode_t returnMode() const override { return mode; }
	int returnCount() const override { return static_cast<int>(static_cast<uint32_t>(mode) >> 24); }
	Mat3x4f lookAtOriginTransform() const override { return raysNode->cumulativeRayTransform().inverse(); }
	const void* fieldData(rgl_field_t f) override;

	void setVelocity(const Vec3f& lin, const Vec3f& ang)
	{
		linVel = lin;
		angVel = ang;
	}
	void setDistortion(bool on) { distortion = on; }
	void setNonHits(float nearD, float farD)
	{
		nearDist = nearD;
		farDist = farD;
	}
	void setMask(const int8_t* m, int count);
	void setDefaultIntensity(float v) { defaultIntensity = v; }
	void setReturnMode(rgl_return_mode_t m)
	{
		if (m != RGL_RETURN_UNKNOWN) {
			mode = m;
		}
	}
	void setBeamDivergence(float h, float v)
	{
		hHalf = h / 2.0f;
		vHalf = v / 2.0f;
	}
	// Traversal statistics of the last configuration (used by the benchmark for byte accounting).
	void traversalStats(int32_t* nodes, int32_t* tris, int32_t* insts);
	void footprint(unsigned long long* keys, uint32_t mask, uint32_t* full);
	void bruteForce(float* t, int32_t* inst, int32_t* prim);
	RaytraceParams buildParams(const SceneView& view) const;
	bool overflowed();
	// Prepares buffers and the scene on `s` and returns this node's launch description (firstBlock unset).
	RaytraceBatchItem batchItem(hipStream_t s);
	// Host copy of the overflow flag of a batched run: slot `index` of `slots`.
	void setOverflowHost(std::shared_ptr<PinnedBuffer> slots, int index)
	{
		overflowSlots = std::move(slots);
		overflowSlot = index;
	}
	// HIP-event timing of the raytrace kernel launches (on the graph stream) while enabled.
	void setTiming(bool enable);
	void getTiming(double* totalMs, int32_t* launches);
	~RaytraceNode() override;

protected:
	void validateImpl() override;
	void enqueueImpl() override;

private:
	std::set<rgl_field_t> findFieldsToCompute();

	std::shared_ptr<IRaysNode> raysNode;
	std::shared_ptr<Scene> scene = Scene::defaultShared();
	std::set<rgl_field_t> fields;
	std::map<rgl_field_t, std::unique_ptr<DeviceBuffer>> data;
	DeviceBuffer defaultRange;
	DeviceBuffer maskBuf;
	int maskCount = 0;
	bool hasMask = false;
	bool distortion = false;
	Vec3f linVel{0, 0, 0}, angVel{0, 0, 0};
	float nearDist = INFINITY, farDist = INFINITY;
	float defaultIntensity = 0.0f;
	rgl_return_mode_t mode = RGL_RETURN_FIRST;
	float hHalf = 0.0f, vHalf = 0.0f;
	PinnedBuffer overflowHost;
	std::shared_ptr<PinnedBuffer> overflowSlots;
	int overflowSlot = 0;
	bool timing = false;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> timingEvents;
	size_t timingUsed = 0;
};

struct TransformPointsNode : IPointsNodeSingleInput
{
	void setParameters(const Mat3x4f& tf) { transform = tf; }
	std::string getName() const override { return "TransformPointsNode"; }
	std::vector<rgl_field_t> requiredFields() const override { return {RGL_FIELD_XYZ_VEC3_F32}; }
	Mat3x4f lookAtOriginTransform() const override { return transform.inverse() * input->lookAtOriginTransform(); }
	const void* fieldData(rgl_field_t f) override;
	BatchStage batchStage() const override { return BatchStage::PointOp; }
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;

protected:
	void enqueueImpl() override;

private:
	Mat3x4f transform = Mat3x4f::identity();
	DeviceBuffer out;
};

struct GaussianNoiseAngularHitpointNode : IPointsNodeSingleInput
{
	GaussianNoiseAngularHitpointNode() { rng.init(2); }
	void setParameters(float mean, float stDev, int axis);
	std::string getName() const override { return "GaussianNoiseAngularHitpointNode"; }
	std::vector<rgl_field_t> requiredFields() const override { return {RGL_FIELD_XYZ_VEC3_F32}; }
	const void* fieldData(rgl_field_t f) override;
	BatchStage batchStage() const override { return BatchStage::PointOp; }
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;
	NoiseStream rng;

protected:
	void validateImpl() override;
	void enqueueImpl() override;

private:
	float mean = 0, stDev = 0;
	int axis = 3;
	bool withDistance = false;
	DeviceBuffer outXyz, outDistance;
};

struct GaussianNoiseDistanceNode : IPointsNodeSingleInput
{
	GaussianNoiseDistanceNode() { rng.init(3); }
	void setParameters(float mean, float base, float rise);
	std::string getName() const override { return "GaussianNoiseDistanceNode"; }
	std::vector<rgl_field_t> requiredFields() const override
	{
		return {RGL_FIELD_XYZ_VEC3_F32, RGL_FIELD_DISTANCE_F32};
	}
	const void* fieldData(rgl_field_t f) override;
	BatchStage batchStage() const override { return BatchStage::PointOp; }
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;
	NoiseStream rng;

protected:
	void enqueueImpl() override;

private:
	float mean = 0, base = 0, rise = 0;
	DeviceBuffer outXyz, outDistance;
};

struct CompactByFieldPointsNode : IPointsNodeSingleInput
{
	void setParameters(rgl_field_t f) { by = f; }
	std::string getName() const override { return "CompactByFieldPointsNode"; }
	BatchStage batchStage() const override { return BatchStage::Compact; }
	std::vector<rgl_field_t> requiredFields() const override { return {by}; }
	bool isDense() const override { return true; }
	const int32_t* countDevice() const override { return count.as<int32_t>(); }
	int hostCount() override;
	const void* fieldData(rgl_field_t f) override;
	// Prepares output buffers and returns this node's compaction description (firstTile unset) for tiles of
	// compaction variant `variant`.
	CompactItem compactItem(FieldCopyList& list, int variant = 0);
	// Host copy of the kept count of a batched run: slot `index` of `slots`.
	void setCountHost(std::shared_ptr<PinnedBuffer> slots, int index)
	{
		countSlots = std::move(slots);
		countSlot = index;
	}

protected:
	void enqueueImpl() override;

private:
	rgl_field_t by = RGL_FIELD_IS_HIT_I32;
	std::map<rgl_field_t, std::unique_ptr<DeviceBuffer>> data;
	DeviceBuffer count, tileCounts;
	PinnedBuffer countHost;
	std::shared_ptr<PinnedBuffer> countSlots;
	int countSlot = 0;
};

struct YieldPointsNode : IPointsNodeSingleInput
{
	void setParameters(const std::vector<rgl_field_t>& f);
	std::string getName() const override { return "YieldPointsNode"; }
	BatchStage batchStage() const override { return prefetch ? BatchStage::PointOp : BatchStage::Passive; }
	std::vector<rgl_field_t> requiredFields() const override { return fields; }
	bool hasField(rgl_field_t f) const override;
	// Prefetch: every run also copies the yielded fields' live points into pinned host memory (one kernel writing
	// over the bus), so results are host-resident when the run completes.
	void setPrefetch(bool on);
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;
	// Host copy of a prefetched field of the last run, or nullptr.
	const void* hostData(rgl_field_t f) const;

protected:
	void enqueueImpl() override;

private:
	std::vector<rgl_field_t> fields;
	bool prefetch = false;
	std::map<rgl_field_t, MappedHostBuffer> host;
	std::set<rgl_field_t> prefetched;
};

struct FormatPointsNode : IPointsNodeSingleInput
{
	void setParameters(const std::vector<rgl_field_t>& f);
	std::string getName() const override { return "FormatPointsNode"; }
	std::vector<rgl_field_t> requiredFields() const override { return fields; }
	bool hasField(rgl_field_t f) const override;
	const void* fieldData(rgl_field_t f) override;
	size_t fieldPointSize(rgl_field_t f) const override;
	BatchStage batchStage() const override { return BatchStage::PointOp; }
	void pointOp(PointOpItem& it, FieldCopyList& layout) override;

protected:
	void enqueueImpl() override;

private:
	FieldCopyList makeLayout();
	std::vector<rgl_field_t> fields;
	DeviceBuffer out;
};

// Radar range x azimuth cell binning with RCS (build-defined; the reference clusters radar hits by scopes instead,
// src/graph/RadarPostprocessPointsNode.cpp:88-130). One output point per non-empty cell, in cell order.
struct RadarBinPointsNode : IPointsNodeSingleInput
{
	void setParameters(float rangeStep, float rangeMax, float azStep, float azMin, float azMax, float raySolidAngle);
	std::string getName() const override { return "RadarBinPointsNode"; }
	std::vector<rgl_field_t> requiredFields() const override
	{
		return {RGL_FIELD_XYZ_VEC3_F32, RGL_FIELD_DISTANCE_F32, RGL_FIELD_INCIDENT_ANGLE_F32, RGL_FIELD_IS_HIT_I32};
	}
	bool isDense() const override { return true; }
	bool hasField(rgl_field_t f) const override;
	std::vector<rgl_field_t> fieldList() const override;
	int capacity() const override { return bp.rangeCells * bp.azCells; }
	const int32_t* countDevice() const override { return count.as<int32_t>(); }
	int hostCount() override;
	rgl_return_mode_t returnMode() const override { return RGL_RETURN_UNKNOWN; }
	int returnCount() const override { return 1; }
	Mat3x4f lookAtOriginTransform() const override { return Mat3x4f::identity(); } // cells are in the sensor frame
	const void* fieldData(rgl_field_t f) override;

protected:
	void enqueueImpl() override;

private:
	RadarBinParams bp{0.2f, 0.017453292f, -1.0471976f, 1e-4f, 1, 1};
	DeviceBuffer sums, counts, keep, count, tileCounts;
	std::map<rgl_field_t, DeviceBuffer> dense, out;
	PinnedBuffer countHost;
};

// Radar detections with the reference's semantics (src/graph/RadarPostprocessPointsNode.cpp): the dense input
// points are clustered by distance scope and separation thresholds; one output point per cluster, carrying the
// fields of the cluster's directional centre point plus RCS (dBsm), POWER (dBm, with noise), NOISE and SNR (dB).
// The noise is this library's seeded stream (tag 6), not the reference's unseeded generator.
struct RadarPostprocessPointsNode : IPointsNodeSingleInput
{
	RadarPostprocessPointsNode() { rng.init(6); }
	void setParameters(const std::vector<RadarScope>& scopes, float azimuthStep, float elevationStep, float frequency,
	                   float powerTransmitted, float deviceGain, float noiseMean, float noiseStDev);
	std::string getName() const override { return "RadarPostprocessPointsNode"; }
	std::vector<rgl_field_t> requiredFields() const override
	{
		return {RGL_FIELD_DISTANCE_F32, RGL_FIELD_AZIMUTH_F32,     RGL_FIELD_ELEVATION_F32,  RGL_FIELD_RADIAL_SPEED_F32,
		        RGL_FIELD_RAY_POSE_MAT3x4_F32, RGL_FIELD_NORMAL_VEC3_F32, RGL_FIELD_XYZ_VEC3_F32};
	}
	bool isDense() const override { return true; }
	bool hasField(rgl_field_t f) const override;
	std::vector<rgl_field_t> fieldList() const override;
	const int32_t* countDevice() const override { return info.as<int32_t>(); }
	int hostCount() override;
	const void* fieldData(rgl_field_t f) override;
	BatchStage batchStage() const override { return BatchStage::RadarPost; }
	// Launch description of this run (advances the noise stream); `gather` receives the centre-point field list,
	// which it.gather must point to in device memory.
	RadarItem radarItem(FieldCopyList& gather);
	// Batched runs: the cluster count and error bits come back in a batch's host slots (index, index + 1).
	void setInfoHost(std::shared_ptr<PinnedBuffer> slots, int index)
	{
		infoSlots = std::move(slots);
		infoSlot = index;
	}
	NoiseStream rng;

protected:
	void validateImpl() override;
	void enqueueImpl() override;

private:
	RadarPostParams params{};
	DeviceBuffer terms, work, centres, power, rcs, noise, snr, info, gatherList;
	std::map<rgl_field_t, DeviceBuffer> gathered;
	PinnedBuffer infoHost;
	std::shared_ptr<PinnedBuffer> infoSlots;
	int infoSlot = 0;
};

struct FromArrayPointsNode : IPointsNode
{
	void setParameters(const void* points, int count, const std::vector<rgl_field_t>& f);
	std::string getName() const override { return "FromArrayPointsNode"; }
	BatchStage batchStage() const override { return BatchStage::Passive; }
	bool isDense() const override { return data.count(RGL_FIELD_IS_HIT_I32) == 0; }
	bool hasField(rgl_field_t f) const override { return data.count(f) != 0; }
	std::vector<rgl_field_t> fieldList() const override;
	int capacity() const override { return width; }
	rgl_return_mode_t returnMode() const override { return RGL_RETURN_UNKNOWN; }
	int returnCount() const override { return 1; }
	const void* fieldData(rgl_field_t f) override;

protected:
	void validateImpl() override;
	void enqueueImpl() override {}

private:
	std::map<rgl_field_t, std::unique_ptr<DeviceBuffer>> data;
	int width = 0;
};

// Registry of live API objects of one kind (handles are raw pointers; src/APIObject.hpp:34-104).
template<typename T>
struct Registry
{
	static std::map<const void*, std::shared_ptr<T>>& map()
	{
		static std::map<const void*, std::shared_ptr<T>> m;
		return m;
	}
	static void add(const std::shared_ptr<T>& p) { map()[p.get()] = p; }
	static bool contains(const void* p) { return map().count(p) != 0; }
	static std::shared_ptr<T> get(const void* p, const char* kind)
	{
		auto it = map().find(p);
		if (it == map().end()) {
			char buf[64];
			std::snprintf(buf, sizeof(buf), "%p", p);
			throw InvalidAPIObject(std::string("RGL API Error: Object does not exist: ") + kind + " " + buf);
		}
		return it->second;
	}
	static void release(const void* p, const char* kind)
	{
		get(p, kind);
		map().erase(p);
	}
};

// Node registry key is the Node* of the most-derived object.
void registerNode(const Node::Ptr& n);
Node::Ptr lookupNode(const void* handle);
bool nodeAlive(const void* handle);
void releaseNode(const Node* n);
std::vector<Node::Ptr> allNodes();

template<typename T>
std::shared_ptr<T> Node::exactlyOneInputOfType() const
{
	std::vector<std::shared_ptr<T>> typed;
	std::string names;
	for (auto& in : inputs) {
		if (auto t = std::dynamic_pointer_cast<T>(in)) {
			typed.push_back(t);
		}
		if (!names.empty()) {
			names += ", ";
		}
		names += in->getName();
	}
	if (typed.size() != 1) {
		throw InvalidPipeline("looked for " + std::string(T::kTypeName) + ", but found [" + names + "]");
	}
	return typed[0];
}

} // namespace rgl
