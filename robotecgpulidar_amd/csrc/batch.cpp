// Batched execution of many independent graphs (one per sensor) on one stream.
//
// The reference runs each graph on its own stream and launches every node's kernels separately
// (src/graph/GraphRunCtx.cpp:144-211, src/gpu/RaytraceNode.cpp + OptiX launch per graph). A scan of a 16-128
// channel lidar is only a few thousand to a few hundred thousand rays, far too few to fill 256 CUs, so on MI355X
// a frame of many sensors is latency-bound when the graphs run one after the other. Here the nodes of all graphs
// are list-scheduled into steps: each step launches every ready node of one stage kind in ONE launch (a raytrace
// launch over every sensor's beams, possibly on different scenes; one elementwise launch for all transform /
// noise / format nodes; one pair of compaction launches). Cheap kinds are drained first and raytracing last, so
// graphs of different shapes (e.g. with and without ray noise) still share one raytrace launch. The launch
// descriptors of the whole batch go to the device in a single copy; kept counts and overflow flags of every graph
// come back in a single copy. Each item runs the same device code as the graph's own kernels with the same
// parameters, so a graph's results do not depend on whether it ran alone or in a batch.
#include <algorithm>
#include <cstring>

#include "graph.hpp"
#include "kernels.hpp"

namespace rgl {

namespace {
constexpr size_t kAlign = 256;

size_t alignUp(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
} // namespace

struct GraphBatch::Plan
{
	struct Level
	{
		std::vector<RaytraceBatchItem> rt, rtm; // single-sample beams; divergent beams (one wave per beam)
		std::vector<RaytraceNode*> rtNodes, rtmNodes;
		std::vector<PointOpItem> op;
		std::vector<int> opLayout; // index into lists, -1 = none
		std::vector<CompactItem> cp;
		std::vector<int> cpList;
		std::vector<CompactByFieldPointsNode*> cpNodes;
		std::vector<RadarItem> rp;
		std::vector<int> rpList;
		std::vector<RadarPostprocessPointsNode*> rpNodes;
		std::vector<Node*> other;
		int rtBlocks = 0, rtmBlocks = 0, opBlocks = 0, cpTiles = 0, rpBlocks = 0;
		size_t rtOff = 0, rtmOff = 0, opOff = 0, cpOff = 0, rpOff = 0, claimOff = 0;
		void clear()
		{
			rt.clear();
			rtNodes.clear();
			rtm.clear();
			rtmNodes.clear();
			op.clear();
			opLayout.clear();
			cp.clear();
			cpList.clear();
			cpNodes.clear();
			rp.clear();
			rpList.clear();
			rpNodes.clear();
			other.clear();
			rtBlocks = rtmBlocks = opBlocks = cpTiles = rpBlocks = 0;
		}
	};
	std::vector<Level> levels; // steps, in launch order
	// Prefetch copies of yield nodes into pinned host memory: nothing in the batch reads them, so they run after every
	// other launch, in chunks of consecutive items, each chunk followed by an event that its nodes' results wait on.
	std::vector<PointOpItem> toHost;
	std::vector<int> toHostLayout;
	std::vector<Node*> toHostNodes;
	std::vector<int> chunkFirst; // first item of each chunk, then the item count
	std::vector<int> chunkBlocks;
	size_t toHostOff = 0;
	std::vector<Node*> readyOp, readyRt, readyCp, readyRp, passive;
	std::vector<FieldCopyList> lists;
	std::vector<const int32_t*> gather;
	std::vector<std::shared_ptr<GraphRunCtx>> ctxs;
	std::vector<Scene*> scenes;
};

GraphBatch& GraphBatch::instance()
{
	// Never destroyed: its streams and buffers must not be released after the HIP runtime shuts down.
	static GraphBatch* batch = new GraphBatch();
	return *batch;
}

std::shared_ptr<PinnedBuffer> GraphBatch::hostSlots(size_t bytes)
{
	for (auto& p : slotPool) {
		if (p.use_count() == 1 && p->capacity() >= bytes) {
			return p;
		}
	}
	auto p = std::make_shared<PinnedBuffer>();
	p->ensure(bytes);
	slotPool.push_back(p);
	return p;
}

std::shared_ptr<Event> GraphBatch::completionEvent()
{
	for (auto& e : eventPool) {
		if (e.use_count() == 1) {
			return e;
		}
	}
	eventPool.push_back(std::make_shared<Event>());
	return eventPool.back();
}

GraphBatch::GraphBatch() = default;

// Folds the elementwise stages that directly follow a single-sample raytrace node into its batched launch: a chain
// of points-transform and distance-noise nodes starting at the raytrace node's only output, continued through nodes
// that have a single output, at most two stages. Every folded node still writes its own output buffers (the results
// of each node stay readable), so what is saved is the re-read of XYZ and DISTANCE and one launch per stage; the
// arithmetic is the shared elementwise code, so results are bit-identical to separate launches.
void GraphBatch::fusePointStages(RaytraceNode* rt, RaytraceBatchItem& it)
{
	it.fusedCount = 0;
	const void* xyz = it.out.xyz;
	const void* distance = it.out.distance;
	const int points = it.p.rayCount * it.p.returnCount;
	Node* prev = rt;
	while (it.fusedCount < 2 && prev->getOutputs().size() == 1) {
		Node* n = prev->getOutputs()[0].get();
		const bool transform = dynamic_cast<TransformPointsNode*>(n) != nullptr;
		const bool noise = dynamic_cast<GaussianNoiseDistanceNode*>(n) != nullptr;
		if ((!transform && !noise) || n->getInputs().size() != 1 || xyz == nullptr || (noise && distance == nullptr)) {
			break;
		}
		PointOpItem& op = it.fused[it.fusedCount];
		op = PointOpItem{};
		FieldCopyList unused;
		n->pointOp(op, unused);
		if (op.in0 != xyz || (noise && op.in1 != distance) || op.count != nullptr || op.capacity != points) {
			throw std::logic_error("batch: a folded point stage does not read the raytrace outputs");
		}
		++it.fusedCount;
		++fusedStages;
		n->batchFused = true;
		xyz = op.out0;
		if (noise) {
			distance = op.out1;
		}
		prev = n;
	}
}

namespace {
constexpr double kTuneMinUnits = 100000.0;
constexpr int64_t kTraceOrderMinLaunchRays = int64_t(1) << 20; // about three occupancy generations of 64-ray waves
// Capacity bytes per prefetch launch: about a third of a millisecond of the host link, so the host starts copying the
// first sensors' points while the last ones are still on the way.
#ifndef RGL_PREFETCH_CHUNK_MIB
#define RGL_PREFETCH_CHUNK_MIB 16
#endif
constexpr size_t kPrefetchChunkBytes = static_cast<size_t>(RGL_PREFETCH_CHUNK_MIB) << 20;
constexpr int kTuneSamples = 3;
} // namespace

int GraphBatch::pick(const Tuner& t) const
{
	if (!autotune) {
		return &t == &rtTune ? forcedRt : forcedCp;
	}
	if (t.chosen >= 0) {
		return t.chosen;
	}
	int best = 0;
	for (int v = 1; v < t.variants; ++v) {
		if (t.samples[v] < t.samples[best]) {
			best = v;
		}
	}
	return best;
}

// Folds completed tuning timings into their tuners and settles a tuner once every variant has enough samples.
void GraphBatch::harvest(bool wait)
{
	for (size_t i = 0; i < pending.size();) {
		PendingTiming& p = pending[i];
		if (wait) {
			RGL_HIP(hipEventSynchronize(p.b));
		} else {
			const hipError_t e = hipEventQuery(p.b);
			if (e == hipErrorNotReady) {
				++i;
				continue;
			}
			RGL_HIP(e);
		}
		float ms = 0.0f;
		RGL_HIP(hipEventElapsedTime(&ms, p.a, p.b));
		Tuner& t = *p.tuner;
		if (t.chosen < 0) {
			t.ns[p.variant] += 1e6 * ms;
			t.units[p.variant] += p.units;
			++t.samples[p.variant];
			bool done = true;
			for (int v = 0; v < t.variants; ++v) {
				done = done && t.samples[v] >= kTuneSamples;
			}
			if (done) {
				int best = 0;
				for (int v = 1; v < t.variants; ++v) {
					if (t.ns[v] / t.units[v] < t.ns[best] / t.units[best]) {
						best = v;
					}
				}
				t.chosen = best;
			}
		}
		tuneEvents.emplace_back(p.a, p.b);
		pending[i] = pending.back();
		pending.pop_back();
	}
}

void GraphBatch::timeStart(hipStream_t s, PendingTiming& p)
{
	if (tuneEvents.empty()) {
		hipEvent_t a, b;
		RGL_HIP(hipEventCreate(&a));
		RGL_HIP(hipEventCreate(&b));
		tuneEvents.emplace_back(a, b);
	}
	p.a = tuneEvents.back().first;
	p.b = tuneEvents.back().second;
	tuneEvents.pop_back();
	RGL_HIP(hipEventRecord(p.a, s));
}

void GraphBatch::setAutotune(bool enable)
{
	synchronize();
	harvest(true);
	autotune = enable;
	forcedRt = forcedCp = 0;
	rtTune = Tuner{gpu::kRaytraceTunedVariants};
	cpTune = Tuner{gpu::kCompactVariants};
}

void GraphBatch::forceVariants(int raytraceVariant, int compactVariant)
{
	if (raytraceVariant < 0 || raytraceVariant >= gpu::kRaytraceVariants || compactVariant < 0 ||
	    compactVariant >= gpu::kCompactVariants) {
		throw std::invalid_argument("no such launch variant");
	}
	setAutotune(false);
	forcedRt = raytraceVariant;
	forcedCp = compactVariant;
}

void GraphBatch::autotuneChoice(int32_t* raytraceVariant, int32_t* compactVariant) const
{
	*raytraceVariant = autotune ? rtTune.chosen : forcedRt;
	*compactVariant = autotune ? cpTune.chosen : forcedCp;
}

void GraphBatch::CapturedGraph::release()
{
	if (exec != nullptr) {
		(void) hipGraphExecDestroy(exec);
		exec = nullptr;
	}
	if (graph != nullptr) {
		(void) hipGraphDestroy(graph);
		graph = nullptr;
	}
}

void GraphBatch::setGraphs(bool enable)
{
	synchronize();
	for (auto& cg : captured) {
		cg.release();
	}
	captured.clear();
	recentSigs.clear();
	graphs = enable;
}

GraphBatch::~GraphBatch()
{
	if (stream_) {
		(void) hipStreamSynchronize(stream_->get());
	}
	for (auto& cg : captured) {
		cg.release();
	}
	for (auto& [a, b] : timingEvents) {
		(void) hipEventDestroy(a);
		(void) hipEventDestroy(b);
	}
	for (auto& p : pending) {
		tuneEvents.emplace_back(p.a, p.b);
	}
	for (auto& [a, b] : tuneEvents) {
		(void) hipEventDestroy(a);
		(void) hipEventDestroy(b);
	}
}

bool GraphBatch::busy() const
{
	if (!lastCompletion) {
		return false;
	}
	const hipError_t e = hipEventQuery(lastCompletion->get());
	if (e == hipErrorNotReady) {
		return true;
	}
	RGL_HIP(e);
	return false;
}

void GraphBatch::synchronize()
{
	if (stream_) {
		stream_->sync();
	}
}

void GraphBatch::setTiming(bool enable)
{
	synchronize();
	timing = enable;
	timingUsed = 0;
}

void GraphBatch::getTiming(double* totalMs, int32_t* launches)
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

void GraphBatch::run(const std::vector<Node::Ptr>& entryNodes)
{
	if (entryNodes.empty()) {
		return;
	}
	if (!stream_) {
		stream_ = std::make_unique<Stream>();
		uploaded = std::make_unique<Event>();
		scratchEvent = std::make_unique<Event>();
		plan = std::make_unique<Plan>();
	}
	Plan& pl = *plan;
	pl.ctxs.clear();
	for (auto& n : entryNodes) {
		if (!n->hasCtx()) {
			GraphRunCtx::createAndAttach(n);
		}
		auto c = n->getCtx();
		if (std::find(pl.ctxs.begin(), pl.ctxs.end(), c) == pl.ctxs.end()) {
			pl.ctxs.push_back(c);
		}
	}
	// Every graph is validated before anything is enqueued, so an invalid graph fails the whole batch cleanly.
	size_t nodeCount = 0;
	for (auto& c : pl.ctxs) {
		for (auto& n : c->validatedOrder()) {
			n->batchLevel = static_cast<int32_t>(n->getInputs().size()); // inputs not yet scheduled
			n->batchFused = false;
			++nodeCount;
		}
	}
	const hipStream_t s = stream_->get();
	uploaded->sync(); // the previous batch's descriptor upload has read the staging buffer
	harvest(false);
	const int rtVariant = pick(rtTune);
	const int cpVariant = pick(cpTune);
	for (auto& c : pl.ctxs) {
		c->beginBatch(s, *scratchEvent);
	}
	std::shared_ptr<Event> countsReady;
	std::vector<std::shared_ptr<Event>> chunkReady;
	try {
		pl.lists.clear();
		pl.gather.clear();
		pl.toHost.clear();
		pl.toHostLayout.clear();
		pl.toHostNodes.clear();
		pl.chunkFirst.clear();
		pl.chunkBlocks.clear();
		pl.scenes.clear();
		pl.readyOp.clear();
		pl.readyRt.clear();
		pl.readyCp.clear();
		pl.readyRp.clear();
		pl.passive.clear();
		int steps = 0;
		auto newStep = [&]() -> Plan::Level& {
			if (pl.levels.size() <= static_cast<size_t>(steps)) {
				pl.levels.emplace_back();
			}
			Plan::Level& L = pl.levels[steps++];
			L.clear();
			return L;
		};
		auto makeReady = [&](Node* n) {
			if (n->batchFused) {
				pl.passive.push_back(n); // its work runs inside its raytrace node's launch
				return;
			}
			switch (n->batchStage()) {
				case BatchStage::Passive:
				case BatchStage::Other: pl.passive.push_back(n); break;
				case BatchStage::PointOp: pl.readyOp.push_back(n); break;
				case BatchStage::Compact: pl.readyCp.push_back(n); break;
				case BatchStage::Raytrace: pl.readyRt.push_back(n); break;
				case BatchStage::RadarPost: pl.readyRp.push_back(n); break;
			}
		};
		auto release = [&](Node* n) {
			for (auto& o : n->getOutputs()) {
				if (--o->batchLevel == 0) {
					makeReady(o.get());
				}
			}
		};
		for (auto& c : pl.ctxs) {
			for (auto& n : c->validatedOrder()) {
				if (n->batchLevel == 0) {
					makeReady(n.get());
				}
			}
		}
		// Launch descriptors; buffers of every node are allocated here, before the first launch.
		size_t scheduled = 0;
		std::vector<Node*> group;
		while (true) {
			while (!pl.passive.empty()) {
				Node* n = pl.passive.back();
				pl.passive.pop_back();
				++scheduled;
				if (n->batchStage() == BatchStage::Other) {
					newStep().other.push_back(n);
				}
				release(n);
			}
			if (!pl.readyOp.empty()) {
				group.swap(pl.readyOp);
			} else if (!pl.readyCp.empty()) {
				group.swap(pl.readyCp);
			} else if (!pl.readyRp.empty()) {
				group.swap(pl.readyRp);
			} else if (!pl.readyRt.empty()) {
				group.swap(pl.readyRt);
			} else {
				break;
			}
			Plan::Level& L = newStep();
			for (Node* n : group) {
				switch (n->batchStage()) {
					case BatchStage::Raytrace: {
						auto* rt = dynamic_cast<RaytraceNode*>(n);
						RaytraceBatchItem it = rt->batchItem(s);
						if (gpu::beamIsMulti(it.p)) {
							it.firstBlock = L.rtmBlocks;
							L.rtmBlocks += gpu::raytraceBlocks(it.p.rayCount, true, 0);
							L.rtm.push_back(it);
							L.rtmNodes.push_back(rt);
						} else {
							fusePointStages(rt, it);
							it.firstBlock = L.rtBlocks;
							L.rtBlocks += gpu::raytraceBlocks(it.p.rayCount, false, rtVariant);
							L.rt.push_back(it);
							L.rtNodes.push_back(rt);
						}
						Scene* sc = &rt->getScene();
						if (std::find(pl.scenes.begin(), pl.scenes.end(), sc) == pl.scenes.end()) {
							pl.scenes.push_back(sc);
						}
						break;
					}
					case BatchStage::PointOp: {
						PointOpItem it{};
						FieldCopyList layout;
						n->pointOp(it, layout);
						if (it.op == kOpToHost) {
							pl.toHostLayout.push_back(static_cast<int>(pl.lists.size()));
							pl.lists.push_back(layout);
							pl.toHost.push_back(it);
							pl.toHostNodes.push_back(n);
							break;
						}
						it.firstBlock = L.opBlocks;
						L.opBlocks += gpu::pointOpBlocks(it.capacity);
						L.op.push_back(it);
						if (it.op == kOpFormat) {
							L.opLayout.push_back(static_cast<int>(pl.lists.size()));
							pl.lists.push_back(layout);
						} else {
							L.opLayout.push_back(-1);
						}
						break;
					}
					case BatchStage::Compact: {
						auto* cn = dynamic_cast<CompactByFieldPointsNode*>(n);
						FieldCopyList list;
						CompactItem it = cn->compactItem(list, cpVariant);
						it.firstTile = L.cpTiles;
						L.cpTiles += it.numTiles;
						L.cp.push_back(it);
						L.cpList.push_back(static_cast<int>(pl.lists.size()));
						L.cpNodes.push_back(cn);
						pl.lists.push_back(list);
						break;
					}
					case BatchStage::RadarPost: {
						auto* rn = dynamic_cast<RadarPostprocessPointsNode*>(n);
						FieldCopyList list;
						RadarItem it = rn->radarItem(list);
						it.firstBlock = L.rpBlocks;
						L.rpBlocks += gpu::radarBlocks(it.capacity);
						L.rp.push_back(it);
						L.rpList.push_back(static_cast<int>(pl.lists.size()));
						L.rpNodes.push_back(rn);
						pl.lists.push_back(list);
						break;
					}
					default: throw std::logic_error("batch: unexpected stage kind");
				}
			}
			// Direction-tile trace order only in launches of several occupancy generations: there it keeps the rays of
			// a wave on the same nodes (64-lidar batch: -1.8% raytrace time), while a launch of about one generation
			// lasts as long as its slowest waves, and grouping costly patches of directions into the same waves
			// lengthens those (single 262k-ray lidar: +17% with the order).
			{
				int64_t launchRays = 0;
				for (const auto& it : L.rt) {
					launchRays += it.p.rayCount;
				}
				if (launchRays < kTraceOrderMinLaunchRays) {
					for (auto& it : L.rt) {
						it.p.order = nullptr;
					}
				}
			}
			scheduled += group.size();
			for (Node* n : group) {
				release(n);
			}
			group.clear();
		}
		if (scheduled != nodeCount) {
			throw std::logic_error("batch: some nodes were never ready");
		}
		const int maxLevel = steps - 1;

		// Prefetch chunks of about kPrefetchChunkBytes of capacity each, in item order (the order of the graphs):
		// a chunk's host copies can start while the link still carries the next chunks.
		size_t chunkBytes = 0;
		for (size_t i = 0; i < pl.toHost.size(); ++i) {
			PointOpItem& it = pl.toHost[i];
			if (pl.chunkFirst.empty() || chunkBytes >= kPrefetchChunkBytes) {
				pl.chunkFirst.push_back(static_cast<int>(i));
				pl.chunkBlocks.push_back(0);
				chunkBytes = 0;
			}
			const FieldCopyList& layout = pl.lists[pl.toHostLayout[i]];
			for (int f = 0; f < layout.count; ++f) {
				chunkBytes += static_cast<size_t>(std::max(it.capacity, 0)) * static_cast<size_t>(layout.f[f].size);
			}
			it.firstBlock = pl.chunkBlocks.back();
			pl.chunkBlocks.back() += gpu::pointOpBlocks(it.capacity);
		}
		pl.chunkFirst.push_back(static_cast<int>(pl.toHost.size()));

		// Host slots: kept counts of compaction nodes, cluster counts and error bits of radar postprocess nodes, then
		// overflow flags of raytrace nodes.
		for (int l = 0; l <= maxLevel; ++l) {
			for (auto& it : pl.levels[l].cp) {
				pl.gather.push_back(it.countOut);
			}
		}
		for (int l = 0; l <= maxLevel; ++l) {
			for (auto& it : pl.levels[l].rp) {
				pl.gather.push_back(it.info);
				pl.gather.push_back(it.info + 1);
			}
		}
		for (int l = 0; l <= maxLevel; ++l) {
			for (auto& it : pl.levels[l].rt) {
				pl.gather.push_back(it.sc.overflow);
			}
			for (auto& it : pl.levels[l].rtm) {
				pl.gather.push_back(it.sc.overflow);
			}
		}

		// One staging image: field lists, gather sources, then per-level item arrays.
		size_t off = 0;
		auto reserve = [&](size_t bytes) {
			const size_t o = alignUp(off);
			off = o + bytes;
			return o;
		};
		const size_t listsOff = reserve(sizeof(FieldCopyList) * pl.lists.size());
		const size_t gatherOff = reserve(sizeof(const int32_t*) * pl.gather.size());
		for (int l = 0; l <= maxLevel; ++l) {
			Plan::Level& L = pl.levels[l];
			L.rtOff = reserve(sizeof(RaytraceBatchItem) * L.rt.size());
			L.claimOff = reserve(sizeof(uint32_t)); // block claims of the persistent raytrace launch, zero at start
			L.rtmOff = reserve(sizeof(RaytraceBatchItem) * L.rtm.size());
			L.opOff = reserve(sizeof(PointOpItem) * L.op.size());
			L.cpOff = reserve(sizeof(CompactItem) * L.cp.size());
			L.rpOff = reserve(sizeof(RadarItem) * L.rp.size());
		}
		pl.toHostOff = reserve(sizeof(PointOpItem) * pl.toHost.size());
		const size_t total = std::max<size_t>(off, 1);
		if (total > items.capacity()) {
			synchronize(); // earlier batches may still read the descriptors being reallocated
		}
		char* dev = static_cast<char*>(items.ensure(total));
		auto* devLists = reinterpret_cast<const FieldCopyList*>(dev + listsOff);
		char* host = static_cast<char*>(staging.ensure(total));
		if (!pl.lists.empty()) {
			std::memcpy(host + listsOff, pl.lists.data(), sizeof(FieldCopyList) * pl.lists.size());
		}
		if (!pl.gather.empty()) {
			std::memcpy(host + gatherOff, pl.gather.data(), sizeof(const int32_t*) * pl.gather.size());
		}
		for (int l = 0; l <= maxLevel; ++l) {
			Plan::Level& L = pl.levels[l];
			for (size_t i = 0; i < L.op.size(); ++i) {
				if (L.opLayout[i] >= 0) {
					L.op[i].format = devLists + L.opLayout[i];
				}
			}
			for (size_t i = 0; i < L.cp.size(); ++i) {
				L.cp[i].fields = devLists + L.cpList[i];
			}
			for (size_t i = 0; i < L.rp.size(); ++i) {
				L.rp[i].gather = devLists + L.rpList[i];
			}
			if (!L.rt.empty()) {
				std::memcpy(host + L.rtOff, L.rt.data(), sizeof(RaytraceBatchItem) * L.rt.size());
			}
			std::memset(host + L.claimOff, 0, sizeof(uint32_t));
			if (!L.rtm.empty()) {
				std::memcpy(host + L.rtmOff, L.rtm.data(), sizeof(RaytraceBatchItem) * L.rtm.size());
			}
			if (!L.op.empty()) {
				std::memcpy(host + L.opOff, L.op.data(), sizeof(PointOpItem) * L.op.size());
			}
			if (!L.cp.empty()) {
				std::memcpy(host + L.cpOff, L.cp.data(), sizeof(CompactItem) * L.cp.size());
			}
			if (!L.rp.empty()) {
				std::memcpy(host + L.rpOff, L.rp.data(), sizeof(RadarItem) * L.rp.size());
			}
		}
		for (size_t i = 0; i < pl.toHost.size(); ++i) {
			pl.toHost[i].format = devLists + pl.toHostLayout[i];
		}
		if (!pl.toHost.empty()) {
			std::memcpy(host + pl.toHostOff, pl.toHost.data(), sizeof(PointOpItem) * pl.toHost.size());
		}
		// Kept counts and overflow flags of every graph come back through one gather and one copy.
		const int g = static_cast<int>(pl.gather.size());
		const size_t slotBytes = sizeof(int32_t) * static_cast<size_t>(g);
		int32_t* slots = nullptr;
		std::shared_ptr<PinnedBuffer> hostBuf;
		if (g > 0) {
			if (slotBytes > slotsDev.capacity()) {
				synchronize();
			}
			slots = slotsDev.ensureT<int32_t>(g);
			hostBuf = hostSlots(slotBytes);
		}

		// Which launches are timed this run (user timing, or an untuned family with a large enough launch).
		bool tuneRt = false, tuneCp = false, anyOther = false;
		for (int l = 0; l <= maxLevel; ++l) {
			const Plan::Level& L = pl.levels[l];
			double rays = 0.0, points = 0.0;
			for (const auto& it : L.rt) {
				rays += it.p.rayCount;
			}
			for (const auto& it : L.cp) {
				points += it.capacity;
			}
			tuneRt = tuneRt || (autotune && rtTune.chosen < 0 && rays >= kTuneMinUnits);
			tuneCp = tuneCp || (autotune && cpTune.chosen < 0 && points >= kTuneMinUnits);
			anyOther = anyOther || !L.other.empty();
		}
		const bool timed = timing || tuneRt || tuneCp;

		// The batch's kernels: stage launches level by level (nodes of one level never depend on each other), then the
		// count gather.
		auto enqueueLaunches = [&](bool eager) {
			for (int l = 0; l <= maxLevel; ++l) {
				Plan::Level& L = pl.levels[l];
				if (!L.rt.empty() || !L.rtm.empty()) {
					const bool userTimed = eager && timing && (L.rtBlocks > 0 || L.rtmBlocks > 0);
					if (userTimed) {
						if (timingUsed == timingEvents.size()) {
							hipEvent_t ea, eb;
							RGL_HIP(hipEventCreate(&ea));
							RGL_HIP(hipEventCreate(&eb));
							timingEvents.emplace_back(ea, eb);
						}
						RGL_HIP(hipEventRecord(timingEvents[timingUsed].first, s));
					}
					if (!L.rt.empty()) {
						double rays = 0.0;
						for (const auto& it : L.rt) {
							rays += it.p.rayCount;
						}
						const bool tune = eager && autotune && rtTune.chosen < 0 && rays >= kTuneMinUnits;
						PendingTiming pt{nullptr, nullptr, &rtTune, rtVariant, rays};
						if (tune) {
							timeStart(s, pt);
						}
						gpu::raytraceBatch(s, reinterpret_cast<const RaytraceBatchItem*>(dev + L.rtOff),
						                   static_cast<int>(L.rt.size()), L.rtBlocks, Scene::bvhFormat(), false, rtVariant,
						                   reinterpret_cast<uint32_t*>(dev + L.claimOff));
						RGL_HIP(hipGetLastError());
						if (tune) {
							RGL_HIP(hipEventRecord(pt.b, s));
							pending.push_back(pt);
						}
					}
					if (!L.rtm.empty()) {
						gpu::raytraceBatch(s, reinterpret_cast<const RaytraceBatchItem*>(dev + L.rtmOff),
						                   static_cast<int>(L.rtm.size()), L.rtmBlocks, Scene::bvhFormat(), true, 0);
						RGL_HIP(hipGetLastError());
					}
					if (userTimed) {
						RGL_HIP(hipEventRecord(timingEvents[timingUsed].second, s));
						++timingUsed;
					}
				}
				if (!L.op.empty()) {
					gpu::pointOpsBatch(s, reinterpret_cast<const PointOpItem*>(dev + L.opOff),
					                   static_cast<int>(L.op.size()), L.opBlocks);
					RGL_HIP(hipGetLastError());
				}
				if (!L.cp.empty()) {
					double points = 0.0;
					for (const auto& it : L.cp) {
						points += it.capacity;
					}
					const bool tune = eager && autotune && cpTune.chosen < 0 && points >= kTuneMinUnits;
					PendingTiming pt{nullptr, nullptr, &cpTune, cpVariant, points};
					if (tune) {
						timeStart(s, pt);
					}
					gpu::compactBatch(s, reinterpret_cast<const CompactItem*>(dev + L.cpOff),
					                  static_cast<int>(L.cp.size()), L.cpTiles, cpVariant);
					RGL_HIP(hipGetLastError());
					if (tune) {
						RGL_HIP(hipEventRecord(pt.b, s));
						pending.push_back(pt);
					}
				}
				if (!L.rp.empty()) {
					gpu::radarPostprocessBatch(s, reinterpret_cast<const RadarItem*>(dev + L.rpOff),
					                           static_cast<int>(L.rp.size()), L.rpBlocks);
					RGL_HIP(hipGetLastError());
				}
				for (Node* n : L.other) {
					n->enqueue();
				}
			}
			if (g > 0) {
				gpu::gatherInts(s, g, reinterpret_cast<const int32_t* const*>(dev + gatherOff), slots);
				RGL_HIP(hipGetLastError());
			}
		};

		RGL_HIP(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, s));
		uploaded->record(s);
		// A batch whose kernel sequence (descriptor buffer, item counts, grid shapes, variants) repeats the previous
		// batch's is captured once into a HIP graph and replayed: one launch instead of one per stage. Only the
		// descriptor contents change between replays, and the kernels read them from the buffer uploaded above.
		std::vector<int64_t> sig;
		if (graphs && !timed && !anyOther) {
			sig = {reinterpret_cast<int64_t>(dev), static_cast<int64_t>(gatherOff), g, reinterpret_cast<int64_t>(slots),
			       rtVariant, cpVariant, Scene::bvhFormat(), maxLevel};
			for (int l = 0; l <= maxLevel; ++l) {
				const Plan::Level& L = pl.levels[l];
				for (int64_t v : {static_cast<int64_t>(L.rt.size()), static_cast<int64_t>(L.rtBlocks),
				                  static_cast<int64_t>(L.rtOff), static_cast<int64_t>(L.claimOff), static_cast<int64_t>(L.rtm.size()),
				                  static_cast<int64_t>(L.rtmBlocks), static_cast<int64_t>(L.rtmOff),
				                  static_cast<int64_t>(L.op.size()), static_cast<int64_t>(L.opBlocks),
				                  static_cast<int64_t>(L.opOff), static_cast<int64_t>(L.cp.size()),
				                  static_cast<int64_t>(L.cpTiles), static_cast<int64_t>(L.cpOff),
				                  static_cast<int64_t>(L.rp.size()), static_cast<int64_t>(L.rpBlocks),
				                  static_cast<int64_t>(L.rpOff)}) {
					sig.push_back(v);
				}
			}
		}
		CapturedGraph* hit = nullptr;
		if (!sig.empty()) {
			for (auto& cg : captured) {
				if (cg.sig == sig) {
					hit = &cg;
				}
			}
		}
		if (hit != nullptr) {
			RGL_HIP(hipGraphLaunch(hit->exec, s));
			hit->lastUse = ++graphClock;
			++graphReplays;
		} else if (!sig.empty() && std::find(recentSigs.begin(), recentSigs.end(), sig) != recentSigs.end()) {
			if (captured.size() >= kCapturedGraphs) {
				auto lru = std::min_element(captured.begin(), captured.end(),
				                            [](const CapturedGraph& a, const CapturedGraph& b) { return a.lastUse < b.lastUse; });
				synchronize(); // the graph being replaced may still run
				lru->release();
				captured.erase(lru);
			}
			CapturedGraph cg;
			cg.sig = sig;
			RGL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			try {
				enqueueLaunches(false);
			} catch (...) {
				hipGraph_t broken = nullptr;
				(void) hipStreamEndCapture(s, &broken);
				if (broken != nullptr) {
					(void) hipGraphDestroy(broken);
				}
				throw;
			}
			RGL_HIP(hipStreamEndCapture(s, &cg.graph));
			const hipError_t e = hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0);
			if (e != hipSuccess) {
				cg.release();
				RGL_HIP(e);
			}
			RGL_HIP(hipGraphLaunch(cg.exec, s));
			cg.lastUse = ++graphClock;
			captured.push_back(std::move(cg));
			++graphCaptures;
		} else {
			enqueueLaunches(true);
			if (!sig.empty()) {
				recentSigs.push_back(std::move(sig));
				if (recentSigs.size() > kRecentSigs) {
					recentSigs.erase(recentSigs.begin());
				}
			}
		}
		if (g > 0) {
			RGL_HIP(hipMemcpyAsync(hostBuf->get(), slots, slotBytes, hipMemcpyDeviceToHost, s));
		}
		if (g > 0) {
			int k = 0;
			for (int l = 0; l <= maxLevel; ++l) {
				for (auto* cn : pl.levels[l].cpNodes) {
					cn->setCountHost(hostBuf, k++);
				}
			}
			for (int l = 0; l <= maxLevel; ++l) {
				for (auto* rn : pl.levels[l].rpNodes) {
					rn->setInfoHost(hostBuf, k);
					k += 2;
				}
			}
			for (int l = 0; l <= maxLevel; ++l) {
				for (auto* rt : pl.levels[l].rtNodes) {
					rt->setOverflowHost(hostBuf, k++);
				}
				for (auto* rt : pl.levels[l].rtmNodes) {
					rt->setOverflowHost(hostBuf, k++);
				}
			}
		}
		if (!pl.toHost.empty()) {
			countsReady = completionEvent();
			countsReady->record(s);
			const auto* toHostItems = reinterpret_cast<const PointOpItem*>(dev + pl.toHostOff);
			for (size_t c = 0; c + 1 < pl.chunkFirst.size(); ++c) {
				const int first = pl.chunkFirst[c];
				gpu::pointOpsBatch(s, toHostItems + first, pl.chunkFirst[c + 1] - first, pl.chunkBlocks[c]);
				RGL_HIP(hipGetLastError());
				chunkReady.push_back(completionEvent());
				chunkReady.back()->record(s);
			}
		}
	} catch (...) {
		auto ev = completionEvent();
		ev->record(s);
		for (Scene* sc : pl.scenes) {
			sc->noteUse(s, ev);
		}
		for (auto& c : pl.ctxs) {
			c->endBatch(s, ev);
		}
		lastCompletion = ev;
		throw;
	}
	auto ev = completionEvent();
	ev->record(s);
	for (Scene* sc : pl.scenes) { // the scenes' structures are read until the batch completes
		sc->noteUse(s, ev);
	}
	for (auto& c : pl.ctxs) {
		for (auto& n : c->validatedOrder()) {
			n->markEnqueued(ev, countsReady, countsReady); // all but the prefetched copies are complete before those run
		}
		c->endBatch(s, ev);
	}
	for (size_t c = 0; c < chunkReady.size(); ++c) {
		for (int i = pl.chunkFirst[c]; i < pl.chunkFirst[c + 1]; ++i) {
			pl.toHostNodes[i]->markEnqueued(ev, chunkReady[c], countsReady);
		}
	}
	lastCompletion = ev;
}

} // namespace rgl
