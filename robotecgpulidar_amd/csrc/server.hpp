// Request server: continuous batching of sensor frames over device-resident scenes.
//
// The reference serves one graph per rgl_graph_run call, each on its own thread and stream
// (src/graph/GraphRunCtx.cpp:57-105); a simulator with many sensors calls it in a loop. This server keeps several
// batches in flight (one GraphBatch and stream per slot). Requests that arrive while batches run wait in a FIFO
// queue; whenever a slot is free, every admissible queued request joins the next batch, so new work overlaps the
// tail of running batches instead of waiting for a fixed batch window. A request is admissible when its graph is
// not already queued in an earlier in-flight batch and every scene it raytraces is resident within the server's
// HBM budget: scenes are admitted on first use and the least recently used idle scene's acceleration structures are
// evicted when the budget would be exceeded (they are rebuilt on the scene's next use).
//
// The server is driven by the client thread (submit / step / poll / wait), so no thread other than the caller's
// touches the library.
#pragma once

#include <chrono>
#include <cstdint>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <set>
#include <unordered_map>
#include <vector>

#include "graph.hpp"

namespace rgl {

struct ServerStats
{
	int64_t submitted = 0;
	int64_t completed = 0;
	int64_t batches = 0;
	int64_t raysCompleted = 0;
	int64_t evictions = 0;
	int64_t deferred = 0;      // times a queued request could not join a batch (graph busy or scene not admissible)
	int64_t failed = 0;        // requests whose scenes alone exceed the HBM budget
	int64_t residentBytes = 0; // acceleration-structure bytes of the admitted scenes
	double p50Ms = 0, p99Ms = 0, meanMs = 0, maxMs = 0;
};

class Server
{
public:
	// slots: batches in flight at once; maxBatchGraphs: graphs per batch; hbmBudget: bytes of scene acceleration
	// structures kept resident (<= 0: unlimited); batching = false runs every request alone on its graph's stream
	// (the plain per-sensor path, one request per slot) under the same queueing and accounting.
	Server(int slots, int maxBatchGraphs, int64_t hbmBudget, bool batching);
	~Server();

	int64_t submit(const Node::Ptr& node);
	// Retires finished batches and launches queued requests into free slots; returns the requests retired.
	int step();
	// 0 queued, 1 running, 2 done, 3 failed (its batch could not be launched). Latency (submit to retirement) is set
	// for done requests.
	int state(int64_t ticket, double* latencyMs) const;
	void wait(int64_t ticket, double* latencyMs);
	ServerStats stats() const;
	void resetStats();
	void drain();

private:
	using Clock = std::chrono::steady_clock;
	struct Request
	{
		int64_t ticket;
		Node::Ptr entry;
		std::vector<Scene*> scenes;
		int64_t rays;
		Clock::time_point submitted;
	};
	struct Slot
	{
		std::unique_ptr<GraphBatch> batch;       // batching mode
		std::unique_ptr<Event> soloDone;          // plain mode: completion of the solo run
		std::vector<Request> requests;
		bool active = false;
		bool busy() const;
	};
	struct Ticket
	{
		int state = 0;
		double latencyMs = 0;
	};

	static constexpr int kAdmitted = 1, kDeferred = 0, kNever = -1;
	int admitScenes(const Request& r, const std::set<Scene*>& pinned);
	void retire(Slot& s);
	static std::vector<Scene*> scenesOf(const Node::Ptr& entry);
	static int64_t raysOf(const Node::Ptr& entry);
	std::set<Scene*> pinnedScenes() const;

	int maxBatchGraphs;
	int64_t budget;
	bool batching;
	int64_t nextTicket = 1;
	std::deque<Request> queue;
	std::vector<Slot> slots;
	std::unordered_map<int64_t, Ticket> tickets;
	std::set<const GraphRunCtx*> busyGraphs; // graphs of requests in active slots
	std::list<Scene*> lru;            // resident scenes, most recent at the front
	std::map<Scene*, int64_t> resident;
	std::map<std::pair<Scene*, int>, int64_t> knownBytes; // built sizes seen at retirement, per node format
	int64_t sceneBytes(Scene* sc, int format) const;
	int64_t residentBytes = 0;
	std::vector<double> latencies;
	ServerStats counters;
};

} // namespace rgl
