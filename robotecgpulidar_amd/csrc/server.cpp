#include "server.hpp"

#include <algorithm>
#include <stdexcept>

namespace rgl {

Server::Server(int slotCount, int maxGraphs, int64_t hbmBudget, bool batch)
    : maxBatchGraphs(maxGraphs), budget(hbmBudget), batching(batch)
{
	if (slotCount < 1 || maxGraphs < 1) {
		throw std::invalid_argument("server needs at least one slot and one graph per batch");
	}
	slots.resize(static_cast<size_t>(slotCount));
	for (Slot& s : slots) {
		if (batching) {
			s.batch = std::make_unique<GraphBatch>();
		} else {
			s.soloDone = std::make_unique<Event>();
		}
	}
}

Server::~Server()
{
	try {
		drain();
	} catch (...) {
	}
}

bool Server::Slot::busy() const
{
	if (!active) {
		return false;
	}
	if (batch) {
		return batch->busy();
	}
	const hipError_t e = hipEventQuery(soloDone->get());
	if (e == hipErrorNotReady) {
		return true;
	}
	RGL_HIP(e);
	return false;
}

std::vector<Scene*> Server::scenesOf(const Node::Ptr& entry)
{
	std::vector<Scene*> out;
	for (auto& n : entry->connectedComponent()) {
		if (auto rt = std::dynamic_pointer_cast<RaytraceNode>(n)) {
			Scene* sc = &rt->getScene();
			if (std::find(out.begin(), out.end(), sc) == out.end()) {
				out.push_back(sc);
			}
		}
	}
	return out;
}

int64_t Server::raysOf(const Node::Ptr& entry)
{
	int64_t rays = 0;
	for (auto& n : entry->connectedComponent()) {
		if (auto r = std::dynamic_pointer_cast<FromMat3x4fRaysNode>(n)) {
			rays += r->rayCount();
		}
	}
	return rays;
}

int64_t Server::submit(const Node::Ptr& node)
{
	if (!node->hasCtx()) {
		GraphRunCtx::createAndAttach(node);
	}
	Request r{nextTicket++, node, scenesOf(node), raysOf(node), Clock::now()};
	tickets[r.ticket] = Ticket{};
	queue.push_back(std::move(r));
	++counters.submitted;
	return queue.back().ticket;
}

// Bytes a scene's structures take once built: the size seen when it last ran, else the scene's own figure (exact when
// built, a conservative estimate otherwise).
int64_t Server::sceneBytes(Scene* sc, int format) const
{
	auto it = knownBytes.find({sc, format});
	return it != knownBytes.end() ? it->second : sc->accelBytes(format);
}

std::set<Scene*> Server::pinnedScenes() const
{
	std::set<Scene*> pinned;
	for (const Slot& s : slots) {
		if (s.active) {
			for (const Request& r : s.requests) {
				pinned.insert(r.scenes.begin(), r.scenes.end());
			}
		}
	}
	return pinned;
}

// Makes the request's scenes resident within the budget, evicting least recently used scenes that no running or
// forming batch uses. Returns kAdmitted, kDeferred (the budget cannot be met now; nothing changed) or kNever (the
// request's scenes alone exceed the budget).
int Server::admitScenes(const Request& r, const std::set<Scene*>& pinned)
{
	const int fmt = Scene::bvhFormat();
	int64_t need = 0;
	for (Scene* sc : r.scenes) {
		if (!resident.count(sc)) {
			need += sceneBytes(sc, fmt);
		}
	}
	if (budget > 0 && need > 0) {
		int64_t own = 0;
		for (Scene* sc : r.scenes) {
			own += resident.count(sc) ? resident[sc] : sceneBytes(sc, fmt);
		}
		if (own > budget) {
			return kNever;
		}
		std::vector<Scene*> victims;
		int64_t freed = 0;
		for (auto it = lru.rbegin(); it != lru.rend() && residentBytes - freed + need > budget; ++it) {
			Scene* v = *it;
			if (!pinned.count(v) && std::find(r.scenes.begin(), r.scenes.end(), v) == r.scenes.end()) {
				victims.push_back(v);
				freed += resident[v];
			}
		}
		if (residentBytes - freed + need > budget) {
			return kDeferred;
		}
		if (!victims.empty()) {
			// Meshes still used by a scene that stays resident keep their structures.
			std::set<Mesh*> keep;
			for (auto& [sc, bytes] : resident) {
				if (std::find(victims.begin(), victims.end(), sc) == victims.end()) {
					for (Mesh* m : sc->meshes()) {
						keep.insert(m);
					}
				}
			}
			for (Scene* sc : r.scenes) {
				for (Mesh* m : sc->meshes()) {
					keep.insert(m);
				}
			}
			RGL_HIP(hipDeviceSynchronize()); // nothing may still read the structures freed below
			for (Scene* v : victims) {
				for (Mesh* m : v->meshes()) {
					if (!keep.count(m)) {
						m->releaseAccel();
					}
				}
				v->releaseEntityAccel();
				residentBytes -= resident[v];
				resident.erase(v);
				lru.remove(v);
				++counters.evictions;
			}
			Scene::markAllDirty(true); // every scene holding pointers to freed structures rebuilds on next use
		}
	}
	for (Scene* sc : r.scenes) {
		auto it = resident.find(sc);
		if (it == resident.end()) {
			const int64_t b = sceneBytes(sc, fmt);
			resident[sc] = b;
			residentBytes += b;
		}
		lru.remove(sc);
		lru.push_front(sc);
	}
	return kAdmitted;
}

void Server::retire(Slot& s)
{
	const auto now = Clock::now();
	for (const Request& r : s.requests) {
		const double ms = std::chrono::duration<double, std::milli>(now - r.submitted).count();
		Ticket& t = tickets[r.ticket];
		t.state = 2;
		t.latencyMs = ms;
		latencies.push_back(ms);
		++counters.completed;
		counters.raysCompleted += r.rays;
		busyGraphs.erase(r.entry->getCtx().get());
	}
	// Structures built by this batch replace their estimates in the accounting.
	const int fmt = Scene::bvhFormat();
	for (const Request& r : s.requests) {
		for (Scene* sc : r.scenes) {
			auto it = resident.find(sc);
			if (it != resident.end()) {
				const int64_t b = sc->accelBytes(fmt);
				residentBytes += b - it->second;
				it->second = b;
				knownBytes[{sc, fmt}] = b;
			}
		}
	}
	s.requests.clear();
	s.active = false;
}

int Server::step()
{
	int retired = 0;
	for (Slot& s : slots) {
		if (s.active && !s.busy()) {
			retired += static_cast<int>(s.requests.size());
			retire(s);
		}
	}
	// Queued requests are spread evenly over the free slots (at most maxBatchGraphs per batch), so that several
	// batches are in flight: the requests of the batch that finishes first return to the queue and run again while
	// the others still occupy the device, instead of every request waiting for one large batch.
	int freeSlots = static_cast<int>(std::count_if(slots.begin(), slots.end(), [](const Slot& x) { return !x.active; }));
	for (Slot& s : slots) {
		if (s.active || queue.empty()) {
			continue;
		}
		std::set<Scene*> pinned = pinnedScenes();
		std::set<const GraphRunCtx*> chosen;
		std::vector<Request> batch;
		const int share = static_cast<int>((queue.size() + static_cast<size_t>(freeSlots) - 1) / static_cast<size_t>(freeSlots));
		--freeSlots;
		const int limit = batching ? std::min(maxBatchGraphs, std::max(share, 1)) : 1;
		for (auto it = queue.begin(); it != queue.end() && static_cast<int>(batch.size()) < limit;) {
			if (!it->entry->hasCtx()) {
				GraphRunCtx::createAndAttach(it->entry);
			}
			const GraphRunCtx* ctx = it->entry->getCtx().get();
			// A graph runs in one batch at a time; later requests for it wait for the earlier ones, in order.
			const int admit = (busyGraphs.count(ctx) || chosen.count(ctx)) ? kDeferred : admitScenes(*it, pinned);
			if (admit == kNever) {
				tickets[it->ticket].state = 3;
				++counters.failed;
				it = queue.erase(it);
				continue;
			}
			if (admit == kDeferred) {
				++counters.deferred;
				chosen.insert(ctx);
				++it;
				continue;
			}
			chosen.insert(ctx);
			pinned.insert(it->scenes.begin(), it->scenes.end());
			batch.push_back(std::move(*it));
			it = queue.erase(it);
		}
		if (batch.empty()) {
			break;
		}
		try {
			if (batching) {
				std::vector<Node::Ptr> entries;
				entries.reserve(batch.size());
				for (const Request& r : batch) {
					entries.push_back(r.entry);
				}
				s.batch->run(entries);
			} else {
				auto ctx = batch.front().entry->getCtx();
				ctx->executeAsync();
				s.soloDone->record(ctx->stream());
			}
		} catch (...) {
			for (const Request& r : batch) {
				tickets[r.ticket].state = 3;
			}
			throw;
		}
		for (const Request& r : batch) {
			tickets[r.ticket].state = 1;
			busyGraphs.insert(r.entry->getCtx().get());
		}
		s.requests = std::move(batch);
		s.active = true;
		++counters.batches;
	}
	return retired;
}

int Server::state(int64_t ticket, double* latencyMs) const
{
	auto it = tickets.find(ticket);
	if (it == tickets.end()) {
		throw std::invalid_argument("unknown server ticket " + std::to_string(ticket));
	}
	if (latencyMs != nullptr) {
		*latencyMs = it->second.latencyMs;
	}
	return it->second.state;
}

void Server::wait(int64_t ticket, double* latencyMs)
{
	int st;
	while ((st = state(ticket, latencyMs)) < 2) {
		step();
	}
	if (st == 3) {
		throw InvalidPipeline("server request " + std::to_string(ticket) + " failed to launch");
	}
}

void Server::drain()
{
	while (!queue.empty() || std::any_of(slots.begin(), slots.end(), [](const Slot& s) { return s.active; })) {
		step();
	}
}

ServerStats Server::stats() const
{
	ServerStats out = counters;
	out.residentBytes = residentBytes;
	if (!latencies.empty()) {
		std::vector<double> v = latencies;
		std::sort(v.begin(), v.end());
		auto pct = [&](double p) {
			const double pos = p * static_cast<double>(v.size() - 1);
			const size_t lo = static_cast<size_t>(pos);
			const size_t hi = std::min(lo + 1, v.size() - 1);
			return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
		};
		out.p50Ms = pct(0.50);
		out.p99Ms = pct(0.99);
		double sum = 0;
		for (double x : v) {
			sum += x;
		}
		out.meanMs = sum / static_cast<double>(v.size());
		out.maxMs = v.back();
	}
	return out;
}

void Server::resetStats()
{
	counters = ServerStats{};
	latencies.clear();
}

} // namespace rgl
