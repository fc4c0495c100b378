#include "hostcopy.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include <unistd.h>

#include "runtime.hpp"

namespace rgl {

namespace {

#ifndef RGL_COPY_WORKERS
#define RGL_COPY_WORKERS 4 // threads beside the caller's (at most a quarter of the host's)
#endif

constexpr size_t kParallelMin = 256 << 10; // below this one memcpy is faster than waking workers
constexpr size_t kChunk = 64 << 10;        // unit of work a thread takes at a time
constexpr size_t kStageChunk = 4 << 20;    // pinned staging chunk of device reads

// Worker threads that split large copies. A job is published by bumping `generation`; workers spin for a while
// after each job (calls come in bursts: every field of every sensor after a frame) and then sleep.
class CopyPool
{
public:
	static CopyPool& instance()
	{
		static CopyPool pool;
		return pool;
	}

	void copy(char* dst, const char* src, size_t bytes)
	{
		// A forked child has no workers (threads do not survive fork): it copies alone.
		if (bytes < kParallelMin || workers.empty() || getpid() != owner) {
			std::memcpy(dst, src, bytes);
			return;
		}
		std::lock_guard<std::mutex> callerLock(callers);
		jobDst = dst;
		jobSrc = src;
		jobBytes = bytes;
		nextChunk.store(0, std::memory_order_relaxed);
		busy.store(static_cast<int>(workers.size()), std::memory_order_relaxed);
		{
			std::lock_guard<std::mutex> lk(m);
			generation.fetch_add(1, std::memory_order_release);
		}
		cv.notify_all();
		work();
		while (busy.load(std::memory_order_acquire) != 0) {
			std::this_thread::yield();
		}
	}

	~CopyPool()
	{
		{
			std::lock_guard<std::mutex> lk(m);
			stop = true;
			generation.fetch_add(1, std::memory_order_release);
		}
		cv.notify_all();
		for (auto& t : workers) {
			t.join();
		}
	}

private:
	CopyPool()
	{
		const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
		const unsigned n = std::min<unsigned>(RGL_COPY_WORKERS, hw / 4);
		// Workers start from the generation before any job: a job published before a worker first runs is still
		// seen as new by it (the caller waits for every worker).
		owner = getpid();
		const uint64_t start = generation.load(std::memory_order_relaxed);
		for (unsigned i = 0; i < n; ++i) {
			workers.emplace_back([this, start] { loop(start); });
		}
	}

	void work()
	{
		const size_t chunks = (jobBytes + kChunk - 1) / kChunk;
		for (size_t c = nextChunk.fetch_add(1, std::memory_order_relaxed); c < chunks;
		     c = nextChunk.fetch_add(1, std::memory_order_relaxed)) {
			const size_t off = c * kChunk;
			std::memcpy(jobDst + off, jobSrc + off, std::min(kChunk, jobBytes - off));
		}
	}

	void loop(uint64_t seen)
	{
		while (true) {
			// Spin up to ~200 us for the next job, then sleep until one is published.
			const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(200);
			uint64_t g;
			while ((g = generation.load(std::memory_order_acquire)) == seen &&
			       std::chrono::steady_clock::now() < until) {
				std::this_thread::yield();
			}
			if (g == seen) {
				std::unique_lock<std::mutex> lk(m);
				cv.wait(lk, [&] { return generation.load(std::memory_order_acquire) != seen; });
				g = generation.load(std::memory_order_acquire);
			}
			seen = g;
			if (stop) {
				return;
			}
			work();
			busy.fetch_sub(1, std::memory_order_release);
		}
	}

	std::vector<std::thread> workers;
	std::mutex callers, m;
	std::condition_variable cv;
	std::atomic<uint64_t> generation{0};
	std::atomic<size_t> nextChunk{0};
	std::atomic<int> busy{0};
	char* jobDst = nullptr;
	const char* jobSrc = nullptr;
	size_t jobBytes = 0;
	bool stop = false;
	pid_t owner = 0;
};

// Pinned double buffer and stream for staged device reads.
struct Staging
{
	PinnedBuffer buf[2];
	std::unique_ptr<Stream> stream;
	std::unique_ptr<Event> done[2];
	std::mutex m;
};

Staging& staging()
{
	// Never destroyed: its pinned memory and stream must not be released after the HIP runtime shuts down.
	static Staging* s = new Staging();
	return *s;
}

} // namespace

void hostCopy(void* dst, const void* src, size_t bytes)
{
	CopyPool::instance().copy(static_cast<char*>(dst), static_cast<const char*>(src), bytes);
}

void deviceToHost(void* dst, const void* src, size_t bytes)
{
	if (bytes < kParallelMin) {
		RGL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
		return;
	}
	Staging& st = staging();
	std::lock_guard<std::mutex> lk(st.m);
	if (!st.stream) {
		st.stream = std::make_unique<Stream>();
		st.done[0] = std::make_unique<Event>();
		st.done[1] = std::make_unique<Event>();
	}
	const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
	auto* out = static_cast<char*>(dst);
	const auto* in = static_cast<const char*>(src);
	auto enqueue = [&](size_t c) {
		const size_t off = c * kStageChunk;
		const size_t n = std::min(kStageChunk, bytes - off);
		void* b = st.buf[c & 1].ensure(kStageChunk);
		RGL_HIP(hipMemcpyAsync(b, in + off, n, hipMemcpyDeviceToHost, st.stream->get()));
		st.done[c & 1]->record(st.stream->get());
	};
	enqueue(0);
	for (size_t c = 0; c < chunks; ++c) {
		if (c + 1 < chunks) {
			enqueue(c + 1); // its buffer's previous chunk (c - 1) was copied out in the last iteration
		}
		st.done[c & 1]->sync();
		const size_t off = c * kStageChunk;
		hostCopy(out + off, st.buf[c & 1].get(), std::min(kStageChunk, bytes - off));
	}
}

} // namespace rgl
