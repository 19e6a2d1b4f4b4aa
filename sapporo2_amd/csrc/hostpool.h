// hostpool.h -- parked host threads for the per-call copies of the engine
// (no reference counterpart: the reference packs and unpacks on one thread,
// sapporohostclass.cpp:613-649, 705-802).
#pragma once

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <system_error>
#include <thread>

#include <unistd.h>

namespace sap {

// Host-side copies of a call (packing the i-particles, unpacking the result
// rows) run on a pool of parked threads once they span a few chunks: a 16384
// particle call moves about 1 MB each way, tens of microseconds on one core.
// A job is a range cut into chunks that the calling thread and the workers
// take from one counter, so the caller never waits for a worker that is slow
// to wake, and a job ends when its last chunk is done.  Workers spin for a
// while after each job (a call's unpack and the next call's pack come close
// together), then sleep on a condition variable.  Threads: SAPPORO2_AMD_HOST_THREADS
// (the caller included; 1 = all copies on the caller), default
// min(HOST_THREADS, cores); spin: SAPPORO2_AMD_HOST_SPIN_US, default HOST_SPIN_US.
constexpr size_t HOST_CHUNK = 1024;
constexpr unsigned HOST_THREADS = 8;
constexpr long HOST_SPIN_US = 300;

class HostPool {
 public:
  // Never destroyed: the workers are detached and run until the process ends.
  static HostPool& get() {
    static HostPool* pool = new HostPool;
    return *pool;
  }
  unsigned threads() const { return nworkers_ + 1; }

  // fn(begin, end) over [0, n) in chunks of `chunk` elements.  Returns false,
  // having run nothing, while another thread's job holds the pool.
  template <typename F>
  bool run(size_t n, size_t chunk, F& fn) {
    if (getpid() != pid_) return false;  // a forked child has none of the workers
    std::unique_lock<std::mutex> own(run_m_, std::try_to_lock);
    if (!own.owns_lock()) return false;
    auto j = std::make_shared<Job>();
    j->n = n;
    j->chunk = chunk;
    j->nchunks = (n + chunk - 1) / chunk;
    j->ctx = &fn;
    j->call = [](void* c, size_t b, size_t e) { (*static_cast<F*>(c))(b, e); };
    std::atomic_store(&job_, j);
    {
      std::lock_guard<std::mutex> lk(m_);
      gen_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
    work(*j);
    while (j->done.load(std::memory_order_acquire) < j->nchunks) _mm_pause();
    return true;
  }

 private:
  struct Job {
    size_t n = 0, chunk = 0, nchunks = 0;
    void* ctx = nullptr;
    void (*call)(void*, size_t, size_t) = nullptr;
    std::atomic<size_t> next{0}, done{0};
  };

  // Takes chunks until none is left.  A worker that arrives after the job's
  // last chunk was taken only reads the counter: the job object stays alive
  // through its shared pointer, and the caller's function is not touched.
  static void work(Job& j) {
    size_t c;
    while ((c = j.next.fetch_add(1, std::memory_order_relaxed)) < j.nchunks) {
      const size_t b = c * j.chunk;
      j.call(j.ctx, b, std::min(j.n, b + j.chunk));
      j.done.fetch_add(1, std::memory_order_release);
    }
  }

  HostPool() {
    unsigned want = std::max(1u, std::min(HOST_THREADS, std::thread::hardware_concurrency()));
    if (const char* v = getenv("SAPPORO2_AMD_HOST_THREADS")) want = (unsigned)std::max(1, atoi(v));
    long spin = HOST_SPIN_US;
    if (const char* v = getenv("SAPPORO2_AMD_HOST_SPIN_US")) spin = std::max(0L, atol(v));
    spin_ = std::chrono::microseconds(spin);
    for (unsigned t = 1; t < want; ++t) {
      try {
        std::thread([this] { loop(); }).detach();
        ++nworkers_;
      } catch (const std::system_error&) {
        break;  // fewer workers; the caller takes the chunks no worker takes
      }
    }
  }

  void loop() {
    uint64_t seen = gen_.load(std::memory_order_acquire);
    for (;;) {
      const auto t0 = std::chrono::steady_clock::now();
      for (unsigned k = 1;; ++k) {
        if (gen_.load(std::memory_order_acquire) != seen) break;
        if ((k & 63) == 0 && std::chrono::steady_clock::now() - t0 > spin_) {
          std::unique_lock<std::mutex> lk(m_);
          cv_.wait(lk, [&] { return gen_.load() != seen; });
        } else {
          _mm_pause();
        }
      }
      seen = gen_.load(std::memory_order_acquire);
      const std::shared_ptr<Job> j = std::atomic_load(&job_);
      if (j) work(*j);
    }
  }

  unsigned nworkers_ = 0;
  const pid_t pid_ = getpid();
  std::mutex run_m_;  // one job at a time
  std::mutex m_;      // sleeping workers
  std::condition_variable cv_;
  std::atomic<uint64_t> gen_{0};
  std::shared_ptr<Job> job_;
  std::chrono::microseconds spin_{HOST_SPIN_US};
};

template <typename F>
void par_for(size_t n, F&& fn) {
  if (n >= 2 * HOST_CHUNK) {
    HostPool& pool = HostPool::get();
    const unsigned T = pool.threads();
    if (T > 1) {
      // about four chunks per thread, so late workers still find some
      const size_t chunk = std::max(HOST_CHUNK, (n + 4 * (size_t)T - 1) / (4 * (size_t)T));
      if (pool.run(n, chunk, fn)) return;
    }
  }
  fn((size_t)0, n);
}

}  // namespace sap
