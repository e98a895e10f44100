// workers.cpp -- the fork-join pool of workers.hpp.
#include "workers.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gpa {
namespace {

thread_local bool tInJob = false;

// One parallelFor call: its tasks are claimed through `next`; a worker that
// wakes late holds this object, finds every task claimed and leaves, so it
// never touches a later job.
struct Job {
  const std::function<void(uint32_t)> *fn = nullptr;
  uint32_t n = 0;
  std::atomic<uint32_t> next{0}, done{0};
  std::mutex errMu;
  uint32_t errIdx = UINT32_MAX;
  std::exception_ptr err;

  void work() {
    const bool was = tInJob;
    tInJob = true;
    for (uint32_t i; (i = next.fetch_add(1)) < n;) {
      try {
        (*fn)(i);
      } catch (...) {
        std::lock_guard<std::mutex> lk(errMu);
        if (i < errIdx) {
          errIdx = i;
          err = std::current_exception();
        }
      }
      done.fetch_add(1, std::memory_order_acq_rel);
    }
    tInJob = was;
  }
};

// Workers that finish a job spin for a short while before sleeping, so the
// several parallel sections of one build find them awake (waking fifteen
// sleeping threads costs tens of microseconds per section).
constexpr auto kSpin = std::chrono::microseconds(300);

inline void cpuRelax() {
#if defined(__x86_64__)
  __builtin_ia32_pause();
#endif
}

class Pool {
 public:
  Pool() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned n = std::min(hw, 16u) - 1;
    for (unsigned i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_.store(true);
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
  }
  uint32_t width() const { return static_cast<uint32_t>(workers_.size()) + 1; }

  // false: the pool is busy with another thread's job (the caller then runs
  // the tasks itself)
  bool run(uint32_t n, const std::function<void(uint32_t)> &f) {
    std::unique_lock<std::mutex> busy(jobMu_, std::try_to_lock);
    if (!busy.owns_lock() || workers_.empty()) return false;
    auto job = std::make_shared<Job>();
    job->fn = &f;
    job->n = n;
    bool wake;
    {
      std::lock_guard<std::mutex> lk(mu_);
      cur_ = job;
      gen_.fetch_add(1, std::memory_order_release);
      wake = sleeping_ > 0;
    }
    if (wake) cv_.notify_all();
    job->work();
    while (job->done.load(std::memory_order_acquire) < n) cpuRelax();
    {
      std::lock_guard<std::mutex> lk(mu_);
      cur_.reset();
    }
    if (job->err) std::rethrow_exception(job->err);
    return true;
  }

 private:
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      // spin, then sleep, until a new job (or the end)
      const auto until = std::chrono::steady_clock::now() + kSpin;
      while (gen_.load(std::memory_order_acquire) == seen && !stop_.load() &&
             std::chrono::steady_clock::now() < until)
        for (int i = 0; i < 64; ++i) cpuRelax();
      std::shared_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        if (gen_.load() == seen && !stop_.load()) {
          ++sleeping_;
          cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
          --sleeping_;
        }
        if (stop_.load()) return;
        seen = gen_.load();
        job = cur_;
      }
      if (job) job->work();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex mu_, jobMu_;
  std::condition_variable cv_;
  std::atomic<bool> stop_{false};
  std::atomic<uint64_t> gen_{0};
  int sleeping_ = 0;  // workers waiting on cv_ (under mu_)
  std::shared_ptr<Job> cur_;
};

Pool &pool() {
  static Pool p;
  return p;
}

}  // namespace

uint32_t parallelWidth() { return pool().width(); }

void parallelFor(uint32_t n, const std::function<void(uint32_t)> &f) {
  if (n == 0) return;
  if (n > 1 && !tInJob && pool().run(n, f)) return;
  // on this thread, with the same error order
  std::exception_ptr err;
  for (uint32_t i = 0; i < n; ++i) {
    try {
      f(i);
    } catch (...) {
      if (!err) err = std::current_exception();
    }
  }
  if (err) std::rethrow_exception(err);
}

}  // namespace gpa
