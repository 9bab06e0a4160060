// A fixed set of host worker threads for the library's byte-parallel host
// work: copies into and out of the pinned staging buffers, comparisons of a
// call's moveouts and weights with the kept ones, and table packing.
// Starting threads per piece of work cost about as much as the copies
// themselves at 16 MiB per staging buffer (a 16 MiB fill took 0.42 ms on
// 8 new threads), so the workers are started once and kept.
//
// Host code only (no HIP): it compiles with g++ for the sanitizer tests.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace bp {

class HostPool {
 public:
  // A pool of `workers` threads (0: every task runs on the calling thread).
  explicit HostPool(unsigned workers) {
    threads_.reserve(workers);
    for (unsigned i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    n_workers_ = workers;
  }
  // Stops and joins the workers (the process-wide pool is never destroyed).
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
  }
  HostPool(const HostPool &) = delete;
  HostPool &operator=(const HostPool &) = delete;

  unsigned workers() const { return n_workers_; }

  // Runs task(i) for every i in [0, n) on the calling thread and on idle
  // workers, and returns when all n have run.  The caller takes tasks
  // itself, so a call completes even when every worker is busy with other
  // callers' tasks.  If tasks throw, the others still run and the first
  // exception is rethrown here.
  void run(size_t n, const std::function<void(size_t)> &task) {
    if (n == 0) return;
    auto job = std::make_shared<Job>();
    job->task = &task;
    job->n = n;
    if (n > 1 && n_workers_ > 0) {
      {
        std::lock_guard<std::mutex> lock(mu_);
        queue_.push_back(job);
      }
      const size_t wake = std::min<size_t>(n - 1, n_workers_);
      for (size_t i = 0; i < wake; ++i) cv_.notify_one();
    }
    work(*job);
    {
      std::unique_lock<std::mutex> lock(job->m);
      job->cv.wait(lock, [&] { return job->finished; });
    }
    if (n > 1 && n_workers_ > 0) {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = std::find(queue_.begin(), queue_.end(), job);
      if (it != queue_.end()) queue_.erase(it);
    }
    if (job->err) std::rethrow_exception(job->err);
  }

  // The process-wide pool: min(16, hardware threads) - 1 workers, started
  // on first use and never destroyed (its threads wait until the process
  // ends; their std::thread objects are never destroyed either).
  static HostPool &shared() {
    static HostPool *pool = new HostPool(
        std::min(16u, std::max(1u, std::thread::hardware_concurrency())) - 1u);
    return *pool;
  }

 private:
  struct Job {
    const std::function<void(size_t)> *task = nullptr;
    size_t n = 0;
    std::atomic<size_t> next{0};
    std::atomic<size_t> done{0};
    std::mutex m;
    std::condition_variable cv;
    bool finished = false;    // guarded by m
    std::exception_ptr err;   // guarded by m
  };

  // Takes and runs tasks of `job` until none is left.  The task pointer is
  // only followed for an index below n, that is, before the job's caller
  // can have returned.
  static void work(Job &job) {
    for (;;) {
      const size_t i = job.next.fetch_add(1, std::memory_order_relaxed);
      if (i >= job.n) return;
      try {
        (*job.task)(i);
      } catch (...) {
        std::lock_guard<std::mutex> lock(job.m);
        if (!job.err) job.err = std::current_exception();
      }
      if (job.done.fetch_add(1, std::memory_order_acq_rel) + 1 == job.n) {
        std::lock_guard<std::mutex> lock(job.m);
        job.finished = true;
        job.cv.notify_all();
      }
    }
  }

  void loop() {
    for (;;) {
      std::shared_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return stop_ || !queue_.empty(); });
        if (queue_.empty()) return;  // stopped
        job = queue_.front();
        if (job->next.load(std::memory_order_relaxed) >= job->n) {
          queue_.pop_front();  // every task taken: nothing left for workers
          continue;
        }
      }
      work(*job);
    }
  }

  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<Job>> queue_;
  bool stop_ = false;  // guarded by mu_
  std::vector<std::thread> threads_;
  unsigned n_workers_ = 0;
};

}  // namespace bp
