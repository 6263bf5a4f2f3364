// A small pool of host threads for the host-pointer entry points: the
// packing of caller planes into pinned staging memory and the copy of
// results back out are plain memcpy work that one thread cannot run at
// PCIe rate.
//
// run(tasks, fn) calls fn(0) .. fn(tasks - 1) across the workers and the
// calling thread and returns when every call has returned; start() and
// finish() split it so the caller can work between them.  Task indices
// are claimed with a compare-and-swap on (generation, index), so a worker
// that wakes late never takes an index of a later run.  Workers spin
// on the generation counter for up to 2 ms before sleeping, since an entry
// point issues its runs within that.
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace vp8g {

class HostPool {
 public:
  explicit HostPool(int workers) {
    for (int i = 0; i < workers; i++) threads_.emplace_back([this] { loop(); });
  }
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
      gen_.fetch_add(1, std::memory_order_release);
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
  }

  int workers() const { return (int)threads_.size(); }

  void run(size_t tasks, const std::function<void(size_t)>& fn) {
    start(tasks, fn);
    finish();
  }

  // start() publishes a run and returns at once (the workers begin on it);
  // finish() has the caller take part and returns when every index has
  // run.  Between the two the caller may do other work; `fn` and what it
  // captures must live until finish() returns, and runs do not nest.
  void start(size_t tasks, const std::function<void(size_t)>& fn) {
    pend_fn_ = &fn;
    pend_tasks_ = tasks;
    pend_live_ = false;
    if (tasks <= 1 || threads_.empty()) return;  // finish() runs them on the caller
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      tasks_ = tasks;
      finished_ = 0;
      pend_g_ = gen_.load(std::memory_order_relaxed) + 1;
      next_.store((uint64_t)pend_g_ << 32, std::memory_order_relaxed);
      gen_.store(pend_g_, std::memory_order_release);
    }
    pend_live_ = true;
    wake_.notify_all();
  }

  void finish() {
    const std::function<void(size_t)>* fn = pend_fn_;
    const size_t tasks = pend_tasks_;
    pend_fn_ = nullptr;
    if (!fn || tasks == 0) return;
    if (!pend_live_) {
      for (size_t i = 0; i < tasks; i++) (*fn)(i);
      return;
    }
    pend_live_ = false;
    const size_t mine = work(pend_g_, fn, tasks);
    std::unique_lock<std::mutex> lk(m_);
    finished_ += mine;
    done_.wait(lk, [&] { return finished_ == tasks_; });
    fn_ = nullptr;
  }

 private:
  // Claims and runs indices of generation g; returns how many it ran.  `fn`
  // is dereferenced only for a claimed index, i.e. while run(g) waits.
  size_t work(uint32_t g, const std::function<void(size_t)>* fn, size_t tasks) {
    size_t ran = 0;
    uint64_t v = next_.load(std::memory_order_acquire);
    for (;;) {
      if ((uint32_t)(v >> 32) != g || (v & 0xFFFFFFFFu) >= tasks) return ran;
      if (next_.compare_exchange_weak(v, v + 1, std::memory_order_acq_rel, std::memory_order_acquire)) {
        (*fn)((size_t)(v & 0xFFFFFFFFu));
        ran++;
        v = next_.load(std::memory_order_acquire);
      }
    }
  }

  void loop() {
    uint32_t seen = 0;
    for (;;) {
      // spin up to 2 ms for the next run before sleeping
      const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(2000);
      while (gen_.load(std::memory_order_acquire) == seen && std::chrono::steady_clock::now() < until)
        std::this_thread::yield();
      const std::function<void(size_t)>* fn;
      size_t tasks;
      uint32_t g;
      {
        std::unique_lock<std::mutex> lk(m_);
        wake_.wait(lk, [&] { return stop_ || gen_.load(std::memory_order_relaxed) != seen; });
        if (stop_) return;
        g = gen_.load(std::memory_order_relaxed);
        seen = g;
        fn = fn_;
        tasks = tasks_;
      }
      if (!fn) continue;  // that run has already completed
      const size_t ran = work(g, fn, tasks);
      if (ran) {
        std::lock_guard<std::mutex> lk(m_);
        finished_ += ran;
        if (finished_ == tasks_) done_.notify_all();
      }
    }
  }

  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable wake_, done_;
  std::atomic<uint32_t> gen_{0};
  std::atomic<uint64_t> next_{0};
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t tasks_ = 0, finished_ = 0;
  bool stop_ = false;
  // the run between start() and finish() (caller side only)
  const std::function<void(size_t)>* pend_fn_ = nullptr;
  size_t pend_tasks_ = 0;
  uint32_t pend_g_ = 0;
  bool pend_live_ = false;
};

}  // namespace vp8g
