// A small pool of host threads for the host-pointer entry points: the
// packing of caller planes into pinned staging memory and the copy of
// results back out are plain memcpy work that one thread cannot run at
// PCIe rate.
//
// run(tasks, fn) calls fn(0) .. fn(tasks - 1) across the workers and the
// calling thread and returns when every call has returned.  Task indices
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
    if (tasks == 0) return;
    if (threads_.empty() || tasks == 1) {
      for (size_t i = 0; i < tasks; i++) fn(i);
      return;
    }
    uint32_t g;
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      tasks_ = tasks;
      finished_ = 0;
      g = gen_.load(std::memory_order_relaxed) + 1;
      next_.store((uint64_t)g << 32, std::memory_order_relaxed);
      gen_.store(g, std::memory_order_release);
    }
    wake_.notify_all();
    const size_t mine = work(g, &fn, tasks);
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
};

}  // namespace vp8g
