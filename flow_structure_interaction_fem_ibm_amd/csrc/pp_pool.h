// Host worker pool of the set-up code (header-only, so that a host
// ThreadSanitizer build can test it): up to 15 worker threads kept for the
// life of the process (set-up runs parallel regions many times per
// geometry; creating threads for each cost more than some regions' work).
// One region runs at a time; the calling thread takes part, and every
// participant takes chunks from the region's counter until none are left.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace pp {

class HostPool {
 public:
  static HostPool &get() {
    static HostPool p;
    return p;
  }
  unsigned size() const { return (unsigned)workers_.size() + 1; }
  // chunk(c) for c in [0, nchunks), returning when all are done
  void run(long nchunks, const std::function<void(long)> &chunk) {
    std::lock_guard<std::mutex> one_region(region_mu_);
    auto r = std::make_shared<Region>();
    r->fn = &chunk;
    r->n = nchunks;
    {
      std::lock_guard<std::mutex> lk(mu_);
      cur_ = r;
      gen_++;
    }
    cv_.notify_all();
    work(*r);
    {
      std::unique_lock<std::mutex> lk(r->m);
      r->cv.wait(lk, [&] { return r->done.load() == r->n; });
    }
    std::lock_guard<std::mutex> lk(mu_);
    cur_.reset();
  }

 private:
  // one parallel region: a worker that wakes late holds the region it saw
  // and finds its chunk counter exhausted
  struct Region {
    const std::function<void(long)> *fn = nullptr;
    long n = 0;
    std::atomic<long> next{0}, done{0};
    std::mutex m;
    std::condition_variable cv;
  };
  HostPool() {
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (unsigned t = 1; t < nt; t++) workers_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      gen_++;
    }
    cv_.notify_all();
    for (auto &w : workers_) w.join();
  }
  static void work(Region &r) {
    long finished = 0;
    for (long c = r.next.fetch_add(1); c < r.n; c = r.next.fetch_add(1)) {
      (*r.fn)(c);
      finished++;
    }
    if (finished && r.done.fetch_add(finished) + finished == r.n) {
      std::lock_guard<std::mutex> lk(r.m);
      r.cv.notify_all();
    }
  }
  void loop() {
    unsigned long seen = 0;
    for (;;) {
      std::shared_ptr<Region> r;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        r = cur_;
      }
      if (r) work(*r);
    }
  }
  std::vector<std::thread> workers_;
  std::mutex region_mu_, mu_;
  std::condition_variable cv_;
  std::shared_ptr<Region> cur_;
  unsigned long gen_ = 0;
  bool stop_ = false;
};

// fn(a, b) over [0, n) split across the pool (contiguous ranges, one per
// participant; serial below 64 items)
inline void pool_parallel(long n, const std::function<void(long, long)> &fn) {
  HostPool &pool = HostPool::get();
  const unsigned nt = pool.size();
  if (n < 64 || nt == 1) { fn(0, n); return; }
  const long parts = std::min<long>(nt, n), chunk = (n + parts - 1) / parts;
  const long nchunks = (n + chunk - 1) / chunk;
  pool.run(nchunks, [&](long c) { fn(c * chunk, std::min(n, (c + 1) * chunk)); });
}

}  // namespace pp
