// ThreadSanitizer driver of the set-up code's host worker pool
// (csrc/pp_pool.h): many parallel regions of varying size from one caller,
// then four callers at once; every index must be visited exactly once.
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "pp_pool.h"

int main() {
  long bad = 0;
  for (int rep = 0; rep < 2000; rep++) {
    const long n = 1 + (rep * 7919L) % 5000;
    std::vector<int> hit(n, 0);
    pp::pool_parallel(n, [&](long a, long b) {
      for (long i = a; i < b; i++) hit[i]++;
    });
    for (long i = 0; i < n; i++) bad += hit[i] != 1;
  }
  std::atomic<long> bad2{0};
  std::vector<std::thread> callers;
  for (int t = 0; t < 4; t++)
    callers.emplace_back([&, t] {
      for (int rep = 0; rep < 300; rep++) {
        const long n = 64 + (rep * 31L + t) % 3000;
        std::vector<int> hit(n, 0);
        pp::pool_parallel(n, [&](long a, long b) {
          for (long i = a; i < b; i++) hit[i]++;
        });
        for (long i = 0; i < n; i++) bad2 += hit[i] != 1;
      }
    });
  for (auto &c : callers) c.join();
  if (bad || bad2) {
    std::printf("pool: %ld + %ld indices not visited exactly once\n", bad, bad2.load());
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
