// The shared TT under the thread sanitizer (g++ -fsanitize=thread -DDDX_SHARED_TT=1, host code
// only): four threads run their own lanes over one TT of 2^9 buckets, taking (deal, strain)
// jobs from one counter, as the shared-TT CPU batch does; every racing TT access is a relaxed
// atomic, so no race may be reported. Prints the tables' checksum.
//   tsan_shared_tt HANDS.bin N
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <thread>
#include <vector>

#include "ddx/machine.hpp"

using namespace ddx;
int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  const long n = atol(argv[2]);
  std::vector<uint64_t> hands(n * 4);
  if (!f || fread(hands.data(), 32, n, f) != (size_t)n) return 2;
  fclose(f);
  const size_t buckets = (size_t)1 << 9;
  std::vector<u32> tags(buckets * kWays + 16, 0);
  std::vector<SharedEntry> ents(buckets * kWays);
  TTView v{tags.data(), (Entry*)ents.data(), (u32)(buckets - 1)};
  std::vector<uint8_t> out(n * 20);
  Batch B{hands.data(), out.data(), nullptr, n * 5, 5, SEED_PLAIN};
  std::atomic<long long> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < 4; t++)
    pool.emplace_back([&] {
      Lane L;
      lane_init(L, v);
      ArrayStack* st = new ArrayStack;
      auto take = [&]() { return next.fetch_add(1); };
      while (L.stage != ST_EXIT) lane_step(L, B, *st, take);
      delete st;
    });
  for (auto& th : pool) th.join();
  unsigned long long s = 0;
  for (long i = 0; i < n * 20; i++) s += out[i];
  printf("sum %llu\n", s);
  return 0;
}
