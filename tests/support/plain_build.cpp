// Test support: the solver machine over a batch of deals, one lane, jobs in order. The test
// compiles it twice: as is, and as the "plain" search with -DDDX_NO_TT -DDDX_NO_QT -DDDX_NO_LT
// -DDDX_NO_QT2 -DDDX_NO_SMALL (no TT, no quick or later tricks, no small-card skip; the
// reference's full-vs-plain check, tests/check.sh:17-24, tests/cmp_tt.cpp).
#include <stdint.h>

#include <vector>

#include "ddx/machine.hpp"

using namespace ddx;

extern "C" uint64_t ddx_plain_run(const uint64_t* hands, int64_t n, int tt_log2, uint8_t* out) {
  const size_t buckets = (size_t)1 << tt_log2;
  std::vector<u32> tags(buckets * kWays + 16, 0);
  std::vector<Entry> ents(buckets * kWays, Entry{0, 0});
  TTView v{(u32*)(((uintptr_t)tags.data() + 63) & ~(uintptr_t)63), ents.data(), (u32)(buckets - 1)};
  Batch B{hands, out, nullptr, n * 5, 5, SEED_PLAIN};
  Lane L;
  lane_init(L, v);
  ArrayStack st;
  long long next = 0;
  auto take = [&]() { return next++; };
  while (L.stage != ST_EXIT) lane_step(L, B, st, take);
  return L.c.nodes;
}
