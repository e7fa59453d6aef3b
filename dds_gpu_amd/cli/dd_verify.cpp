// dd_verify: the solver with every TT hit re-checked by a search of the same position and
// target that uses no TT (the reference's verify_one built with -DDD_VERIFY,
// tests/verify_one.cpp:1-15, src/dd.h:651-663). Built from the header-only core with
// DDX_VERIFY = 1; CPU only.
//   dd_verify N E S W STRAIN [TT_LOG2=14]     one deal: hands as hex masks (bit suit*16 + rank)
//   dd_verify K DEALS [SEED=1] [TT_LOG2=10]   DEALS random K-card endings, all 5 strains
// Prints "BAD TT HIT: ..." for every contradicted hit, then the tricks (one deal) and
// "verified H TT hits, B bad". Exit 0 = no bad hit, 1 = some, 2 = usage.
#define DDX_VERIFY 1
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "ddx/deal.hpp"
#include "ddx/machine.hpp"

using namespace ddx;

namespace {

struct Solver {
  std::vector<u32> tags;
  std::vector<Entry> ents;
  Lane L;
  ArrayStack st;
  explicit Solver(int tt_log2) {
    const size_t buckets = (size_t)1 << tt_log2;
    tags.assign(buckets * kWays + 16, 0);
    ents.assign(buckets * kWays, Entry{0, 0});
    TTView v;
    v.tags = (u32*)(((uintptr_t)tags.data() + 63) & ~(uintptr_t)63);
    v.ents = ents.data();
    v.mask = (u32)(buckets - 1);
    lane_init(L, v);
  }
  // The four leads of one strain (one TT generation), as the batch solver runs a job.
  void solve_strain(const u64 h[4], int strain, uint8_t out[20]) {
    const Batch B{h, out, nullptr, 5, 5, SEED_PLAIN};
    auto none = [&]() -> long long { return B.jobs; };
    lane_load_job(L, B, strain);
    while (L.stage != ST_JOB && L.stage != ST_EXIT) lane_step(L, B, st, none);
  }
};

bool valid(const u64 h[4]) {
  u64 all = 0;
  const int k = popc(h[0]);
  for (int i = 0; i < 4; i++) {
    if ((h[i] & ~kAllSuits) || (all & h[i]) || popc(h[i]) != k || k == 0) return false;
    all |= h[i];
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 6) {
    u64 h[4];
    for (int i = 0; i < 4; i++) h[i] = strtoull(argv[1 + i], nullptr, 16);
    const int strain = atoi(argv[5]);
    const int tt_log2 = argc > 6 ? atoi(argv[6]) : 14;
    if (!valid(h) || strain < 0 || strain > 4 || tt_log2 < kMinLog2 || tt_log2 > 24) {
      fprintf(stderr, "usage: %s N E S W STRAIN [TT_LOG2=14] (hex hands of equal length)\n", argv[0]);
      return 2;
    }
    Solver s(tt_log2);
    uint8_t out[20] = {0};
    s.solve_strain(h, strain, out);
    printf("tricks N E S W: %d %d %d %d\n", out[4 * strain], out[4 * strain + 1], out[4 * strain + 2],
           out[4 * strain + 3]);
    printf("verified %llu TT hits, %llu bad\n", (unsigned long long)s.L.c.verified,
           (unsigned long long)s.L.c.bad);
    return s.L.c.bad ? 1 : 0;
  }
  if (argc >= 3) {
    const int k = atoi(argv[1]);
    const long long n = atoll(argv[2]);
    const u64 seed = argc > 3 ? strtoull(argv[3], nullptr, 10) : 1;
    const int tt_log2 = argc > 4 ? atoi(argv[4]) : 10;
    if (k < 1 || k > 13 || n < 1 || tt_log2 < kMinLog2 || tt_log2 > 24) {
      fprintf(stderr, "usage: %s K DEALS [SEED=1] [TT_LOG2=10]\n", argv[0]);
      return 2;
    }
    Solver s(tt_log2);
    for (long long d = 0; d < n; d++) {
      u64 h[4];
      random_deal(seed, d, k, h);
      uint8_t out[20];
      for (int strain = 0; strain < 5; strain++) s.solve_strain(h, strain, out);
    }
    printf("%lld deals of %d cards: verified %llu TT hits, %llu bad\n", n, k,
           (unsigned long long)s.L.c.verified, (unsigned long long)s.L.c.bad);
    return s.L.c.bad ? 1 : 0;
  }
  fprintf(stderr, "usage: %s N E S W STRAIN [TT_LOG2=14] | %s K DEALS [SEED=1] [TT_LOG2=10]\n", argv[0],
          argv[0]);
  return 2;
}
