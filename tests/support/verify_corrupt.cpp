// Host harness for the TT-hit verifier (DDX_VERIFY): solves strain jobs lead by lead and, when
// asked, overwrites the bounds of every entry of the current generation after the first lead
// (data only, the solver's code untouched), so later leads hit entries that lie. The verifier
// must stay silent on the clean runs and report the lies.
#define DDX_VERIFY 1
#include <stdio.h>

#include <vector>

#include "ddx/machine.hpp"

using namespace ddx;

extern "C" long long ddx_verify_run(const uint64_t* hands, long long n, int tt_log2, int corrupt,
                                    uint8_t* out, unsigned long long* verified,
                                    unsigned long long* bad) {
  const size_t buckets = (size_t)1 << tt_log2;
  std::vector<u32> tags(buckets * kWays + 16, 0);
  std::vector<Entry> ents(buckets * kWays, Entry{0, 0});
  TTView v;
  v.tags = (u32*)(((uintptr_t)tags.data() + 63) & ~(uintptr_t)63);
  v.ents = ents.data();
  v.mask = (u32)(buckets - 1);
  Lane L;
  lane_init(L, v);
  ArrayStack st;
  const Batch B{hands, out, nullptr, n * 5, 5, SEED_PLAIN};
  auto none = [&]() -> long long { return B.jobs; };
  for (long long job = 0; job < n * 5; job++) {
    lane_load_job(L, B, job);
    bool done = false;
    while (L.stage != ST_JOB && L.stage != ST_EXIT) {
      lane_step(L, B, st, none);
      if (corrupt && !done && L.decl > 0) {  // the first lead is answered: make its entries lie
        for (size_t b = 0; b < buckets; b++)
          for (int w = 0; w < kWays; w++) {
            const u32 t = v.tags[b * kWays + w];
            if (!t || ((t >> 2) & 63) != L.c.gen) continue;
            Entry& e = v.ents[b * kWays + ((t >> 8) & 15)];
            e.w0 &= ~(0xFFull << kQBits);  // lb = ub = 0: "NS take no more tricks"
          }
        done = true;
      }
    }
  }
  *verified = L.c.verified;
  *bad = L.c.bad;
  return (long long)L.c.nodes;
}
