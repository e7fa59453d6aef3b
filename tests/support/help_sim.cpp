// Test support: the batch solver's tail helpers (dds_gpu_amd/csrc/ddx_help.hpp) run on the CPU
// by stepping `lanes` solver lanes round-robin in one thread -- the same machine, the same
// helper code, host atomics replaced by plain operations. Tests compile this with g++ and
// compare the tables with the plain solver's, and read how much splitting happened.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../dds_gpu_amd/csrc/ddx_help.hpp"

using namespace ddx;

// stats[0] rounds, [1] records reserved, [2] records taken, [3] nodes, [4] lead answers
// written, [5] max lanes idle at once, [6] the round at which the batch queue ran empty
extern "C" int ddx_help_sim(const uint64_t* hands, int64_t n, int lanes, int tt_log2, int jpd,
                            uint64_t spawn_nodes, int64_t pool, uint8_t* out, int64_t* stats) {
  const size_t buckets = (size_t)1 << tt_log2;
  std::vector<u32> tags((size_t)lanes * buckets * kWays, 0);
  std::vector<Entry> ents((size_t)lanes * buckets * kWays);
  std::vector<Lane> L(lanes);
  std::vector<ArrayStack> st(lanes);
  std::vector<HelpLane> S(lanes);
  std::vector<HelpRec> recs((size_t)pool);
  std::vector<u32> done((size_t)pool, 0), lead_done((size_t)n * 20, 0);
  for (auto& r : recs) r.state = RS_WRITING;
  unsigned long long ctr[3] = {0, 0, 0}, leads_done = 0, next = 0;
  Batch B{hands, out, nullptr, n * jpd, jpd, SEED_PLAIN};
  B.lead_done = lead_done.data();
  B.leads_done = &leads_done;
  HelpCtx H{recs.data(), done.data(), ctr, pool, n * 20, spawn_nodes};
  memset(out, 0xFF, (size_t)n * 20);
  for (int l = 0; l < lanes; l++) {
    TTView tt{tags.data() + (size_t)l * buckets * kWays, ents.data() + (size_t)l * buckets * kWays,
              (u32)(buckets - 1)};
    lane_init(L[l], tt);
    help_lane_init(S[l]);
  }
  auto take = [&]() -> long long { return (long long)next++; };
  int64_t rounds = 0, max_idle = 0, drained = -1;
  for (;;) {
    bool any = false;
    for (int l = 0; l < lanes; l++) {
      Lane& x = L[l];
      if (x.stage == ST_EXIT) continue;
      any = true;
      if ((rounds & 7) == 0) help_checks(x, B, st[l], H, S[l], &next);
      help_cheap(x, B, st[l], H, S[l], take);
      if (x.stage == ST_ENTER) lane_enter(x, B, st[l]);
    }
    rounds++;
    if (drained < 0 && (long long)next >= B.jobs) drained = rounds;
    if ((int64_t)ctr[2] > max_idle) max_idle = (int64_t)ctr[2];
    if (!any) break;
    if (rounds > (int64_t)4e9) return -1;
  }
  uint64_t nodes = 0;
  for (auto& x : L) nodes += x.c.nodes;
  stats[0] = rounds;
  stats[1] = (int64_t)ctr[0];
  stats[2] = (int64_t)ctr[1];
  stats[3] = (int64_t)nodes;
  stats[4] = (int64_t)leads_done;
  stats[5] = max_idle;
  stats[6] = drained;
  return 0;
}
