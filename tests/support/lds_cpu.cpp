// Test support: the solver machine with the LDS tag-line cache's logic (dds_gpu_amd/csrc/
// ddx_lds.hpp) run on the host over a plain array. The cache is write-through and private
// to its lane, so tables and node counts must equal the plain solver's.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../dds_gpu_amd/csrc/ddx_lds.hpp"
#include "ddx/machine.hpp"

using namespace ddx;

static u32 g_cache[kLdsStride];

struct HostLdsOps : SeqOps {
  static int probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& ent_reads) {
    return tt_probe(tt, k, need, rel, hint, ent_reads, LdsLines{g_cache});
  }
  static void store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 gen) {
    tt_store(tt, k, rel, lb, ub, best, gen, LdsLines{g_cache});
  }
  static void clear_tags(const TTView& tt) {
    SeqOps::clear_tags(tt);
    lds_reset(g_cache, false);
  }
};

// One lane over the batch, jobs in order; with_cache selects the ops. Returns nodes; hits out.
extern "C" uint64_t ddx_lds_run(const uint64_t* hands, int64_t n, int tt_log2, int jpd, int with_cache,
                                uint8_t* out, uint64_t* hits) {
  const size_t buckets = (size_t)1 << tt_log2;
  std::vector<u32> tags(buckets * kWays, 0);
  std::vector<Entry> ents(buckets * kWays);
  Batch B{hands, out, nullptr, n * jpd, jpd, SEED_PLAIN};
  Lane L;
  lane_init(L, TTView{tags.data(), ents.data(), (u32)(buckets - 1)});
  ArrayStack st;
  long long next = 0;
  auto take = [&]() { return next++; };
  lds_reset(g_cache, true);
  if (with_cache)
    while (L.stage != ST_EXIT) lane_step<HostLdsOps>(L, B, st, take);
  else
    while (L.stage != ST_EXIT) lane_step<SeqOps>(L, B, st, take);
  *hits = g_cache[4 + 16 * kLdsLines];
  return L.c.nodes;
}
