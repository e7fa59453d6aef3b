// LDS cache of TT tag lines for the one-lane-per-search kernel.
//
// A search's tag-line accesses repeat at short range: the node that probes bucket b at its
// entry stores into b when it finishes, and siblings hash to the same buckets. Replaying the
// solver's probe and store sequence on the CPU gives these hit rates for a small per-search
// cache (full deals, TT_LOG2 11): 1 line 17%, 3 lines FIFO 28%, 8 lines LRU 46%, 64 lines LRU
// 76%. LDS is 160 KB per CU; at 768 lanes per CU (three waves per SIMD) that is 213 bytes per
// lane, room for three lines. Each lane therefore keeps the last three tag lines it loaded,
// write-through: a store updates memory and the cached copy, and only the lane itself ever
// writes its TT, so the copy is never stale. A generation wrap clears the TT's tags and the
// cache with them. Entries (16 B each, read only for matching tags) stay in memory.
//
// Per lane, 53 dwords (odd, so lanes of a wave at the same offset hit different banks):
//   [0] next line to replace (FIFO), [1..3] bucket of each line (~0 = empty),
//   [4..51] three lines of 16 tags, [52] hits.
// The cache logic (LdsLines, lds_reset) also compiles for the host, where tests run it over a
// plain array and check that the solver's results and node counts do not change.
#pragma once
#include "ddx/search.hpp"

namespace ddx {

constexpr int kLdsLines = 3;
constexpr int kLdsStride = 53;
constexpr int kLdsBlock = 256;  // threads per block the LDS array is sized for

DDX_FN void lds_reset(u32* c, bool counters) {
  c[0] = 0;
#pragma unroll
  for (int i = 0; i < kLdsLines; i++) c[1 + i] = ~0u;
  if (counters) c[4 + 16 * kLdsLines] = 0;
}

struct LdsLines {
  u32* c;
  DDX_M void load(const TTView& tt, u32 b, TagLine& L) const {
    int hit = -1;
#pragma unroll
    for (int i = 0; i < kLdsLines; i++) hit = c[1 + i] == b ? i : hit;
    if (hit >= 0) {
      const u32* q = c + 4 + 16 * hit;
#pragma unroll
      for (int t = 0; t < kWays; t++) L.t[t] = q[t];
      c[4 + 16 * kLdsLines]++;
      return;
    }
    load_line(tt, b, L);
    const u32 v = c[0];
    c[1 + v] = b;
    u32* q = c + 4 + 16 * v;
#pragma unroll
    for (int t = 0; t < kWays; t++) q[t] = L.t[t];
    c[0] = v + 1 == kLdsLines ? 0 : v + 1;
  }
  DDX_M void store(const TTView& tt, u32 b, const TagLine& L) const {
    store_line(tt, b, L);
#pragma unroll
    for (int i = 0; i < kLdsLines; i++) {
      if (c[1 + i] != b) continue;
      u32* q = c + 4 + 16 * i;
#pragma unroll
      for (int t = 0; t < kWays; t++) q[t] = L.t[t];
    }
  }
};

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
__device__ __forceinline__ u32* lds_lane() {
  __shared__ u32 cache[kLdsBlock * kLdsStride];
  return cache + threadIdx.x * kLdsStride;
}

// SeqOps with the tag lines going through the lane's LDS cache.
struct LdsOps : SeqOps {
  __device__ __forceinline__ static int probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint,
                                              u64& ent_reads) {
    return tt_probe(tt, k, need, rel, hint, ent_reads, LdsLines{lds_lane()});
  }
  __device__ __forceinline__ static void store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best,
                                               u32 gen) {
    tt_store(tt, k, rel, lb, ub, best, gen, LdsLines{lds_lane()});
  }
  __device__ __forceinline__ static void clear_tags(const TTView& tt) {
    SeqOps::clear_tags(tt);
    lds_reset(lds_lane(), false);
  }
};
#endif

}  // namespace ddx
