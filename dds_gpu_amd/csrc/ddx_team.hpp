// Team-parallel node operations for gfx950: 16 lanes (a quarter of a wave64) run one search.
//
// Every lane of a team holds the same search state and follows the same control flow; only
// the data-parallel parts of a node are spread over the lanes:
//   TT probe / store  lane t reads (and rewrites) tag t of the bucket, so the 64-byte tag
//                     line moves in one coalesced access; lane w owns entry way w (it alone
//                     loads and stores it), so every lane only re-reads memory it wrote itself
//   move ordering     lane j weighs the j-th candidate card, then a 16-lane bitonic sort
//                     orders (weight, card) descending -- the order of the sequential
//                     insertion sort
// Each op follows the rules of its sequential version (include/ddx/tt.hpp, search.hpp), and the
// tables are identical. Node counts are not quite: over the 4,096 8-card endings at TT_LOG2 10
// the team kernel enters 68,056,658-68,059,338 nodes, different from run to run and always a
// little fewer than the 68,059,954 that one lane per search enters (the reference's count).
// These ops restated lane by lane on the host (tests/support/team_emu.cpp) enter exactly the
// reference's nodes, so the variation comes from how the kernel executes, not from the ops'
// logic; its cause is not identified.
// Requirement: whenever these run, all 16 lanes of the team are active with identical
// arguments.
#pragma once
#include <hip/hip_runtime.h>

#include "ddx/search.hpp"

namespace ddx {

struct Team16 {
  static constexpr int kSize = 16;

  __device__ __forceinline__ static int lane() { return (int)(__lane_id() & 15); }
  __device__ __forceinline__ static u32 ballot(bool p) {
    return (u32)((__ballot(p) >> (__lane_id() & 48)) & 0xFFFFull);
  }
  __device__ __forceinline__ static int bcast(int v, int src) { return __shfl(v, src, 16); }
  __device__ __forceinline__ static u32 bcast(u32 v, int src) { return (u32)__shfl((int)v, src, 16); }
  __device__ __forceinline__ static u64 bcast(u64 v, int src) {
    return (u64)__shfl((unsigned long long)v, src, 16);
  }
  __device__ __forceinline__ static long long bcast(long long v, int src) { return __shfl(v, src, 16); }
  __device__ __forceinline__ static u32 or32(u32 v) {
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) v |= (u32)__shfl_xor((int)v, m, 16);
    return v;
  }
  __device__ __forceinline__ static u64 or64(u64 v) {
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) v |= (u64)__shfl_xor((unsigned long long)v, m, 16);
    return v;
  }
  __device__ __forceinline__ static u32 min32(u32 v) {
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
      const u32 o = (u32)__shfl_xor((int)v, m, 16);
      v = o < v ? o : v;
    }
    return v;
  }

  // tt_probe, position-parallel: the first decisive matching entry in tag order decides;
  // otherwise the last matching entry with a best card gives the hint.
  __device__ static int probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint,
                              u64& ent_reads) {
    const int t = lane();
    const u32 my = tt.tags[(size_t)k.bucket * kWays + t];
    const bool tm = (my & ~0xF00u) == k.tag;
    const u32 m = ballot(tm);
    if (!m) return -1;
    const int my_way = (int)((my >> 8) & 15);
    const u32 ways = or32(tm ? 1u << my_way : 0u);
    int flags = 0, best = 63;  // bit 0 valid, 1 decisive, 2 true, 3 has a hint
    u64 r = 0;
    if ((ways >> t) & 1) {
      const Entry e = tt.ents[(size_t)k.bucket * kWays + t];
      if ((e.w0 & kQMask) == k.q && entry_matches(k, e.w0, e.w1, r)) {
        const int lb = (int)(e.w0 >> kQBits) & 15, ub = (int)(e.w0 >> (kQBits + 4)) & 15;
        best = (int)(e.w0 >> (kQBits + 8)) & 63;
        const bool dec = lb >= need || ub < need;
        flags = 1 | (dec ? 2 : 0) | (lb >= need ? 4 : 0) | (!dec && best != 63 ? 8 : 0);
      }
    }
    const int f_at = bcast(flags, my_way);  // the flags of this position's entry
    const u32 dm = ballot(tm && (f_at & 2));
    if (dm) {
      const int pd = __builtin_ctz(dm);
      const int wd = bcast(my_way, pd);
      ent_reads += (u64)__builtin_popcount(m & ((2u << pd) - 1));
      rel = bcast(r, wd);
      return (bcast(flags, wd) & 4) ? 1 : 0;
    }
    ent_reads += (u64)__builtin_popcount(m);
    const u32 hm = ballot(tm && (f_at & 8));
    if (hm) {
      const int wh = bcast(my_way, 31 - __builtin_clz(hm));
      hint = bcast(best, wh);
    }
    return -1;
  }

  // tt_store, position-parallel: tighten the first exact-key entry in tag order, else take a
  // way (least 2 * left + age, other generations first, oldest on ties) and shift the line.
  __device__ static void store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best,
                               u32 gen) {
    u64 w1, hi;
    if (!pack_owners(k, rel, w1, hi)) return;
    const int t = lane();
    u32* line = tt.tags + (size_t)k.bucket * kWays;
    Entry* base = tt.ents + (size_t)k.bucket * kWays;
    const u32 my = line[t];
    const bool tm = (my & ~0xF00u) == k.tag;
    const int my_way = (int)((my >> 8) & 15);
    if (ballot(tm)) {
      const u32 ways = or32(tm ? 1u << my_way : 0u);
      int exact = 0;
      u64 w0 = 0;
      if ((ways >> t) & 1) {
        const Entry e = base[t];
        w0 = e.w0;
        exact = (w0 & kQMask) == k.q && e.w1 == w1 && (w0 >> 53 << 53) == hi;
      }
      const u32 em = ballot(tm && bcast(exact, my_way));
      if (em) {
        const int we = bcast(my_way, __builtin_ctz(em));
        if (t == we) {
          const int olb = (int)(w0 >> kQBits) & 15, oub = (int)(w0 >> (kQBits + 4)) & 15;
          const int obest = (int)(w0 >> (kQBits + 8)) & 63;
          if (olb > lb) lb = olb;
          if (oub < ub) ub = oub;
          if (best < 0 && obest != 63) best = obest;
          base[t].w0 = pack_w0(k.q, lb, ub, best, hi);
        }
        return;
      }
    }
    int pos = kWays - 1, way;
    if (bcast(my, kWays - 1) != 0) {
      const int v = ((my >> 2) & 63) != gen ? -1 : 2 * (int)((my >> 12) & 15) + kWays - 1 - t;
      const u32 kmin = min32(((u32)(v + 1) << 4) | (u32)(15 - t));
      pos = 15 - (int)(kmin & 15);
      way = (int)((bcast(my, pos) >> 8) & 15);
    } else {
      way = __builtin_ctz(~ballot(my != 0));
    }
    const u32 prev = bcast(my, t > 0 ? t - 1 : 0);
    line[t] = t == 0 ? (k.tag | ((u32)way << 8)) : (t <= pos ? prev : my);
    if (t == way) {
      Entry v;
      v.w0 = pack_w0(k.q, lb, ub, best, hi);
      v.w1 = w1;
      base[t] = v;
    }
  }

  // gen_moves, candidate-parallel.
  __device__ static void moves(const Pos& p, int seat, int hint, const Top& tp, Frame& f) {
    const int t = lane();
    const u64 own = hand_of(p, seat);
    u64 legal = own;
    if (p.nplayed) {
      const u64 fs = own & suit_mask(trick_card(p.trick, 0) >> 4);
      if (fs) legal = fs;
    }
    const u64 all = all_cards(p) | trick_cards(p);
    Follow fw;
    if (p.nplayed) follow_init(fw, p.hand, p.trump, p.leader, p.trick, p.nplayed);
    u64 g = own, pass = ~(all & ~own) & kAllSuits;
    g |= pass & (g >> 1);
    pass &= pass >> 1;
    g |= pass & (g >> 2);
    pass &= pass >> 2;
    g |= pass & (g >> 4);
    pass &= pass >> 4;
    g |= pass & (g >> 8);
    const u64 rest = legal & ~(g >> 1);
    const int n = popc(rest);
    u32 key = 0;
    if (t < n) {
      int c = 0;  // the (t+1)-th highest candidate
#pragma unroll
      for (int s = 32; s; s >>= 1) c += popc(rest >> (c + s)) >= t + 1 ? s : 0;
      const bool hinted = hint >= 0 && equivalent(own, all, hint, c);
      int sc;
      if (p.nplayed == 0) sc = lead_weight(p.hand, tp, p.trump, p.left, seat, c, hinted);
      else sc = hinted ? 1000 : follow_weight(fw, p.hand, p.trump, p.nplayed, seat, c);
      key = ((u32)(sc + 8192) << 6) | (u32)c;
    }
#pragma unroll
    for (int kk = 2; kk <= 16; kk <<= 1)
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) {
        const u32 o = (u32)__shfl_xor((int)key, j, 16);
        const bool take_max = ((t & j) == 0) == ((t & kk) == 0);
        key = take_max ? (o > key ? o : key) : (o < key ? o : key);
      }
    f.mv = or64(t < 10 ? (u64)(key & 63) << (6 * t) : 0);
    f.mv2 = or32(t >= 10 && t < 13 ? (key & 63) << (6 * (t - 10)) : 0u);
    f.n = n;
  }

  // Generation wrap: lane t clears tag t of every bucket (one 64-byte line per store).
  __device__ static void clear_tags(const TTView& tt) {
    const int t = lane();
    const size_t nb = (size_t)tt.mask + 1;
    for (size_t b = 0; b < nb; b++) tt.tags[b * kWays + t] = 0;
  }
};

// One lane per search (the reference's mapping) or a team of 16 lanes per search: the ops and
// the team-wide job counter / broadcast used by the kernels.
template <int TEAM>
struct TeamOf;

// Team size of the wide kernels: 16, or 1 in the signature-entry and shared-TT builds (Team16's
// TT code reads and writes exact entries), whose open calls then reject team = 16.
constexpr int kWideTeam = DDX_TT_SIG || DDX_SHARED_TT ? 1 : 16;
template <>
struct TeamOf<1> {
  using Ops = SeqOps;
  __device__ static bool leader() { return true; }
  template <class V>
  __device__ static V bcast(V v, int) { return v; }
  __device__ static long long take(unsigned long long* next) { return (long long)atomicAdd(next, 1ull); }
};
template <>
struct TeamOf<16> {
  using Ops = Team16;
  __device__ static bool leader() { return Team16::lane() == 0; }
  __device__ static int bcast(int v, int src) { return Team16::bcast(v, src); }
  __device__ static long long bcast(long long v, int src) { return Team16::bcast(v, src); }
  __device__ static long long take(unsigned long long* next) {
    long long j = 0;
    if (Team16::lane() == 0) j = (long long)atomicAdd(next, 1ull);
    return Team16::bcast(j, 0);
  }
};

}  // namespace ddx
