// Host emulation of the 16-lane team ops of the GPU kernel (dds_gpu_amd/csrc/ddx_team.hpp): each
// op is restated lane by lane -- arrays indexed by the lane t of the team stand for its
// registers, loops over t for its ballots, broadcasts and butterfly reductions -- and plugged
// into the same machine as the sequential ops, so the search runs exactly as a team would
// if every memory access took effect in program order.
//
//   ddx_team_emu_solve(hands, n, tt_log2, out, job_nodes, ops, order)
//     one search over every (deal, strain) job, like ddx_cpu_solve_batch with one thread, with
//     the sequential ops (ops 0) or the emulated team ops (ops 1), jobs in index order (order 0)
//     or last job first (order 1); job_nodes gets the nodes of every job.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "ddx/machine.hpp"

using namespace ddx;

namespace {

constexpr int T = 16;

u32 ctz32(u32 x) { return (u32)__builtin_ctz(x); }

struct TeamEmu {
  static int probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& ent_reads) {
    u32 my[T], m = 0;
    bool tm[T];
    int my_way[T];
    for (int t = 0; t < T; t++) {
      my[t] = tt.tags[(size_t)k.bucket * kWays + t];
      tm[t] = (my[t] & ~0xF00u) == k.tag;
      m |= (u32)tm[t] << t;
      my_way[t] = (int)((my[t] >> 8) & 15);
    }
    if (!m) return -1;
    u32 ways = 0;
    for (int t = 0; t < T; t++) ways |= tm[t] ? 1u << my_way[t] : 0u;
    int flags[T], best[T];
    u64 r[T];
    for (int t = 0; t < T; t++) {
      flags[t] = 0;
      best[t] = 63;
      r[t] = 0;
      if ((ways >> t) & 1) {
        const Entry e = tt.ents[(size_t)k.bucket * kWays + t];
        if ((e.w0 & kQMask) == k.q && entry_matches(k, e.w0, e.w1, r[t])) {
          const int lb = (int)(e.w0 >> kQBits) & 15, ub = (int)(e.w0 >> (kQBits + 4)) & 15;
          best[t] = (int)(e.w0 >> (kQBits + 8)) & 63;
          const bool dec = lb >= need || ub < need;
          flags[t] = 1 | (dec ? 2 : 0) | (lb >= need ? 4 : 0) | (!dec && best[t] != 63 ? 8 : 0);
        }
      }
    }
    u32 dm = 0, hm = 0;
    for (int t = 0; t < T; t++) {
      const int f_at = flags[my_way[t]];
      dm |= (u32)(tm[t] && (f_at & 2)) << t;
      hm |= (u32)(tm[t] && (f_at & 8)) << t;
    }
    if (dm) {
      const int pd = (int)ctz32(dm), wd = my_way[pd];
      ent_reads += (u64)__builtin_popcount(m & ((2u << pd) - 1));
      rel = r[wd];
      return (flags[wd] & 4) ? 1 : 0;
    }
    ent_reads += (u64)__builtin_popcount(m);
    if (hm) hint = best[my_way[31 - __builtin_clz(hm)]];
    return -1;
  }

  static void store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 gen) {
    u64 w1, hi;
    if (!pack_owners(k, rel, w1, hi)) return;
    u32* line = tt.tags + (size_t)k.bucket * kWays;
    Entry* base = tt.ents + (size_t)k.bucket * kWays;
    u32 my[T], any = 0;
    bool tm[T];
    int my_way[T];
    for (int t = 0; t < T; t++) {
      my[t] = line[t];
      tm[t] = (my[t] & ~0xF00u) == k.tag;
      any |= (u32)tm[t] << t;
      my_way[t] = (int)((my[t] >> 8) & 15);
    }
    if (any) {
      u32 ways = 0;
      for (int t = 0; t < T; t++) ways |= tm[t] ? 1u << my_way[t] : 0u;
      int exact[T];
      u64 w0[T];
      for (int t = 0; t < T; t++) {
        exact[t] = 0;
        w0[t] = 0;
        if ((ways >> t) & 1) {
          const Entry e = base[t];
          w0[t] = e.w0;
          exact[t] = (w0[t] & kQMask) == k.q && e.w1 == w1 && (w0[t] >> 53 << 53) == hi;
        }
      }
      u32 em = 0;
      for (int t = 0; t < T; t++) em |= (u32)(tm[t] && exact[my_way[t]]) << t;
      if (em) {
        const int we = my_way[ctz32(em)];
        const int olb = (int)(w0[we] >> kQBits) & 15, oub = (int)(w0[we] >> (kQBits + 4)) & 15;
        const int obest = (int)(w0[we] >> (kQBits + 8)) & 63;
        if (olb > lb) lb = olb;
        if (oub < ub) ub = oub;
        if (best < 0 && obest != 63) best = obest;
        base[we].w0 = pack_w0(k.q, lb, ub, best, hi);
        return;
      }
    }
    int pos = kWays - 1, way;
    if (my[kWays - 1] != 0) {
      u32 kmin = ~0u;
      for (int t = 0; t < T; t++) {
        const int v = ((my[t] >> 2) & 63) != gen ? -1 : 2 * (int)((my[t] >> 12) & 15) + kWays - 1 - t;
        const u32 key = ((u32)(v + 1) << 4) | (u32)(15 - t);
        kmin = key < kmin ? key : kmin;
      }
      pos = 15 - (int)(kmin & 15);
      way = (int)((my[pos] >> 8) & 15);
    } else {
      u32 nz = 0;
      for (int t = 0; t < T; t++) nz |= (u32)(my[t] != 0) << t;
      way = (int)ctz32(~nz);
    }
    u32 next[T];
    for (int t = 0; t < T; t++) {
      const u32 prev = my[t > 0 ? t - 1 : 0];
      next[t] = t == 0 ? (k.tag | ((u32)way << 8)) : (t <= pos ? prev : my[t]);
    }
    for (int t = 0; t < T; t++) line[t] = next[t];
    Entry v;
    v.w0 = pack_w0(k.q, lb, ub, best, hi);
    v.w1 = w1;
    base[way] = v;
  }

  static void moves(const Pos& p, int seat, int hint, const Top& tp, Frame& f) {
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
    u32 key[T];
    for (int t = 0; t < T; t++) {
      key[t] = 0;
      if (t < n) {
        int c = 0;
        for (int s = 32; s; s >>= 1) c += popc(rest >> (c + s)) >= t + 1 ? s : 0;
        const bool hinted = hint >= 0 && equivalent(own, all, hint, c);
        int sc;
        if (p.nplayed == 0) sc = lead_weight(p.hand, tp, p.trump, p.left, seat, c, hinted);
        else sc = hinted ? 1000 : follow_weight(fw, p.hand, p.trump, p.nplayed, seat, c);
        key[t] = ((u32)(sc + 8192) << 6) | (u32)c;
      }
    }
    for (int kk = 2; kk <= 16; kk <<= 1)
      for (int j = kk >> 1; j > 0; j >>= 1) {
        u32 next[T];
        for (int t = 0; t < T; t++) {
          const u32 o = key[t ^ j];
          const bool take_max = ((t & j) == 0) == ((t & kk) == 0);
          next[t] = take_max ? (o > key[t] ? o : key[t]) : (o < key[t] ? o : key[t]);
        }
        memcpy(key, next, sizeof key);
      }
    u64 mv = 0;
    u32 mv2 = 0;
    for (int t = 0; t < T; t++) {
      if (t < 10) mv |= (u64)(key[t] & 63) << (6 * t);
      if (t >= 10 && t < 13) mv2 |= (key[t] & 63) << (6 * (t - 10));
    }
    f.mv = mv;
    f.mv2 = mv2;
    f.n = n;
  }

  static void clear_tags(const TTView& tt) { SeqOps::clear_tags(tt); }
};

}  // namespace

template <class X>
static int solve(const uint64_t* hands, int64_t n, int tt_log2, uint8_t* out, uint64_t* job_nodes, int order) {
  const size_t buckets = (size_t)1 << tt_log2;
  std::vector<u32> tags(buckets * kWays + 16, 0);
  std::vector<Entry> ents(buckets * kWays);
  TTView v;
  v.tags = (u32*)(((uintptr_t)tags.data() + 63) & ~(uintptr_t)63);
  v.ents = ents.data();
  v.mask = (u32)(buckets - 1);
  Lane L;
  lane_init(L, v);
  ArrayStack* st = new ArrayStack;
  Batch B{hands, out, nullptr, (long long)n * 5, 5, 0};
  memset(out, 0xFF, (size_t)n * 20);
  long long cur = -1;
  uint64_t base = 0;
  long long taken = 0;
  auto take = [&]() -> long long {
    if (cur >= 0 && cur < B.jobs) job_nodes[cur] = L.c.nodes - base;
    const long long i = taken++;
    cur = i >= B.jobs ? B.jobs : order ? B.jobs - 1 - i : i;  // order 1: last job first
    base = L.c.nodes;
    return cur;
  };
  while (L.stage != ST_EXIT) lane_step<X>(L, B, *st, take);
  delete st;
  return 0;
}

// ops 0: the sequential ops, 1: the team ops emulated lane by lane.
extern "C" int ddx_team_emu_solve(const uint64_t* hands, int64_t n, int tt_log2, uint8_t* out,
                                  uint64_t* job_nodes, int ops, int order) {
  return ops ? solve<TeamEmu>(hands, n, tt_log2, out, job_nodes, order)
             : solve<SeqOps>(hands, n, tt_log2, out, job_nodes, order);
}
