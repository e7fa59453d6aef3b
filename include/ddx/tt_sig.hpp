// ddx core, part 2b: 12-byte signature TT entries (build flag DDX_TT_SIG = 1; reference
// -DDD_TT_SIG, src/dd.h:74-85, 352-437). Included by tt.hpp, whose tag line, victim rule and
// helpers these entries share.
#pragma once
#include "tt.hpp"

#if DDX_TT_SIG  // empty in the other builds

namespace ddx {

// ---- Signature entries -----------------------------------------------------------------
// A 12-byte entry keeps a 64-bit signature of its key -- the owner codes of the relevant
// cards and the suit lengths -- instead of the key itself, so it has no kMaxRel limit; it
// is not strictly exact: a wrong hit needs two keys with the same bucket, the same tag and
// the same signature. Tag line, victim rule and visiting order are the exact entries'.
//   sig (u64) | meta (u32): lb (4) | ub (4) | best card (6, 63 = none) | 0 (2) |
//                           relevant-card counts per suit n0..n3 (4 bits each, bits 16..31)
struct alignas(4) SigEntry {
  u32 lo, hi, meta;
  DDX_M u64 sig() const { return (u64)hi << 32 | lo; }
};

DDX_FN SigEntry* sig_bucket(const TTView& tt, u32 b) {
  return (SigEntry*)((char*)tt.ents + (size_t)b * kBucketEntryBytes);
}

// Signature of a key: the suit lengths, then the owner codes of suits 0-1 and 2-3 (27 bits
// each, leading 1 included), chained through fmix64. The tag is not hashed: a probe only
// compares entries whose tag (leader, generation, tricks left) already matches.
DDX_FN u64 key_sig(const u64* p, u64 lens) {
  u64 h = fmix64(lens ^ 0x2545F4914F6CDD1Dull);
  h = fmix64(h ^ (p[0] | p[1] << 27));
  return fmix64(h + (p[2] | p[3] << 27));
}

// The owner codes of the top counts[s] cards per suit (counts: 4 bits per suit); false when a
// suit has fewer cards left than the entry covers.
DDX_FN bool sig_codes(const Key& k, u32 counts, u64* p) {
  bool fit = true;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int n = (int)(counts >> (4 * s)) & 15, len = code_len(k.code[s]);
    fit = fit && n <= len;
    p[s] = n <= len ? k.code[s] >> (2 * (len - n)) : 1;
  }
  return fit;
}

DDX_FN u32 sig_meta(u32 counts, int lb, int ub, int best) {
  return (u32)lb | (u32)ub << 4 | (u32)(best < 0 ? 63 : best) << 8 | counts << 16;
}

template <int N>
struct SigQueue {
  SigEntry e[N];
  DDX_M void load(const SigEntry* base, u64 ways) {
#pragma unroll
    for (int j = 0; j < N; j++) e[j] = base[(ways >> (4 * j)) & 15];
  }
  DDX_M SigEntry pop() {
    const SigEntry f = e[0];
#pragma unroll
    for (int j = 0; j + 1 < N; j++) e[j] = e[j + 1];
    return f;
  }
};

// Bounds lookup with signature entries (reference dd_tt_probe under -DDD_TT_SIG,
// src/dd.h:369-398): matching tags newest first; an entry answers when every suit still holds
// the cards it covers and its signature equals that of the current owners of those cards.
// Consecutive entries with the same counts share one signature computation.
template <class Lines = GlobalLines>
DDX_FN int tt_probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& ent_reads,
                    const Lines& lines = Lines()) {
  TagLine L;
  lines.load(tt, k.bucket, L);
  u64 ways;
  int n = match_ways(L, k.tag, ways);
  const SigEntry* base = sig_bucket(tt, k.bucket);
  u32 last = ~0u;
  u64 sig = 0;
  bool fit = false;
  while (n > 0) {
    SigQueue<kProbeBatch> q;
    q.load(base, ways);
    const int cnt = n < kProbeBatch ? n : kProbeBatch;
    for (int j = 0; j < cnt; j++) {
      const SigEntry e = q.pop();
      ent_reads++;
      const u32 counts = e.meta >> 16;
      if (counts != last) {
        u64 p[4];
        last = counts;
        fit = sig_codes(k, counts, p);
        sig = fit ? key_sig(p, k.lens) : 0;
      }
      if (!fit || e.sig() != sig) continue;
      const int lb = (int)(e.meta & 15), ub = (int)(e.meta >> 4) & 15;
      if (lb >= need || ub < need) {
        u64 r = 0;
#pragma unroll
        for (int s = 0; s < 4; s++) r |= top_n(k.all, s, (int)(counts >> (4 * s)) & 15);
        rel = r;
        return lb >= need ? 1 : 0;
      }
      const int b = (int)(e.meta >> 8) & 63;
      if (b != 63) hint = b;
    }
    ways >>= 4 * kProbeBatch;
    n -= kProbeBatch;
  }
  return -1;
}

// Store with signature entries (reference dd_tt_store under -DDD_TT_SIG, src/dd.h:400-434):
// an entry with the same signature and counts is tightened in place, else a way is taken by
// the victim rule of the exact entries. Any number of relevant cards is stored.
template <class Lines = GlobalLines>
DDX_FN void tt_store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 gen,
                     const Lines& lines = Lines()) {
  u64 p[4];
  u32 counts = 0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    p[s] = relevant_code(k.code[s], k.all, rel, s);
    counts |= (u32)code_len(p[s]) << (4 * s);
  }
  const u64 sig = key_sig(p, k.lens);
  TagLine L;
  lines.load(tt, k.bucket, L);
  SigEntry* base = sig_bucket(tt, k.bucket);
  u64 ways;
  int n = match_ways(L, k.tag, ways);
  while (n > 0) {
    SigQueue<kStoreBatch> q;
    q.load(base, ways);
    const int cnt = n < kStoreBatch ? n : kStoreBatch;
    for (int j = 0; j < cnt; j++) {
      const SigEntry e = q.pop();
      const int w = (int)(ways & 15);
      ways >>= 4;
      if (e.sig() == sig && (e.meta >> 16) == counts) {
        const int olb = (int)(e.meta & 15), oub = (int)(e.meta >> 4) & 15, obest = (int)(e.meta >> 8) & 63;
        if (olb > lb) lb = olb;
        if (oub < ub) ub = oub;
        if (best < 0 && obest != 63) best = obest;
        base[w].meta = sig_meta(counts, lb, ub, best);
        return;
      }
    }
    ways >>= 4 * (kStoreBatch - cnt);
    n -= kStoreBatch;
  }
  const int way = take_way(L, k.tag, gen);
  lines.store(tt, k.bucket, L);
  SigEntry v;
  v.lo = (u32)sig;
  v.hi = (u32)(sig >> 32);
  v.meta = sig_meta(counts, lb, ub, best);
  base[way] = v;
}

}  // namespace ddx

#endif  // DDX_TT_SIG
