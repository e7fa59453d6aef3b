// Host harness: the TT probe and store, which read the matching entries of a tag line in
// batches, against a restatement that reads them one at a time in tag order (reference
// dd_tt_probe / dd_tt_store, src/dd.h:458-532), on random tag lines with 0..16 matching tags
// and entries that match the position, differ in q, or differ in the owner pattern.
#include <stdint.h>
#include <string.h>

#include <random>

#include "ddx/deal.hpp"
#include "ddx/tt.hpp"

using namespace ddx;

namespace {

int probe_one_by_one(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& reads) {
  const u32* t = tt.tags + (size_t)k.bucket * kWays;
  const Entry* base = tt.ents + (size_t)k.bucket * kWays;
  for (int i = 0; i < kWays; i++) {
    if ((t[i] & ~0xF00u) != k.tag) continue;
    const Entry e = base[(t[i] >> 8) & 15];
    reads++;
    if ((e.w0 & kQMask) != k.q) continue;
    u64 r;
    if (!entry_matches(k, e.w0, e.w1, r)) continue;
    const int lb = (int)(e.w0 >> kQBits) & 15, ub = (int)(e.w0 >> (kQBits + 4)) & 15;
    if (lb >= need || ub < need) {
      rel = r;
      return lb >= need ? 1 : 0;
    }
    const int b = (int)(e.w0 >> (kQBits + 8)) & 63;
    if (b != 63) hint = b;
  }
  return -1;
}

void store_one_by_one(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 gen) {
  u64 w1, hi;
  if (!pack_owners(k, rel, w1, hi)) return;
  TagLine L;
  load_line(tt, k.bucket, L);
  Entry* base = tt.ents + (size_t)k.bucket * kWays;
  for (int i = 0; i < kWays; i++) {
    if ((L.t[i] & ~0xF00u) != k.tag) continue;
    Entry* e = base + ((L.t[i] >> 8) & 15);
    const u64 w0 = e->w0;
    if ((w0 & kQMask) == k.q && e->w1 == w1 && (w0 >> 53 << 53) == hi) {
      int olb = (int)(w0 >> kQBits) & 15, oub = (int)(w0 >> (kQBits + 4)) & 15;
      const int obest = (int)(w0 >> (kQBits + 8)) & 63;
      if (olb > lb) lb = olb;
      if (oub < ub) ub = oub;
      if (best < 0 && obest != 63) best = obest;
      e->w0 = pack_w0(k.q, lb, ub, best, hi);
      return;
    }
  }
  const int way = take_way(L, k.tag, gen);
  store_line(tt, k.bucket, L);
  Entry v;
  v.w0 = pack_w0(k.q, lb, ub, best, hi);
  v.w1 = w1;
  base[way] = v;
}

struct alignas(64) Bucket {
  u32 tags[kWays];
  Entry ents[kWays];
};

}  // namespace

// Runs `cases` random lines; returns the number of disagreements. covered[n] counts the lines
// with n matching tags.
extern "C" long long ddx_tt_batch_check(long long cases, unsigned seed, long long* covered) {
  std::mt19937_64 rng(seed);
  long long bad = 0;
  for (long long c = 0; c < cases; c++) {
    const int kcards = 2 + (int)(rng() % 12);
    u64 h[4];
    random_deal(seed, c, kcards, h);
    const int leader = (int)(rng() % 4);
    const u32 gen = 1 + (u32)(rng() % 63);
    Key k;
    make_key(h, leader, kcards, gen, 0, k);
    Bucket b;
    memset(&b, 0, sizeof b);
    TTView tt{b.tags, b.ents, 0};
    int way_of[kWays];
    for (int i = 0; i < kWays; i++) way_of[i] = i;
    for (int i = kWays - 1; i > 0; i--) {
      const int j = (int)(rng() % (i + 1));
      const int x = way_of[i];
      way_of[i] = way_of[j];
      way_of[j] = x;
    }
    const int filled = (int)(rng() % (kWays + 1));
    const int pm = (int)(rng() % 5);  // 0..4 quarters of the tags match
    int matches = 0;
    for (int i = 0; i < filled; i++) {
      const u32 w = (u32)way_of[i];
      const bool match = (int)(rng() % 4) < pm;
      const u32 other_gen = 1 + (gen % 63);
      b.tags[i] = (match ? k.tag : (k.tag ^ (1u + (u32)(rng() % 3) * 0x10000u) ^ ((rng() & 1) ? (other_gen ^ gen) << 2 : 0))) |
                  (w << 8);
      if ((b.tags[i] & ~0xF00u) == k.tag) matches++;
      // The entry behind the tag: the position's own class, a q collision, or another pattern.
      const u64 rel = rng() & k.all;
      u64 w1, hi;
      const int kind = (int)(rng() % 3);
      if (!pack_owners(k, rel, w1, hi)) {
        w1 = rng();
        hi = 0;
      }
      const int lb = (int)(rng() % (kcards + 1)), ub = lb + (int)(rng() % (kcards + 1 - lb));
      const int best = (rng() & 1) ? (int)(rng() % 64) : -1;
      Entry e;
      e.w0 = pack_w0(kind == 1 ? (k.q ^ 1) : k.q, lb, ub, best == 63 ? -1 : best, hi);
      e.w1 = kind == 2 ? (w1 ^ (1ull << (16 + rng() % 40))) : w1;
      b.ents[w] = e;
    }
    covered[matches]++;
    for (int need = 0; need <= kcards + 1; need++) {
      u64 r1 = 0, r2 = 0, n1 = 0, n2 = 0;
      int h1 = -1, h2 = -1;
      const int a = tt_probe(tt, k, need, r1, h1, n1);
      const int e = probe_one_by_one(tt, k, need, r2, h2, n2);
      if (a != e || h1 != h2 || n1 != n2 || (a >= 0 && r1 != r2)) bad++;
    }
    // Stores: the batched and the one-by-one store on copies of the line.
    Bucket b1 = b, b2 = b;
    TTView t1{b1.tags, b1.ents, 0}, t2{b2.tags, b2.ents, 0};
    const u64 rel = rng() & k.all;
    const int lb = (int)(rng() % (kcards + 1)), ub = lb + (int)(rng() % (kcards + 1 - lb));
    const int best = (rng() & 1) ? (int)(rng() % 52) : -1;
    tt_store(t1, k, rel, lb, ub, best, gen);
    store_one_by_one(t2, k, rel, lb, ub, best, gen);
    if (memcmp(&b1, &b2, sizeof b1) != 0) bad++;
  }
  return bad;
}
