// ddx core, part 2: the partition-search transposition table (TT).
//
// One TT serves one solver lane. It holds 2^log2 buckets; bucket b is
//   - a 64-byte tag line: 16 x u32 tags, newest first, at tags + 16*b, and
//   - 16 exact 16-byte entries at ents + 16*b.
// The two arrays are separate so a probe that finds no matching tag reads exactly one
// 64-byte line, and tag lines of one lane are packed 64 KiB apart from the entries (their
// generation clear then touches one contiguous range).
//
// Tag (u32), 0 = empty:  key-hash bits 48..63 (16) | tricks left (4) | way (4) |
//                        generation (6) | leader (2)
// Entry word 0:          q = key-hash bits 9..47 (39) | lb (4) | ub (4) |
//                        best card (6, 63 = none) | owner bits 48..58 (11)
// Entry word 1:          relevant-card counts per suit n0..n3 (4 bits each) |
//                        owner bits 0..47
// Owners: 2 bits per relevant card, suit 0 first, high card first.
// These formats, the victim rule and the hash are the reference's (src/dd.h:86-111,
// 330-350, 439-532) so the two solvers keep identical TT contents for the same job order;
// the code is ours and written for register-only execution on gfx950.
//
// The bucket index is hash bits 0..log2-1, the tag keeps bits 48..63 and the entry bits
// 9..47; since key_hash() is a bijection, bucket + tag + q fix the suit lengths, which is
// why entries need no length field. This requires log2 >= 9.
//
// Build flag DDX_TT_SIG = 1 (reference -DDD_TT_SIG, src/dd.h:74-85, 352-437) swaps the exact
// entries for 12-byte signature entries (tt_sig.hpp). Build flag DDX_SHARED_TT = 1 (reference
// -DDD_SHARED_TT, src/dd.h:66-73, 246-323) makes one TT serve every search, with 32-byte
// checksummed entries (tt_shared.hpp).
#pragma once
#include "bits.hpp"

#ifndef DDX_TT_SIG
#define DDX_TT_SIG 0
#endif
#ifndef DDX_SHARED_TT
#define DDX_SHARED_TT 0
#endif
#if DDX_TT_SIG && DDX_SHARED_TT
#error "DDX_TT_SIG and DDX_SHARED_TT are separate builds"
#endif

namespace ddx {

constexpr int kWays = 16;
constexpr int kGenBits = 6;
constexpr int kQBits = 39;
constexpr int kMinLog2 = 9;
constexpr int kMaxRel = 29;  // relevant cards one entry can hold
constexpr u64 kQMask = (1ull << kQBits) - 1;

struct alignas(16) Entry {
  u64 w0, w1;
};

// Bytes of one bucket's entries: 16 exact 16-byte entries, 16 signature entries of 12, or 16
// shared-TT entries of 32.
constexpr int kBucketEntryBytes = DDX_TT_SIG ? kWays * 12 : DDX_SHARED_TT ? kWays * 32 : kWays * 16;

// One lane's TT. `tags` is 64-byte aligned.
struct TTView {
  u32* tags;
  Entry* ents;
  u32 mask;  // buckets - 1
};

// Everything a trick-start node needs for its probe and, later, its store.
struct Key {
  u64 code[4];  // owner code per suit
  u64 all;      // all remaining cards
  u64 q;        // 39 hash bits kept in the entry
  u32 bucket;
  u32 tag;      // way bits 0
#if DDX_TT_SIG || DDX_SHARED_TT
  u64 lens;     // suit lengths: signature and shared entries keep them (exact ones imply them by q)
#endif
};

DDX_FN void make_key(const u64* h, int leader, int left, u32 gen, u32 mask, Key& k) {
  suit_codes(h, k.code);
  u64 lens = lengths(h);
  k.all = h[0] | h[1] | h[2] | h[3];
  u64 x = key_hash(lens, leader);
  k.bucket = (u32)x & mask;
  k.q = (x >> 9) & kQMask;
  k.tag = ((u32)(x >> 32) & 0xFFFF0000u) | ((u32)left << 12) | (gen << 2) | (u32)leader;
#if DDX_TT_SIG || DDX_SHARED_TT
  k.lens = lens;
#endif
}

// Word 1 and the top bits of word 0 for an entry covering the relevant cards `rel`; false
// when more than kMaxRel cards are relevant (such results are not stored).
DDX_FN bool pack_owners(const Key& k, u64 rel, u64& w1, u64& hi) {
  u64 counts = 0, owners = 0;
  int total = 0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    u64 p = relevant_code(k.code[s], k.all, rel, s);
    int n = code_len(p);
    if (total + n > kMaxRel) return false;
    counts |= (u64)n << (4 * s);
    owners |= (p ^ (1ull << (2 * n))) << (2 * total);
    total += n;
  }
  w1 = counts | (owners << 16);
  hi = (owners >> 48) << 53;
  return true;
}

DDX_FN u64 pack_w0(u64 q, int lb, int ub, int best, u64 hi) {
  return q | ((u64)lb << kQBits) | ((u64)ub << (kQBits + 4)) |
         ((u64)(best < 0 ? 63 : best) << (kQBits + 8)) | hi;
}

// Does entry (w0, w1) describe the position with owner codes k.code? Each suit's stored
// pattern must equal the current top n cards of that suit.
DDX_FN bool entry_matches(const Key& k, u64 w0, u64 w1, u64& rel) {
  u64 owners = (w1 >> 16) | ((w0 >> 53) << 48);
  int off = 0;
  bool ok = true;
  u64 r = 0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    int n = (int)(w1 >> (4 * s)) & 15;
    int len = code_len(k.code[s]);
    u64 p = (owners >> off & ((1ull << (2 * n)) - 1)) | (1ull << (2 * n));
    ok = ok && n <= len && (k.code[s] >> (2 * (len - n))) == p;
    off += 2 * n;
    r |= top_n(k.all, s, n);
  }
  rel = r;
  return ok;
}

// Loads the 16 tags of a bucket (one 64-byte line: four 16-byte loads on the GPU).
struct TagLine {
  u32 t[kWays];
};
DDX_FN void load_line(const TTView& tt, u32 b, TagLine& L) {
  const Q4* p = (const Q4*)(tt.tags + (size_t)b * kWays);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    Q4 v = p[i];
    L.t[4 * i + 0] = v.x;
    L.t[4 * i + 1] = v.y;
    L.t[4 * i + 2] = v.z;
    L.t[4 * i + 3] = v.w;
  }
}
DDX_FN void store_line(const TTView& tt, u32 b, const TagLine& L) {
  Q4* p = (Q4*)(tt.tags + (size_t)b * kWays);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    Q4 v;
    v.x = L.t[4 * i + 0];
    v.y = L.t[4 * i + 1];
    v.z = L.t[4 * i + 2];
    v.w = L.t[4 * i + 3];
    p[i] = v;
  }
}

// Where tag lines come from: straight from the TT in memory (the default), or through a
// per-search cache (the GPU kernel's LDS line cache) with the same interface.
struct GlobalLines {
  DDX_M void load(const TTView& tt, u32 b, TagLine& L) const { load_line(tt, b, L); }
  DDX_M void store(const TTView& tt, u32 b, const TagLine& L) const { store_line(tt, b, L); }
};

DDX_FN u32 tag_at(const TagLine& L, int pos) {  // run-time position, by masks (see seat_hand)
  u32 v = 0;
#pragma unroll
  for (int i = 0; i < kWays; i++) v |= L.t[i] & (0u - (u32)(i == pos));
  return v;
}

// Ways of the tags that equal `tag` apart from their way bits, newest first, 4 bits each
// (the first in bits 0..3); returns how many there are.
DDX_FN int match_ways(const TagLine& L, u32 tag, u64& ways) {
  u64 w = 0;
  int n = 0;
#pragma unroll
  for (int i = 0; i < kWays; i++) {
    const bool hit = (L.t[i] & ~0xF00u) == tag;
    w |= hit ? (u64)((L.t[i] >> 8) & 15) << (4 * n) : 0;
    n += hit ? 1 : 0;
  }
  ways = w;
  return n;
}

// The entries of the next N ways of `ways`, read with their loads issued together (one
// memory round trip for N dependent ones); ways past the matching ones read way 0, which is
// never examined. pop() returns the front entry and moves the others up one place (no
// run-time array index, so the N stay in registers).
template <int N>
struct EntryQueue {
  Entry e[N];
  DDX_M void load(const Entry* base, u64 ways) {
#pragma unroll
    for (int j = 0; j < N; j++) e[j] = base[(ways >> (4 * j)) & 15];
  }
  DDX_M Entry pop() {
    const Entry f = e[0];
#pragma unroll
    for (int j = 0; j + 1 < N; j++) e[j] = e[j + 1];
    return f;
  }
};
// Entries read together by a probe and by a store. Measured on gfx950 (one lane per search,
// full deals): 4 and 2 give 14% more nodes/s than one at a time; wider batches in the store
// cost registers (the store also keeps the tag line live) for no gain.
constexpr int kProbeBatch = 4;
constexpr int kStoreBatch = 2;

// Bounds lookup: 1 = NS reach `need` more tricks, 0 = they do not (rel = the relevant top
// cards of the entry), -1 = unknown; then `hint` may get a stored best card. Matching tags
// are visited newest first; a later matching entry's best card replaces an earlier one
// (reference dd_tt_probe, src/dd.h:458-493). The entries are read kProbeBatch at a time and still
// examined one by one in that order, so the result and `ent_reads` (entries examined) are
// those of a one-by-one read.
#if !DDX_TT_SIG && !DDX_SHARED_TT
template <class Lines = GlobalLines>
DDX_FN int tt_probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& ent_reads,
                    const Lines& lines = Lines()) {
  TagLine L;
  lines.load(tt, k.bucket, L);
  u64 ways;
  int n = match_ways(L, k.tag, ways);
  const Entry* base = tt.ents + (size_t)k.bucket * kWays;
  while (n > 0) {
    EntryQueue<kProbeBatch> q;
    q.load(base, ways);
    const int cnt = n < kProbeBatch ? n : kProbeBatch;
    for (int j = 0; j < cnt; j++) {
      const Entry e = q.pop();
      ent_reads++;
      if ((e.w0 & kQMask) != k.q) continue;
      u64 r;
      if (!entry_matches(k, e.w0, e.w1, r)) continue;
      int lb = (int)(e.w0 >> kQBits) & 15, ub = (int)(e.w0 >> (kQBits + 4)) & 15;
      if (lb >= need || ub < need) {
        rel = r;
        return lb >= need ? 1 : 0;
      }
      int b = (int)(e.w0 >> (kQBits + 8)) & 63;
      if (b != 63) hint = b;
    }
    ways >>= 4 * kProbeBatch;
    n -= kProbeBatch;
  }
  return -1;
}

#endif  // !DDX_TT_SIG && !DDX_SHARED_TT

// Takes a way for a new entry and moves its tag to the front of the line. A full line gives
// up the tag with the least 2 * tricks-left + age rank (15 = newest ... 0 = oldest), tags of
// other generations first (value -1), the oldest on ties; the tags in front of it move back
// one place and the new entry reuses its way. A line that is not full has ways 0..n-1 in use
// and gets way n (reference dd_tt_slot, src/dd.h:330-350).
DDX_FN int take_way(TagLine& L, u32 tag, u32 gen) {
  int pos = kWays - 1, way;
  if (L.t[kWays - 1]) {
    int least = 99;
#pragma unroll
    for (int p = kWays - 1; p >= 0; p--) {
      u32 t = L.t[p];
      int v = ((t >> 2) & 63) != gen ? -1 : 2 * (int)((t >> 12) & 15) + kWays - 1 - p;
      if (v < least) {
        least = v;
        pos = p;
      }
    }
    way = (int)(tag_at(L, pos) >> 8) & 15;
  } else {
    u32 used = 0;
#pragma unroll
    for (int i = 0; i < kWays; i++) used |= (u32)(L.t[i] != 0) << i;
    way = __builtin_ctz(~used);
  }
  // Shift positions 0..pos-1 back by one, new tag in front.
#pragma unroll
  for (int i = kWays - 1; i > 0; i--) L.t[i] = i <= pos ? L.t[i - 1] : L.t[i];
  L.t[0] = tag | ((u32)way << 8);
  return way;
}

// Stores bounds [lb, ub] for the position class (k, rel). An entry with the same key is
// tightened in place (its best card kept when none is given); otherwise a way is taken.
// Classes with more than kMaxRel relevant cards are not stored (reference dd_tt_store,
// src/dd.h:500-532).
#if !DDX_TT_SIG && !DDX_SHARED_TT
template <class Lines = GlobalLines>
DDX_FN void tt_store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 gen,
                     const Lines& lines = Lines()) {
  u64 w1, hi;
  if (!pack_owners(k, rel, w1, hi)) return;
  TagLine L;
  lines.load(tt, k.bucket, L);
  Entry* base = tt.ents + (size_t)k.bucket * kWays;
  u64 ways;
  int n = match_ways(L, k.tag, ways);
  while (n > 0) {  // matching entries kStoreBatch at a time, examined newest first
    EntryQueue<kStoreBatch> q;
    q.load(base, ways);
    const int cnt = n < kStoreBatch ? n : kStoreBatch;
    for (int j = 0; j < cnt; j++) {
      const Entry e = q.pop();
      const int w = (int)(ways & 15);
      ways >>= 4;
      const u64 w0 = e.w0;
      if ((w0 & kQMask) == k.q && e.w1 == w1 && (w0 >> 53 << 53) == hi) {
        int olb = (int)(w0 >> kQBits) & 15, oub = (int)(w0 >> (kQBits + 4)) & 15;
        int obest = (int)(w0 >> (kQBits + 8)) & 63;
        if (olb > lb) lb = olb;
        if (oub < ub) ub = oub;
        if (best < 0 && obest != 63) best = obest;
        base[w].w0 = pack_w0(k.q, lb, ub, best, hi);
        return;
      }
    }
    ways >>= 4 * (kStoreBatch - cnt);
    n -= kStoreBatch;
  }
  int way = take_way(L, k.tag, gen);
  lines.store(tt, k.bucket, L);
  Entry* e = base + way;
  Entry v;
  v.w0 = pack_w0(k.q, lb, ub, best, hi);
  v.w1 = w1;
  *e = v;
}
#endif  // !DDX_TT_SIG && !DDX_SHARED_TT

#if DDX_TT_SIG || DDX_SHARED_TT
// 64-bit finalizer of MurmurHash3 (fmix64, public domain): every input bit reaches every
// output bit.
DDX_FN u64 fmix64(u64 x) {
  x ^= x >> 33;
  x *= 0xFF51AFD7ED558CCDull;
  x ^= x >> 33;
  x *= 0xC4CEB9FE1A85EC53ull;
  return x ^ (x >> 33);
}
#endif

}  // namespace ddx

// The other entry formats, each a build of its own.
#if DDX_TT_SIG
#include "tt_sig.hpp"
#endif
#if DDX_SHARED_TT
#include "tt_shared.hpp"
#endif
