// ddx core, part 2c: one lock-free TT shared by every search (build flag DDX_SHARED_TT = 1;
// reference -DDD_SHARED_TT, src/dd.h:66-73, 246-323). Included by tt.hpp.
#pragma once
#include "tt.hpp"

#if DDX_SHARED_TT  // empty in the other builds

namespace ddx {

// ---- Shared TT -----------------------------------------------------------------------------
// One TT for every search, with no locks (reference -DDD_SHARED_TT). Entries carry their whole
// key -- the owner codes of the relevant cards, the exact suit lengths, the leader and the
// trump suit (in the tag's generation field: trump + 2; there are no generations, so an entry
// stays valid for every deal with that trump) -- and a 32-bit checksum, written word by word:
// a reader that meets a half-written entry sees a checksum mismatch and treats it as a miss,
// and every whole entry, however old, is a true bound. Tags are not kept in age order (moving
// them would race): tag i belongs to entry i, and a new entry overwrites one way in place --
// the first empty one, else one picked by the node count.
//   w0: owner codes of suits 0 and 1 (27 bits each, leading 1 included) | leader (2) << 54 |
//       trump + 2 (6) << 56
//   w1: owner codes of suits 2 and 3
//   w2: suit lengths
//   w3: lb (4) | ub (4) << 4 | best card (6, 63 = none) << 8 | checksum (32) << 32
struct alignas(32) SharedEntry {
  u64 w0, w1, w2, w3;
};

DDX_FN SharedEntry* shared_bucket(const TTView& tt, u32 b) {
  return (SharedEntry*)((char*)tt.ents + (size_t)b * kBucketEntryBytes);
}

// Racing accesses are relaxed atomics (agent scope on the GPU): each word is read exactly once,
// so the checksum a reader checks covers the very words it then uses; stale or torn entries
// fail the key and checksum checks.
template <class T>
DDX_FN T sh_load(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#endif
}
template <class T>
DDX_FN void sh_store(T* p, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  __atomic_store_n(p, v, __ATOMIC_RELAXED);
#endif
}

DDX_FN u64 shared_check(u64 w0, u64 w1, u64 w2, u64 w3) {
  const u64 h = fmix64(w0 ^ fmix64(w1 ^ (fmix64(w2 + 0x8CB92BA72F3D8DD7ull) + (w3 & 0x3FFF))));
  return h & 0xFFFFFFFF00000000ull;
}

DDX_FN u64 shared_meta(int lb, int ub, int best) {
  return (u64)lb | (u64)ub << 4 | (u64)(best < 0 ? 63 : best) << 8;
}

// Reads entry e; false when its words do not form a whole entry.
DDX_FN bool shared_read(const SharedEntry* e, u64& w0, u64& w1, u64& w2, u64& w3) {
  w0 = sh_load(&e->w0);
  w1 = sh_load(&e->w1);
  w2 = sh_load(&e->w2);
  w3 = sh_load(&e->w3);
  return (w3 & 0xFFFFFFFF00000000ull) == shared_check(w0, w1, w2, w3);
}

// Bounds lookup in the shared TT (reference dd_tt_probe under -DDD_SHARED_TT, src/dd.h:263-289):
// every way whose tag matches, in way order; an entry answers when it is whole, its trump,
// leader and suit lengths are the position's, and each suit's stored code is the top of the
// suit's current code.
template <class Lines = GlobalLines>
DDX_FN int tt_probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& ent_reads,
                    const Lines& = Lines()) {
  const u32* tags = tt.tags + (size_t)k.bucket * kWays;
  const SharedEntry* base = shared_bucket(tt, k.bucket);
  const u64 who = (u64)(k.tag & 3) << 54 | (u64)((k.tag >> 2) & 63) << 56;
  for (int i = 0; i < kWays; i++) {
    if (sh_load(tags + i) != k.tag) continue;
    u64 w0, w1, w2, w3;
    ent_reads++;
    if (!shared_read(base + i, w0, w1, w2, w3)) continue;
    if ((w0 & (0xFFull << 54)) != who || w2 != k.lens) continue;
    const u64 code[4] = {w0 & 0x7FFFFFF, (w0 >> 27) & 0x7FFFFFF, w1 & 0x7FFFFFF, (w1 >> 27) & 0x7FFFFFF};
    bool match = true;
    int n[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      n[s] = code_len(code[s]);
      const int len = code_len(k.code[s]);
      match = match && n[s] <= len && (k.code[s] >> (2 * (len - n[s]))) == code[s];
    }
    if (!match) continue;
    const int lb = (int)(w3 & 15), ub = (int)(w3 >> 4) & 15;
    if (lb >= need || ub < need) {
      u64 r = 0;
#pragma unroll
      for (int s = 0; s < 4; s++) r |= top_n(k.all, s, n[s]);
      rel = r;
      return lb >= need ? 1 : 0;
    }
    const int b = (int)(w3 >> 8) & 63;
    if (b != 63) hint = b;
  }
  return -1;
}

// Store in the shared TT (reference dd_tt_store under -DDD_SHARED_TT, src/dd.h:291-323): a whole
// entry with the same key is tightened in place (its bounds word rewritten with a new
// checksum); else the first empty way, or way (nodes ^ nodes >> 7) & 15 when none is empty,
// gets the entry. `nodes` is the searching context's node count.
template <class Lines = GlobalLines>
DDX_FN void tt_store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 nodes,
                     const Lines& = Lines()) {
  u64 c[4];
#pragma unroll
  for (int s = 0; s < 4; s++) c[s] = relevant_code(k.code[s], k.all, rel, s);
  const u64 w0 = c[0] | c[1] << 27 | (u64)(k.tag & 3) << 54 | (u64)((k.tag >> 2) & 63) << 56;
  const u64 w1 = c[2] | c[3] << 27;
  u32* tags = tt.tags + (size_t)k.bucket * kWays;
  SharedEntry* base = shared_bucket(tt, k.bucket);
  int victim = -1;
  for (int i = 0; i < kWays; i++) {
    const u32 t = sh_load(tags + i);
    if (t == 0 && victim < 0) victim = i;
    if (t != k.tag) continue;
    u64 e0, e1, e2, e3;
    if (!shared_read(base + i, e0, e1, e2, e3) || e0 != w0 || e1 != w1 || e2 != k.lens) continue;
    const int olb = (int)(e3 & 15), oub = (int)(e3 >> 4) & 15, obest = (int)(e3 >> 8) & 63;
    if (olb > lb) lb = olb;
    if (oub < ub) ub = oub;
    if (best < 0 && obest != 63) best = obest;
    const u64 m = shared_meta(lb, ub, best);
    sh_store(&base[i].w3, m | shared_check(w0, w1, k.lens, m));
    return;
  }
  if (victim < 0) victim = (int)((nodes ^ (nodes >> 7)) & (kWays - 1));
  const u64 m = shared_meta(lb, ub, best);
  SharedEntry* e = base + victim;
  sh_store(tags + victim, 0u);
  sh_store(&e->w0, w0);
  sh_store(&e->w1, w1);
  sh_store(&e->w2, k.lens);
  sh_store(&e->w3, m | shared_check(w0, w1, k.lens, m));
  sh_store(tags + victim, k.tag);
}

}  // namespace ddx

#endif  // DDX_SHARED_TT
