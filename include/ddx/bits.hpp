// ddx core, part 1: card bit primitives and trick-start position keys.
//
// Header-only; every function compiles for the host (g++/clang) and for gfx950 device code
// (hipcc). Everything is 64-bit integer arithmetic on registers: no tables, no small local
// arrays with run-time indices (those would be placed in scratch memory on the GPU).
//
// Card encoding (shared with hsperr/dds_gpu, src/dd.h:13-14): card = suit * 16 + rank, a bit
// of a uint64_t hand mask; suit 0..3 = S,H,D,C; rank 0..12 = 2..A. Seats 0..3 = N,E,S,W.
// Strains 0..4 = C,D,H,S,NT.
#pragma once
#include <stddef.h>
#include <stdint.h>

// DDX_FN: free functions; DDX_M: member functions.
#if defined(__HIPCC__)
#define DDX_FN __host__ __device__ __forceinline__
#define DDX_M __host__ __device__ __forceinline__
#else
#define DDX_FN static inline __attribute__((always_inline))
#define DDX_M inline __attribute__((always_inline))
#endif

namespace ddx {

typedef uint64_t u64;
typedef uint32_t u32;

// 16-byte word group: one global_load_dwordx4 / global_store_dwordx4 on gfx950.
struct alignas(16) Q4 {
  u32 x, y, z, w;
};

// Bit counts (reference src/dd.h:26-46).
DDX_FN int popc(u64 x) { return __builtin_popcountll(x); }
DDX_FN int msb(u64 x) { return 63 - __builtin_clzll(x); }  // x != 0
DDX_FN int lsb(u64 x) { return __builtin_ctzll(x); }       // x != 0
DDX_FN u64 bit(int card) { return 1ull << card; }

// All 13 cards of suit s (reference DD_SUIT, src/dd.h:48).
DDX_FN u64 suit_mask(int s) { return 0x1FFFull << (16 * s); }
constexpr u64 kAllSuits = 0x1FFF1FFF1FFF1FFFull;

// Seat of `card` from the E, S and W masks alone: 0 when none of them holds it
// (reference dd_owner, src/dd.h:51-53).
DDX_FN int owner(const u64* h, int card) {
  return (int)(((h[1] >> card) & 1) | (((h[2] >> card) & 1) << 1) | (((h[3] >> card) & 1) * 3));
}

// Hand of a seat chosen at run time, by masks. Indexing the 4-hand array with a run-time
// seat -- or a select between its elements, which the compiler turns back into a load from a
// run-time address -- would force the hands (and the structs holding them) out of registers
// into scratch memory on the GPU.
DDX_FN u64 seat_hand(const u64* h, int seat) {
  return (h[0] & (0ull - (u64)(seat == 0))) | (h[1] & (0ull - (u64)(seat == 1))) |
         (h[2] & (0ull - (u64)(seat == 2))) | (h[3] & (0ull - (u64)(seat == 3)));
}

// Trick-winner order: same suit -> higher rank, else only a trump beats
// (reference dd_beats, src/dd.h:537-541).
DDX_FN bool beats(int card, int best, int trump) {
  int s = card >> 4, bs = best >> 4;
  return s == bs ? card > best : s == trump;
}

// ---------------------------------------------------------------------------------------
// Suit lengths, packed: nibble (hand h, suit s) at bit 60 - 16h - 4s (hand 0 / suit 0 in the
// top nibble; reference dd_lengths, src/dd.h:179-189). Four 16-bit-lane SWAR popcounts; a
// multiply folds one hand's four counts into 16 bits.
DDX_FN u64 lengths(const u64* h) {
  u64 k = 0;
#pragma unroll
  for (int p = 0; p < 4; p++) {
    u64 x = h[p] - ((h[p] >> 1) & 0x5555555555555555ull);
    x = (x & 0x3333333333333333ull) + ((x >> 2) & 0x3333333333333333ull);
    x = (x + (x >> 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x + (x >> 8)) & 0x00FF00FF00FF00FFull;
    k = (k << 16) | ((x * 0x1000010000100001ull) >> 48);
  }
  return k;
}
DDX_FN int len_of(u64 lens, int h, int s) { return (int)((lens >> (60 - 16 * h - 4 * s)) & 15); }

// Position hash. It must be a bijection of lens ^ (leader << 62) so that bucket index + tag
// + the 39 stored q bits identify the suit lengths exactly (entries do not store them).
// Odd multiplies and xor-shifts are invertible. Same mixing steps as the reference
// (src/dd.h:191-196) so that both solvers use the same buckets; the constants are the
// 64-bit golden ratio and the SplitMix64 multiplier.
DDX_FN u64 key_hash(u64 lens, int leader) {
  u64 x = (lens ^ ((u64)leader << 62)) * 0x9E3779B97F4A7C15ull;
  x ^= x >> 29;
  x *= 0xBF58476D1CE4E5B9ull;
  return x ^ (x >> 32);
}

// Packs, for every suit separately (16-bit lanes), the bits of a and b selected by m down
// to the low end of the lane, keeping their order. Parallel-suffix compress (Hacker's
// Delight 7-4) with the carries cut at lane boundaries (reference dd_compress,
// src/dd.h:131-148).
DDX_FN void lane_compress(u64& a, u64& b, u64 m) {
  const u64 keep1 = 0xFFFEFFFEFFFEFFFEull, keep2 = 0xFFFCFFFCFFFCFFFCull,
            keep4 = 0xFFF0FFF0FFF0FFF0ull, keep8 = 0xFF00FF00FF00FF00ull;
  a &= m;
  b &= m;
  u64 zeros = (~m << 1) & keep1;  // per bit: count of zeros of m below it (parity stages)
#pragma unroll
  for (int i = 0; i < 4; i++) {
    u64 p = zeros ^ ((zeros << 1) & keep1);
    p ^= (p << 2) & keep2;
    p ^= (p << 4) & keep4;
    p ^= (p << 8) & keep8;
    u64 move = p & m;
    int sh = 1 << i;
    m = (m ^ move) | (move >> sh);
    u64 t = a & move;
    a = (a ^ t) | (t >> sh);
    t = b & move;
    b = (b ^ t) | (t >> sh);
    zeros &= ~p;
  }
}

// Interleaves, in each 32-bit half, bits 0..15 (to even positions) with bits 16..31 (to odd
// positions). Four delta swaps (reference dd_shuffle, src/dd.h:151-160).
DDX_FN u64 zip16(u64 x) {
  u64 t = (x ^ (x >> 8)) & 0x0000FF000000FF00ull;
  x ^= t ^ (t << 8);
  t = (x ^ (x >> 4)) & 0x00F000F000F000F0ull;
  x ^= t ^ (t << 4);
  t = (x ^ (x >> 2)) & 0x0C0C0C0C0C0C0C0Cull;
  x ^= t ^ (t << 2);
  t = (x ^ (x >> 1)) & 0x2222222222222222ull;
  return x ^ t ^ (t << 1);
}

// Owner code of every suit: the suit's remaining cards, high to low, as 2-bit owners
// (bit 0 = E or W, bit 1 = S or W) under a leading 1 bit, i.e. code = 1 << 2*len | owners
// (reference dd_suit_codes, src/dd.h:165-175). Computed without a loop over cards:
// compress the two owner planes per suit, then interleave them.
DDX_FN void suit_codes(const u64* h, u64* code) {
  u64 all = h[0] | h[1] | h[2] | h[3];
  u64 b0 = h[1] | h[3], b1 = h[2] | h[3];
  lane_compress(b0, b1, all);
  const u64 lo = 0x0000FFFF0000FFFFull;
  u64 s02 = zip16((b0 & lo) | ((b1 & lo) << 16));
  u64 s13 = zip16(((b0 >> 16) & lo) | (b1 & ~lo));
  code[0] = (s02 & 0x3FFFFFF) | (1ull << (2 * popc(all & suit_mask(0))));
  code[1] = (s13 & 0x3FFFFFF) | (1ull << (2 * popc(all & suit_mask(1))));
  code[2] = ((s02 >> 32) & 0x3FFFFFF) | (1ull << (2 * popc(all & suit_mask(2))));
  code[3] = ((s13 >> 32) & 0x3FFFFFF) | (1ull << (2 * popc(all & suit_mask(3))));
}

// Number of cards in an owner code (reference dd_pattern_len, src/dd.h:217).
DDX_FN int code_len(u64 code) { return msb(code) >> 1; }

// The n highest remaining cards of suit s as a mask: the lowest rank r with n cards at or
// above it, found by a 4-step binary search on the popcount (r = 15 for n = 0; reference
// dd_top_cards, src/dd.h:200-205).
DDX_FN u64 top_n(u64 all, int s, int n) {
  u64 a = (all >> (16 * s)) & 0x1FFF;
  int r = 0;
#pragma unroll
  for (int k = 8; k; k >>= 1) r += popc(a >> (r + k)) >= n ? k : 0;
  return (a >> r << r) << (16 * s);
}

// Owner code of the relevant top cards of suit s: every remaining card at or above the
// lowest relevant card of the suit (just the leading 1 when the suit has no relevant card;
// reference dd_pattern, src/dd.h:208-215).
DDX_FN u64 relevant_code(u64 code, u64 all, u64 rel, int s) {
  u64 r = rel & suit_mask(s);
  if (!r) return 1;
  int low = lsb(r);
  int len = popc(all & suit_mask(s));
  int n = popc(all & suit_mask(s) & ~((1ull << low) - 1));
  return code >> (2 * (len - n));
}

// Two cards of one hand are interchangeable when they share a suit and no other hand holds
// a card between them (reference dd_equal, src/dd.h:554-559).
DDX_FN bool equivalent(u64 own, u64 all, int a, int b) {
  if ((a >> 4) != (b >> 4)) return false;
  int lo = a < b ? a : b, hi = a < b ? b : a;
  u64 between = ((1ull << hi) - 1) & ~((2ull << lo) - 1);
  return !(all & ~own & between);
}

// Strain C,D,H,S,NT (0..4) -> trump suit 3,2,1,0 or -1 (reference dd_strain_suit,
// src/dd.h:893).
DDX_FN int strain_trump(int strain) { return strain == 4 ? -1 : 3 - strain; }

}  // namespace ddx
