// ddx core, part 6: seeded deal generation.
//
// The reference deals with std::shuffle over mt19937_64 (gpu/gpu_main.cu:154-163,
// gpu/stream_main.cu:116-122), whose algorithm is implementation-defined, so this project
// fixes its own: deal `index` of stream `seed` is an independent SplitMix64 sequence
// (state = mix(seed ^ mix(index + C))), an unbiased Lemire bounded draw per step, and a
// Fisher-Yates shuffle of the 52 cards (i/13)*16 + i%13; card i of the shuffled deck goes to
// hand i%4 for the first 4k cards (k cards per hand; k = 13 is a full deal), as in the
// reference. Deals are random-access by index, so any slice can be regenerated anywhere.
#pragma once
#include "bits.hpp"

namespace ddx {

DDX_FN u64 mix64(u64 z) {  // SplitMix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct DealRng {
  u64 s;
  DDX_M u64 next() {
    s += 0x9E3779B97F4A7C15ull;
    return mix64(s);
  }
  // Uniform in [0, n), n <= 2^32 (Lemire, with rejection: exactly uniform).
  DDX_M u32 below(u32 n) {
    u64 m = (next() >> 32) * (u64)n;
    u32 l = (u32)m;
    if (l < n) {
      u32 thr = (0u - n) % n;
      while (l < thr) {
        m = (next() >> 32) * (u64)n;
        l = (u32)m;
      }
    }
    return (u32)(m >> 32);
  }
};

DDX_FN void random_deal(u64 seed, long long index, int k, u64* h) {
  DealRng r;
  r.s = mix64(seed ^ mix64((u64)index + 0x632BE59BD9B4E019ull));
  int cards[52];
  for (int i = 0; i < 52; i++) cards[i] = (i / 13) * 16 + i % 13;
  for (int i = 51; i > 0; i--) {
    int j = (int)r.below((u32)(i + 1));
    int t = cards[i];
    cards[i] = cards[j];
    cards[j] = t;
  }
  h[0] = h[1] = h[2] = h[3] = 0;
  for (int i = 0; i < 4 * k; i++) h[i % 4] |= 1ull << cards[i];
}

}  // namespace ddx
