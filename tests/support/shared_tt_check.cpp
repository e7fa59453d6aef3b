// Host check of the shared TT's torn-entry protection (include/ddx/tt_shared.hpp, DDX_SHARED_TT): an
// entry stored and probed back answers; the same entry with any one of its four words replaced
// by the word of another entry (what a reader sees when it meets a half-written way), or with
// its bounds changed behind the checksum, is a miss. Built and run by tests/test_shared_tt.py:
//   g++ -O2 -std=c++17 -DDDX_SHARED_TT=1 -Iinclude tests/support/shared_tt_check.cpp
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "ddx/tt.hpp"

using namespace ddx;

static u64 rng_state = 12345;
static u64 rnd() {  // SplitMix64
  u64 z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void random_hands(int k, u64* h) {
  int cards[52];
  for (int i = 0; i < 52; i++) cards[i] = (i / 13) * 16 + i % 13;
  for (int i = 51; i > 0; i--) {
    int j = (int)(rnd() % (u64)(i + 1)), t = cards[i];
    cards[i] = cards[j];
    cards[j] = t;
  }
  for (int p = 0; p < 4; p++) {
    h[p] = 0;
    for (int i = 0; i < k; i++) h[p] |= 1ull << cards[p * k + i];
  }
}

int main() {
  const int log2 = 9;
  std::vector<u32> tags(((size_t)1 << log2) * kWays + 16, 0);
  std::vector<SharedEntry> ents(((size_t)1 << log2) * kWays);
  TTView tt;
  tt.tags = tags.data();
  tt.ents = (Entry*)ents.data();
  tt.mask = (1u << log2) - 1;
  int answered = 0, torn_missed = 0, cases = 0;
  SharedEntry prev{};
  bool have_prev = false;
  for (int trial = 0; trial < 2000; trial++) {
    u64 h[4];
    const int k = 3 + (int)(rnd() % 11);
    random_hands(k, h);
    const int leader = (int)(rnd() & 3), trump = (int)(rnd() % 5) - 1;
    Key key;
    make_key(h, leader, k, (u32)(trump + 2), tt.mask, key);
    const u64 rel = key.all & ~(rnd() & rnd());  // most cards relevant
    const int lb = (int)(rnd() % (u64)(k + 1)), ub = lb + (int)(rnd() % (u64)(k - lb + 1));
    tt_store(tt, key, rel, lb, ub, -1, (u32)trial);
    int way = -1;
    for (int i = 0; i < kWays; i++)
      if (tt.tags[(size_t)key.bucket * kWays + i] == key.tag) way = i;
    if (way < 0) return printf("stored entry not found\n"), 1;
    SharedEntry* e = shared_bucket(tt, key.bucket) + way;
    const int need = lb > 0 ? lb : ub + 1;  // a need the stored bounds decide
    if (need > k) continue;
    cases++;
    u64 r = 0, reads = 0;
    int hint = -1;
    if (tt_probe(tt, key, need, r, hint, reads) != (lb >= need ? 1 : 0)) return printf("whole entry: no answer\n"), 1;
    answered++;
    const SharedEntry good = *e;
    if (have_prev) {
      for (int w = 0; w < 4; w++) {  // one word from another entry: a torn read
        SharedEntry t = good;
        u64* tw = &t.w0 + w;
        const u64* pw = &prev.w0 + w;
        if (*tw == *pw) continue;
        *tw = *pw;
        *e = t;
        if (tt_probe(tt, key, need, r, hint, reads) != -1) return printf("torn word %d answered\n", w), 1;
        torn_missed++;
      }
    }
    SharedEntry t = good;  // bounds changed without a new checksum
    t.w3 ^= 1;
    *e = t;
    if (tt_probe(tt, key, need, r, hint, reads) != -1) return printf("changed bounds answered\n"), 1;
    *e = good;
    prev = good;
    have_prev = true;
  }
  printf("shared TT: %d of %d stored entries answered, %d torn reads missed\n", answered, cases, torn_missed);
  return answered == cases && torn_missed > 3 * cases ? 0 : 1;
}
