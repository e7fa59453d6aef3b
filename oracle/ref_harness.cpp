// Golden-vector writer around the reference's header-only CPU solver (test infrastructure).
//
// Built by oracle/Makefile against the reference sources where they lie (-I$(REF)/src);
// only its outputs are committed, as fixtures under tests/golden/.
//
//   ref_harness tables HANDS.bin N TT_LOG2 TABLES.bin NODES.bin [SEEDS.bin]
//       One context, jobs in the order (deal 0 strain 0..4, deal 1 strain 0..4, ...):
//       dd_solve_strain (src/dd.h:926-939), or dd_solve_strain_guess (src/dd.h:915-924)
//       when SEEDS.bin (n x 20 uint8 NS-trick guesses, [strain][declarer]) is given.
//       Writes n x 20 uint8 tables ([strain C,D,H,S,NT][declarer N,E,S,W]) and n x 5
//       uint64 node counts (DDCtx::nodes per job).
//   ref_harness funcs N SEED OUT.bin
//       N random positions with the per-function outputs of the bound, ordering and
//       key helpers (record layout: FUNC_FIELDS int64 words, see below).
//   ref_harness time HANDS.bin N THREADS TT_LOG2
//       Multi-threaded dd_solve_strain over the deals, timed like cpu/cpu_main.cpp:
//       prints "deals threads seconds deals_per_s nodes".
#include <atomic>
#include <chrono>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <thread>
#include <vector>

#include "dd.h"

static std::vector<uint64_t> read_u64(const char* path, size_t count) {
  std::vector<uint64_t> v(count);
  FILE* f = fopen(path, "rb");
  if (!f || fread(v.data(), 8, count, f) != count) {
    fprintf(stderr, "cannot read %zu words from %s\n", count, path);
    exit(1);
  }
  fclose(f);
  return v;
}

static void write_all(const char* path, const void* p, size_t bytes) {
  FILE* f = fopen(path, "wb");
  if (!f || fwrite(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot write %s\n", path);
    exit(1);
  }
  fclose(f);
}

static int mode_tables(int argc, char** argv) {
  if (argc < 7) return 2;
  long n = atol(argv[3]);
  int tt_log2 = atoi(argv[4]);
  std::vector<uint64_t> hands = read_u64(argv[2], n * 4);
  std::vector<uint8_t> seeds;
  if (argc > 7) {
    seeds.resize(n * 20);
    FILE* f = fopen(argv[7], "rb");
    if (!f || fread(seeds.data(), 1, n * 20, f) != (size_t)(n * 20)) return 1;
    fclose(f);
  }
  std::vector<DDBucket> tt((size_t)1 << tt_log2);
  DDCtx c{};
  c.tt = tt.data();
  c.tt_mask = (1u << tt_log2) - 1;
  std::vector<uint8_t> tables(n * 20);
  std::vector<uint64_t> nodes(n * 5);
  for (long i = 0; i < n; i++) {
    for (int st = 0; st < 5; st++) {
      uint64_t before = c.nodes;
      int out4[4];
      if (!seeds.empty()) {
        int ns4[4];
        for (int d = 0; d < 4; d++) ns4[d] = seeds[i * 20 + st * 4 + d];
        dd_solve_strain_guess(c, &hands[i * 4], st, ns4, out4);
      } else {
        dd_solve_strain(c, &hands[i * 4], st, out4);
      }
      for (int d = 0; d < 4; d++) tables[i * 20 + st * 4 + d] = (uint8_t)out4[d];
      nodes[i * 5 + st] = c.nodes - before;
    }
  }
  write_all(argv[5], tables.data(), tables.size());
  write_all(argv[6], nodes.data(), nodes.size() * 8);
  return 0;
}

static int mode_time(int argc, char** argv) {
  if (argc < 6) return 2;
  long n = atol(argv[3]);
  int threads = atoi(argv[4]), tt_log2 = atoi(argv[5]);
  std::vector<uint64_t> hands = read_u64(argv[2], n * 4);
  std::vector<int> got(n * 20);
  std::atomic<long> next{0};
  std::atomic<uint64_t> nodes{0};
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&] {
      std::vector<DDBucket> tt((size_t)1 << tt_log2);
      DDCtx c{};
      c.tt = tt.data();
      c.tt_mask = (1u << tt_log2) - 1;
      for (long job; (job = next++) < n * 5;)
        dd_solve_strain(c, &hands[(job / 5) * 4], job % 5, &got[(job / 5) * 20 + (job % 5) * 4]);
      nodes += c.nodes;
    });
  for (auto& th : pool) th.join();
  double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("%ld %d %.6f %.4f %llu\n", n, threads, s, n / s, (unsigned long long)nodes.load());
  return 0;
}

// --- per-function vectors -------------------------------------------------------------
// Record of FUNC_FIELDS int64 words:
//   0-3 hands now (cards of the current trick removed)   4 trump   5 leader   6 nplayed
//   7-9 trick[0..2] (-1 unused)   10 ns_won   11 left (tricks not finished)   12 target
//   13 hint card (-1 none)
//   trick start (nplayed 0), else 0:
//   14-17 suit codes   18 lengths   19 dd_hash(lengths, leader)
//   20 quick tricks q   21 its win-rank mask   22 decided   23 later-tricks verdict
//   (ns_lead ? !dd_later_min : dd_later_max)   24 its win-rank mask
//   2nd hand (nplayed 1), else 0:   25 dd_quick_tricks_2nd verdict   26 its mask
//   27 move count   28-40 moves best-first (-1 pad)   41-53 weight of each move
enum { FUNC_FIELDS = 54 };

struct Rng {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int below(int n) { return (int)((next() >> 33) % (uint64_t)n); }
};

static int mode_funcs(int argc, char** argv) {
  if (argc < 5) return 2;
  long n = atol(argv[2]);
  Rng rng{strtoull(argv[3], 0, 10)};
  std::vector<int64_t> rec(n * FUNC_FIELDS, 0);
  for (long r = 0; r < n; r++) {
    int64_t* o = &rec[r * FUNC_FIELDS];
    int k = 2 + rng.below(12);  // 2..13 cards per hand
    int cards[52];
    for (int i = 0; i < 52; i++) cards[i] = (i / 13) * 16 + i % 13;
    for (int i = 51; i > 0; i--) {
      int j = rng.below(i + 1);
      int t = cards[i];
      cards[i] = cards[j];
      cards[j] = t;
    }
    DDCtx c{};
    for (int i = 0; i < 4 * k; i++) c.hand[i % 4] |= 1ull << cards[i];
    c.trump = dd_strain_suit(rng.below(5));
    c.leader = rng.below(4);
    c.left = k;
    c.ns_won = rng.below(14 - k);
    int need = 1 + rng.below(k);
    int target = c.ns_won + need;
    int np = rng.below(4);
    c.nplayed = 0;
    uint64_t start_hand[4] = {c.hand[0], c.hand[1], c.hand[2], c.hand[3]};
    for (int p = 0; p < np; p++) {  // random legal cards into the current trick
      int seat = (c.leader + p) & 3;
      uint64_t legal = c.hand[seat];
      if (p) {
        uint64_t f = legal & DD_SUIT(c.trick[0] >> 4);
        if (f) legal = f;
      }
      int m = rng.below(dd_popc(legal));
      uint64_t x = legal;
      for (int i = 0; i < m; i++) x &= x - 1;
      int card = dd_lsb(x);
      c.hand[seat] ^= 1ull << card;
      c.trick[c.nplayed++] = card;
    }
    int seat = (c.leader + np) & 3;
    int hint = -1;
    if (rng.below(2)) {
      uint64_t x = c.hand[seat];
      int m = rng.below(dd_popc(x));
      for (int i = 0; i < m; i++) x &= x - 1;
      hint = dd_lsb(x);
    }
    for (int h = 0; h < 4; h++) o[h] = (int64_t)c.hand[h];
    o[4] = c.trump;
    o[5] = c.leader;
    o[6] = np;
    for (int i = 0; i < 3; i++) o[7 + i] = i < np ? c.trick[i] : -1;
    o[10] = c.ns_won;
    o[11] = c.left;
    o[12] = target;
    o[13] = hint;
    DDTop t;
    memset(&t, 0, sizeof t);
    if (np == 0) {
      uint64_t code[4];
      dd_suit_codes(c, code);
      uint64_t lens = dd_lengths(c.hand);
      for (int s = 0; s < 4; s++) o[14 + s] = (int64_t)code[s];
      o[18] = (int64_t)lens;
      o[19] = (int64_t)dd_hash(lens, c.leader);
      dd_top_init(t, c.hand, lens);
      bool ns_lead = (c.leader & 1) == 0;
      int cut_ns = need, cut_ew = c.left - need + 1;
      uint64_t wr = 0;
      bool decided;
      int q = dd_quick_tricks(c.hand, t, c.trump, c.leader, ns_lead ? cut_ns : cut_ew,
                              ns_lead ? cut_ew : cut_ns, wr, decided);
      o[20] = q;
      o[21] = (int64_t)wr;
      o[22] = decided;
      wr = 0;
      bool lt = ns_lead ? !dd_later_min(t, c.trump, c.leader, c.ns_won, c.left, target, wr)
                        : dd_later_max(t, c.trump, c.leader, c.ns_won, c.left, target, wr);
      o[23] = lt;
      o[24] = (int64_t)wr;
    } else if (np == 1) {
      uint64_t th[4] = {c.hand[0], c.hand[1], c.hand[2], c.hand[3]};
      th[c.leader] |= 1ull << c.trick[0];
      DDTop t2;
      dd_top_init(t2, th, dd_lengths(th));
      int second = (c.leader + 1) & 3;
      int need2 = target - c.ns_won;
      int cutoff = (second & 1) == 0 ? need2 : c.left - need2 + 1;
      uint64_t wr = 0;
      o[25] = dd_quick_tricks_2nd(c.hand, t2, c.trump, second, c.trick[0], cutoff, wr);
      o[26] = (int64_t)wr;
    }
    (void)start_hand;
    int moves[13];
    int nm = dd_moves(c, seat, hint, &t, moves);
    o[27] = nm;
    uint64_t all = c.hand[0] | c.hand[1] | c.hand[2] | c.hand[3];
    for (int i = 0; i < c.nplayed; i++) all |= 1ull << c.trick[i];
    DDFollow f;
    if (np) dd_follow_init(f, c.hand, c.trump, c.leader, c.trick, np);
    for (int i = 0; i < 13; i++) {
      o[28 + i] = i < nm ? moves[i] : -1;
      if (i >= nm) continue;
      int card = moves[i];
      int sc = np == 0 ? dd_lead_weight(c.hand, t, c.trump, c.left, seat, card,
                                        hint >= 0 && dd_equal(c.hand[seat], all, hint, card))
                       : dd_follow_weight(f, c.hand, c.trump, np, seat, card);
      if (np && hint >= 0 && dd_equal(c.hand[seat], all, hint, card)) sc = 1000;
      o[41 + i] = sc;
    }
  }
  write_all(argv[4], rec.data(), rec.size() * 8);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && !strcmp(argv[1], "tables")) return mode_tables(argc, argv);
  if (argc >= 2 && !strcmp(argv[1], "funcs")) return mode_funcs(argc, argv);
  if (argc >= 2 && !strcmp(argv[1], "time")) return mode_time(argc, argv);
  fprintf(stderr, "usage: %s tables|funcs|time ...\n", argv[0]);
  return 2;
}
