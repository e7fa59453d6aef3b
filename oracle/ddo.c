/* ddo: the CPU oracle of the double-dummy path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the checker;
 * the product (libddx) never links, loads or calls it.
 *
 * What it restates from the reference, and how each part is pinned:
 *   - trick rules: follow suit if able; the highest trump wins, else the highest card of the
 *     led suit (reference dd_beats/dd_winning/dd_play, src/dd.h:537-549, 739-756);
 *   - the NS trick count of a lead as an exact search over "can NS take >= t?" targets
 *     (reference dd_ns_tricks, src/dd.h:877-891); pinned by the unpruned minimax below on
 *     endings of up to 5 cards, by the reference CPU solver's tables (oracle/_ref) on 5- to
 *     8-card endings and by the DDS tables of the reference fixture data/deals_20k.npy;
 *   - the table convention: declarer d's leader is d+1, E/W entries are 13 - NS tricks
 *     (reference src/dd.h:915-939), pinned by the same fixtures;
 *   - Pgx decoding (reference src/npy.h:47-63), pinned by the DDS fixture (13 cards per hand,
 *     disjoint hands, tables equal to the solver's);
 *   - the deal generator of the product (a spec of this project, not the reference's
 *     std::shuffle), pinned by agreement with the product and by uniformity statistics.
 * The search itself is textbook and shares no heuristic with the reference: a null-window
 * alpha-beta over whole tricks with an exact-position cache of proven bounds (positions are
 * keyed by all four hands, the leader and the trump suit, so cached values are facts), and
 * only the provably safe pruning of equivalent cards (two cards of one hand with no other
 * remaining card between them, cards of the current trick counting as remaining).
 */
#include "ddo.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;

#define SUIT(s) (0x1FFFull << (16 * (s)))
#define ALL_SUITS 0x1FFF1FFF1FFF1FFFull

static int popc(u64 x) { return __builtin_popcountll(x); }
static int lsb(u64 x) { return __builtin_ctzll(x); }

/* ---- deal generator ---------------------------------------------------------------- */
static u64 mix(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static u64 rng_next(u64* s) {
  *s += 0x9E3779B97F4A7C15ull;
  return mix(*s);
}
static u32 rng_below(u64* s, u32 n) {
  u64 m = (rng_next(s) >> 32) * (u64)n;
  if ((u32)m < n) {
    u32 thr = (u32)(0u - n) % n;
    while ((u32)m < thr) m = (rng_next(s) >> 32) * (u64)n;
  }
  return (u32)(m >> 32);
}
void ddo_random_deal(uint64_t seed, int64_t index, int k, uint64_t h[4]) {
  u64 s = mix(seed ^ mix((u64)index + 0x632BE59BD9B4E019ull));
  int deck[52];
  for (int i = 0; i < 52; i++) deck[i] = 16 * (i / 13) + i % 13;
  for (int i = 51; i >= 1; i--) {
    int j = (int)rng_below(&s, (u32)i + 1);
    int t = deck[i];
    deck[i] = deck[j];
    deck[j] = t;
  }
  h[0] = h[1] = h[2] = h[3] = 0;
  for (int i = 0; i < 4 * k; i++) h[i & 3] |= 1ull << deck[i];
}

/* ---- Pgx ----------------------------------------------------------------------------- */
void ddo_pgx_hands(const int32_t keys[4], uint64_t h[4]) {
  h[0] = h[1] = h[2] = h[3] = 0;
  for (int s = 0; s < 4; s++)
    for (int k = 0; k < 13; k++) {  /* k = 0 is the ace, k = 1..12 the deuce..king */
      int seat = ((u32)keys[s] >> (24 - 2 * k)) & 3;
      int rank = k == 0 ? 12 : k - 1;
      h[seat] |= 1ull << (16 * s + rank);
    }
}
int ddo_pgx_tricks(const int32_t values[4], int seat, int st) {
  return ((u32)values[seat] >> (16 - 4 * st)) & 15;
}

/* ---- exact search ---------------------------------------------------------------------- */
typedef struct {
  u64 h[4];
  u32 meta; /* leader | trump+1 << 2 | valid << 8 */
  int8_t lb, ub;
} Slot;

#define CACHE_LOG2 19

typedef struct {
  u64 h[4];
  int trump;
  Slot* cache;
} Search;

/* Does card a beat card b (b played earlier)? */
static int wins_over(int a, int b, int trump) {
  if ((a >> 4) == (b >> 4)) return a > b;
  return (a >> 4) == trump;
}

static Slot* slot_of(Search* s, int leader) {
  u64 x = s->h[0] * 0x9E3779B97F4A7C15ull ^ s->h[1] * 0xC2B2AE3D27D4EB4Full ^
          s->h[2] * 0x165667B19E3779F9ull ^ s->h[3] * 0x27D4EB2F165667C5ull ^
          (u64)(leader * 4 + s->trump + 1);
  x = mix(x);
  return &s->cache[x & ((1u << CACHE_LOG2) - 1)];
}

static int trick_start(Search* s, int leader, int need);

/* Ply i of the current trick (cards c[0..i-1] already played). Returns whether NS take at
 * least `need` of the tricks from this trick on. */
static int ply(Search* s, int leader, int i, int* c, int need) {
  int seat = (leader + i) & 3;
  u64 own = s->h[seat], legal = own;
  if (i > 0 && (own & SUIT(c[0] >> 4))) legal = own & SUIT(c[0] >> 4);
  u64 present = s->h[0] | s->h[1] | s->h[2] | s->h[3];
  for (int j = 0; j < i; j++) present |= 1ull << c[j];
  int ns = (seat & 1) == 0;
  for (u64 rest = legal; rest; rest &= rest - 1) {
    int x = lsb(rest);
    /* The next higher remaining card of the suit is our own: x is equivalent to it. */
    u64 above = present & SUIT(x >> 4) & ~((2ull << x) - 1);
    if (above && ((own >> lsb(above)) & 1)) continue;
    s->h[seat] ^= 1ull << x;
    c[i] = x;
    int r;
    if (i == 3) {
      int w = 0;
      for (int j = 1; j < 4; j++)
        if (wins_over(c[j], c[w], s->trump)) w = j;
      int winner = (leader + w) & 3;
      r = trick_start(s, winner, need - ((winner & 1) == 0));
    } else {
      r = ply(s, leader, i + 1, c, need);
    }
    s->h[seat] ^= 1ull << x;
    if (ns && r) return 1;
    if (!ns && !r) return 0;
  }
  return ns ? 0 : 1;
}

static int trick_start(Search* s, int leader, int need) {
  int left = popc(s->h[leader]);
  if (need <= 0) return 1;
  if (need > left) return 0;
  Slot* sl = slot_of(s, leader);
  u32 meta = (u32)leader | (u32)(s->trump + 1) << 2 | 1u << 8;
  int hit = sl->meta == meta && !memcmp(sl->h, s->h, sizeof sl->h);
  if (hit) {
    if (sl->lb >= need) return 1;
    if (sl->ub < need) return 0;
  }
  int c[4];
  int r = ply(s, leader, 0, c, need);
  if (!hit) {
    memcpy(sl->h, s->h, sizeof sl->h);
    sl->meta = meta;
    sl->lb = 0;
    sl->ub = (int8_t)left;
  }
  if (r && sl->lb < need) sl->lb = (int8_t)need;
  if (!r && sl->ub > need - 1) sl->ub = (int8_t)(need - 1);
  return r;
}

static int valid_deal(const u64 h[4]) {
  int k = popc(h[0]);
  u64 all = 0;
  if (k < 1 || k > 13) return 0;
  for (int i = 0; i < 4; i++) {
    if ((h[i] & ~ALL_SUITS) || (all & h[i]) || popc(h[i]) != k) return 0;
    all |= h[i];
  }
  return 1;
}

static int ns_tricks_with(Search* s, const u64 h[4], int trump, int leader) {
  memcpy(s->h, h, sizeof s->h);
  s->trump = trump;
  int lo = 0, hi = popc(h[0]);
  while (lo < hi) {
    int t = (lo + hi + 1) / 2;
    if (trick_start(s, leader, t)) lo = t;
    else hi = t - 1;
  }
  return lo;
}

int ddo_ns_tricks(const uint64_t h[4], int trump, int leader) {
  if (!valid_deal(h) || trump < -1 || trump > 3 || leader < 0 || leader > 3) return -1;
  Search s;
  s.cache = (Slot*)calloc((size_t)1 << CACHE_LOG2, sizeof(Slot));
  if (!s.cache) return -1;
  int r = ns_tricks_with(&s, h, trump, leader);
  free(s.cache);
  return r;
}

/* ---- unpruned minimax ---------------------------------------------------------------- */
static int brute(u64* h, int trump, int leader, int i, int* c) {
  if (i == 0 && popc(h[leader]) == 0) return 0;
  int seat = (leader + i) & 3;
  u64 legal = h[seat];
  if (i > 0 && (h[seat] & SUIT(c[0] >> 4))) legal = h[seat] & SUIT(c[0] >> 4);
  int ns = (seat & 1) == 0, best = ns ? -1 : 99;
  for (u64 rest = legal; rest; rest &= rest - 1) {
    int x = lsb(rest), v;
    h[seat] ^= 1ull << x;
    c[i] = x;
    if (i == 3) {
      int w = 0;
      for (int j = 1; j < 4; j++)
        if (wins_over(c[j], c[w], trump)) w = j;
      int winner = (leader + w) & 3;
      int cc[4];
      v = ((winner & 1) == 0) + brute(h, trump, winner, 0, cc);
    } else {
      v = brute(h, trump, leader, i + 1, c);
    }
    h[seat] ^= 1ull << x;
    if (ns ? v > best : v < best) best = v;
  }
  return best;
}

int ddo_ns_tricks_brute(const uint64_t h[4], int trump, int leader) {
  if (!valid_deal(h) || popc(h[0]) > 5) return -1;
  u64 hh[4] = {h[0], h[1], h[2], h[3]};
  int c[4];
  return brute(hh, trump, leader, 0, c);
}

/* ---- tables ------------------------------------------------------------------------------ */
static int table_with(Search* s, const u64 h[4], int use_brute, uint8_t out[20]) {
  if (!valid_deal(h)) return -1;
  for (int st = 0; st < 5; st++) {
    int trump = st == 4 ? -1 : 3 - st;
    for (int d = 0; d < 4; d++) {
      int leader = (d + 1) & 3;
      int ns = use_brute ? ddo_ns_tricks_brute(h, trump, leader) : ns_tricks_with(s, h, trump, leader);
      if (ns < 0) return -1;
      out[st * 4 + d] = (uint8_t)((d & 1) == 0 ? ns : 13 - ns);
    }
  }
  return 0;
}

int ddo_table(const uint64_t h[4], int use_brute, uint8_t out[20]) {
  Search s;
  s.cache = use_brute ? NULL : (Slot*)calloc((size_t)1 << CACHE_LOG2, sizeof(Slot));
  if (!use_brute && !s.cache) return -1;
  int r = table_with(&s, h, use_brute, out);
  free(s.cache);
  return r;
}

typedef struct {
  const u64* hands;
  int64_t n;
  uint8_t* out;
  atomic_llong next;
  atomic_int bad;
} Work;

static void* worker(void* arg) {
  Work* w = (Work*)arg;
  Search s;
  s.cache = (Slot*)calloc((size_t)1 << CACHE_LOG2, sizeof(Slot));
  if (!s.cache) {
    atomic_store(&w->bad, 1);
    return NULL;
  }
  for (;;) {
    long long i = atomic_fetch_add(&w->next, 1);
    if (i >= w->n) break;
    if (table_with(&s, w->hands + 4 * i, 0, w->out + 20 * i) != 0) atomic_store(&w->bad, 1);
  }
  free(s.cache);
  return NULL;
}

int ddo_tables(const uint64_t* hands, int64_t n, int threads, uint8_t* out) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  Work w;
  w.hands = hands;
  w.n = n;
  w.out = out;
  atomic_init(&w.next, 0);
  atomic_init(&w.bad, 0);
  pthread_t th[256];
  for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &w);
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  return atomic_load(&w.bad) ? -1 : 0;
}
