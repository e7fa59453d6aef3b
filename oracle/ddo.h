/* ddo: the CPU oracle of the double-dummy path (test infrastructure only).
 *
 * Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker,
 * never by the product. It restates WHAT the reference computes -- the trick-taking rules,
 * the NS-trick count of a lead and the 20-entry table convention -- with textbook exact
 * search (see ddo.c), not HOW it searches.
 */
#ifndef DDO_H
#define DDO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Deal `index` of the seeded stream `seed`, k cards per hand (the deal-generator spec of the
 * product, restated independently: SplitMix64, Lemire bounded draws, Fisher-Yates). */
void ddo_random_deal(uint64_t seed, int64_t index, int k, uint64_t h[4]);

/* Pgx key/value decoding (reference src/npy.h:47-63): hands from keys[4]; tricks of
 * declarer `seat` in strain `st` from values[4]. */
void ddo_pgx_hands(const int32_t keys[4], uint64_t h[4]);
int ddo_pgx_tricks(const int32_t values[4], int seat, int st);

/* Exact NS tricks with `leader` on lead, trump suit 0..3 or -1 (exact null-window search over
 * targets with an exact-position memo; reference semantics of dd_ns_tricks,
 * src/dd.h:877-891). Returns -1 on a malformed deal. */
int ddo_ns_tricks(const uint64_t h[4], int trump, int leader);

/* Same by unpruned minimax (reference tests/test_small.cpp:12-46); for k <= 5 only. */
int ddo_ns_tricks_brute(const uint64_t h[4], int trump, int leader);

/* The 20-entry table [strain C,D,H,S,NT][declarer N,E,S,W]: declarer d's leader is d+1; N/S
 * entries are NS tricks, E/W entries 13 - NS tricks, also for endings (reference
 * src/dd.h:915-939). brute != 0 uses the unpruned minimax. Returns 0, or -1 on a bad deal. */
int ddo_table(const uint64_t h[4], int brute, uint8_t out[20]);

/* Tables of n deals on `threads` threads. Returns 0 or -1. */
int ddo_tables(const uint64_t* hands, int64_t n, int threads, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
