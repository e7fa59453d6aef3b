// ddx core, part 4: the null-window search and its resumable stage machine.
//
// The search answers "can NS take >= target tricks?" with an explicit-stack depth-first
// search; the trick count of a lead comes from a sequence of such searches over targets
// (binary search, or stepping from a seed), all sharing one TT (reference src/dd.h:537-939,
// src/dd_wave.h). Every search also returns `rel`, the cards whose ranks decided it
// (partition search), which the TT stores.
//
// One machine serves the CPU and the GPU. It keeps the frame of the deepest expanded node
// ("top") in registers and writes it to the lane's stack only when a child expands, so a
// child decided at entry (TT hit, quick tricks, trivial bound) costs no stack traffic. The
// stack itself is a template parameter: a plain array on the CPU, the lane's slice of HBM
// on the GPU.
#pragma once
#include "bounds.hpp"
#include "tt.hpp"

namespace ddx {

// ---------------------------------------------------------------------------------------
// Position.
struct Pos {
  u64 hand[4];
  u32 trick;  // cards of the current trick, byte i = i-th card played
  int trump;  // 0..3, -1 = notrump
  int leader, nplayed, ns_won, left;  // left = tricks not finished
};

DDX_FN int trick_card(u32 trick, int i) { return (int)((trick >> (8 * i)) & 63); }
DDX_FN u64 all_cards(const Pos& p) { return p.hand[0] | p.hand[1] | p.hand[2] | p.hand[3]; }
DDX_FN u64 trick_cards(const Pos& p) {
  u64 m = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
    if (i < p.nplayed) m |= bit(trick_card(p.trick, i));
  return m;
}
// Index (play order) of the card winning a complete trick.
DDX_FN int trick_winner(u32 trick, int trump) {
  int w = 0, wc = trick_card(trick, 0);
#pragma unroll
  for (int i = 1; i < 4; i++) {
    int c = trick_card(trick, i);
    if (beats(c, wc, trump)) {
      w = i;
      wc = c;
    }
  }
  return w;
}
DDX_FN u64 hand_of(const Pos& p, int seat) { return seat_hand(p.hand, seat); }
DDX_FN void toggle(Pos& p, int seat, u64 m) {
#pragma unroll
  for (int i = 0; i < 4; i++) p.hand[i] ^= seat == i ? m : 0;
}

// ---------------------------------------------------------------------------------------
// Frames. The hot frame lives in registers; the packed form (32 bytes) on the stack.
struct Frame {
  u64 acc;   // relevant cards collected from the children
  u64 mv;    // moves 0..9, 6 bits each, best first
  u32 mv2;   // moves 10..12
  u32 save;  // the trick this ply completed (4 cards), for undo
  u32 low_tried;  // per suit (nibble s): 1 + rank of the lowest card tried, 0 = none
  u32 nodes0;     // the context's node count / 16 when the frame was made, modulo 2^24
  int n, i, best, result, seat, need, lb, ub;
};
struct alignas(16) PackedFrame {
  u64 acc, mv;
  u32 mv2;  // moves 10..12 (bits 0..17) | n (18..21) | i (22..25) | seat (26..27) | result (28)
  u32 save;
  u32 meta;  // low_tried (0..15) | best + 1 (16..22) | need (23..26)
  u32 bounds;  // lb (0..3) | ub (4..7) | nodes0 (8..31)
};

DDX_FN PackedFrame pack_frame(const Frame& f) {
  PackedFrame q;
  q.acc = f.acc;
  q.mv = f.mv;
  q.mv2 = f.mv2 | ((u32)f.n << 18) | ((u32)f.i << 22) | ((u32)f.seat << 26) | ((u32)f.result << 28);
  q.save = f.save;
  q.meta = f.low_tried | ((u32)(f.best + 1) << 16) | ((u32)f.need << 23);
  q.bounds = (u32)f.lb | ((u32)f.ub << 4) | (f.nodes0 << 8);
  return q;
}
DDX_FN Frame unpack_frame(const PackedFrame& q) {
  Frame f;
  f.acc = q.acc;
  f.mv = q.mv;
  f.mv2 = q.mv2 & 0x3FFFF;
  f.n = (int)((q.mv2 >> 18) & 15);
  f.i = (int)((q.mv2 >> 22) & 15);
  f.seat = (int)((q.mv2 >> 26) & 3);
  f.result = (int)((q.mv2 >> 28) & 1);
  f.save = q.save;
  f.low_tried = q.meta & 0xFFFF;
  f.best = (int)((q.meta >> 16) & 127) - 1;
  f.need = (int)((q.meta >> 23) & 15);
  f.lb = (int)(q.bounds & 15);
  f.ub = (int)((q.bounds >> 4) & 15);
  f.nodes0 = q.bounds >> 8;
  return f;
}
DDX_FN int move_at(const Frame& f, int i) {
  return i < 10 ? (int)((f.mv >> (6 * i)) & 63) : (int)((f.mv2 >> (6 * (i - 10))) & 63);
}

// DDX_VERIFY = 1 (host builds only, the dd_verify tool) re-checks every TT hit with a search of
// the same position and target that uses no TT (reference -DDD_VERIFY, src/dd.h:651-663).
#ifndef DDX_VERIFY
#define DDX_VERIFY 0
#endif

// Per-lane solver context: its TT, the generation of the current solve, counters.
struct Ctx {
  TTView tt;
  u32 gen;  // 1..63; tags of other generations count as empty
  u64 nodes;
  u64 probes, hits;
  u64 ent_reads, stores;  // TT entries read by probes; TT stores (algorithmic HBM bytes)
#if DDX_VERIFY
  int no_tt;              // this context's searches neither probe nor store
  u64 verified, bad;      // TT hits re-checked; hits the TT-free search contradicted
#endif
};

// Whether the context leaves the TT alone (only a verifying search does).
#if DDX_VERIFY
DDX_FN bool tt_off(const Ctx& c) { return c.no_tt != 0; }
struct Pos;
inline void verify_hit(Ctx& c, const Pos& p, int target, int hit, u64 rel);
#else
DDX_FN bool tt_off(const Ctx&) { return false; }
#endif

// ---------------------------------------------------------------------------------------
// Move generation: one candidate per run of equivalent cards, best first by the DDS
// weights; the TT's best card (or one equal to it) is favoured (reference dd_moves,
// src/dd.h:562-611). Candidates arrive high card first and each is pushed through a chain
// of 13 max/min exchanges, which keeps the list sorted by (weight, card) descending -- the
// order of the reference's stable insertion sort -- in registers.
DDX_FN void gen_moves(const Pos& p, int seat, int hint, const Top& t, Frame& f) {
  const u64 own = hand_of(p, seat);
  u64 legal = own;
  if (p.nplayed) {
    u64 fs = own & suit_mask(trick_card(p.trick, 0) >> 4);
    if (fs) legal = fs;
  }
  const u64 all = all_cards(p) | trick_cards(p);
  Follow fw;
  if (p.nplayed) follow_init(fw, p.hand, p.trump, p.leader, p.trick, p.nplayed);
  // A card whose next higher remaining card is its own hand's is skipped: flood every own
  // card down through ranks no other hand holds (ranks 13..15 stop the flood).
  u64 g = own, pass = ~(all & ~own) & kAllSuits;
  g |= pass & (g >> 1);
  pass &= pass >> 1;
  g |= pass & (g >> 2);
  pass &= pass >> 2;
  g |= pass & (g >> 4);
  pass &= pass >> 4;
  g |= pass & (g >> 8);
  u64 rest = legal & ~(g >> 1);
  u32 k[13];
#pragma unroll
  for (int q = 0; q < 13; q++) k[q] = 0;
  int n = 0;
  while (rest) {
    const int card = msb(rest);
    rest ^= bit(card);
    const bool hinted = hint >= 0 && equivalent(own, all, hint, card);
    int sc;
    if (p.nplayed == 0) sc = lead_weight(p.hand, t, p.trump, p.left, seat, card, hinted);
    else sc = hinted ? 1000 : follow_weight(fw, p.hand, p.trump, p.nplayed, seat, card);
    u32 key = ((u32)(sc + 8192) << 6) | (u32)card;  // weights stay far above -8192
#pragma unroll
    for (int q = 0; q < 13; q++) {
      u32 hi = k[q] > key ? k[q] : key;
      key = k[q] > key ? key : k[q];
      k[q] = hi;
    }
    n++;
  }
  u64 mv = 0;
  u32 mv2 = 0;
#pragma unroll
  for (int q = 0; q < 10; q++) mv |= (u64)(k[q] & 63) << (6 * q);
#pragma unroll
  for (int q = 10; q < 13; q++) mv2 |= (k[q] & 63) << (6 * (q - 10));
  f.mv = mv;
  f.mv2 = mv2;
  f.n = n;
}

// The data-parallel pieces of a node, as a policy: SeqOps runs them on one thread (the CPU, and
// the GPU kernel with one lane per search); the GPU team kernel supplies a version that spreads
// them over the lanes of a team. Everything else in a node runs the same code either way.
struct SeqOps {
  DDX_M static int probe(const TTView& tt, const Key& k, int need, u64& rel, int& hint, u64& ent_reads) {
    return tt_probe(tt, k, need, rel, hint, ent_reads);
  }
  DDX_M static void store(const TTView& tt, const Key& k, u64 rel, int lb, int ub, int best, u32 gen) {
    tt_store(tt, k, rel, lb, ub, best, gen);
  }
  DDX_M static void moves(const Pos& p, int seat, int hint, const Top& t, Frame& f) {
    gen_moves(p, seat, hint, t, f);
  }
  DDX_M static void clear_tags(const TTView& tt) {
    Q4 z = {0, 0, 0, 0};
    Q4* q = (Q4*)tt.tags;
    const size_t n = ((size_t)tt.mask + 1) * (kWays / 4);
    for (size_t i = 0; i < n; i++) q[i] = z;
  }
};

// New generation for a new solve; after 63 solves every tag of the TT is cleared
// (reference dd_next_gen, src/dd.h:864-874). The shared TT has no generations: its entries stay
// valid for every deal with the same trump.
template <class X = SeqOps>
DDX_FN void next_generation(Ctx& c) {
  if (DDX_SHARED_TT) return;
  c.gen++;
  if (c.gen >= (1u << kGenBits)) {
    X::clear_tags(c.tt);
    c.gen = 1;
  }
}

// What a tag's generation field holds: the solve's generation, or, in the shared TT, the trump
// suit + 2 (reference DD_KEY, src/dd.h:223-227).
DDX_FN u32 tag_key(const Ctx& c, const Pos& p) { return DDX_SHARED_TT ? (u32)(p.trump + 2) : c.gen; }
// The store's last input: the generation (a full tag line gives up other generations first), or,
// in the shared TT, the node count (which picks the way to overwrite when none is empty).
DDX_FN u32 store_aux(const Ctx& c) { return DDX_SHARED_TT ? (u32)c.nodes : c.gen; }

// Compile-time switches used only by the tests' "plain" build: no TT, no quick tricks, no
// later tricks, no second-hand quick tricks, no small-card skip (reference DD_NO_*,
// tests/check.sh:18).
#ifndef DDX_NO_TT
#define DDX_NO_TT 0
#endif
#ifndef DDX_NO_QT
#define DDX_NO_QT 0
#endif
#ifndef DDX_NO_LT
#define DDX_NO_LT 0
#endif
#ifndef DDX_NO_QT2
#define DDX_NO_QT2 0
#endif
#ifndef DDX_NO_SMALL
#define DDX_NO_SMALL 0
#endif

// Where enter() decided a node (profiling).
enum { EX_TRIVIAL, EX_LAST, EX_TT, EX_QT, EX_LT, EX_QT2, EX_MOVES, EX_COUNT };

// Node entry at the position p (reference dd_enter, src/dd.h:626-737). Returns 1/0 when the
// node is decided at once (`rel` = its relevant cards), else -1 with `f` holding the
// ordered moves. Splits into the trick-start part (TT, quick tricks, later tricks) and the
// rest so that the GPU can schedule them as separate stages.
template <class X = SeqOps>
DDX_FN int enter_lead(Ctx& c, const Pos& p, int target, Frame& f, u64& rel, int& exit_at) {
  c.nodes++;
  rel = 0;
  exit_at = EX_TRIVIAL;
  if (p.ns_won >= target) return 1;
  if (p.ns_won + p.left < target) return 0;
  const int need = target - p.ns_won;
  const u64 all = all_cards(p);
  if (p.left == 1) {
    // Last trick: the highest trump, else the highest card of the led suit, wins.
    const int s = p.trump >= 0 && (all & suit_mask(p.trump)) ? p.trump
                                                            : msb(hand_of(p, p.leader)) >> 4;
    const u64 in = all & suit_mask(s);
    const int top = msb(in);
    if (popc(in) >= 2) rel = bit(top);
    exit_at = EX_LAST;
    return (owner(p.hand, top) & 1) == 0 ? 1 : 0;
  }
  Key key;
  make_key(p.hand, p.leader, p.left, tag_key(c, p), c.tt.mask, key);
  int hint = -1;
  if (!DDX_NO_TT && !tt_off(c)) {
    int hit = X::probe(c.tt, key, need, rel, hint, c.ent_reads);
    c.probes++;
    c.hits += hit >= 0;
    exit_at = EX_TT;
#if DDX_VERIFY
    if (hit >= 0) verify_hit(c, p, target, hit, rel);
#endif
    if (hit >= 0) return hit;
  }
  Top t;
  top_init(t, p.hand, lengths(p.hand));
  const bool ns_lead = (p.leader & 1) == 0;
  const int cut_ns = need, cut_ew = p.left - need + 1;
  u64 wr = 0;
  bool decided;
  int q = quick_tricks(p.hand, t, p.trump, p.leader, ns_lead ? cut_ns : cut_ew,
                       ns_lead ? cut_ew : cut_ns, wr, decided);
  if (DDX_NO_QT) decided = false;
  if (decided) {
    const bool val = ns_lead ? q != 0 : q == 0;
    if (!tt_off(c)) {
      X::store(c.tt, key, wr, val ? need : 0, val ? p.left : need - 1, -1, store_aux(c));
      c.stores++;
    }
    rel = wr;
    exit_at = EX_QT;
    return val ? 1 : 0;
  }
  wr = 0;
  const bool lt = ns_lead ? !later_min(t, p.trump, p.leader, p.ns_won, p.left, target, wr)
                          : later_max(t, p.trump, p.leader, p.ns_won, p.left, target, wr);
  if (!DDX_NO_LT && lt) {
    const bool val = !ns_lead;
    if (!tt_off(c)) {
      X::store(c.tt, key, wr, val ? need : 0, val ? p.left : need - 1, -1, store_aux(c));
      c.stores++;
    }
    rel = wr;
    exit_at = EX_LT;
    return val ? 1 : 0;
  }
  X::moves(p, p.leader, hint, t, f);
  f.seat = p.leader;
  f.result = (p.leader & 1) ? 1 : 0;  // the side to move has not yet reached its goal
  f.need = need;
  f.lb = 0;
  f.ub = p.left;
  f.i = 0;
  f.best = -1;
  f.acc = 0;
  f.low_tried = 0;
  f.nodes0 = (u32)(c.nodes >> 4) & 0xFFFFFFu;
  exit_at = EX_MOVES;
  return -1;
}

template <class X = SeqOps>
DDX_FN int enter_follow(Ctx& c, const Pos& p, int target, Frame& f, u64& rel, int& exit_at) {
  c.nodes++;
  Top t;
  if (p.nplayed == 1 && !DDX_NO_QT2) {
    u64 th[4] = {p.hand[0], p.hand[1], p.hand[2], p.hand[3]};
    const int c0 = trick_card(p.trick, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) th[i] |= i == p.leader ? bit(c0) : 0;
    top_init(t, th, lengths(th));
    const int second = (p.leader + 1) & 3;
    const int need2 = target - p.ns_won;
    const int cutoff = (second & 1) == 0 ? need2 : p.left - need2 + 1;
    u64 wr;
    if (quick_tricks_2nd(p.hand, t, p.trump, second, c0, cutoff, wr)) {
      rel = wr;
      exit_at = EX_QT2;
      return (second & 1) == 0 ? 1 : 0;
    }
  }
  const int seat = (p.leader + p.nplayed) & 3;
  X::moves(p, seat, -1, t, f);
  f.seat = seat;
  f.result = (seat & 1) ? 1 : 0;
  f.need = 0;
  f.lb = 0;
  f.ub = 0;
  f.i = 0;
  f.best = -1;
  f.acc = 0;
  f.low_tried = 0;
  f.nodes0 = (u32)(c.nodes >> 4) & 0xFFFFFFu;
  exit_at = EX_MOVES;
  return -1;
}

template <class X = SeqOps>
DDX_FN int enter(Ctx& c, const Pos& p, int target, Frame& f, u64& rel, int& exit_at) {
  return p.nplayed == 0 ? enter_lead<X>(c, p, target, f, rel, exit_at)
                        : enter_follow<X>(c, p, target, f, rel, exit_at);
}

// Plays `card` for the frame's seat; a fourth card settles the trick (reference dd_play,
// src/dd.h:739-756).
DDX_FN void play(Pos& p, Frame& f, int card) {
  toggle(p, f.seat, bit(card));
  p.trick = (p.trick & ~(0xFFu << (8 * p.nplayed))) | ((u32)card << (8 * p.nplayed));
  p.nplayed++;
  if (p.nplayed == 4) {
    const int w = (p.leader + trick_winner(p.trick, p.trump)) & 3;
    f.save = p.trick;
    p.leader = w;
    p.nplayed = 0;
    p.left--;
    p.ns_won += (w & 1) == 0;
  }
}

// Undoes play(); a trick won by rank (two or more cards of the winning suit) makes its
// winning card relevant (reference dd_unplay, src/dd.h:758-769).
DDX_FN void unplay(Pos& p, const Frame& f, int card, u64& rel) {
  if (p.nplayed == 0) {
    p.trick = f.save;
    p.ns_won -= (p.leader & 1) == 0;
    p.left++;
    p.nplayed = 4;
    p.leader = (f.seat + 1) & 3;  // the fourth hand is the leader's right-hand opponent
    const int wi = trick_winner(f.save, p.trump);
    const int wc = trick_card(f.save, wi);
    int same = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) same += (trick_card(f.save, i) >> 4) == (wc >> 4);
    if (same >= 2) rel |= bit(wc);
  }
  p.nplayed--;
  toggle(p, f.seat, bit(card));
}

// All moves of f are done (or one cut off): widen its relevant cards and, at a trick start,
// store the result (reference dd_finish, src/dd.h:772-797). Only the top card of a run of
// equivalent cards was tried, so when the lowest relevant card of a suit heads such a run
// the rest of the run must be relevant too -- else a TT entry could match a position where
// another hand's card splits the run.
template <class X = SeqOps>
DDX_FN u64 finish(Ctx& c, const Pos& p, const Frame& f) {
  u64 acc = f.acc;
  const u64 own = hand_of(p, f.seat);
  const u64 all = all_cards(p) | trick_cards(p);
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const u64 r = acc & suit_mask(s);
    if (!r) continue;
    const int low = lsb(r);
    if (!((own >> low) & 1)) continue;
    // The own cards directly below `low` in the order of the remaining cards.
    const u64 below = all & suit_mask(s) & (bit(low) - 1);
    const u64 other = below & ~own;
    acc |= other ? below & ~((2ull << msb(other)) - 1) : below;
  }
  if (p.nplayed == 0 && !DDX_NO_TT && !tt_off(c)) {
    Key key;
    make_key(p.hand, p.leader, p.left, tag_key(c, p), c.tt.mask, key);
    int lb = f.lb, ub = f.ub;
    if (f.result) lb = f.need;
    else ub = f.need - 1;
    X::store(c.tt, key, acc, lb, ub, f.best, store_aux(c));
    c.stores++;
  }
  return acc;
}

// Whether `card` is a small card f need not try: below every relevant card of its suit
// collected so far, with a small card of that suit already tried.
DDX_FN bool small_skip(const Frame& f, int card) {
  const int s = card >> 4, r = card & 15;
  const u64 rs = f.acc & suit_mask(s);
  const int cut = rs ? (lsb(rs) & 15) : 16;  // ranks below `cut` are small
  const int lt = (int)((f.low_tried >> (4 * s)) & 15);  // 1 + lowest tried rank
  return !DDX_NO_SMALL && r < cut && lt != 0 && lt - 1 < cut;
}

// The next move of f to try, skipping small cards (below every relevant card of their suit)
// once one small card of that suit was tried; -1 when none is left.
DDX_FN int next_move(Frame& f) {
  while (f.i < f.n) {
    const int card = move_at(f, f.i);
    if (small_skip(f, card)) {
      f.i++;
      continue;
    }
    const int s = card >> 4, r = card & 15;
    const int lt = (int)((f.low_tried >> (4 * s)) & 15);
    if (lt == 0 || r + 1 < lt) f.low_tried = (f.low_tried & ~(15u << (4 * s))) | ((u32)(r + 1) << (4 * s));
    return card;
  }
  return -1;
}

}  // namespace ddx
