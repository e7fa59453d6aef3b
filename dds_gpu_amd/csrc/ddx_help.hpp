// Helpers for the batch tail: splitting a long null-window search into parallel subsearches.
//
// A batch of full deals is heavy-tailed: most leads cost 1e5..1e6 nodes, a few cost 1e8 and
// more. Once the job queue is empty the lanes that hold those few leads decide the finish
// time while every other lane idles. With helpers, a lane whose current null-window search
// ("can N/S reach t tricks?") has run for a while publishes, when lanes are idle, one
// subsearch per move of that search's root: the position after the move, same target. Idle
// lanes take them and may split them again. The root's answer is the OR (N/S to move) or AND
// (E/W to move) of its children's answers, so the first child that says yes (OR) or no (AND)
// decides it. Results are kept as trick values clamped to the window [t-1, t] -- yes = t,
// no = t-1 -- so the OR/AND is a max/min, and a record can also carry a wider window (a whole
// lead, [0, tricks left]).
//
// The lane that published keeps searching. A holder record collects its children; whichever
// finishes first -- the lane or the children -- claims the holder (a first-writer-wins flag),
// and the slower side drops its work: the lane polls its holder and, when the children have
// answered, abandons its own search and continues its binary search with their answer;
// children poll their parent and their lead. With jobs of five leads, the job's leads that
// have not started are published as whole-lead subproblems too. Every answer is exact: each
// null-window answer is exact whoever computes it, and a lead's answer is written once, by the
// first claimant of its flag.
//
// Records live in one pool that is also the queue: a lane reserves k consecutive records
// (one atomic add), writes them, and then marks each ready (or skip: the pool's overflow, and
// holders, which are never solved by a helper). Takers claim the next index with a
// compare-and-swap and never wait inside a round for a record being written -- the writer may
// be a lane of the same wave -- they retry on their next round.
//
// Everything here runs one lane per search (the per-lane kernel).
#pragma once
// The machine's helper hooks must be compiled in: include this header before any other ddx
// header in the translation unit.
#if defined(DDX_HELP) && !DDX_HELP
#error "ddx_help.hpp needs DDX_HELP=1: include it before the other ddx headers"
#endif
#undef DDX_HELP
#define DDX_HELP 1
#include "ddx/machine.hpp"

namespace ddx {

enum { REC_LEAD = 0, REC_CHILD = 1, REC_HOLDER = 2 };
enum { RS_WRITING = 0, RS_READY = 1, RS_SKIP = 2 };

struct alignas(64) HelpRec {
  u64 hand[4];
  u32 trick;
  int8_t trump, leader, nplayed, ns_won, left, lo, hi, type;
  int parent;  // the holder this child's value feeds (REC_CHILD), else -1
  int lead;    // answer index 20 * deal + 4 * strain + declarer
  int acc;     // holder: max / min of its children's values so far, the value once claimed
  int pending;  // holder: children not yet finished
  u32 state;   // RS_*
};

struct HelpCtx {
  HelpRec* recs;
  u32* done;                 // per record: claimed (value known)
  unsigned long long* ctr;   // [0] records reserved, [1] records taken, [2] lanes idle
  long long cap;             // records in the pool
  long long total_leads;     // 20 * deals
  u64 spawn_nodes;           // a unit must have run this many nodes before it splits
};

// Per-lane helper bookkeeping.
struct HelpLane {
  long long claim;  // a claimed record not yet ready, or -1
  int holder;       // holder of the current search's children, or -1
  u32 serial;       // the lane's search serial when last checked
  u64 nodes0;       // lane nodes at that check (start of the current search, roughly)
  bool spawned;     // the current search was split
  bool idle;        // counted in ctr[2]
  bool main_done;   // the batch's job queue is empty
};

DDX_FN void help_lane_init(HelpLane& S) {
  S.claim = -1;
  S.holder = -1;
  S.serial = ~0u;
  S.nodes0 = 0;
  S.spawned = false;
  S.idle = false;
  S.main_done = false;
}

DDX_FN unsigned long long atomic_ld(const unsigned long long* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return *p;
#endif
}
DDX_FN unsigned long long atomic_add(unsigned long long* p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  const unsigned long long o = *p;
  *p += v;
  return o;
#endif
}
DDX_FN bool atomic_cas(unsigned long long* p, unsigned long long expect, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicCAS(p, expect, v) == expect;
#else
  if (*p != expect) return false;
  *p = v;
  return true;
#endif
}
DDX_FN void fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  __threadfence();
#endif
}
DDX_FN void set_state(u32* p, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
#else
  *p = v;
#endif
}
DDX_FN u32 get_state(const u32* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
#else
  return *p;
#endif
}
// Folds v into a holder's running max / min; true when the holder's value is now known: v
// reaches the mover's best in the window (decisive), or it was the last child.
DDX_FN bool fold_child(HelpRec& P, bool take_max, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (take_max) atomicMax(&P.acc, v);
  else atomicMin(&P.acc, v);
  __threadfence();
  if (v == (take_max ? P.hi : P.lo)) return true;
  const bool last = atomicSub(&P.pending, 1) == 1;
  __threadfence();
  return last;
#else
  P.acc = take_max ? (v > P.acc ? v : P.acc) : (v < P.acc ? v : P.acc);
  if (v == (take_max ? P.hi : P.lo)) return true;
  return --P.pending == 0;
#endif
}
DDX_FN int load_acc(const HelpRec& P) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(&P.acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return P.acc;
#endif
}

// N/S is to move at the position: its value is the max over the moves, else the min.
DDX_FN bool ns_to_move(int leader, int nplayed) { return ((leader + nplayed) & 1) == 0; }

DDX_FN void help_complete_lead(const Batch& B, long long lead, int ns) {
  if (!claim_flag(B.lead_done + lead)) return;
  B.out[lead] = (uint8_t)((lead & 1) == 0 ? ns : 13 - ns);  // declarer = lead % 4
  count_one(B.leads_done);
}

// Record `x` has value v: claim it; a lead writes its answer, a child feeds its holder, and a
// holder whose value is now known is claimed (its owner lane polls it). An answer is written
// once, by the first claimant.
DDX_FN void help_complete(const HelpCtx& H, const Batch& B, int x, int v) {
  if (!claim_flag(H.done + x)) return;
  HelpRec& R = H.recs[x];
  if (R.type == REC_LEAD) {
    help_complete_lead(B, R.lead, v);
    return;
  }
  if (R.type == REC_CHILD) {
    HelpRec& P = H.recs[R.parent];
    if (fold_child(P, ns_to_move(P.leader, P.nplayed), v)) claim_flag(H.done + R.parent);
  }
}

DDX_FN void rec_set_pos(HelpRec& R, const Pos& p) {
#pragma unroll
  for (int i = 0; i < 4; i++) R.hand[i] = p.hand[i];
  R.trick = p.trick;
  R.trump = (int8_t)p.trump;
  R.leader = (int8_t)p.leader;
  R.nplayed = (int8_t)p.nplayed;
  R.ns_won = (int8_t)p.ns_won;
  R.left = (int8_t)p.left;
}

// The root position of lead (deal, strain, declarer) of the batch.
DDX_FN Pos lead_root(const Batch& B, long long deal, int trump, int decl) {
  Pos p;
#pragma unroll
  for (int i = 0; i < 4; i++) p.hand[i] = B.hands[4 * deal + i];
  p.trick = 0;
  p.trump = trump;
  p.leader = (decl + 1) & 3;
  p.nplayed = 0;
  p.ns_won = 0;
  p.left = popc(p.hand[0]);
  return p;
}

// Publishes the root moves of the lane's current null-window search (target t) as subsearches
// with window [t-1, t], under a holder record; for a five-lead job also the leads that have
// not started. Needs depth >= 1 (the root expanded). Returns the holder, or -1.
template <class Stack>
DDX_FN int help_spawn(const HelpCtx& H, const Batch& B, const Lane& L, Stack& st) {
  const Frame f0 = L.depth == 1 ? L.top : st.load(0);
  const long long lead = 20 * L.deal + 4 * L.strain + L.decl;
  const int extra_leads = L.rec < 0 ? L.last_decl - L.decl : 0;
  Pos root;
  if (L.rec < 0) {
    root = lead_root(B, L.deal, L.p.trump, L.decl);
  } else {
    const HelpRec& R = H.recs[L.rec];
#pragma unroll
    for (int i = 0; i < 4; i++) root.hand[i] = R.hand[i];
    root.trick = R.trick;
    root.trump = R.trump;
    root.leader = R.leader;
    root.nplayed = R.nplayed;
    root.ns_won = R.ns_won;
    root.left = R.left;
  }
  const long long k = 1 + (long long)f0.n + extra_leads;
  const long long base = (long long)atomic_add(H.ctr, (unsigned long long)k);
  if (base + k > H.cap) {  // pool full: release what was reserved
    for (long long i = base; i < base + k && i < H.cap; i++) set_state(&H.recs[i].state, RS_SKIP);
    return -1;
  }
  const int t = L.target;
  const int holder = (int)base;
  HelpRec& V = H.recs[holder];
  rec_set_pos(V, root);
  V.lo = (int8_t)(t - 1);
  V.hi = (int8_t)t;
  V.type = REC_HOLDER;
  V.parent = -1;
  V.lead = (int)lead;
  V.acc = ns_to_move(root.leader, root.nplayed) ? t - 1 : t;
  V.pending = f0.n;
  long long at = base + 1;
  for (int i = 0; i < f0.n; i++) {
    Pos c = root;
    Frame tmp;
    tmp.seat = (root.leader + root.nplayed) & 3;
    play(c, tmp, move_at(f0, i));
    HelpRec& C = H.recs[at++];
    rec_set_pos(C, c);
    C.lo = (int8_t)(t - 1);
    C.hi = (int8_t)t;
    C.type = REC_CHILD;
    C.parent = holder;
    C.lead = (int)lead;
    C.acc = 0;
    C.pending = 0;
  }
  const long long first_lead = at;
  for (int d = L.decl + 1; d <= L.decl + extra_leads; d++) {
    HelpRec& R = H.recs[at++];
    const Pos p = lead_root(B, L.deal, L.p.trump, d);
    rec_set_pos(R, p);
    R.lo = 0;
    R.hi = (int8_t)p.left;
    R.type = REC_LEAD;
    R.parent = -1;
    R.lead = (int)(20 * L.deal + 4 * L.strain + d);
    R.acc = 0;
    R.pending = 0;
  }
  fence();  // the holder's fields before any child can be taken and finish
  set_state(&V.state, RS_SKIP);
  for (long long i = base + 1; i < first_lead; i++) set_state(&H.recs[i].state, RS_READY);
  for (long long i = first_lead; i < at; i++)
    set_state(&H.recs[i].state, flag_set(B.lead_done + H.recs[i].lead) ? RS_SKIP : RS_READY);
  return holder;
}

// Takes the next ready subproblem without waiting; -1 if none this round.
DDX_FN int help_take(const HelpCtx& H, HelpLane& S) {
  if (S.claim < 0) {
    const long long tail = (long long)atomic_ld(H.ctr + 1);
    long long lim = (long long)atomic_ld(H.ctr);
    if (lim > H.cap) lim = H.cap;
    if (tail >= lim) return -1;
    if (!atomic_cas(H.ctr + 1, (unsigned long long)tail, (unsigned long long)tail + 1)) return -1;
    S.claim = tail;
  }
  const u32 s = get_state(&H.recs[S.claim].state);
  if (s == RS_WRITING) return -1;  // its writer finishes within its own round
  const int r = (int)S.claim;
  S.claim = -1;
  return s == RS_READY ? r : -1;
}

// Starts solving record r: its position, its window, plain bisection.
template <class X = SeqOps>
DDX_FN void help_load(Lane& L, const HelpCtx& H, int r) {
  const HelpRec& R = H.recs[r];
#pragma unroll
  for (int i = 0; i < 4; i++) L.p.hand[i] = R.hand[i];
  L.p.trick = R.trick;
  L.p.trump = R.trump;
  L.p.leader = R.leader;
  L.p.nplayed = R.nplayed;
  L.p.ns_won = R.ns_won;
  L.p.left = R.left;
  L.rec = r;
  L.deal = R.lead / 20;
  L.strain = (R.lead % 20) / 4;
  L.decl = R.lead % 4;
  L.last_decl = L.decl;
  L.lo = R.lo;
  L.hi = R.hi;
  L.searches = 2;  // bisection from the start
  L.target = (L.lo + L.hi + 1) / 2;
  L.depth = 0;
  L.serial++;
  next_generation<X>(L.c);
  L.stage = L.lo < L.hi ? ST_ENTER : ST_UNIT;
}

// Closes the lane's holder (its search ended or was dropped): children still running see
// the claim and stop.
DDX_FN void help_close(const HelpCtx& H, HelpLane& S) {
  if (S.holder >= 0) claim_flag(H.done + S.holder);
  S.holder = -1;
}

// Cheap stages of one lane back to back (the wavefront kernel's cheap chain), with subproblem
// work once the batch's queue is empty. Returns with the lane at ENTER, EXIT, or at JOB when
// nothing could be taken this round (the lane is then idle).
template <class X = SeqOps, class Stack, class Take>
DDX_FN void help_cheap(Lane& L, const Batch& B, Stack& st, const HelpCtx& H, HelpLane& S, Take take) {
  while (L.stage != ST_ENTER && L.stage != ST_EXIT) {
    if (L.stage == ST_JOB) {
      if (!S.main_done) {
        const long long job = take();
        if (job < B.jobs) {
          L.rec = -1;
          lane_load_job<X>(L, B, job);
          continue;
        }
        S.main_done = true;
      }
      const int r = help_take(H, S);
      if (r >= 0) {
        if (S.idle) {
          atomic_add(H.ctr + 2, ~0ull);
          S.idle = false;
        }
        help_load<X>(L, H, r);
        continue;
      }
      if ((long long)atomic_ld(B.leads_done) >= H.total_leads) {
        if (S.idle) atomic_add(H.ctr + 2, ~0ull);
        S.idle = false;
        L.stage = ST_EXIT;
        return;
      }
      if (!S.idle) {
        atomic_add(H.ctr + 2, 1ull);
        S.idle = true;
      }
      return;
    } else if (L.stage == ST_ROOT) {
      lane_root(L, B);
    } else if (L.stage == ST_NEXT) {
      lane_next<X>(L, B, st);
    } else if (L.stage == ST_RET) {
      lane_ret<X>(L, B, st);
    } else {  // ST_UNIT
      help_close(H, S);
      help_complete(H, B, L.rec, L.lo);
      L.rec = -1;
      L.stage = ST_JOB;
    }
  }
}

// Restores the root position of the lane's unit after its search was dropped midway.
DDX_FN void help_restore_root(Lane& L, const Batch& B, const HelpCtx& H) {
  if (L.rec < 0) {
    const Pos p = lead_root(B, L.deal, L.p.trump, L.decl);
    L.p = p;
  } else {
    const HelpRec& R = H.recs[L.rec];
#pragma unroll
    for (int i = 0; i < 4; i++) L.p.hand[i] = R.hand[i];
    L.p.trick = R.trick;
    L.p.leader = R.leader;
    L.p.nplayed = R.nplayed;
    L.p.ns_won = R.ns_won;
    L.p.left = R.left;
  }
  L.depth = 0;
}

// Periodic checks of a lane in a search: drop the unit if someone else has answered it; take
// the answer of the current search from its helpers; or split the search when lanes are idle
// and it has run for a while. `queue_next` is the batch queue's counter.
template <class Stack>
DDX_FN void help_checks(Lane& L, const Batch& B, Stack& st, const HelpCtx& H, HelpLane& S,
                        const unsigned long long* queue_next) {
  if (S.serial != L.serial) {  // a new search since the last check
    help_close(H, S);
    S.serial = L.serial;
    S.nodes0 = L.c.nodes;
    S.spawned = false;
  }
  if (L.stage != ST_ENTER && L.stage != ST_NEXT && L.stage != ST_RET) return;
  bool gone;
  if (L.rec < 0) {
    gone = flag_set(B.lead_done + 20 * L.deal + 4 * L.strain + L.decl);
  } else {
    const HelpRec& R = H.recs[L.rec];
    gone = flag_set(H.done + L.rec) || (R.parent >= 0 && flag_set(H.done + R.parent)) ||
           flag_set(B.lead_done + R.lead);
  }
  if (gone) {
    help_close(H, S);
    L.depth = 0;
    if (L.rec < 0) {  // restore the deal's hands, go to the next lead
#pragma unroll
      for (int i = 0; i < 4; i++) L.p.hand[i] = B.hands[4 * L.deal + i];
      lane_next_lead(L);
    } else {
      L.rec = -1;
      L.stage = ST_JOB;
    }
    return;
  }
  if (S.holder >= 0 && flag_set(H.done + S.holder)) {  // the helpers answered this search
    fence();
    const int v = load_acc(H.recs[S.holder]);
    S.holder = -1;
    help_restore_root(L, B, H);
    L.v = v >= L.target ? 1 : 0;
    lane_search_done(L, B);  // the next target, or the unit's answer
    return;
  }
  if (S.spawned || L.depth < 1 || L.c.nodes - S.nodes0 < H.spawn_nodes) return;
  if ((long long)atomic_ld(queue_next) < B.jobs) return;  // the batch queue still has jobs
  const long long backlog = (long long)atomic_ld(H.ctr) - (long long)atomic_ld(H.ctr + 1);
  if ((long long)atomic_ld(H.ctr + 2) <= backlog) return;  // idle lanes have work waiting
  S.spawned = true;
  S.holder = help_spawn(H, B, L, st);
}

}  // namespace ddx
