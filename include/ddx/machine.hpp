// ddx core, part 5: the resumable solver machine shared by the CPU and the GPU.
//
// A lane runs jobs; a job is one (deal, strain) -- its four leads share one TT generation --
// or one (deal, strain, declarer) when jobs_per_deal is 20 (reference src/dd_wave.h:17-147).
// Each lead's NS trick count comes from a sequence of null-window searches; how the next
// target is chosen is the seed policy:
//   PLAIN  binary search over [0, 13], first target 7 then the previous lead's answer
//          (reference dd_ns_tricks / dd_solve_strain, src/dd.h:877-891, 926-939)
//   STEP   from a given seed, step up or down one trick at a time
//          (reference dd_ns_tricks_from / dd_solve_strain_guess, src/dd.h:898-924)
//   CERT   probe the seed g, then step one trick at a time in the direction the answers
//          point, at most twice more; a right seed costs two searches (the certificate "at
//          least g, not g+1"), a seed off by one or two three; after three searches, binary
//          search on what is left. Seeds come from the caller (e.g. the trick-count net).
//   CERTP  CERT seeded with the previous lead's answer (7 for the first); the no-net
//          ablation of CERT.
// Every policy keeps the invariant answer in [lo, hi] and stops at lo == hi, so all of them
// return the exact trick count; seeds only change how many searches it takes.
//
// Stages: JOB (take a job), ROOT (start a lead), ENTER (enter a node), NEXT (play the next
// move of the top frame), RET (a child returned), EXIT; with splitting (split.hpp) also WAIT
// (the top frame's untried moves were published as tasks; wait for their join) and HELP (no
// job is left: take published tasks).
#pragma once
#include "search.hpp"
#include "split.hpp"

#if DDX_VERIFY
#include <stdio.h>
#endif

#ifndef DDX_PER_LEADER_TT
#define DDX_PER_LEADER_TT 0
#endif

namespace ddx {

enum { ST_JOB, ST_ROOT, ST_ENTER, ST_NEXT, ST_RET, ST_EXIT, ST_WAIT, ST_HELP, ST_COUNT };

enum { SEED_PLAIN = 0, SEED_STEP = 1, SEED_CERT = 2, SEED_CERTP = 3 };

// The batch a lane works on. Deal d's hands are hands[4d..4d+3]; its results go to
// out[20d + 4*strain + declarer] as the declarer's tricks (NS count for N/S declarers,
// 13 - NS count for E/W, also for endings, as the reference does: src/dd.h:922,937).
struct Batch {
  const u64* hands;
  uint8_t* out;
  const uint8_t* seeds;  // NS-trick seeds in the same layout as out, or null
  long long jobs;        // deals * jobs_per_deal
  int jobs_per_deal;     // 5 (per strain) or 20 (per lead)
  int seed_mode;
  // Splitting of long searches once the job queue is empty (split.hpp); off while
  // sp.tasks is null, and then every split path folds away at compile time in kernels and
  // host code that leave it so.
  SplitView sp = {};
};

struct Lane {
  Ctx c;
  Pos p;
  Frame top;  // frame of the deepest expanded node, when depth > 0
  int depth;  // expanded nodes on the current path; frames 0..depth-2 are on the stack
  int stage;
  int target, lo, hi, searches;
  int v;  // result of the node that just finished
  u64 r;  // its relevant cards
  long long deal;
  int strain, decl, last_decl, prev_ns, last_true;
  int exit_at;  // where the last enter() decided (profiling)
  u32 jobs_done;
  u32 serial;  // null-window searches started (identifies the current search)
  // Splitting (split.hpp):
  long long job;     // batch job of the current work (a task inherits its owner's)
  long long tt_job;  // job whose positions the TT's current generation holds
  long long ticket;  // task slot this idle lane waits for, or -1
  int task;          // task being solved, or -1 for a job of the batch
  int task_join;     // that task's join
  int wait_join;     // join the top frame waits for (stage WAIT)
  int may_split;     // lanes were waiting for work at the last look (refreshed by the caller)
  u32 splits;        // nodes this lane split
};

DDX_FN void lane_init(Lane& L, TTView tt) {
  L.c.tt = tt;
  L.c.gen = 0;
  L.c.nodes = L.c.probes = L.c.hits = L.c.ent_reads = L.c.stores = 0;
#if DDX_VERIFY
  L.c.no_tt = 0;
  L.c.verified = L.c.bad = 0;
#endif
  L.depth = 0;
  L.stage = ST_JOB;
  L.v = 0;
  L.r = 0;
  L.deal = -1;
  L.strain = L.decl = L.last_decl = 0;
  L.prev_ns = 7;
  L.last_true = 0;
  L.target = L.lo = L.hi = L.searches = 0;
  L.exit_at = 0;
  L.jobs_done = 0;
  L.serial = 0;
  L.job = L.tt_job = L.ticket = -1;
  L.task = L.task_join = L.wait_join = -1;
  L.may_split = 0;
  L.splits = 0;
}

DDX_FN void start_search(Lane& L) {
  L.p.leader = (L.decl + 1) & 3;
  L.p.nplayed = 0;
  L.p.ns_won = 0;
  L.p.left = popc(L.p.hand[0]);
  L.p.trick = 0;
  L.depth = 0;
  L.serial++;
  L.stage = ST_ENTER;
}

// Loads job `job` of the batch into the lane (reference dd_wave_step JOB, src/dd_wave.h:72-90).
template <class X = SeqOps>
DDX_FN void lane_load_job(Lane& L, const Batch& B, long long job) {
  const int jpd = B.jobs_per_deal;
  L.deal = job / jpd;
  const int sub = (int)(job % jpd);
  const u64* h = B.hands + 4 * L.deal;
#pragma unroll
  for (int i = 0; i < 4; i++) L.p.hand[i] = h[i];
  L.strain = jpd == 5 ? sub : sub / 4;
  L.decl = jpd == 5 ? 0 : sub % 4;
  L.last_decl = jpd == 5 ? 3 : L.decl;
  L.p.trump = strain_trump(L.strain);
  next_generation<X>(L.c);
  L.job = L.tt_job = job;
  L.prev_ns = 7;
  L.stage = ST_ROOT;
}

// ROOT: first target of this lead's search sequence.
DDX_FN void lane_root(Lane& L, const Batch& B) {
  const int left = popc(L.p.hand[0]);
  L.lo = 0;
  L.searches = 0;
  int g;
  if (B.seed_mode == SEED_PLAIN) {
    L.hi = 13;
    g = L.prev_ns;
    L.target = g < 1 ? 1 : g > 13 ? 13 : g;
  } else {
    L.hi = left;
    if (B.seed_mode == SEED_CERTP) g = L.prev_ns;
    else g = B.seeds ? (int)B.seeds[20 * L.deal + 4 * L.strain + L.decl] : 7;
    g = g < 0 ? 0 : g > left ? left : g;
    L.target = g == 0 ? 1 : g;  // "at least 0" holds without a search
  }
  start_search(L);
}

// The root search finished with L.v: narrow [lo, hi]; pick the next target, or record the
// answer and move to the next lead or job (reference dd_wave_node_done, src/dd_wave.h:42-65).
DDX_FN void lane_search_done(Lane& L, const Batch& B) {
  if (B.sp.tasks && L.task >= 0) {  // a task: its value goes to the owner's join
    const SplitTask& t = B.sp.tasks[L.task];
    split_report(B.sp, L.task_join, L.v, t.goal, L.r | t.extra, t.card);
    L.task = L.task_join = -1;
    L.stage = ST_HELP;
    return;
  }
  L.searches++;
  if (L.v) L.lo = L.target;
  else L.hi = L.target - 1;
  if (L.lo < L.hi) {
    const bool adjacent = B.seed_mode == SEED_STEP || (B.seed_mode >= SEED_CERT && L.searches < 3);
    if (adjacent) L.target = L.v ? L.lo + 1 : L.hi;
    else L.target = (L.lo + L.hi + 1) / 2;
    start_search(L);
    return;
  }
  const int ns = L.lo;
  const long long idx = 20 * L.deal + 4 * L.strain + L.decl;
  B.out[idx] = (uint8_t)((L.decl & 1) == 0 ? ns : 13 - ns);
  L.decl++;
  if (L.decl <= L.last_decl) {
    L.prev_ns = ns < 1 ? 1 : ns;
    // Measurement switch: the leads of a strain share no TT entries (reference
    // -DDD_PER_LEADER_TT in dd_solve_strain, src/dd.h:932-934).
    if (DDX_PER_LEADER_TT && B.seed_mode == SEED_PLAIN) next_generation(L.c);
    L.stage = ST_ROOT;
    return;
  }
  L.jobs_done++;
  L.stage = ST_JOB;
}

// ENTER: enter the node at the current position; `st` is the lane's frame stack
// (st.store(i, f) / st.load(i)).
template <class X = SeqOps, class Stack>
DDX_FN void lane_enter(Lane& L, const Batch& B, Stack& st) {
  Frame f;
  u64 rel = 0;
  const int e = enter<X>(L.c, L.p, L.target, f, rel, L.exit_at);
  if (e >= 0) {
    L.v = e;
    L.r = rel;
    if (L.depth == 0) lane_search_done(L, B);
    else L.stage = ST_RET;
    return;
  }
  if (L.depth > 0) st.store(L.depth - 1, L.top);
  L.top = f;
  L.depth++;
  L.stage = ST_NEXT;
}

// The top frame finished with (L.v, L.r).
template <class X = SeqOps, class Stack>
DDX_FN void lane_node_done(Lane& L, const Batch& B, Stack& st) {
  L.depth--;
  if (L.depth == 0) {
    lane_search_done(L, B);
    return;
  }
  L.top = st.load(L.depth - 1);
  L.stage = ST_RET;
}

// NEXT: play the next move of the top frame, or finish it.
template <class X = SeqOps, class Stack>
DDX_FN void lane_next(Lane& L, const Batch& B, Stack& st) {
  const int card = next_move(L.top);
  if (card >= 0) {
    play(L.p, L.top, card);
    L.stage = ST_ENTER;
    return;
  }
  L.v = L.top.result;
  L.r = finish<X>(L.c, L.p, L.top);
  lane_node_done<X>(L, B, st);
}

// Publishes the top frame's untried moves as tasks and sends the lane to WAIT (split.hpp),
// when lanes are waiting for work, the children searched so far took at least 16 * min_nodes16
// nodes (so the rest are likely big too), the node has at least min_left tricks to go and at
// least min_moves moves are left. Called at RET after a child that did not cut off, with L.p the
// frame's position. A move is left out when the small-card rule would skip it now: that test
// only uses children already searched (published moves do not count as tried, since their
// relevant cards are not known yet). Returns whether it split.
DDX_FN bool lane_split(Lane& L, const Batch& B) {
  const SplitView& s = B.sp;
  Frame& f = L.top;
  if (L.p.left < s.min_left || f.n - f.i < s.min_moves) return false;
  if ((((u32)(L.c.nodes >> 4) - f.nodes0) & 0xFFFFFFu) < s.min_nodes16) return false;
  int k = 0;
  for (int i = f.i; i < f.n; i++) k += small_skip(f, move_at(f, i)) ? 0 : 1;
  if (k < s.min_moves || !split_starving(s)) return false;
  const long long j = (long long)sp_add(s.ctr + 2, 1);
  if (j >= s.join_cap) return false;
  const long long base = (long long)sp_add(s.ctr + 0, (unsigned long long)k);
  if (base + k > s.task_cap) {
    // Out of slots: lanes may hold tickets for the reserved ones, so mark them skipped.
    for (long long q = base; q < s.task_cap && q < base + k; q++) sp_store_rel(&s.tasks[q].state, SLOT_SKIP);
    return false;
  }
  SplitJoin& J = s.joins[j];
  J.acc = 0;
  J.cut_rel = 0;
  J.cut_card = -1;
  J.pending = k;
  J.parent = L.task >= 0 ? L.task_join : -1;
  J.state = JOIN_OPEN;
  const int8_t goal = (int8_t)((f.seat & 1) == 0 ? 1 : 0);
  int q = 0;
  for (int i = f.i; i < f.n; i++) {
    const int card = move_at(f, i);
    if (small_skip(f, card)) continue;
    Pos c = L.p;
    Frame g = f;
    play(c, g, card);
    u64 extra = 0;
    Pos back = c;
    unplay(back, g, card, extra);
    SplitTask& t = s.tasks[base + q];
#pragma unroll
    for (int h = 0; h < 4; h++) t.hand[h] = c.hand[h];
    t.extra = extra;
    t.trick = c.trick;
    t.join = (int)j;
    t.job = (int)L.job;
    t.trump = (int8_t)c.trump;
    t.leader = (int8_t)c.leader;
    t.nplayed = (int8_t)c.nplayed;
    t.ns_won = (int8_t)c.ns_won;
    t.left = (int8_t)c.left;
    t.target = (int8_t)L.target;
    t.card = (int8_t)card;
    t.goal = goal;
    sp_store_rel(&t.state, SLOT_READY);  // publishes the task and, before it, the join
    q++;
  }
  f.i = f.n;
  L.wait_join = (int)j;
  L.stage = ST_WAIT;
  L.splits++;
  sp_add(s.ctr + 3, 1);
  return true;
}

// RET: the child of the top frame returned (L.v, L.r): undo its move; a result that reaches
// the mover's goal cuts off, else collect its relevant cards and try the next move
// (reference dd_search child_done, src/dd.h:840-854), or split the rest off (split.hpp).
template <class X = SeqOps, class Stack>
DDX_FN void lane_ret(Lane& L, const Batch& B, Stack& st) {
  Frame& f = L.top;
  const int card = move_at(f, f.i);
  unplay(L.p, f, card, L.r);
  if (L.v == ((f.seat & 1) == 0 ? 1 : 0)) {
    f.result = L.v;
    f.acc = L.r;
    f.best = card;
    L.r = finish<X>(L.c, L.p, f);
    lane_node_done<X>(L, B, st);
    return;
  }
  f.acc |= L.r;
  f.i++;
  L.stage = ST_NEXT;
  if (B.sp.tasks && L.may_split) lane_split(L, B);
}

// WAIT: the top frame's remaining moves are tasks. A child that cut off decides the frame
// (best move and relevant cards are that child's); when every task has reported without a
// cutoff, the frame continues as if it had tried them itself (their relevant cards added), and
// NEXT finds no move left.
template <class X = SeqOps, class Stack>
DDX_FN void lane_wait(Lane& L, const Batch& B, Stack& st) {
  SplitJoin& J = B.sp.joins[L.wait_join];
  const u32 state = sp_load_acq(&J.state);
  Frame& f = L.top;
  if (state == JOIN_CUT) {
    f.result = (f.seat & 1) == 0 ? 1 : 0;
    f.acc = J.cut_rel;
    f.best = J.cut_card;
    L.wait_join = -1;
    L.v = f.result;
    L.r = finish<X>(L.c, L.p, f);
    lane_node_done<X>(L, B, st);
    return;
  }
  if (state == JOIN_OPEN && sp_load_acq(&J.pending) == 0) {
    f.acc |= sp_load64(&J.acc);
    L.wait_join = -1;
    L.stage = ST_NEXT;
  }
}

// Drops the lane's task (its join chain was cut off): the stack is abandoned, and a join the
// lane waited for is left to its own tasks, which find the same cut above them.
DDX_FN void lane_drop_task(Lane& L) {
  L.task = L.task_join = L.wait_join = -1;
  L.depth = 0;
  L.stage = ST_HELP;
}

// Whether the lane's task lies under a join that was cut off.
DDX_FN bool lane_task_moot(const Lane& L, const Batch& B) {
  return B.sp.tasks && L.task >= 0 && split_chain_cut(B.sp, L.task_join);
}

// HELP: an idle lane with ticket L.ticket (handed out by the caller) takes the task in that
// slot once it is written; a skipped slot or a task already moot returns the ticket. Returns
// true when the lane took a task (stage ENTER).
template <class X = SeqOps>
DDX_FN bool lane_help(Lane& L, const Batch& B) {
  if (L.ticket < 0 || L.ticket >= B.sp.task_cap) return false;  // beyond the pool: stays idle
  SplitTask& t = B.sp.tasks[L.ticket];
  const u32 state = sp_load_acq(&t.state);
  if (state == SLOT_EMPTY) return false;
  const long long slot = L.ticket;
  L.ticket = -1;
  if (state == SLOT_SKIP) return false;
  if (split_chain_cut(B.sp, t.join)) {
    sp_add(B.sp.ctr + 5, 1);
    return false;
  }
#pragma unroll
  for (int h = 0; h < 4; h++) L.p.hand[h] = t.hand[h];
  L.p.trick = t.trick;
  L.p.trump = t.trump;
  L.p.leader = t.leader;
  L.p.nplayed = t.nplayed;
  L.p.ns_won = t.ns_won;
  L.p.left = t.left;
  L.target = t.target;
  L.task = (int)slot;
  L.task_join = t.join;
  L.job = t.job;
  if (L.job != L.tt_job) {  // positions of another job: a fresh TT generation
    next_generation<X>(L.c);
    L.tt_job = L.job;
  }
  L.depth = 0;
  L.serial++;
  L.stage = ST_ENTER;
  sp_add(B.sp.ctr + 4, 1);
  return true;
}

// One stage of the machine. `take` returns the next job index (>= jobs: none left).
template <class X = SeqOps, class Stack, class Take>
DDX_FN void lane_step(Lane& L, const Batch& B, Stack& st, Take take) {
  switch (L.stage) {
    case ST_JOB: {
      long long job = take();
      if (job >= B.jobs) {
        L.stage = B.sp.tasks ? ST_HELP : ST_EXIT;
        return;
      }
      lane_load_job<X>(L, B, job);
      return;
    }
    case ST_ROOT:
      lane_root(L, B);
      return;
    case ST_ENTER:
      lane_enter<X>(L, B, st);
      return;
    case ST_NEXT:
      lane_next<X>(L, B, st);
      return;
    case ST_RET:
      lane_ret<X>(L, B, st);
      return;
    default:
      return;
  }
}

// A frame stack in ordinary memory (CPU, tests).
struct ArrayStack {
  PackedFrame f[52];
  DDX_M void store(int i, const Frame& x) { f[i] = pack_frame(x); }
  DDX_M Frame load(int i) const { return unpack_frame(f[i]); }
};


#if DDX_VERIFY
// Re-checks a TT hit: the same position and target searched by the same machine with no TT
// (neither probes nor stores; quick tricks, later tricks and the move order as usual). One
// search decides it: [lo, hi] = [target - 1, target]. A contradiction is counted and printed
// in the reference's form (src/dd.h:651-663).
inline void verify_hit(Ctx& c, const Pos& p, int target, int hit, u64 rel) {
  Lane V;
  lane_init(V, c.tt);
  V.c.no_tt = 1;
  V.c.verified = V.c.bad = 0;
  V.c.gen = c.gen;
  V.p = p;
  V.target = target;
  V.lo = target - 1;
  V.hi = target;
  V.deal = 0;
  V.strain = V.decl = V.last_decl = 0;
  V.stage = ST_ENTER;
  uint8_t out[20];
  const Batch B{p.hand, out, nullptr, 1, 5, SEED_PLAIN};
  ArrayStack st;
  auto none = [&]() -> long long { return B.jobs; };
  while (V.stage != ST_JOB && V.stage != ST_EXIT) lane_step(V, B, st, none);
  const bool truth = V.lo >= target;
  c.verified++;
  if (truth != (hit != 0)) {
    c.bad++;
    printf("BAD TT HIT: leader %d need %d left %d ns_won %d hit %d truth %d rel %llx\n", p.leader,
           target - p.ns_won, p.left, p.ns_won, hit, truth ? 1 : 0, (unsigned long long)rel);
    for (int h = 0; h < 4; h++) printf("  hand %d %llx\n", h, (unsigned long long)p.hand[h]);
  }
}
#endif

}  // namespace ddx
