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
// move of the top frame), RET (a child returned), EXIT, and UNIT (a helper subproblem --
// a position with a trick window, see dds_gpu_amd/csrc/ddx_help.hpp -- has its value in lo).
#pragma once
#include "search.hpp"

#if DDX_VERIFY
#include <stdio.h>
#endif

// DDX_HELP = 1 compiles the hooks for tail helpers (dds_gpu_amd/csrc/ddx_help.hpp, which sets
// it); without it they fold away and the machine is exactly the plain one.
#ifndef DDX_HELP
#define DDX_HELP 0
#endif

namespace ddx {

enum { ST_JOB, ST_ROOT, ST_ENTER, ST_NEXT, ST_RET, ST_EXIT, ST_UNIT, ST_COUNT };

// First-writer-wins flag and counter for answers that helpers may also produce (device
// atomics; a host build is single-threaded per flag).
DDX_FN bool claim_flag(u32* f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicCAS(f, 0u, 1u) == 0u;
#else
  if (*f) return false;
  *f = 1;
  return true;
#endif
}
DDX_FN void count_one(unsigned long long* c) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd(c, 1ull);
#else
  ++*c;
#endif
}
DDX_FN bool flag_set(const u32* f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
#else
  return *f != 0;
#endif
}
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
#if DDX_HELP
  // With helpers: one flag per answer (same layout as out); the first to finish a lead -- its
  // own lane or the helpers -- claims it, writes out and counts it in *leads_done.
  u32* lead_done = nullptr;
  unsigned long long* leads_done = nullptr;
#endif
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
  int rec;     // helper subproblem being solved, or -1 for the batch's own leads
  u32 serial;  // null-window searches started (identifies the current search)
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
  L.rec = -1;
  L.serial = 0;
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
  L.prev_ns = 7;
  L.stage = ST_ROOT;
}

// Moves on to the job's next lead, or to the next job.
DDX_FN void lane_next_lead(Lane& L) {
  L.decl++;
  if (L.decl <= L.last_decl) {
    L.stage = ST_ROOT;
    return;
  }
  L.jobs_done++;
  L.stage = ST_JOB;
}

// ROOT: first target of this lead's search sequence (a lead that helpers already answered is
// skipped).
DDX_FN void lane_root(Lane& L, const Batch& B) {
#if DDX_HELP
  if (B.lead_done && flag_set(B.lead_done + 20 * L.deal + 4 * L.strain + L.decl)) {
    lane_next_lead(L);
    return;
  }
#endif
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
  L.searches++;
  if (L.v) L.lo = L.target;
  else L.hi = L.target - 1;
  if (L.lo < L.hi) {
    const bool sub = DDX_HELP && L.rec >= 0;
    const bool adjacent =
        !sub && (B.seed_mode == SEED_STEP || (B.seed_mode >= SEED_CERT && L.searches < 3));
    if (adjacent) L.target = L.v ? L.lo + 1 : L.hi;
    else L.target = (L.lo + L.hi + 1) / 2;
    if (sub) {  // a subproblem: the search unwound to its root position; search again
      L.depth = 0;
      L.serial++;
      L.stage = ST_ENTER;
    } else {
      start_search(L);
    }
    return;
  }
  if (DDX_HELP && L.rec >= 0) {
    L.stage = ST_UNIT;
    return;
  }
  const int ns = L.lo;
  const long long idx = 20 * L.deal + 4 * L.strain + L.decl;
#if DDX_HELP
  if (!B.lead_done || claim_flag(B.lead_done + idx)) {
    B.out[idx] = (uint8_t)((L.decl & 1) == 0 ? ns : 13 - ns);
    if (B.leads_done) count_one(B.leads_done);
  }
#else
  B.out[idx] = (uint8_t)((L.decl & 1) == 0 ? ns : 13 - ns);
#endif
  L.decl++;
  if (L.decl <= L.last_decl) {
    L.prev_ns = ns < 1 ? 1 : ns;
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

// RET: the child of the top frame returned (L.v, L.r): undo its move; a result that reaches
// the mover's goal cuts off, else collect its relevant cards and try the next move
// (reference dd_search child_done, src/dd.h:840-854).
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
}

// One stage of the machine. `take` returns the next job index (>= jobs: none left).
template <class X = SeqOps, class Stack, class Take>
DDX_FN void lane_step(Lane& L, const Batch& B, Stack& st, Take take) {
  switch (L.stage) {
    case ST_JOB: {
      long long job = take();
      if (job >= B.jobs) {
        L.stage = ST_EXIT;
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
