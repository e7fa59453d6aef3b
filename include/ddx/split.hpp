// ddx core, part 6: splitting long searches over idle lanes (the batch's tail).
//
// A batch of full deals is heavy-tailed: most (deal, strain) jobs cost 1e5..1e6 nodes, a few
// cost 1e8 and more, nearly all of it in one or two null-window searches ("can N/S take t
// tricks?") of a lead. Once the job queue is empty, the lanes that hold those searches
// decide when the batch ends while every other lane idles. Splitting lets the idle lanes
// share such a search:
//
//   * Young-brothers-wait rule. A node's children are searched one by one as usual; when a
//     child returns without deciding the node (no cutoff) and lanes are idle, the lane that
//     owns the node publishes every child it has not tried yet as a task and waits. A node
//     whose first child returns without a cutoff usually needs all its children (an "all"
//     node of the proof), so the tasks are rarely wasted; a node whose first child cuts off
//     is never split.
//   * A task is a whole null-window search from the position after its move, with the
//     owner's target. An idle lane takes it, searches it with its own TT and stack (and may
//     split it again), and reports (value, relevant cards) to the task's join record.
//   * Join. The first task whose value is the mover's goal cuts the node off: it claims the
//     join, and the owner resumes with that child as the best move and its relevant cards.
//     Otherwise every task reports "no cutoff" with its relevant cards OR-ed into the join;
//     when the last one has reported, the owner resumes with the union, as if it had tried
//     the moves itself. A cut-off join makes the tasks still queued or running below it
//     pointless: lanes find that by walking the chain of joins above their task and drop
//     their work.
//
// Every answer stays exact: each task's value is the exact value of its position, and the
// owner combines them with the same OR/AND the sequential search applies (the relevant-card
// sets may differ from the sequential search's, both are sound). Node counts differ from the
// sequential search once a node is split; splitting starts only when the job queue is empty,
// so jobs before the tail keep the reference's node-for-node order.
//
// Pool layout (device memory, one per batch launch): tasks[] and joins[] are linear arrays
// with counters ctr[0] = task slots reserved by publishers, ctr[1] = tickets handed to idle
// lanes (an idle lane takes ticket k and waits until slot k is written), ctr[2] = joins used.
// ctr[1] > ctr[0] means lanes are waiting for work: that is the signal to split. A publisher
// reserves its slots with one atomic add and marks each slot ready with a release store, so a
// lane never waits inside a round for a slot written by another lane of its own wave.
#pragma once
#include "bits.hpp"

namespace ddx {

enum { SLOT_EMPTY = 0, SLOT_READY = 1, SLOT_SKIP = 2 };
enum { JOIN_OPEN = 0, JOIN_CLAIMED = 1, JOIN_CUT = 2 };

// One published child search (64 bytes).
struct alignas(64) SplitTask {
  u64 hand[4];  // position after the child's move
  u64 extra;    // relevant cards the owner's undo of the move adds (a trick won by rank)
  u32 trick;
  int join;     // join record the result goes to
  int job;      // batch job the position belongs to (the lanes' TT generations follow jobs)
  int8_t trump, leader, nplayed, ns_won, left, target, card, goal;  // goal: value that cuts off
  u32 state;    // SLOT_*
};

// The owner's side of a split node (32 bytes).
struct alignas(32) SplitJoin {
  u64 acc;       // union of the relevant cards of children that did not cut off
  u64 cut_rel;   // relevant cards of the child that cut off (with the undo's addition)
  int parent;    // join of the task the owner was solving, -1 for a job of the batch
  int pending;   // children that have not reported "no cutoff"
  u32 state;     // JOIN_*
  int cut_card;  // move of the child that cut off
};

// The pool as the lanes see it; tasks == nullptr switches splitting off.
struct SplitView {
  SplitTask* tasks = nullptr;
  SplitJoin* joins = nullptr;
  unsigned long long* ctr = nullptr;  // [0] slots reserved, [1] tickets, [2] joins used,
                                      // [3] splits, [4] tasks taken, [5] tasks dropped
  long long task_cap = 0, join_cap = 0;
  u32 min_nodes16 = 1024;  // split only nodes whose searched children took >= 16 * this many nodes
  int min_left = 3;        // ... with at least this many tricks left
  int min_moves = 2;       // ... and at least this many untried moves
  long long max_open_jobs = 1ll << 62;  // splits start when at most this many jobs are unfinished
};

// Memory-order helpers: agent-scope atomics on the GPU; the host simulation runs its lanes on
// one thread, where plain accesses suffice but the GCC atomics keep the same shape.
DDX_FN unsigned long long sp_add(unsigned long long* p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}
DDX_FN unsigned long long sp_load(const unsigned long long* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#endif
}
DDX_FN u32 sp_load_acq(const u32* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_load_n(p, __ATOMIC_ACQUIRE);
#endif
}
DDX_FN int sp_load_acq(const int* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_load_n(p, __ATOMIC_ACQUIRE);
#endif
}
DDX_FN void sp_store_rel(u32* p, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
#else
  __atomic_store_n(p, v, __ATOMIC_RELEASE);
#endif
}
DDX_FN bool sp_cas(u32* p, u32 expect, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_compare_exchange_n(p, &expect, v, false, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED);
#endif
}
DDX_FN void sp_or(u64* p, u64 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_fetch_or((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
#else
  __atomic_fetch_or(p, v, __ATOMIC_RELAXED);
#endif
}
// Decrement with release order: the caller's earlier writes (the OR of its relevant cards) are
// visible to whoever sees the new count.
DDX_FN void sp_dec_rel(int* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_fetch_add(p, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
#else
  __atomic_fetch_add(p, -1, __ATOMIC_RELEASE);
#endif
}
DDX_FN u64 sp_load64(const u64* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (u64)__hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#endif
}

// Lanes are waiting for tasks that nobody has published yet.
DDX_FN bool split_starving(const SplitView& s) { return sp_load(s.ctr + 1) > sp_load(s.ctr + 0); }

// Whether the join chain above `join` has been cut off (the work below it is moot).
DDX_FN bool split_chain_cut(const SplitView& s, int join) {
  for (int guard = 0; join >= 0 && guard < 64; guard++) {
    if (sp_load_acq(&s.joins[join].state) != JOIN_OPEN) return true;
    join = s.joins[join].parent;
  }
  return false;
}

// A child reports: `v` its value, `rel` its relevant cards seen from the owner's node.
DDX_FN void split_report(const SplitView& s, int join, int v, int goal, u64 rel, int card) {
  SplitJoin& J = s.joins[join];
  if (v == goal) {
    if (sp_cas(&J.state, JOIN_OPEN, JOIN_CLAIMED)) {
      J.cut_rel = rel;
      J.cut_card = card;
      sp_store_rel(&J.state, JOIN_CUT);
    }
    return;
  }
  sp_or(&J.acc, rel);
  sp_dec_rel(&J.pending);
}

}  // namespace ddx
