// Host simulation of the batch kernel with splitting (include/ddx/split.hpp): `lanes` solver
// lanes advance round-robin, one round = the lane's cheap stages up to its next node entry,
// then that entry (the batch kernel's schedule), so a job's length in rounds is about the
// number of nodes on its critical path. Lanes that find the job queue empty take tickets and
// help with published tasks exactly as on the GPU; each lane has its own TT and stack.
//
//   ddx_split_sim(hands, n, lanes, tt_log2, jobs_per_deal, min_left, min_nodes, task_cap, out, stats)
//   stats: [0] rounds until every job was done, [1] nodes, [2] splits, [3] tasks taken,
//          [4] tasks dropped (moot when taken), [5] task slots used, [6] joins used,
//          [7] rounds at which the job queue ran empty, [8] tasks abandoned while running
// min_left < 0 switches splitting off (the plain kernel).
#include <stdint.h>
#include <string.h>

#include <vector>

#include "ddx/machine.hpp"

using namespace ddx;

extern "C" int ddx_split_sim(const uint64_t* hands, int64_t n, int lanes, int tt_log2, int jpd,
                             int min_left, int min_nodes, int64_t task_cap, uint8_t* out, int64_t* stats) {
  const size_t buckets = (size_t)1 << tt_log2;
  std::vector<u32> tags((size_t)lanes * buckets * kWays + 16, 0);
  std::vector<Entry> ents((size_t)lanes * buckets * kWays);
  u32* tag0 = (u32*)(((uintptr_t)tags.data() + 63) & ~(uintptr_t)63);
  std::vector<Lane> L((size_t)lanes);
  std::vector<ArrayStack> st((size_t)lanes);
  for (int i = 0; i < lanes; i++) {
    TTView v;
    v.tags = tag0 + (size_t)i * buckets * kWays;
    v.ents = ents.data() + (size_t)i * buckets * kWays;
    v.mask = (u32)(buckets - 1);
    lane_init(L[i], v);
  }
  const bool split = min_left >= 0;
  std::vector<SplitTask> tasks(split ? (size_t)task_cap : 0);
  std::vector<SplitJoin> joins(split ? (size_t)task_cap : 0);
  unsigned long long ctr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  Batch B{hands, out, nullptr, (long long)n * jpd, jpd, 0};
  if (split) {
    B.sp.tasks = tasks.data();
    B.sp.joins = joins.data();
    B.sp.ctr = ctr;
    B.sp.task_cap = task_cap;
    B.sp.join_cap = task_cap;
    B.sp.min_left = min_left;
    B.sp.min_nodes16 = (u32)(min_nodes / 16);
  }
  memset(out, 0xFF, (size_t)n * 20);
  long long next = 0, done = 0, empty_at = -1, abandoned = 0;
  auto take = [&]() -> long long { return next < B.jobs ? next++ : B.jobs; };
  long long round = 0;
  for (; done < B.jobs; round++) {
    if (empty_at < 0 && next >= B.jobs) empty_at = round;
    for (int i = 0; i < lanes && done < B.jobs; i++) {
      Lane& l = L[i];
      if (l.stage == ST_EXIT) continue;
      const u32 jobs0 = l.jobs_done;
      if (split && (round & 63) == 0)
        l.may_split = split_starving(B.sp) && B.jobs - done <= B.sp.max_open_jobs;
      if (split && (round & 255) == 0 && lane_task_moot(l, B)) {
        lane_drop_task(l);
        abandoned++;
      }
      if (l.stage == ST_HELP) {
        if (l.ticket < 0) l.ticket = (long long)sp_add(ctr + 1, 1);
        if (!lane_help(l, B)) continue;
      }
      if (l.stage == ST_WAIT) {
        lane_wait(l, B, st[i]);
        if (l.stage == ST_WAIT) continue;
      }
      while (l.stage == ST_JOB || l.stage == ST_ROOT || l.stage == ST_NEXT || l.stage == ST_RET) {
        lane_step(l, B, st[i], take);
        if (l.stage == ST_WAIT || l.stage == ST_HELP) break;
      }
      if (l.stage == ST_ENTER) lane_enter(l, B, st[i]);
      done += l.jobs_done - jobs0;
    }
  }
  uint64_t nodes = 0;
  for (int i = 0; i < lanes; i++) nodes += L[i].c.nodes;
  stats[0] = round;
  stats[1] = (int64_t)nodes;
  stats[2] = (int64_t)ctr[3];
  stats[3] = (int64_t)ctr[4];
  stats[4] = (int64_t)ctr[5];
  stats[5] = (int64_t)(ctr[0] < (unsigned long long)task_cap ? ctr[0] : task_cap);
  stats[6] = (int64_t)ctr[2];
  stats[7] = empty_at;
  stats[8] = abandoned;
  return 0;
}
