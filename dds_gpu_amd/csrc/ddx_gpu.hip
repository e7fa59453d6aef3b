// GPU side of the ddx C ABI: the wavefront solver kernel for gfx950 and its host driver.
//
// One thread = one solver lane with a private TT slice and frame stack in HBM; lanes take
// (deal, strain) jobs from one atomic counter until none is left (persistent grid). The
// machine is the one the CPU runs (include/ddx/machine.hpp), so results are bit-identical.
//
// Per-lane device memory (L lanes, B = 2^tt_log2 buckets):
//   tags   L * B * 64 B    tag lines, lane-major (lane l's TT is one contiguous range)
//   ents   L * B * 256 B   entries, lane-major
//   stack  L * 52 * 32 B   packed frames, lane-major: a lane's neighbouring depths share
//                          128-byte lines, so push/pop pairs hit the same line
//   gen    L * 4 B         TT generation of each lane, kept across launches so tags of a
//                          previous call never match (no per-call clear of the tag lines)
//
// Teams: with team = 16, sixteen lanes run one search together (ddx_team.hpp): the TT tag
// line is probed and rewritten in one coalesced 64-byte access and the move weights are
// computed one candidate per lane, so a search advances about five times faster (measured
// under full load) while four searches share a wave. Storage is then per search, not per
// lane.
//
// Wavefront schedules (one wave64 = 64 lanes in lockstep; stages diverge per lane):
//   PRIORITY  each round runs one stage for the lanes in it: RET > NEXT > ROOT > JOB > ENTER
//             (the reference's order, gpu/gpu_main.cu:118-132)
//   CHAIN     each lane runs its cheap stages (JOB/ROOT/NEXT/RET) back to back until it
//             reaches ENTER; then the wave enters nodes for all lanes together
//   SPLIT     as CHAIN, but the chain also enters follow nodes (2nd-4th hand, cheap), so
//             trick-start nodes (TT probe, quick/later tricks) gather and run with the whole
//             wave active and their HBM probes issued together
// CHAIN and SPLIT run a lane's stages in passes of RET, NEXT, JOB, ROOT (and, for SPLIT, a
// follow node's ENTER), the order in which a lane usually meets them.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <new>
#include <type_traits>

#include "ddx_lanes.hpp"
#include "ddx_lds.hpp"
#include "ddx_team.hpp"

using namespace ddx;

struct KernelArgs {
  const u64* hands;
  uint8_t* out;
  const uint8_t* seeds;
  long long jobs;
  int jobs_per_deal;
  int seed_mode;
  LaneView lanes;
  unsigned long long* counters;  // [0] next job, [1] nodes, [2] probes, [3] hits, [4] jobs done,
                                 // [5] finished jobs so far, [8] nodes of the finished jobs so far
                                 // [5] progress (jobs done, updated while running),
                                 // [6] TT entries read, [7] TT stores
  volatile int* stop;            // host-mapped; nonzero = stop
  unsigned long long* wave_prog;  // host-mapped, 2 per wave: finished jobs, nodes up to them
  unsigned long long* prof;      // rounds [0..7], lanes [8..15], exits [16..22], cycles [24..31]
  SplitView sp;                  // task pool of the split kernels (ddx/split.hpp)
};

// Lanes of a wave that may help with published tasks (lanes 0, 16, 32, 48): tasks land on
// different waves -- a wave runs its lanes' searches one branch at a time, so four helpers in
// one wave share its issue slots, and up to 64 would share them sixty-four ways.
constexpr int kHelperStride = 16;

// Splits start once at most lanes * DDX_SPLIT_OPEN_NUM / DDX_SPLIT_OPEN_DEN jobs of the batch are
// unfinished: by default a quarter of the helpers (four per wave); other values build another
// gate for measurement.
#ifndef DDX_SPLIT_OPEN_NUM
#define DDX_SPLIT_OPEN_NUM 1
#endif
#ifndef DDX_SPLIT_OPEN_DEN
#define DDX_SPLIT_OPEN_DEN 64
#endif

// Publishes the wave's finished jobs and the nodes its searches spent up to them to its two
// slots of host memory (plain stores by one lane), so the host reads progress without any GPU
// work -- a copy would have to wait for a free CU while the persistent kernel holds them all.
// Called with the whole wave converged; `leader` marks the lanes that own a search's counters.
__device__ __forceinline__ void publish_wave_progress(unsigned long long* wave_prog, bool leader, u32 jobs,
                                                      u64 nodes) {
  unsigned long long j = leader ? jobs : 0, n = leader ? nodes : 0;
#pragma unroll
  for (int m = 32; m; m >>= 1) {
    j += __shfl_xor(j, m);
    n += __shfl_xor(n, m);
  }
  if (__lane_id() == 0) {
    const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    wave_prog[2 * w] = j;
    wave_prog[2 * w + 1] = n;
  }
}

// The split kernels' part of a round, before the cheap stages:
//   * every 64 rounds: whether lanes wait for work and at most max_open_jobs jobs are unfinished
//     (splits are tried only then: while many jobs run, a split would idle its owner with
//     little to gain), and whether every job of the batch is done (lanes without one of them
//     left then exit);
//   * every 256 rounds: a lane whose task lies under a cut-off join drops it;
//   * a lane waiting for its join checks it;
//   * idle lanes -- at most four per wave -- draw tickets (one atomic add per wave) and take the
//     task in their slot once it is written.
// Nothing here waits inside the round: a lane that cannot go on retries next round.
template <class Stack>
__device__ __forceinline__ void split_round(Lane& L, const Batch& B, Stack& st, const KernelArgs& a,
                                            unsigned rounds) {
  if ((rounds & 63) == 0) {
    const unsigned long long done = sp_load(a.counters + 5);
    L.may_split = split_starving(B.sp) && (unsigned long long)a.jobs - done <= (unsigned long long)B.sp.max_open_jobs;
    const bool all_done = done >= (unsigned long long)a.jobs;
    if (all_done && (L.stage == ST_HELP || L.task >= 0)) L.stage = ST_EXIT;
  }
  if ((rounds & 255) == 0 && L.stage != ST_EXIT && lane_task_moot(L, B)) lane_drop_task(L);
  if (L.stage == ST_WAIT) lane_wait<SeqOps>(L, B, st);
  const bool want = L.stage == ST_HELP && L.ticket < 0 && (__lane_id() % kHelperStride) == 0;
  const unsigned long long m = __ballot(want);
  if (m) {
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if ((int)__lane_id() == leader) base = sp_add(B.sp.ctr + 1, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    if (want) L.ticket = (long long)(base + __popcll(m & ((1ull << __lane_id()) - 1)));
  }
  if (L.stage == ST_HELP) lane_help<SeqOps>(L, B);
}

template <int SCHED, bool PROF, int TEAM, bool LDS = false, bool SPLIT = false>
__device__ __forceinline__ void solve_body(const KernelArgs& a) {
  using TM = TeamOf<TEAM>;
  using X = typename std::conditional<LDS, LdsOps, typename TM::Ops>::type;
  if (LDS) lds_reset(lds_lane(), true);
  const long long lane = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / TEAM;  // the search
  Lane L;
  GpuStack st;
  lane_attach(a.lanes, lane, L, st);
  // Without SPLIT the task pool is a compile-time null, and every split path folds away.
  const Batch B{a.hands, a.out, a.seeds, a.jobs, a.jobs_per_deal, a.seed_mode, SPLIT ? a.sp : SplitView{}};
  unsigned long long* next = a.counters;
  // Each search's first job is assigned so that consecutive jobs go to different waves (job j
  // to slot j / W of wave j % W, W waves): a batch with fewer jobs than searches then puts one
  // or a few searches in each wave instead of filling the first waves, whose lanes would run
  // their divergent searches one branch at a time. Later jobs come from the atomic counter,
  // after the S statically assigned ones. Each job's search is the same wherever it runs.
  const long long S = (long long)gridDim.x * blockDim.x / TEAM, per_wave = 64 / TEAM;
  bool fresh = true;
  auto take = [&]() -> long long {
    if (fresh) {
      fresh = false;
      const long long j = (lane % per_wave) * (S / per_wave) + lane / per_wave;
      if (j < a.jobs) return j;
    }
    return S + TM::take(next);
  };
  // Profile (PROF only; wave-uniform counters kept by every lane, lane 0 reports):
  // kinds 0..5 = stage rounds of the PRIORITY schedule, 6 = cheap chains (with the follow
  // nodes' ENTER under SPLIT), 2 = (trick-start) ENTER in the other schedules.
  unsigned long long p_rounds[8], p_lanes[8], p_cycles[8];
  unsigned p_exit[EX_COUNT];
  if (PROF) {
    for (int i = 0; i < 8; i++) p_rounds[i] = p_lanes[i] = p_cycles[i] = 0;
    for (int i = 0; i < EX_COUNT; i++) p_exit[i] = 0;
  }
  auto prof_round = [&](int kind, bool active, long long t0) {
    if (PROF) {
      unsigned long long m = __ballot(active);
      p_rounds[kind]++;
      p_lanes[kind] += __popcll(m);
      p_cycles[kind] += (unsigned long long)(clock64() - t0);
    }
  };
  unsigned rounds = 0;
  u32 reported = 0;  // finished jobs already added to the progress counter
  u64 reported_nodes = 0;

  // Cheap stages of one lane, back to back, until it must enter a node or has no work. Each
  // pass runs the stages in the order a lane usually meets them -- RET, NEXT, JOB, ROOT -- so
  // that a child's return and the next move, or a search's end and the next one's start, take
  // one pass rather than one per stage. Under SPLIT a pass also enters follow nodes (2nd-4th
  // hand), so the chain runs until every lane waits at a trick start (or has no work).
  // With SPLIT a lane also leaves the chain when it waits for a join (WAIT) or has no job left
  // (HELP); the round loop serves both.
  auto cheap_chain = [&](unsigned& passes) {
    while (L.stage != ST_EXIT && (L.stage != ST_ENTER || (SCHED == DDX_SCHED_SPLIT && L.p.nplayed != 0))) {
      if (SPLIT && (L.stage == ST_WAIT || L.stage == ST_HELP)) break;
      if (PROF) passes++;
      if (L.stage == ST_RET) lane_ret<X>(L, B, st);
      if (L.stage == ST_NEXT) lane_next<X>(L, B, st);
      if (L.stage == ST_JOB) {
        const long long job = take();
        if (job >= a.jobs) L.stage = SPLIT ? ST_HELP : ST_EXIT;
        else lane_load_job<X>(L, B, job);
      }
      if (L.stage == ST_ROOT) lane_root(L, B);
      if (SCHED == DDX_SCHED_SPLIT && L.stage == ST_ENTER && L.p.nplayed != 0) {
        lane_enter<X>(L, B, st);
        if (PROF) p_exit[L.exit_at]++;
      }
    }
  };

  for (;;) {
    // The stop flag changes under the kernel: one lane reads it for the whole team.
    if ((++rounds & 4095) == 0 && TM::bcast((int)*a.stop, 0)) L.stage = ST_EXIT;
    if ((rounds & 255) == 0) publish_wave_progress(a.wave_prog, TM::leader(), L.jobs_done, reported_nodes);
    if (L.jobs_done != reported) {  // progress for the host (one atomic per finished job)
      if (TM::leader()) {
        atomicAdd(a.counters + 5, (unsigned long long)(L.jobs_done - reported));
        atomicAdd(a.counters + 8, (unsigned long long)(L.c.nodes - reported_nodes));
      }
      reported = L.jobs_done;
      reported_nodes = L.c.nodes;
    }
    if (SPLIT) split_round(L, B, st, a, rounds);
    if (SCHED == DDX_SCHED_THREAD) {
      // No wave schedule: each lane runs its own next stage, whichever it is, so a round
      // executes every stage some lane of the wave is in (the reference's per-thread kernel).
      if (L.stage != ST_EXIT) lane_step<X>(L, B, st, take);
      if (__all(L.stage == ST_EXIT)) break;
    } else if (SCHED == DDX_SCHED_PRIORITY) {
      if (__all(L.stage == ST_EXIT)) break;
      int pick;
      if (__any(L.stage == ST_RET)) pick = ST_RET;
      else if (__any(L.stage == ST_NEXT)) pick = ST_NEXT;
      else if (__any(L.stage == ST_ROOT)) pick = ST_ROOT;
      else if (__any(L.stage == ST_JOB)) pick = ST_JOB;
      else pick = ST_ENTER;
      const bool run = L.stage == pick;
      const long long t0 = PROF ? clock64() : 0;
      if (run) {
        lane_step<X>(L, B, st, take);
        if (PROF && pick == ST_ENTER) p_exit[L.exit_at]++;
      }
      prof_round(pick, run, t0);
    } else {
      long long t0 = PROF ? clock64() : 0;
      unsigned passes = 0;
      cheap_chain(passes);
      if (PROF) {
        // Each pass runs for the lanes still in the chain: the wave's rounds are the most passes
        // any lane made, its active lanes the passes of all lanes.
        unsigned most = passes, all = passes;
#pragma unroll
        for (int m = 32; m; m >>= 1) {
          const unsigned o = __shfl_xor(most, m);
          most = o > most ? o : most;
          all += __shfl_xor(all, m);
        }
        p_rounds[6] += most;
        p_lanes[6] += all;
        p_cycles[6] += (unsigned long long)(clock64() - t0);
      }
      if (__all(L.stage == ST_EXIT)) break;
      // A wave with no node to enter only polls (its lanes wait for joins or for work): sleep
      // a little so that idle waves do not crowd the memory system with those polls.
      if (SPLIT && !__any(L.stage == ST_ENTER)) __builtin_amdgcn_s_sleep(16);
      // The trick-start nodes (SPLIT), or every node (CHAIN), entered with the wave gathered.
      const bool run = L.stage == ST_ENTER;
      t0 = PROF ? clock64() : 0;
      if (run) {
        lane_enter<X>(L, B, st);
        if (PROF) p_exit[L.exit_at]++;
      }
      prof_round(ST_ENTER, run, t0);
    }
  }
  publish_wave_progress(a.wave_prog, TM::leader(), L.jobs_done, reported_nodes);
  if (!TM::leader()) return;
  a.lanes.gen[lane] = L.c.gen;
  atomicAdd(a.counters + 1, (unsigned long long)L.c.nodes);
  atomicAdd(a.counters + 2, (unsigned long long)L.c.probes);
  atomicAdd(a.counters + 3, (unsigned long long)L.c.hits);
  atomicAdd(a.counters + 4, (unsigned long long)L.jobs_done);
  atomicAdd(a.counters + 6, (unsigned long long)L.c.ent_reads);
  atomicAdd(a.counters + 7, (unsigned long long)L.c.stores);
  if (LDS) atomicAdd(a.counters + 9, (unsigned long long)lds_lane()[4 + 16 * kLdsLines]);
  if (PROF) {
    if ((threadIdx.x & 63) == 0) {  // (a team leader; lane 0 of the wave)
      for (int i = 0; i < 8; i++) {
        atomicAdd(a.prof + i, p_rounds[i]);
        atomicAdd(a.prof + 8 + i, p_lanes[i]);
        atomicAdd(a.prof + 24 + i, p_cycles[i]);
      }
    }
    for (int i = 0; i < EX_COUNT; i++)
      if (p_exit[i]) atomicAdd(a.prof + 16 + i, (unsigned long long)p_exit[i]);
  }
}

// The solver at the compiler's register allocation (2 waves per SIMD: ~230 VGPRs one lane per
// search, ~210 for teams) ...
template <int SCHED, bool PROF, int TEAM>
__global__ __launch_bounds__(256) void solve_kernel(KernelArgs a) {
  solve_body<SCHED, PROF, TEAM>(a);
}
// ... or held to W waves per SIMD (168 VGPRs for 3, 128 for 4): more waves to hide memory
// latency, at the price of spilling the coldest values to scratch.
template <int TEAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void solve_kernel_w3(KernelArgs a) {
  solve_body<DDX_SCHED_SPLIT, false, TEAM>(a);
}
template <int TEAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void solve_kernel_w4(KernelArgs a) {
  solve_body<DDX_SCHED_SPLIT, false, TEAM>(a);
}
// One lane per search with splitting of the tail (ddx/split.hpp), at the compiler's allocation
// or held to 3 waves per SIMD.
__global__ __launch_bounds__(256) void solve_kernel_split(KernelArgs a) {
  solve_body<DDX_SCHED_SPLIT, false, 1, false, true>(a);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void solve_kernel_split_w3(KernelArgs a) {
  solve_body<DDX_SCHED_SPLIT, false, 1, false, true>(a);
}
// One lane per search with the LDS cache of tag lines (ddx_lds.hpp), at the compiler's
// allocation or held to 3 waves per SIMD.
__global__ __launch_bounds__(256) void solve_kernel_lds(KernelArgs a) {
  solve_body<DDX_SCHED_SPLIT, false, 1, true>(a);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void solve_kernel_lds_w3(KernelArgs a) {
  solve_body<DDX_SCHED_SPLIT, false, 1, true>(a);
}

namespace ddx {

__global__ void fill_kernel(Q4* p, size_t n, u32 v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  Q4 q = {v, v, v, v};
  for (; i < n; i += stride) p[i] = q;
}
__global__ void fill_bytes_kernel(uint8_t* p, size_t n, uint8_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int launch_fill(void* p, size_t bytes, u32 v, hipStream_t s) {
  size_t n = bytes / 16;
  if (n) {
    int grid = (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, s, (Q4*)p, n, v);
    HIPCHECK(hipGetLastError());
  }
  return DDX_OK;
}

int lanes_alloc(LaneStore& s, long long lanes, int tt_log2) {
  const size_t L = (size_t)lanes, buckets = (size_t)1 << tt_log2;
  const size_t tts = DDX_SHARED_TT ? 1 : L;  // TTs: one per search, or one shared
  const size_t tag_bytes = tts * buckets * 64, ent_bytes = tts * buckets * kBucketEntryBytes;
  const size_t stack_bytes = L * kStackDepth * sizeof(PackedFrame), gen_bytes = (L * 4 + 15) / 16 * 16;
  s.lanes = lanes;
  s.tt_log2 = tt_log2;
  if (hipMalloc(&s.tags, tag_bytes) != hipSuccess || hipMalloc(&s.ents, ent_bytes) != hipSuccess ||
      hipMalloc(&s.stacks, stack_bytes) != hipSuccess || hipMalloc(&s.gen, gen_bytes) != hipSuccess) {
    lanes_free(s);
    ddx_set_error("cannot allocate %.2f GB of lane state for %lld lanes at TT_LOG2 %d",
                  (tag_bytes + ent_bytes + stack_bytes) / 1e9, lanes, tt_log2);
    return DDX_E_NOMEM;
  }
  s.bytes = tag_bytes + ent_bytes + stack_bytes + gen_bytes;
  int rc;
  if ((rc = launch_fill(s.tags, tag_bytes, 0, nullptr)) != DDX_OK ||
      (rc = launch_fill(s.gen, gen_bytes, 0, nullptr)) != DDX_OK) {
    lanes_free(s);
    return rc;
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    lanes_free(s);
    ddx_set_error("clearing the lanes' TT failed");
    return DDX_E_HIP;
  }
  return DDX_OK;
}

void lanes_free(LaneStore& s) {
  if (s.tags) (void)hipFree(s.tags);
  if (s.ents) (void)hipFree(s.ents);
  if (s.stacks) (void)hipFree(s.stacks);
  if (s.gen) (void)hipFree(s.gen);
  s.tags = nullptr;
  s.ents = nullptr;
  s.stacks = nullptr;
  s.gen = nullptr;
  s.bytes = 0;
}

}  // namespace ddx

// ---------------------------------------------------------------------------------------
// Host driver.
struct ddx_gpu {
  ddx_gpu_config cfg;
  LaneStore lanes;
  unsigned long long* counters = nullptr;
  unsigned long long* prof = nullptr;
  int* h_stop = nullptr;  // host-mapped
  int* d_stop = nullptr;
  unsigned long long* h_prog = nullptr;  // host-mapped wave progress slots (2 per wave)
  unsigned long long* d_prog = nullptr;
  // staging for ddx_gpu_solve (host buffers in, host buffers out)
  u64* d_hands = nullptr;
  uint8_t* d_out = nullptr;
  uint8_t* d_seeds = nullptr;
  int64_t cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipStream_t stream = nullptr;  // stream of the last launch
  hipStream_t own = nullptr;     // non-blocking stream for calls that pass none
  hipStream_t side = nullptr;    // non-blocking stream for progress reads
  hipGraphExec_t gexec = nullptr;  // captured batch (config.graph)
  // Task pool of the split kernels (config.split_min_nodes): cleared at every launch.
  SplitTask* sp_tasks = nullptr;
  SplitJoin* sp_joins = nullptr;
  unsigned long long* sp_ctr = nullptr;
  long long sp_cap = 0;
  int64_t g_n = -1;
  const void* g_hands = nullptr;
  const void* g_out = nullptr;
  const void* g_seeds = nullptr;
  hipStream_t g_stream = nullptr;
  std::chrono::steady_clock::time_point t0;
  double last_ms = 0;
  int64_t last_jobs = 0;
  uint64_t bytes = 0;
};

extern "C" void ddx_gpu_default_config(ddx_gpu_config* cfg) {
  cfg->device = 0;
  cfg->lanes = 65536;
  cfg->tt_log2 = 10;
  cfg->jobs_per_deal = 5;
  cfg->schedule = DDX_SCHED_CHAIN;
  cfg->seed_mode = DDX_SEED_PLAIN;
  cfg->block = 256;
  cfg->profile = 0;
  cfg->team = 1;
  cfg->graph = 0;
  cfg->waves_per_simd = 0;
  cfg->lds_lines = 0;
  cfg->split_min_nodes = 0;
}

static void free_all(ddx_gpu* h) {
  lanes_free(h->lanes);
  if (h->counters) (void)hipFree(h->counters);
  if (h->prof) (void)hipFree(h->prof);
  if (h->h_stop) (void)hipHostFree(h->h_stop);
  if (h->h_prog) (void)hipHostFree(h->h_prog);
  if (h->d_hands) (void)hipFree(h->d_hands);
  if (h->d_out) (void)hipFree(h->d_out);
  if (h->d_seeds) (void)hipFree(h->d_seeds);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->own) (void)hipStreamDestroy(h->own);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->sp_tasks) (void)hipFree(h->sp_tasks);
  if (h->sp_joins) (void)hipFree(h->sp_joins);
  if (h->sp_ctr) (void)hipFree(h->sp_ctr);
}


extern "C" int ddx_gpu_open(const ddx_gpu_config* cfg_in, ddx_gpu** out) {
  if (!out) return DDX_E_ARG;
  ddx_gpu_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else ddx_gpu_default_config(&cfg);
  if ((DDX_TT_SIG || DDX_SHARED_TT) && cfg.team != 1) {
    ddx_set_error("ddx_gpu_open: the %s build runs one lane per search (team 1)",
                  DDX_TT_SIG ? "signature-entry" : "shared-TT");
    return DDX_E_ARG;
  }
  if (DDX_SHARED_TT && (cfg.split_min_nodes || cfg.lds_lines)) {
    ddx_set_error("ddx_gpu_open: the shared-TT build has no splitting and no LDS line cache");
    return DDX_E_ARG;
  }
  if (cfg.tt_log2 < kMinLog2 || cfg.tt_log2 > kMaxTTLog2 || cfg.lanes < 64 || cfg.lanes % 64 ||
      (cfg.jobs_per_deal != 5 && cfg.jobs_per_deal != 20) || cfg.schedule < 0 ||
      cfg.schedule > 3 || (cfg.schedule == DDX_SCHED_THREAD && cfg.profile) || cfg.seed_mode < 0 ||
      cfg.seed_mode > 3 || cfg.block < 64 ||
      cfg.block > 256 || cfg.block % 64 || cfg.lanes % cfg.block || (cfg.team != 1 && cfg.team != kWideTeam) ||
      (cfg.waves_per_simd != 0 &&
       (cfg.waves_per_simd < 3 || cfg.waves_per_simd > 4 || cfg.schedule != DDX_SCHED_SPLIT || cfg.profile)) ||
      (cfg.lds_lines != 0 && (cfg.lds_lines != 3 || cfg.team != 1 || cfg.schedule != DDX_SCHED_SPLIT ||
                              cfg.profile || cfg.waves_per_simd == 4)) ||
      cfg.split_min_nodes < 0 ||
      (cfg.split_min_nodes > 0 && (cfg.team != 1 || cfg.schedule != DDX_SCHED_SPLIT || cfg.profile ||
                                   cfg.lds_lines || cfg.waves_per_simd == 4))) {
    ddx_set_error("ddx_gpu_open: bad config (lanes %lld must be a multiple of block %d and of 64, "
                  "TT_LOG2 %d in [9, %d], jobs_per_deal 5 or 20, team 1 or 16; splitting needs one "
                  "lane per search, the split schedule, no profile, no LDS cache, 2 or 3 waves per SIMD)",
                  (long long)cfg.lanes, cfg.block, cfg.tt_log2, kMaxTTLog2);
    return DDX_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg.device) {
    ddx_set_error("ddx_gpu_open: no HIP device %d", cfg.device);
    return DDX_E_NODEV;
  }
  HIPCHECK(hipSetDevice(cfg.device));
  ddx_gpu* h = new (std::nothrow) ddx_gpu;
  if (!h) return DDX_E_NOMEM;
  h->cfg = cfg;
  auto fail = [&](const char* what) {
    ddx_set_error("ddx_gpu_open: cannot allocate %s", what);
    free_all(h);
    delete h;
    return DDX_E_NOMEM;
  };
  int rc = lanes_alloc(h->lanes, cfg.lanes / cfg.team, cfg.tt_log2);
  if (rc != DDX_OK) {
    delete h;
    return rc;
  }
  if (hipMalloc(&h->counters, 16 * 8) != hipSuccess) return fail("counters");
  if (cfg.split_min_nodes > 0) {
    // Eight task slots per lane: a split publishes at most 12, and nodes are split only while
    // lanes wait for work.
    h->sp_cap = cfg.lanes * 8;
    if (hipMalloc(&h->sp_tasks, (size_t)h->sp_cap * sizeof(SplitTask)) != hipSuccess ||
        hipMalloc(&h->sp_joins, (size_t)h->sp_cap * sizeof(SplitJoin)) != hipSuccess ||
        hipMalloc(&h->sp_ctr, 8 * 8) != hipSuccess)
      return fail("split task pool");
  }
  if (hipMalloc(&h->prof, 32 * 8) != hipSuccess) return fail("profile counters");
  if (hipHostMalloc((void**)&h->h_stop, 64, hipHostMallocMapped) != hipSuccess) return fail("stop flag");
  *h->h_stop = 0;
  if (hipHostGetDevicePointer((void**)&h->d_stop, h->h_stop, 0) != hipSuccess) return fail("stop flag");
  const size_t prog_bytes = (size_t)(cfg.lanes / 64) * 16;
  if (hipHostMalloc((void**)&h->h_prog, prog_bytes, hipHostMallocMapped) != hipSuccess) return fail("progress slots");
  memset(h->h_prog, 0, prog_bytes);
  if (hipHostGetDevicePointer((void**)&h->d_prog, h->h_prog, 0) != hipSuccess) return fail("progress slots");
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return fail("events");
  if (hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess)
    return fail("streams");
  h->bytes = h->lanes.bytes + (uint64_t)h->sp_cap * (sizeof(SplitTask) + sizeof(SplitJoin));
  if ((rc = launch_fill(h->prof, 32 * 8, 0, nullptr)) != DDX_OK || hipDeviceSynchronize() != hipSuccess) {
    free_all(h);
    delete h;
    return rc != DDX_OK ? rc : DDX_E_HIP;
  }
  *out = h;
  return DDX_OK;
}

extern "C" void ddx_gpu_close(ddx_gpu* h) {
  if (!h) return;
  (void)hipSetDevice(h->cfg.device);
  (void)hipDeviceSynchronize();
  free_all(h);
  delete h;
}

extern "C" uint64_t ddx_gpu_device_bytes(const ddx_gpu* h) { return h ? h->bytes : 0; }
extern "C" void ddx_gpu_request_stop(ddx_gpu* h) {
  if (h && h->h_stop) __atomic_store_n(h->h_stop, 1, __ATOMIC_SEQ_CST);
}
extern "C" double ddx_gpu_last_kernel_ms(const ddx_gpu* h) { return h ? h->last_ms : 0; }

template <int S, bool P>
static void launch_solve(const KernelArgs& a, int grid, int block, int team, hipStream_t s) {
  if (team == 16) hipLaunchKernelGGL((solve_kernel<S, P, kWideTeam>), dim3(grid), dim3(block), 0, s, a);
  else hipLaunchKernelGGL((solve_kernel<S, P, 1>), dim3(grid), dim3(block), 0, s, a);
}

// Queues one batch on stream s: clear the counters, mark every answer unsolved (0xFF), run the
// solver. Kernels only (no memset nodes), so the sequence can be captured in a graph.
static int enqueue_solve(ddx_gpu* h, const KernelArgs& a, int64_t n, hipStream_t s) {
  int rc;
  if ((rc = launch_fill(h->counters, 128, 0, s)) != DDX_OK) return rc;
  if ((rc = launch_fill(h->d_prog, (size_t)(h->cfg.lanes / 64) * 16, 0, s)) != DDX_OK) return rc;
  if (h->sp_tasks) {  // every slot SLOT_EMPTY, the counters zero
    if ((rc = launch_fill(h->sp_ctr, 8 * 8, 0, s)) != DDX_OK) return rc;
    if ((rc = launch_fill(h->sp_tasks, (size_t)h->sp_cap * sizeof(SplitTask), 0, s)) != DDX_OK) return rc;
  }
  if (n) {
    int grid = (int)((20 * n + 255) / 256 < 65536 ? (20 * n + 255) / 256 : 65536);
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(grid), dim3(256), 0, s, a.out, (size_t)(20 * n),
                       (uint8_t)0xFF);
    HIPCHECK(hipGetLastError());
  }
  const int block = h->cfg.block;
  const int grid = (int)(h->cfg.lanes / block);
  const bool P = h->cfg.profile != 0;
  const int W = h->cfg.waves_per_simd;
  if (h->sp_tasks) {  // one lane per search, split schedule, no profile (checked at open)
    if (W == 3) hipLaunchKernelGGL(solve_kernel_split_w3, dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL(solve_kernel_split, dim3(grid), dim3(block), 0, s, a);
    HIPCHECK(hipGetLastError());
    return DDX_OK;
  }
  if (h->cfg.lds_lines) {  // one lane per search, split schedule, no profile (checked at open)
    if (W == 3) hipLaunchKernelGGL(solve_kernel_lds_w3, dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL(solve_kernel_lds, dim3(grid), dim3(block), 0, s, a);
    HIPCHECK(hipGetLastError());
    return DDX_OK;
  }
  if (W == 3 || W == 4) {  // split schedule, no profile (checked at open)
    if (W == 3 && h->cfg.team == 16) hipLaunchKernelGGL(solve_kernel_w3<kWideTeam>, dim3(grid), dim3(block), 0, s, a);
    else if (W == 3) hipLaunchKernelGGL(solve_kernel_w3<1>, dim3(grid), dim3(block), 0, s, a);
    else if (h->cfg.team == 16) hipLaunchKernelGGL(solve_kernel_w4<kWideTeam>, dim3(grid), dim3(block), 0, s, a);
    else hipLaunchKernelGGL(solve_kernel_w4<1>, dim3(grid), dim3(block), 0, s, a);
    HIPCHECK(hipGetLastError());
    return DDX_OK;
  }
  switch (h->cfg.schedule * 2 + (P ? 1 : 0)) {
    case 0: launch_solve<DDX_SCHED_PRIORITY, false>(a, grid, block, h->cfg.team, s); break;
    case 1: launch_solve<DDX_SCHED_PRIORITY, true>(a, grid, block, h->cfg.team, s); break;
    case 2: launch_solve<DDX_SCHED_CHAIN, false>(a, grid, block, h->cfg.team, s); break;
    case 3: launch_solve<DDX_SCHED_CHAIN, true>(a, grid, block, h->cfg.team, s); break;
    case 4: launch_solve<DDX_SCHED_SPLIT, false>(a, grid, block, h->cfg.team, s); break;
    case 5: launch_solve<DDX_SCHED_SPLIT, true>(a, grid, block, h->cfg.team, s); break;
    default: launch_solve<DDX_SCHED_THREAD, false>(a, grid, block, h->cfg.team, s); break;
  }
  HIPCHECK(hipGetLastError());
  return DDX_OK;
}

extern "C" int ddx_gpu_solve_async(ddx_gpu* h, const uint64_t* d_hands, int64_t n,
                                   const uint8_t* d_seeds, uint8_t* d_out, void* stream) {
  if (!h || n < 0 || (n && (!d_hands || !d_out)) ||
      ((h->cfg.seed_mode == DDX_SEED_STEP || h->cfg.seed_mode == DDX_SEED_CERT) && !d_seeds && n)) {
    ddx_set_error("ddx_gpu_solve_async: bad arguments");
    return DDX_E_ARG;
  }
  HIPCHECK(hipSetDevice(h->cfg.device));
  hipStream_t s = stream ? (hipStream_t)stream : h->own;
  h->stream = s;
  __atomic_store_n(h->h_stop, 0, __ATOMIC_SEQ_CST);
  const long long jobs = (long long)n * h->cfg.jobs_per_deal;
  h->last_jobs = jobs;
  KernelArgs a;
  a.hands = d_hands;
  a.out = d_out;
  a.seeds = d_seeds;
  a.jobs = jobs;
  a.jobs_per_deal = h->cfg.jobs_per_deal;
  a.seed_mode = h->cfg.seed_mode;
  a.lanes = lanes_view(h->lanes);
  a.counters = h->counters;
  a.stop = h->d_stop;
  a.wave_prog = h->d_prog;
  a.prof = h->prof;
  a.sp = SplitView{};
  if (h->sp_tasks) {
    a.sp.tasks = h->sp_tasks;
    a.sp.joins = h->sp_joins;
    a.sp.ctr = h->sp_ctr;
    a.sp.task_cap = a.sp.join_cap = h->sp_cap;
    a.sp.min_nodes16 = (u32)((h->cfg.split_min_nodes + 15) / 16);
    a.sp.max_open_jobs = h->cfg.lanes * DDX_SPLIT_OPEN_NUM / DDX_SPLIT_OPEN_DEN;
  }
  h->t0 = std::chrono::steady_clock::now();
  HIPCHECK(hipEventRecord(h->ev0, s));
  if (!h->cfg.graph) {
    const int rc = enqueue_solve(h, a, n, s);
    if (rc != DDX_OK) return rc;
  } else {
    // One graph per (buffers, size, stream): fill kernels + solve kernel, replayed on reuse.
    if (!h->gexec || h->g_n != n || h->g_hands != d_hands || h->g_out != d_out ||
        h->g_seeds != d_seeds || h->g_stream != s) {
      if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
      h->gexec = nullptr;
      hipGraph_t graph;
      HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_solve(h, a, n, s);
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (rc != DDX_OK) return rc;
      HIPCHECK(e);
      e = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHECK(e);
      h->g_n = n;
      h->g_hands = d_hands;
      h->g_out = d_out;
      h->g_seeds = d_seeds;
      h->g_stream = s;
    }
    HIPCHECK(hipGraphLaunch(h->gexec, s));
  }
  HIPCHECK(hipEventRecord(h->ev1, s));
  return DDX_OK;
}

extern "C" int ddx_gpu_wait(ddx_gpu* h, ddx_stats* stats) {
  if (!h) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->cfg.device));
  HIPCHECK(hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  unsigned long long c[16];  // the whole counter block (16 x 8 bytes)
  HIPCHECK(hipMemcpy(c, h->counters, sizeof c, hipMemcpyDeviceToHost));
  if (stats) {
    stats->nodes = c[1];
    stats->tt_probes = c[2];
    stats->tt_hits = c[3];
    stats->jobs_done = c[4];
    stats->tt_entry_reads = c[6];
    stats->tt_stores = c[7];
    stats->tt_line_cache_hits = c[9];
    stats->splits = stats->tasks_taken = 0;
    if (h->sp_ctr) {
      unsigned long long sc[8];
      HIPCHECK(hipMemcpy(sc, h->sp_ctr, sizeof sc, hipMemcpyDeviceToHost));
      stats->splits = sc[3];
      stats->tasks_taken = sc[4];
    }
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - h->t0).count();
  }
  if ((int64_t)c[4] < h->last_jobs) {
    ddx_set_error("stopped after %llu of %lld jobs", c[4], (long long)h->last_jobs);
    return DDX_E_STOPPED;
  }
  return DDX_OK;
}

static int ensure_staging(ddx_gpu* h, int64_t n) {
  if (n <= h->cap) return DDX_OK;
  if (h->d_hands) (void)hipFree(h->d_hands);
  if (h->d_out) (void)hipFree(h->d_out);
  if (h->d_seeds) (void)hipFree(h->d_seeds);
  h->d_hands = nullptr;
  h->d_out = h->d_seeds = nullptr;
  h->cap = 0;
  HIPCHECK(hipMalloc(&h->d_hands, (size_t)n * 32));
  HIPCHECK(hipMalloc(&h->d_out, (size_t)n * 20));
  HIPCHECK(hipMalloc(&h->d_seeds, (size_t)n * 20));
  h->cap = n;
  return DDX_OK;
}

extern "C" int ddx_gpu_solve(ddx_gpu* h, const uint64_t* hands, int64_t n, const uint8_t* seeds,
                             uint8_t* out, ddx_stats* stats) {
  if (!h || n < 0 || (n && (!hands || !out))) {
    ddx_set_error("ddx_gpu_solve: bad arguments");
    return DDX_E_ARG;
  }
  for (int64_t i = 0; i < n; i++)
    if (!ddx_check_deal(hands + 4 * i)) {
      ddx_set_error("ddx_gpu_solve: deal %lld is not a valid deal", (long long)i);
      return DDX_E_ARG;
    }
  HIPCHECK(hipSetDevice(h->cfg.device));
  int rc = ensure_staging(h, n > 0 ? n : 1);
  if (rc != DDX_OK) return rc;
  if (n) HIPCHECK(hipMemcpy(h->d_hands, hands, (size_t)n * 32, hipMemcpyHostToDevice));
  const bool want_seeds = h->cfg.seed_mode == DDX_SEED_STEP || h->cfg.seed_mode == DDX_SEED_CERT;
  if (want_seeds) {
    if (!seeds) {
      ddx_set_error("ddx_gpu_solve: this seed mode needs seeds");
      return DDX_E_ARG;
    }
    if (n) HIPCHECK(hipMemcpy(h->d_seeds, seeds, (size_t)n * 20, hipMemcpyHostToDevice));
  }
  rc = ddx_gpu_solve_async(h, h->d_hands, n, want_seeds ? h->d_seeds : nullptr, h->d_out, h->own);
  if (rc != DDX_OK) return rc;
  int wrc = ddx_gpu_wait(h, stats);
  if (wrc != DDX_OK && wrc != DDX_E_STOPPED) return wrc;
  if (n) HIPCHECK(hipMemcpy(out, h->d_out, (size_t)n * 20, hipMemcpyDeviceToHost));
  return wrc;
}

// Jobs finished so far by the running (or last) solve; safe to call while it runs.
extern "C" int ddx_gpu_progress(ddx_gpu* h, uint64_t* jobs_done) {
  if (!h || !jobs_done) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->cfg.device));
  uint64_t nodes;
  return ddx_gpu_progress_nodes(h, jobs_done, &nodes);
}

extern "C" int ddx_gpu_progress_nodes(ddx_gpu* h, uint64_t* jobs_done, uint64_t* nodes) {
  if (!h || !jobs_done || !nodes) return DDX_E_ARG;
  // Host memory written by the kernel's waves: no GPU call, so it answers while the kernel
  // occupies every CU. Each slot is one aligned 8-byte store, so reads never tear.
  const volatile unsigned long long* p = h->h_prog;
  uint64_t j = 0, n = 0;
  for (int64_t w = 0; w < h->cfg.lanes / 64; w++) {
    j += p[2 * w];
    n += p[2 * w + 1];
  }
  *jobs_done = j;
  *nodes = n;
  return DDX_OK;
}

// Per-stage profile of the last solves (config.profile = 1).
extern "C" int ddx_gpu_profile(ddx_gpu* h, uint64_t* out32, int reset) {
  if (!h || !out32) return DDX_E_ARG;
  HIPCHECK(hipSetDevice(h->cfg.device));
  HIPCHECK(hipMemcpy(out32, h->prof, 32 * 8, hipMemcpyDeviceToHost));
  if (reset) {
    int rc = launch_fill(h->prof, 32 * 8, 0, nullptr);
    if (rc != DDX_OK) return rc;
    HIPCHECK(hipDeviceSynchronize());
  }
  return DDX_OK;
}
