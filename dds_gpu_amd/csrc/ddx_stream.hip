// Streaming solver: one persistent kernel fed through a ring of pinned host slots, results
// published as soon as a deal is solved (out of input order). Replaces the reference's
// dd_stream (gpu/stream_main.cu:38-114 kernel, 124-345 host).
//
// Host memory (pinned, mapped, coherent):
//   hands[ring][4], seeds[ring][20], out[ring][20]   slot contents
//   work[work_ring]   slot of the i-th queued deal (slots are freed out of order)
//   comp[ring]        completion ring: slot + 1 of each finished deal, 0 = empty
//   ctl               produced (deals queued), eof, stop
// Device memory: the job counter, comp_head, per-slot finished-job counts.
//
// Ordering: the host writes a slot and work[p], then releases produced = p + 1; a lane
// acquires `produced` (system scope) before it reads them. A lane that finishes a job fences
// (system scope) before it counts the job in deal_done[slot] (device-scope atomic); the lane
// that completes the deal resets the count, fences, takes a position from comp_head (device
// memory) and release-stores slot + 1 into comp[] in host memory. There are no atomics on
// host memory, which PCIe platforms need not support.
#include <hip/hip_runtime.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <mutex>
#include <new>
#include <vector>

#include "ddx_lanes.hpp"
#include "ddx_team.hpp"

using namespace ddx;

struct StreamCtl {
  long long produced;
  int eof;
  int stop;
  long long pad[6];
};

struct StreamArgs {
  LaneView lanes;
  const u64* hands;
  const uint8_t* seeds;
  uint8_t* out;
  const u32* work;
  u32* comp;
  volatile StreamCtl* ctl;
  unsigned long long* counters;  // [0] next job, [1] comp_head, [2] nodes, [3] probes, [4] hits, [5] jobs
  u32* deal_done;                // per slot
  long long ring, work_ring;
  int jobs_per_deal, seed_mode;
};

// Searches of a wave run their divergent branches one at a time, so a search alone in its wave
// advances several times faster than one of four. When the ring holds little work, jobs are
// spread over waves: the first search slot of every wave keeps a job reserved in advance, the
// other slots reserve one only while produced jobs are waiting unreserved. Under full load
// every slot is busy either way. DDX_STREAM_SPREAD=0 builds the plain first-come order.
#ifndef DDX_STREAM_SPREAD
#define DDX_STREAM_SPREAD 1
#endif
// Rounds between reads of the host's counters while some lane of the wave waits for input
// (a power of two); DDX_STREAM_POLL builds another interval for measurement.
#ifndef DDX_STREAM_POLL
#define DDX_STREAM_POLL 256
#endif

__device__ __forceinline__ long long load_relaxed_sys(volatile long long* p) {
  return __hip_atomic_load((long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int load_relaxed_sys(volatile int* p) {
  return __hip_atomic_load((int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// TEAM = 1: one lane per search; 16: a 16-lane team per search. Values the host changes while
// the kernel runs (produced, eof, stop) live in host memory: one lane of the wave reads them,
// at most every DDX_STREAM_POLL rounds and only while some lane waits for input, and
// broadcasts them to the wave. The reads are relaxed; only when they show something new (more
// deals, the end) does that lane add the system-scope acquire that makes the host's slot
// contents visible. An acquire at system scope invalidates the caches the whole GPU shares,
// TT lines included, so doing it on every poll of every wave slowed the searches most when the
// load is light: at half load, polling every 16 rounds with an acquire each time gave a median
// latency of 15.7 s, every 256 rounds 4.3 s. The lanes of a team never disagree (they see one
// broadcast value).
template <int TEAM>
__device__ __forceinline__ void stream_body(const StreamArgs& a) {
  using TM = TeamOf<TEAM>;
  using X = typename TM::Ops;
  const long long lane = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / TEAM;
  Lane L;
  GpuStack st;
  lane_attach(a.lanes, lane, L, st);
  const int jpd = a.jobs_per_deal;
  const Batch B{a.hands, a.out, a.seeds, (long long)1 << 62, jpd, a.seed_mode};
  long long pending = -1;  // job reserved, maybe not yet produced
  u32 published = 0;
  unsigned rounds = 0;
  long long seen_produced = 0;  // the wave's last view of the host's counters
  int seen_end = 0;
  for (;;) {
    ++rounds;
    if (((rounds & 1023) == 0 && !(seen_end & 2)) ||
        ((rounds & (DDX_STREAM_POLL - 1)) == 0 &&
         __any(L.stage == ST_JOB && (pending < 0 || pending / jpd >= seen_produced)))) {
      long long v = 0;
      int e = 0;
      if (__lane_id() == 0) {  // end flags first: the host sets eof only after its last produced
        e = (load_relaxed_sys(&a.ctl->eof) ? 1 : 0) | (load_relaxed_sys(&a.ctl->stop) ? 2 : 0);
        if (e) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // eof before the last produced
        v = load_relaxed_sys(&a.ctl->produced);
        if (v != seen_produced) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // slots, work[]
      }
      seen_produced = __shfl(v, 0);
      seen_end = __shfl(e, 0);
    }
    if (seen_end & 2) L.stage = ST_EXIT;
    // Cheap stages until the search must enter a node, needs a job, or is done. (The batch
    // kernel's fused passes, follow-node ENTER inside the chain, measured 14% lower capacity
    // here with 16-lane teams: 77.6 against 90.1 deals/s.)
    while (L.stage == ST_ROOT || L.stage == ST_NEXT || L.stage == ST_RET) {
      if (L.stage == ST_ROOT) lane_root(L, B);
      else if (L.stage == ST_NEXT) lane_next<X>(L, B, st);
      else lane_ret<X>(L, B, st);
    }
    if (L.jobs_done != published) {  // a job just finished: count it, maybe publish the deal
      published = L.jobs_done;
      const long long slot = L.deal;
      if (TM::leader()) {
        __threadfence_system();
        if (atomicAdd(&a.deal_done[slot], 1u) == (u32)jpd - 1) {
          atomicExch(&a.deal_done[slot], 0u);
          __threadfence_system();
          const unsigned long long pos = atomicAdd(&a.counters[1], 1ull);
          __hip_atomic_store(&a.comp[pos % a.ring], (u32)slot + 1, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_SYSTEM);
        }
      }
    }
    if (L.stage == ST_JOB) {
      if (pending < 0) {
        int may = 1;
        if (DDX_STREAM_SPREAD && lane % (64 / TEAM) != 0) {
          const unsigned long long taken =
              __hip_atomic_load(&a.counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          may = TM::bcast((int)((long long)taken < seen_produced * jpd), 0);
        }
        if (may) pending = TM::take(&a.counters[0]);
      }
      // A search without a reservation (a spread slot while no job waits) loads nothing: the
      // test must not see -1 / jpd, which truncates to 0.
      if (pending >= 0 && pending / jpd < seen_produced) {
        const long long p = pending / jpd;
        const long long slot = a.work[p % a.work_ring];
        lane_load_job<X>(L, B, slot * jpd + pending % jpd);  // to ROOT; runs next round
        pending = -1;
      } else if (seen_end) {
        L.stage = ST_EXIT;
      }
    }
    // Every ballot below runs with the whole wave (no divergent continue above), so __all
    // and __any see all 64 lanes.
    if (__all(L.stage == ST_EXIT)) break;
    if (__any(L.stage == ST_ENTER && L.p.nplayed != 0)) {
      if (L.stage == ST_ENTER && L.p.nplayed != 0) lane_enter<X>(L, B, st);
      continue;
    }
    if (__any(L.stage == ST_ENTER)) {
      if (L.stage == ST_ENTER) lane_enter<X>(L, B, st);
      continue;
    }
    if (!__any(L.stage == ST_ROOT)) __builtin_amdgcn_s_sleep(64);  // all waiting for input
  }
  if (!TM::leader()) return;
  a.lanes.gen[lane] = L.c.gen;
  atomicAdd(&a.counters[2], (unsigned long long)L.c.nodes);
  atomicAdd(&a.counters[3], (unsigned long long)L.c.probes);
  atomicAdd(&a.counters[4], (unsigned long long)L.c.hits);
  atomicAdd(&a.counters[5], (unsigned long long)L.jobs_done);
}

// At the compiler's register allocation, or held to 3 waves per SIMD (more searches resident).
template <int TEAM>
__global__ __launch_bounds__(256) void stream_kernel(StreamArgs a) {
  stream_body<TEAM>(a);
}
template <int TEAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void stream_kernel_w3(StreamArgs a) {
  stream_body<TEAM>(a);
}

// ---------------------------------------------------------------------------------------
struct ddx_stream {
  ddx_gpu_config cfg;
  LaneStore lanes;
  long long ring = 0, work_ring = 0;
  u64* hands = nullptr;  // host, pinned + mapped
  uint8_t* seeds = nullptr;
  uint8_t* out = nullptr;
  u32* work = nullptr;
  u32* comp = nullptr;
  StreamCtl* ctl = nullptr;
  unsigned long long* counters = nullptr;  // device
  u32* deal_done = nullptr;                // device
  hipStream_t stream = nullptr;
  std::mutex mu;  // free slots, produced
  std::vector<u32> free_slots;
  std::vector<int64_t> slot_id;
  long long produced = 0, consumed = 0;
  unsigned long long tail = 0;
  bool launched = false, finished = false;
  std::chrono::steady_clock::time_point t0;
};

template <class T>
static int host_alloc(T** p, size_t bytes) {
  if (hipHostMalloc((void**)p, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    *p = nullptr;
    return DDX_E_NOMEM;
  }
  memset(*p, 0, bytes);
  return DDX_OK;
}

template <class T>
static T* dev_ptr(T* host) {
  void* d = nullptr;
  if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) return nullptr;
  return (T*)d;
}

static void stream_free(ddx_stream* s) {
  lanes_free(s->lanes);
  if (s->hands) (void)hipHostFree(s->hands);
  if (s->seeds) (void)hipHostFree(s->seeds);
  if (s->out) (void)hipHostFree(s->out);
  if (s->work) (void)hipHostFree(s->work);
  if (s->comp) (void)hipHostFree(s->comp);
  if (s->ctl) (void)hipHostFree(s->ctl);
  if (s->counters) (void)hipFree(s->counters);
  if (s->deal_done) (void)hipFree(s->deal_done);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

extern "C" int ddx_stream_open(const ddx_gpu_config* cfg_in, int ring_slots, ddx_stream** out) {
  if (!out || ring_slots < 1 || ring_slots > (1 << 22)) {
    ddx_set_error("ddx_stream_open: ring_slots must be in [1, 2^22]");
    return DDX_E_ARG;
  }
  ddx_gpu_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else ddx_gpu_default_config(&cfg);
  if ((DDX_TT_SIG || DDX_SHARED_TT) && cfg.team != 1) {
    ddx_set_error("ddx_stream_open: the %s build runs one lane per search (team 1)",
                  DDX_TT_SIG ? "signature-entry" : "shared-TT");
    return DDX_E_ARG;
  }
  if (cfg.tt_log2 < kMinLog2 || cfg.tt_log2 > kMaxTTLog2 || cfg.lanes < 64 || cfg.lanes % 256 ||
      (cfg.jobs_per_deal != 5 && cfg.jobs_per_deal != 20) || cfg.seed_mode < 0 || cfg.seed_mode > 3 ||
      (cfg.team != 1 && cfg.team != kWideTeam) || (cfg.waves_per_simd != 0 && cfg.waves_per_simd != 3) ||
      cfg.split_min_nodes != 0) {
    ddx_set_error("ddx_stream_open: bad config (lanes multiple of 256, TT_LOG2 in [9, %d], "
                  "jobs_per_deal 5 or 20; the stream kernel does not split searches)", kMaxTTLog2);
    return DDX_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg.device) {
    ddx_set_error("ddx_stream_open: no HIP device %d", cfg.device);
    return DDX_E_NODEV;
  }
  HIPCHECK(hipSetDevice(cfg.device));
  ddx_stream* s = new (std::nothrow) ddx_stream;
  if (!s) return DDX_E_NOMEM;
  s->cfg = cfg;
  s->ring = ring_slots;
  s->work_ring = 1;
  while (s->work_ring < 4 * s->ring) s->work_ring <<= 1;
  int rc = lanes_alloc(s->lanes, cfg.lanes / cfg.team, cfg.tt_log2);
  if (rc != DDX_OK) {
    delete s;
    return rc;
  }
  auto fail = [&](const char* what) {
    ddx_set_error("ddx_stream_open: cannot allocate %s", what);
    stream_free(s);
    delete s;
    return DDX_E_NOMEM;
  };
  if (host_alloc(&s->hands, s->ring * 32) || host_alloc(&s->seeds, s->ring * 20) ||
      host_alloc(&s->out, s->ring * 20) || host_alloc(&s->work, s->work_ring * 4) ||
      host_alloc(&s->comp, s->ring * 4) || host_alloc(&s->ctl, sizeof(StreamCtl)))
    return fail("pinned ring");
  if (hipMalloc(&s->counters, 64) != hipSuccess ||
      hipMalloc(&s->deal_done, ((s->ring * 4 + 15) / 16) * 16) != hipSuccess)
    return fail("device counters");
  // Non-blocking: the persistent kernel must not be waited on by work on the null stream
  // (other libraries' synchronous copies in the same process would otherwise never return).
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return fail("a HIP stream");
  if ((rc = launch_fill(s->counters, 64, 0, s->stream)) != DDX_OK ||
      (rc = launch_fill(s->deal_done, ((s->ring * 4 + 15) / 16) * 16, 0, s->stream)) != DDX_OK) {
    stream_free(s);
    delete s;
    return rc;
  }
  s->free_slots.reserve(s->ring);
  for (long long i = s->ring - 1; i >= 0; i--) s->free_slots.push_back((u32)i);
  s->slot_id.assign(s->ring, -1);
  StreamArgs a;
  a.lanes = lanes_view(s->lanes);
  a.hands = dev_ptr(s->hands);
  a.seeds = dev_ptr(s->seeds);
  a.out = dev_ptr(s->out);
  a.work = dev_ptr(s->work);
  a.comp = dev_ptr(s->comp);
  a.ctl = dev_ptr(s->ctl);
  a.counters = s->counters;
  a.deal_done = s->deal_done;
  a.ring = s->ring;
  a.work_ring = s->work_ring;
  a.jobs_per_deal = cfg.jobs_per_deal;
  a.seed_mode = cfg.seed_mode;
  if (!a.hands || !a.seeds || !a.out || !a.work || !a.comp || !a.ctl) return fail("mapped views");
  s->t0 = std::chrono::steady_clock::now();
  const dim3 grid((unsigned)(cfg.lanes / 256));
  if (cfg.waves_per_simd == 3) {
    if (cfg.team == 16) hipLaunchKernelGGL(stream_kernel_w3<kWideTeam>, grid, dim3(256), 0, s->stream, a);
    else hipLaunchKernelGGL(stream_kernel_w3<1>, grid, dim3(256), 0, s->stream, a);
  } else if (cfg.team == 16) {
    hipLaunchKernelGGL(stream_kernel<kWideTeam>, grid, dim3(256), 0, s->stream, a);
  } else {
    hipLaunchKernelGGL(stream_kernel<1>, grid, dim3(256), 0, s->stream, a);
  }
  if (hipGetLastError() != hipSuccess) {
    ddx_set_error("ddx_stream_open: kernel launch failed");
    stream_free(s);
    delete s;
    return DDX_E_HIP;
  }
  s->launched = true;
  *out = s;
  return DDX_OK;
}

extern "C" int ddx_stream_push(ddx_stream* s, const uint64_t hands[4], const uint8_t* seeds20,
                               int64_t id) {
  if (!s || !hands || !ddx_check_deal(hands)) {
    ddx_set_error("ddx_stream_push: bad deal");
    return DDX_E_ARG;
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (s->finished) {
    ddx_set_error("ddx_stream_push: stream already finished");
    return DDX_E_ARG;
  }
  if (s->free_slots.empty()) return DDX_E_FULL;
  const u32 slot = s->free_slots.back();
  s->free_slots.pop_back();
  memcpy(s->hands + 4 * (size_t)slot, hands, 32);
  if (seeds20) memcpy(s->seeds + 20 * (size_t)slot, seeds20, 20);
  memset(s->out + 20 * (size_t)slot, 0xFF, 20);
  s->slot_id[slot] = id;
  s->work[s->produced % s->work_ring] = slot;
  s->produced++;
  __atomic_store_n(&s->ctl->produced, s->produced, __ATOMIC_RELEASE);
  return DDX_OK;
}

extern "C" int ddx_stream_poll(ddx_stream* s, int max, int64_t* ids, uint8_t* tables, int* got) {
  if (!s || max < 0 || (max && (!ids || !tables)) || !got) return DDX_E_ARG;
  int n = 0;
  while (n < max) {
    u32* e = &s->comp[s->tail % s->ring];
    const u32 v = __atomic_load_n(e, __ATOMIC_ACQUIRE);
    if (!v) break;
    const u32 slot = v - 1;
    __atomic_store_n(e, 0u, __ATOMIC_RELAXED);
    s->tail++;
    memcpy(tables + 20 * (size_t)n, s->out + 20 * (size_t)slot, 20);
    {
      std::lock_guard<std::mutex> g(s->mu);
      ids[n] = s->slot_id[slot];
      s->free_slots.push_back(slot);
      s->consumed++;
    }
    n++;
  }
  *got = n;
  return DDX_OK;
}

extern "C" int64_t ddx_stream_in_flight(ddx_stream* s) {
  if (!s) return 0;
  std::lock_guard<std::mutex> g(s->mu);
  return s->produced - s->consumed;
}

extern "C" int ddx_stream_finish(ddx_stream* s) {
  if (!s) return DDX_E_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  s->finished = true;
  __atomic_store_n(&s->ctl->eof, 1, __ATOMIC_RELEASE);
  return DDX_OK;
}

extern "C" int ddx_stream_close(ddx_stream* s, ddx_stats* stats) {
  if (!s) return DDX_E_ARG;
  if (stats) memset(stats, 0, sizeof *stats);
  (void)hipSetDevice(s->cfg.device);
  int rc = DDX_OK;
  if (s->launched) {
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (!s->finished || s->consumed < s->produced) __atomic_store_n(&s->ctl->stop, 1, __ATOMIC_RELEASE);
      __atomic_store_n(&s->ctl->eof, 1, __ATOMIC_RELEASE);
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) {
      ddx_set_error("ddx_stream_close: the stream kernel failed");
      rc = DDX_E_HIP;
    }
    if (stats && rc == DDX_OK) {
      unsigned long long c[8];
      if (hipMemcpy(c, s->counters, sizeof c, hipMemcpyDeviceToHost) == hipSuccess) {
        stats->nodes = c[2];
        stats->tt_probes = c[3];
        stats->tt_hits = c[4];
        stats->jobs_done = c[5];
      }
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
    }
  }
  stream_free(s);
  delete s;
  return rc;
}
