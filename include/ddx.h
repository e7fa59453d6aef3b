/* ddx: exact double-dummy solver for the AMD Instinct MI355X (gfx950) -- C ABI.
 *
 * The reference (hsperr/dds_gpu) exposes its solver as a header-only C++ API and three
 * command lines; it has no FFI. These entry points are what a binding would need to replace
 * that surface: plain pointers and sizes, status codes, no C++ or torch types. The library
 * never calls exit(); every failure returns a negative status and leaves a message for
 * ddx_last_error(). Host buffers belong to the caller; device TT and lane state belong to
 * the handle that allocated them. A handle is used by one host thread at a time.
 *
 * Encoding (reference src/dd.h:13-14): card = suit*16 + rank as a bit of a uint64_t hand;
 * suits 0..3 = S,H,D,C; ranks 0..12 = 2..A; seats 0..3 = N,E,S,W; strains 0..4 =
 * C,D,H,S,NT. A table is 20 uint8 tricks [strain][declarer N,E,S,W]; declarer E/W values are
 * 13 - NS tricks, also for endings with fewer than 13 cards (reference src/dd.h:922,937).
 */
#ifndef DDX_H
#define DDX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DDX_OK = 0,
  DDX_E_ARG = -1,     /* invalid argument (sizes, TT_LOG2 < 9, bad deal) */
  DDX_E_NOMEM = -2,   /* host or device allocation failed */
  DDX_E_HIP = -3,     /* HIP runtime error */
  DDX_E_NODEV = -4,   /* no GPU, or the GPU code is missing */
  DDX_E_STOPPED = -5, /* ddx_gpu_request_stop() ended the run; unsolved results are 0xFF */
  DDX_E_IO = -6,      /* file format or I/O error */
  DDX_E_FULL = 1      /* not an error: the stream's ring is full, poll and push again */
};

/* Seed policies for the trick-count search of each lead. */
enum {
  DDX_SEED_PLAIN = 0, /* binary search, 7 then the previous lead's answer (dd_solve_strain) */
  DDX_SEED_STEP = 1,  /* step from seeds[] one trick at a time (dd_solve_strain_guess) */
  DDX_SEED_CERT = 2,  /* probe seed, then seed +- 1, then binary search (learned seeds) */
  DDX_SEED_CERTP = 3  /* DDX_SEED_CERT seeded with the previous lead's answer (no seeds[]) */
};

/* Counters of one solve call. */
typedef struct ddx_stats {
  uint64_t nodes;        /* search nodes entered (reference DDCtx::nodes) */
  uint64_t tt_probes;    /* trick-start TT probes (reference DDCtx::st_probe) */
  uint64_t tt_hits;      /* probes that decided the node (reference DDCtx::st_hit) */
  uint64_t jobs_done;    /* jobs finished */
  double seconds;        /* wall time of the solve, device work included */
  uint64_t tt_entry_reads; /* 16-byte TT entries read by probes */
  uint64_t tt_stores;    /* TT stores (each rewrites a 64-byte tag line and one entry) */
  uint64_t tt_line_cache_hits; /* tag lines served by the LDS line cache (config.lds_lines) */
  uint64_t splits;       /* search nodes split over idle lanes (config.split_min_nodes) */
  uint64_t tasks_taken;  /* published child searches that idle lanes took */
} ddx_stats;

/* Message of the last failure on this thread ("" if none). */
const char* ddx_last_error(void);

/* Library version string; it names the TT entry format the library was built with (exact
 * 16-byte entries, or 12-byte signature entries: the reference's -DDD_TT_SIG build). */
const char* ddx_version(void);

/* ---- deals and files ------------------------------------------------------------------ */

/* Deals [first, first + n) of the seeded stream `seed`, k cards per hand (1..13), into
 * hands[4*n] (N,E,S,W). Replaces the reference's random_deals / random_deal
 * (gpu/gpu_main.cu:154-163, gpu/stream_main.cu:116-122), whose std::shuffle is not portable. */
int ddx_random_deals(uint64_t seed, int64_t first, int64_t n, int k, uint64_t* hands);

/* Rows [start, start + n) of a Pgx DDS-table .npy file: hands[4*n] and, when tables is not
 * NULL, the stored tables[20*n]. n < 0 reads to the end; *n_read gets the count.
 * Replaces pgx_load / pgx_hands / pgx_tricks (src/npy.h:15-63). */
int ddx_pgx_read(const char* path, int64_t start, int64_t n, uint64_t* hands, uint8_t* tables,
                 int64_t* n_read);
/* Number of rows of a Pgx file. */
int ddx_pgx_rows(const char* path, int64_t* rows);
/* Writes deals and tables (tables may be NULL: zeros) as a Pgx file. */
int ddx_pgx_write(const char* path, const uint64_t* hands, const uint8_t* tables, int64_t n);

/* ---- CPU solver ------------------------------------------------------------------------ */

/* One CPU solver context with its own TT of 2^tt_log2 buckets (tt_log2 >= 9), like the
 * reference's caller-owned DDCtx + TT (cpu/cpu_main.cpp:43-51). Single-threaded. */
typedef struct ddx_cpu ddx_cpu;
int ddx_cpu_create(int tt_log2, ddx_cpu** out);
void ddx_cpu_destroy(ddx_cpu* ctx);

/* Tricks for declarers N,E,S,W in one strain; the four leads share the TT
 * (reference dd_solve_strain, src/dd.h:926-939). */
int ddx_cpu_solve_strain(ddx_cpu* ctx, const uint64_t hands[4], int strain, int out4[4]);
/* Same, stepping from a guess of NS tricks per declarer
 * (reference dd_solve_strain_guess, src/dd.h:915-924). */
int ddx_cpu_solve_strain_guess(ddx_cpu* ctx, const uint64_t hands[4], int strain,
                               const int ns4[4], int out4[4]);
/* NS tricks with `leader` on lead, binary search from `guess`, in the strain and deal of the
 * last solve_strain call (reference dd_ns_tricks, src/dd.h:877-891). */
int ddx_cpu_ns_tricks(ddx_cpu* ctx, const uint64_t hands[4], int strain, int leader, int guess,
                      int* ns);
/* Search nodes entered by this context so far (reference DDCtx::nodes). */
uint64_t ddx_cpu_nodes(const ddx_cpu* ctx);

/* Jobs of n deals on `threads` host threads, each with its own TT of 2^tt_log2 buckets and
 * an atomic job counter (reference dd_cpu, cpu/cpu_main.cpp:42-74). jobs_per_deal is 5 (one
 * job per strain, leads share the TT) or 20 (one per lead). seeds: n*20 NS-trick seeds for
 * DDX_SEED_STEP / DDX_SEED_CERT, else NULL. With threads == 1 the jobs run in index order
 * on one context and job_nodes (if not NULL, n*jobs_per_deal entries) gets the nodes of
 * each job. */
int ddx_cpu_solve_batch(const uint64_t* hands, int64_t n, int threads, int tt_log2,
                        int jobs_per_deal, int seed_mode, const uint8_t* seeds, uint8_t* out,
                        uint64_t* job_nodes, ddx_stats* stats);

/* ---- GPU solver (MI355X) --------------------------------------------------------------- */

typedef struct ddx_gpu_config {
  int device;           /* HIP device ordinal */
  int64_t lanes;        /* concurrent solver lanes, each with its own TT (default 65536) */
  int tt_log2;          /* buckets per lane TT, >= 9 (default 10) */
  int jobs_per_deal;    /* 5 or 20 */
  int schedule;         /* stage schedule of the batch kernel's wavefront loop, see DDX_SCHED_*
                           (the stream kernel has a loop of its own) */
  int seed_mode;        /* DDX_SEED_* */
  int block;            /* threads per block (multiple of 64; default 256) */
  int profile;          /* 1: collect per-stage counters (slower) */
  int team;             /* lanes per search: 1 (one search per lane, as the reference) or 16
                           (a team of 16 lanes per search: lanes/16 searches, each with its own
                           TT, about five times faster per search under full load) */
  int graph;            /* 1: launch each batch as a captured HIP graph (clear + solve kernels),
                           re-captured only when the buffers, size or stream change */
  int waves_per_simd;   /* 0: the compiler's register allocation (2 waves per SIMD); 3 or 4: a
                           kernel held to that many waves per SIMD (fewer VGPRs, some spilled to
                           scratch); split schedule only, no profile */
  int lds_lines;        /* 0, or 3: each lane keeps its last 3 TT tag lines in LDS (write-through;
                           one lane per search, split schedule, no profile, waves_per_simd 0/3) */
  int split_min_nodes;  /* 0: off. Otherwise, once the job queue is empty, a search node whose
                           children searched so far took at least this many nodes publishes its
                           untried moves as tasks for idle lanes (young-brothers-wait; exact).
                           One lane per search, split schedule, no profile, waves_per_simd 0/3,
                           no LDS cache. No counterpart in the reference, whose tail runs on the
                           lanes that hold it (README.md:189-190). */
} ddx_gpu_config;

enum {
  DDX_SCHED_PRIORITY = 0, /* one stage per round: RET > NEXT > ROOT > JOB > ENTER (reference) */
  DDX_SCHED_CHAIN = 1,    /* cheap stages run back to back per lane, then ENTER for all */
  DDX_SCHED_SPLIT = 2,    /* as CHAIN, with follow nodes entered inside the chain, so that
                             trick-start nodes are entered with the whole wave gathered */
  DDX_SCHED_THREAD = 3    /* no wave schedule: every lane runs its own next stage each round, as
                             the reference's per-thread solve_kernel (gpu/gpu_main.cu:33-50) runs
                             dd_solve_strain in each thread (no profile) */
};

typedef struct ddx_gpu ddx_gpu;

/* Default configuration (device 0, 65536 lanes, TT_LOG2 10, 5 jobs per deal, team 1). */
void ddx_gpu_default_config(ddx_gpu_config* cfg);
/* Allocates the lanes' TTs and stacks on the device; replaces the reference's per-run
 * cudaMalloc of TT and DDWave state (gpu/gpu_main.cu:205-252). */
int ddx_gpu_open(const ddx_gpu_config* cfg, ddx_gpu** out);
void ddx_gpu_close(ddx_gpu* h);
/* Device bytes held by the handle. */
uint64_t ddx_gpu_device_bytes(const ddx_gpu* h);

/* Solves n deals: host hands[4n] in, host out[20n] (0xFF = unsolved after a stop). seeds:
 * host n*20 NS-trick seeds for DDX_SEED_STEP/CERT, else NULL. Replaces the reference's
 * wave_kernel batch run (gpu/gpu_main.cu:61-152, 231-274). */
int ddx_gpu_solve(ddx_gpu* h, const uint64_t* hands, int64_t n, const uint8_t* seeds,
                  uint8_t* out, ddx_stats* stats);
/* Same with device-resident buffers (HIP device pointers), on HIP stream `stream` (NULL = a
 * non-blocking stream of the handle); returns after the work is queued. Then ddx_gpu_wait(). */
int ddx_gpu_solve_async(ddx_gpu* h, const uint64_t* d_hands, int64_t n, const uint8_t* d_seeds,
                        uint8_t* d_out, void* stream);
/* Waits for the last async solve and fills stats. */
int ddx_gpu_wait(ddx_gpu* h, ddx_stats* stats);
/* Asks a running solve to stop soon (reference TIME_LIMIT stop flag, gpu/gpu_main.cu:90,264);
 * callable from another thread. */
void ddx_gpu_request_stop(ddx_gpu* h);
/* Jobs finished so far by the running or last solve (callable while it runs). */
int ddx_gpu_progress(ddx_gpu* h, uint64_t* jobs_done);
/* The same, plus the nodes its searches spent up to their last finished job (for a live node
 * rate over a window of a running solve). No counterpart in the reference, which reports
 * nodes only per finished run. */
int ddx_gpu_progress_nodes(ddx_gpu* h, uint64_t* jobs_done, uint64_t* nodes);
/* Duration in ms of the last solve kernel, from HIP events on its stream. */
double ddx_gpu_last_kernel_ms(const ddx_gpu* h);
/* Per-stage counters of the last solves (config.profile = 1), per wave64: out32[0..7] rounds by
 * kind (0 JOB, 1 ROOT, 2 ENTER -- trick-start ENTER under SPLIT, every ENTER under CHAIN --,
 * 3 NEXT, 4 RET, 6 cheap-stage chains, with the follow nodes' ENTER under SPLIT; 7 unused),
 * out32[8..15] active lanes summed over those rounds,
 * out32[16..22] where ENTER decided nodes (trivial, last trick, TT, quick tricks, later tricks,
 * 2nd-hand quick tricks, moves), out32[24..31] shader-clock cycles by kind. reset != 0 clears. */
int ddx_gpu_profile(ddx_gpu* h, uint64_t* out32, int reset);

/* ---- trick-count net (MI355X) ------------------------------------------------------------ */

/* An MLP predicting NS tricks per (deal, strain, leader): 217 inputs (card owners, strain,
 * leader) -> 256 ReLU -> 256 ReLU -> 14 classes, BF16 weights and activations with FP32
 * accumulation on the matrix cores, FP32 master weights and Adam. Its predictions are the seeds
 * of DDX_SEED_CERT. No counterpart in the reference (its dd_solve_strain_guess takes a guess,
 * src/dd.h:915-924, but nothing produces one). */
typedef struct ddx_mlp ddx_mlp;
int ddx_mlp_create(int device, uint64_t init_seed, ddx_mlp** out);
void ddx_mlp_destroy(ddx_mlp* h);
/* Length of the FP32 parameter vector (W1t[256][224], b1[256], W2t[256][256], b2[256],
 * W3t[16][256], b3[16]; padding entries stay zero). */
int64_t ddx_mlp_param_count(void);
/* get != 0: copy the parameters out to buf; get == 0: load them from buf. */
int ddx_mlp_params(ddx_mlp* h, float* buf, int get);
/* Gradients of the mean cross-entropy over the 20 samples of each of n deals (host buffers:
 * hands[4n], tables[20n] as declarer tricks); loss_acc[0] = mean loss, [1] = accuracy. */
int ddx_mlp_grads(ddx_mlp* h, const uint64_t* hands, const uint8_t* tables, int64_t n, float* grads,
                  float* loss_acc);
/* Adam training, `epochs` passes over n deals in shuffled batches of batch_deals; loss_acc gets
 * the last epoch's mean loss and accuracy. */
int ddx_mlp_train(ddx_mlp* h, const uint64_t* hands, const uint8_t* tables, int64_t n, int epochs,
                  int batch_deals, float lr, uint64_t shuffle_seed, float* loss_acc);
/* Seeds (most likely NS tricks) for n deals: device buffers d_hands[4n] -> d_seeds[20n], ordered
 * on HIP stream `stream` (NULL: synchronous). */
int ddx_mlp_predict(ddx_mlp* h, const uint64_t* d_hands, int64_t n, uint8_t* d_seeds, void* stream);
/* Same with host buffers. */
int ddx_mlp_predict_host(ddx_mlp* h, const uint64_t* hands, int64_t n, uint8_t* seeds);
/* ddx_mlp_predict that allocates nothing, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops). No counterpart in the reference. The activations live in
 * a fixed device scratch of the library (13.1 MB per GPU, held until the process ends), so n is
 * at most 256 deals (DDX_E_ARG beyond). Kernels run on `stream` (NULL: the null stream, and the
 * call returns when the seeds are written). Every call on one GPU shares that scratch: calls
 * that overlap on a GPU race, and the caller orders them. */
int ddx_mlp_predict_noalloc(const ddx_mlp* h, const uint64_t* d_hands, int64_t n, uint8_t* d_seeds,
                            void* stream);
int ddx_mlp_save(ddx_mlp* h, const char* path);
int ddx_mlp_load(ddx_mlp* h, const char* path);

/* ---- streaming solver (MI355X) --------------------------------------------------------- */

/* One persistent kernel fed through `ring_slots` pinned host slots; each deal's table is
 * published as soon as it is solved, in any order. Replaces dd_stream's kernel and ring
 * (gpu/stream_main.cu:38-114, 146-164). cfg->jobs_per_deal 20 keeps more lanes busy than 5
 * when ring_slots * jobs_per_deal is near the lane count. */
typedef struct ddx_stream ddx_stream;
int ddx_stream_open(const ddx_gpu_config* cfg, int ring_slots, ddx_stream** out);
/* Queues one deal (hands N,E,S,W; seeds20 = NS-trick seeds [strain][declarer] for the seeded
 * modes, else NULL) under a caller id. Returns DDX_OK, or DDX_E_FULL when every slot is in
 * use (poll, then push again). Safe to call from one producer thread while another polls. */
int ddx_stream_push(ddx_stream* s, const uint64_t hands[4], const uint8_t* seeds20, int64_t id);
/* Takes up to `max` finished deals: ids[i] and tables[20*i] (the 28-byte output record of
 * dd_stream, gpu/stream_main.cu:4-8, is id + table). Never blocks; *got = how many. */
int ddx_stream_poll(ddx_stream* s, int max, int64_t* ids, uint8_t* tables, int* got);
/* Deals pushed and not yet polled. */
int64_t ddx_stream_in_flight(ddx_stream* s);
/* No more pushes: the kernel ends once every queued deal is solved. */
int ddx_stream_finish(ddx_stream* s);
/* Stops the kernel (unsolved deals are dropped unless ddx_stream_finish was called and all
 * results were polled), waits for it and frees everything. */
int ddx_stream_close(ddx_stream* s, ddx_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* DDX_H */
