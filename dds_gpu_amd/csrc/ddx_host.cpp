// Host side of the ddx C ABI: errors, deals, Pgx files and the CPU solver.
//
// The CPU solver runs the same machine as the GPU kernel (include/ddx/machine.hpp) with
// its frame stack in ordinary memory and one TT per thread, so CPU and GPU results and,
// for a fixed job order, node counts are identical.
#include <atomic>
#include <chrono>
#include <new>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <thread>
#include <vector>

#include "ddx.h"
#include "ddx/deal.hpp"
#include "ddx/machine.hpp"
#include "ddx/pgx.hpp"
#include "ddx_internal.h"

using namespace ddx;

static thread_local std::string g_err;

void ddx_set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

extern "C" const char* ddx_last_error(void) { return g_err.c_str(); }
extern "C" const char* ddx_version(void) {
  return DDX_TT_SIG      ? "ddx 0.1 (gfx950, 12-byte signature TT entries)"
         : DDX_SHARED_TT ? "ddx 0.1 (gfx950, one shared lock-free TT, 32-byte checksummed entries)"
                         : "ddx 0.1 (gfx950, exact TT entries)";
}

extern "C" int ddx_random_deals(uint64_t seed, int64_t first, int64_t n, int k, uint64_t* hands) {
  if (n < 0 || first < 0 || k < 1 || k > 13 || (!hands && n)) {
    ddx_set_error("ddx_random_deals: bad arguments (n=%lld k=%d)", (long long)n, k);
    return DDX_E_ARG;
  }
  for (int64_t i = 0; i < n; i++) random_deal(seed, first + i, k, hands + 4 * i);
  return DDX_OK;
}

extern "C" int ddx_pgx_rows(const char* path, int64_t* rows) {
  PgxFile f;
  std::string err;
  if (!pgx_read(path, 0, 0, &f, &err)) {
    ddx_set_error("%s", err.c_str());
    return DDX_E_IO;
  }
  // pgx_read with count 0 validated the header; read the total from the file size.
  FILE* fp = fopen(path, "rb");
  if (!fp) return DDX_E_IO;
  unsigned char pre[12];
  size_t got = fread(pre, 1, 12, fp);
  long hlen = pre[8] | (pre[9] << 8), off = 10;
  if (got == 12 && pre[6] >= 2) {
    hlen |= ((long)pre[10] << 16) | ((long)pre[11] << 24);
    off = 12;
  }
  fseek(fp, 0, SEEK_END);
  long size = ftell(fp);
  fclose(fp);
  *rows = (size - off - hlen) / 32;
  return DDX_OK;
}

extern "C" int ddx_pgx_read(const char* path, int64_t start, int64_t n, uint64_t* hands,
                            uint8_t* tables, int64_t* n_read) {
  PgxFile f;
  std::string err;
  if (!pgx_read(path, start, n, &f, &err)) {
    ddx_set_error("%s: %s", path, err.c_str());
    return DDX_E_IO;
  }
  for (long long i = 0; i < f.n; i++) {
    if (hands) pgx_hands(f, i, hands + 4 * i);
    if (tables)
      for (int st = 0; st < 5; st++)
        for (int seat = 0; seat < 4; seat++)
          tables[i * 20 + st * 4 + seat] = (uint8_t)pgx_tricks(f, i, seat, st);
  }
  if (n_read) *n_read = f.n;
  return DDX_OK;
}

extern "C" int ddx_pgx_write(const char* path, const uint64_t* hands, const uint8_t* tables,
                             int64_t n) {
  if (!pgx_write(path, hands, tables, n)) {
    ddx_set_error("cannot write %s", path);
    return DDX_E_IO;
  }
  return DDX_OK;
}

// Checks that a deal is well formed: disjoint hands of equal size, only real cards.
bool ddx_check_deal(const uint64_t* h) {
  uint64_t all = 0;
  int k = popc(h[0]);
  if (k < 1 || k > 13) return false;
  for (int i = 0; i < 4; i++) {
    if ((h[i] & ~kAllSuits) || (all & h[i]) || popc(h[i]) != k) return false;
    all |= h[i];
  }
  return true;
}

// ---------------------------------------------------------------------------------------
// CPU solver context.
struct ddx_cpu {
  std::vector<u32> tags;
  std::vector<Entry> ents;
  Lane lane;
  ArrayStack stack;
  uint64_t last_hands[4];
  int last_strain;
};

static bool make_tt(int tt_log2, std::vector<u32>& tags, std::vector<Entry>& ents, TTView& v) {
  size_t buckets = (size_t)1 << tt_log2;
  try {
    tags.assign(buckets * kWays + 16, 0);
    ents.assign(buckets * (kBucketEntryBytes / sizeof(Entry)) + 4, Entry{0, 0});
  } catch (const std::bad_alloc&) {
    return false;
  }
  // 64-byte align the tag lines and the entries.
  uintptr_t p = (uintptr_t)tags.data(), q = (uintptr_t)ents.data();
  v.tags = (u32*)((p + 63) & ~(uintptr_t)63);
  v.ents = (Entry*)((q + 63) & ~(uintptr_t)63);
  v.mask = (u32)(buckets - 1);
  return true;
}

extern "C" int ddx_cpu_create(int tt_log2, ddx_cpu** out) {
  if (!out || tt_log2 < kMinLog2 || tt_log2 > 30) {
    ddx_set_error("ddx_cpu_create: TT_LOG2 must be in [%d, 30], got %d", kMinLog2, tt_log2);
    return DDX_E_ARG;
  }
  ddx_cpu* c = new (std::nothrow) ddx_cpu;
  if (!c) return DDX_E_NOMEM;
  TTView v;
  if (!make_tt(tt_log2, c->tags, c->ents, v)) {
    delete c;
    ddx_set_error("ddx_cpu_create: cannot allocate the TT");
    return DDX_E_NOMEM;
  }
  lane_init(c->lane, v);
  memset(c->last_hands, 0, sizeof c->last_hands);
  c->last_strain = -1;
  *out = c;
  return DDX_OK;
}

extern "C" void ddx_cpu_destroy(ddx_cpu* ctx) { delete ctx; }
extern "C" uint64_t ddx_cpu_nodes(const ddx_cpu* ctx) { return ctx ? ctx->lane.c.nodes : 0; }

// Runs the lane until it asks for another job.
template <class Stack>
static void run_until_job(Lane& L, const Batch& B, Stack& st) {
  auto none = [&]() -> long long { return B.jobs; };
  while (L.stage != ST_JOB && L.stage != ST_EXIT) lane_step(L, B, st, none);
}

static int solve_strain_common(ddx_cpu* ctx, const uint64_t hands[4], int strain,
                               const int* ns4, int out4[4]) {
  if (!ctx || !hands || !out4 || strain < 0 || strain > 4 || !ddx_check_deal(hands)) {
    ddx_set_error("ddx_cpu_solve_strain: bad arguments");
    return DDX_E_ARG;
  }
  uint8_t out[20], seeds[20];
  memset(out, 0, sizeof out);
  memset(seeds, 0, sizeof seeds);
  if (ns4)
    for (int d = 0; d < 4; d++) seeds[strain * 4 + d] = (uint8_t)(ns4[d] < 0 ? 0 : ns4[d] > 13 ? 13 : ns4[d]);
  Batch B{hands, out, ns4 ? seeds : nullptr, 5, 5, ns4 ? SEED_STEP : SEED_PLAIN};
  Lane& L = ctx->lane;
  lane_load_job(L, B, strain);
  run_until_job(L, B, ctx->stack);
  for (int d = 0; d < 4; d++) out4[d] = out[strain * 4 + d];
  memcpy(ctx->last_hands, hands, sizeof ctx->last_hands);
  ctx->last_strain = strain;
  return DDX_OK;
}

extern "C" int ddx_cpu_solve_strain(ddx_cpu* ctx, const uint64_t hands[4], int strain, int out4[4]) {
  return solve_strain_common(ctx, hands, strain, nullptr, out4);
}

extern "C" int ddx_cpu_solve_strain_guess(ddx_cpu* ctx, const uint64_t hands[4], int strain,
                                          const int ns4[4], int out4[4]) {
  if (!ns4) return DDX_E_ARG;
  return solve_strain_common(ctx, hands, strain, ns4, out4);
}

extern "C" int ddx_cpu_ns_tricks(ddx_cpu* ctx, const uint64_t hands[4], int strain, int leader,
                                 int guess, int* ns) {
  if (!ctx || !hands || !ns || strain < 0 || strain > 4 || leader < 0 || leader > 3 ||
      !ddx_check_deal(hands)) {
    ddx_set_error("ddx_cpu_ns_tricks: bad arguments");
    return DDX_E_ARG;
  }
  Lane& L = ctx->lane;
  // A new deal or strain needs a new TT generation, or entries of the old one could match.
  if (memcmp(hands, ctx->last_hands, sizeof ctx->last_hands) != 0 || strain != ctx->last_strain) {
    next_generation(L.c);
    memcpy(ctx->last_hands, hands, sizeof ctx->last_hands);
    ctx->last_strain = strain;
  }
  uint8_t out[20] = {0};
  Batch B{hands, out, nullptr, 5, 5, SEED_PLAIN};
  for (int i = 0; i < 4; i++) L.p.hand[i] = hands[i];
  L.p.trump = strain_trump(strain);
  L.deal = 0;
  L.strain = strain;
  L.decl = (leader + 3) & 3;
  L.last_decl = L.decl;
  L.prev_ns = guess;
  L.stage = ST_ROOT;
  run_until_job(L, B, ctx->stack);
  int v = out[strain * 4 + L.last_decl];
  *ns = (L.last_decl & 1) == 0 ? v : 13 - v;
  return DDX_OK;
}

extern "C" int ddx_cpu_solve_batch(const uint64_t* hands, int64_t n, int threads, int tt_log2,
                                   int jobs_per_deal, int seed_mode, const uint8_t* seeds,
                                   uint8_t* out, uint64_t* job_nodes, ddx_stats* stats) {
  if (!hands || !out || n < 0 || threads < 1 || tt_log2 < kMinLog2 || tt_log2 > 30 ||
      (jobs_per_deal != 5 && jobs_per_deal != 20) || seed_mode < 0 || seed_mode > 3 ||
      ((seed_mode == SEED_STEP || seed_mode == SEED_CERT) && !seeds)) {
    ddx_set_error("ddx_cpu_solve_batch: bad arguments");
    return DDX_E_ARG;
  }
  for (int64_t i = 0; i < n; i++)
    if (!ddx_check_deal(hands + 4 * i)) {
      ddx_set_error("ddx_cpu_solve_batch: deal %lld is not a valid deal", (long long)i);
      return DDX_E_ARG;
    }
  const long long jobs = (long long)n * jobs_per_deal;
  Batch B{hands, out, seeds, jobs, jobs_per_deal, seed_mode};
  std::atomic<long long> next{0};
  std::atomic<uint64_t> nodes{0}, probes{0}, hits{0}, ent_reads{0}, stores{0};
  std::atomic<int> failed{0};
  auto t0 = std::chrono::steady_clock::now();
  // The shared-TT build gives all threads one TT as large as their private TTs together
  // (threads rounded up to a power of two; reference cpu/cpu_main.cpp:44-51).
  int shared_log2 = tt_log2;
  while (DDX_SHARED_TT && shared_log2 < 30 && (1 << (shared_log2 - tt_log2)) < threads) shared_log2++;
  std::vector<u32> shared_tags;
  std::vector<Entry> shared_ents;
  TTView shared_view{};
  if (DDX_SHARED_TT && !make_tt(shared_log2, shared_tags, shared_ents, shared_view)) {
    ddx_set_error("ddx_cpu_solve_batch: cannot allocate a TT of 2^%d buckets", shared_log2);
    return DDX_E_NOMEM;
  }
  auto worker = [&](bool record) {
    std::vector<u32> tags;
    std::vector<Entry> ents;
    TTView v = shared_view;
    if (!DDX_SHARED_TT && !make_tt(tt_log2, tags, ents, v)) {
      failed = 1;
      return;
    }
    Lane L;
    lane_init(L, v);
    ArrayStack* st = new ArrayStack;
    long long cur = -1;
    uint64_t base = 0;
    auto take = [&]() -> long long {
      if (record && cur >= 0 && cur < jobs) job_nodes[cur] = L.c.nodes - base;
      cur = next.fetch_add(1);
      base = L.c.nodes;
      return cur;
    };
    while (L.stage != ST_EXIT) lane_step(L, B, *st, take);
    delete st;
    nodes += L.c.nodes;
    probes += L.c.probes;
    hits += L.c.hits;
    ent_reads += L.c.ent_reads;
    stores += L.c.stores;
  };
  if (threads == 1) {
    worker(job_nodes != nullptr);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(worker, false);
    for (auto& th : pool) th.join();
  }
  if (failed) {
    ddx_set_error("ddx_cpu_solve_batch: cannot allocate a TT of 2^%d buckets", tt_log2);
    return DDX_E_NOMEM;
  }
  if (stats) {
    stats->nodes = nodes;
    stats->tt_probes = probes;
    stats->tt_hits = hits;
    stats->jobs_done = (uint64_t)jobs;
    stats->tt_entry_reads = ent_reads;
    stats->tt_stores = stores;
    stats->tt_line_cache_hits = 0;
    stats->splits = stats->tasks_taken = 0;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return DDX_OK;
}
