// dd_gpu: batch solve on the MI355X and check the answers.
//   dd_gpu K DEALS [TT_LOG2] [CHECK]      random K-card endings; the first CHECK deals are
//                                         re-solved on the CPU (all cores) and compared
//   dd_gpu FILE.npy START DEALS [TT_LOG2] full deals from a Pgx file, checked against it
// Same argv, output lines and exit codes as the reference's gpu/gpu_main.cu:165-342
// (0 = every checked answer equal; 1 = a mismatch, load failure or GPU error; 2 = usage or
// TT_LOG2 < 9). Environment:
//   THREADS     solver lanes (default 32768; clamped to the jobs, rounded up to 256)
//   TIME_LIMIT  seconds; then the lanes stop and unfinished answers are skipped in the check
//   SCHEDULE    0 priority (the reference's), 1 chain (CHEAP_LOOP=1), 2 split (default)
//   PER_LEAD=1  one job per lead (20 per deal) instead of per strain
//   SEEDS=FILE  n x 20 uint8 NS-trick seeds for the seeded search (DDX_SEED_CERT)
//   MODEL=FILE  a trained trick-count net (dd_train): its predictions seed DDX_SEED_CERT; the
//               prediction time is counted in the reported time
//   TEAM=16     sixteen lanes per search (THREADS / 16 searches)
//   WAVES=3|4   hold the kernel to that many waves per SIMD (split schedule); the tuned
//               configuration for full deals is THREADS=196608 WAVES=3 with TT_LOG2 11
//   LDS=3       keep each lane's last 3 TT tag lines in LDS (one lane per search)
//   WAVE=0      the per-thread kernel: no wave schedule, each lane runs its own next stage
//               (DDX_SCHED_THREAD; the reference's WAVE=0 solve_kernel). Unset or 1: the
//               wavefront kernel with SCHEDULE
//   STACK_KB    accepted for compatibility: frames live in HBM, there is no stack limit to set
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "ddx.h"

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s K DEALS [TT_LOG2] [CHECK] | %s FILE START DEALS [TT_LOG2]\n", argv[0],
            argv[0]);
    return 2;
  }
  const bool file_mode = strstr(argv[1], ".npy") != nullptr;
  std::vector<uint64_t> hands;
  std::vector<uint8_t> want;
  long long n;
  int tt_log2, check = 0;
  if (file_mode) {
    long long start = atoll(argv[2]);
    n = argc > 3 ? atoll(argv[3]) : 0;
    tt_log2 = argc > 4 ? atoi(argv[4]) : 10;
    hands.resize(n * 4);
    want.resize(n * 20);
    int64_t got = 0;
    if (n <= 0 || ddx_pgx_read(argv[1], start, n, hands.data(), want.data(), &got) != DDX_OK || got != n) {
      fprintf(stderr, "cannot load\n");
      return 1;
    }
  } else {
    int k = atoi(argv[1]);
    n = atoll(argv[2]);
    tt_log2 = argc > 3 ? atoi(argv[3]) : 10;
    check = argc > 4 ? atoi(argv[4]) : 200;
    if (k < 1 || k > 13 || n < 1) {
      fprintf(stderr, "K must be 1..13 and DEALS >= 1\n");
      return 2;
    }
    hands.resize(n * 4);
    ddx_random_deals(12345, 0, n, k, hands.data());
  }
  if (tt_log2 < 9) {
    fprintf(stderr, "TT_LOG2 must be >= %d\n", 9);
    return 2;
  }
  ddx_gpu_config cfg;
  ddx_gpu_default_config(&cfg);
  cfg.jobs_per_deal = getenv("PER_LEAD") && atoi(getenv("PER_LEAD")) ? 20 : 5;
  cfg.team = getenv("TEAM") ? atoi(getenv("TEAM")) : 1;
  cfg.waves_per_simd = getenv("WAVES") ? atoi(getenv("WAVES")) : 0;
  cfg.lds_lines = getenv("LDS") ? atoi(getenv("LDS")) : 0;
  const long long jobs = n * cfg.jobs_per_deal;
  long long threads = getenv("THREADS") ? atoll(getenv("THREADS")) : 32768;
  if (threads > jobs * cfg.team) threads = jobs * cfg.team;  // spare searches would find no job
  threads = (threads + 255) / 256 * 256;
  cfg.lanes = threads;
  cfg.tt_log2 = tt_log2;
  cfg.schedule = DDX_SCHED_SPLIT;
  if (getenv("CHEAP_LOOP") && atoi(getenv("CHEAP_LOOP"))) cfg.schedule = DDX_SCHED_CHAIN;
  if (getenv("SCHEDULE")) cfg.schedule = atoi(getenv("SCHEDULE"));
  if (getenv("WAVE") && !atoi(getenv("WAVE"))) cfg.schedule = DDX_SCHED_THREAD;
  std::vector<uint8_t> seeds;
  if (getenv("SEEDS")) {
    seeds.resize(n * 20);
    FILE* f = fopen(getenv("SEEDS"), "rb");
    if (!f || fread(seeds.data(), 20, n, f) != (size_t)n) {
      fprintf(stderr, "cannot read SEEDS\n");
      return 1;
    }
    fclose(f);
    cfg.seed_mode = DDX_SEED_CERT;
  }
  double predict_s = 0;
  if (getenv("MODEL")) {
    ddx_mlp* net = nullptr;
    seeds.resize(n * 20);
    const auto p0 = now_s();
    if (ddx_mlp_create(0, 1, &net) != DDX_OK || ddx_mlp_load(net, getenv("MODEL")) != DDX_OK ||
        ddx_mlp_predict_host(net, hands.data(), n, seeds.data()) != DDX_OK) {
      fprintf(stderr, "%s\n", ddx_last_error());
      return 1;
    }
    ddx_mlp_destroy(net);
    predict_s = now_s() - p0;
    cfg.seed_mode = DDX_SEED_CERT;
    printf("learned seeds from %s: %.3f s\n", getenv("MODEL"), predict_s);
  }
  printf("%lld threads, ", threads);
  printf("%lld deals, %lld solves, TT %.0f KB/solve, %.2f GB total\n", n, jobs,
         320.0 * (1 << tt_log2) / 1024.0, 320.0 * (double)threads * (1 << tt_log2) / 1e9);
  if (cfg.schedule == DDX_SCHED_THREAD) printf("per-thread kernel\n");
  else printf("wavefront kernel, schedule %d\n", cfg.schedule);
  fflush(stdout);
  ddx_gpu* h = nullptr;
  if (ddx_gpu_open(&cfg, &h) != DDX_OK) {
    fprintf(stderr, "%s\n", ddx_last_error());
    return 1;
  }
  const double limit = getenv("TIME_LIMIT") ? atof(getenv("TIME_LIMIT")) : 1e18;
  std::vector<uint8_t> got(n * 20);
  std::atomic<bool> running{true};
  ddx_stats stats;
  int rc = DDX_OK;
  std::thread solver([&] {
    rc = ddx_gpu_solve(h, hands.data(), n, seeds.empty() ? nullptr : seeds.data(), got.data(), &stats);
    running = false;
  });
  const double t0 = now_s();
  double next_print = 10;
  bool stopped = false;
  while (running) {
    std::this_thread::sleep_for(std::chrono::milliseconds(100));
    const double el = now_s() - t0;
    if (el < next_print) continue;
    next_print += 10;
    uint64_t fin = 0;
    ddx_gpu_progress(h, &fin);
    printf("  %.0fs: %llu of %lld jobs done (%.1f deals/s)\n", el, (unsigned long long)fin, jobs,
           fin / (double)cfg.jobs_per_deal / el);
    if (el > limit && !stopped) {
      ddx_gpu_request_stop(h);
      stopped = true;
      printf("  time limit: stopping, unfinished jobs are skipped in the check\n");
    }
    fflush(stdout);
  }
  solver.join();
  if (rc != DDX_OK && rc != DDX_E_STOPPED) {
    fprintf(stderr, "%s\n", ddx_last_error());
    ddx_gpu_close(h);
    return 1;
  }
  const double sec = stats.seconds + predict_s;
  if (stopped)
    printf("finished jobs: %llu in %.1f s = %.1f full-job deals/s\n", (unsigned long long)stats.jobs_done,
           sec, stats.jobs_done / (double)cfg.jobs_per_deal / sec);
  long long solved = 0;
  for (long long i = 0; i < n; i++) {
    bool all = true;
    for (int j = 0; j < 20; j++) all = all && got[i * 20 + j] != 0xFF;
    solved += all;
  }
  printf("GPU: %.3fs, %.0f deals/s (%.0f strain-solves/s), %.0f nodes/solve, %lld whole tables\n", sec,
         solved / sec, stats.jobs_done / sec * (cfg.jobs_per_deal == 5 ? 1 : 0.25),
         (double)stats.nodes / (stats.jobs_done ? stats.jobs_done : 1), solved);
  ddx_gpu_close(h);

  long long bad = 0, checked = 0;
  if (file_mode) {
    for (long long i = 0; i < n * 20; i++) {
      if (got[i] == 0xFF) continue;
      bad += got[i] != want[i];
      checked++;
    }
  } else {
    const long long m = check < n ? check : n;
    std::vector<uint8_t> cpu(m * 20);
    ddx_stats cs;
    const int nt = (int)std::thread::hardware_concurrency();
    if (m > 0) {
      if (ddx_cpu_solve_batch(hands.data(), m, nt, 12, 5, DDX_SEED_PLAIN, nullptr, cpu.data(), nullptr,
                              &cs) != DDX_OK) {
        fprintf(stderr, "%s\n", ddx_last_error());
        return 1;
      }
      printf("CPU (%d threads, same code): %.0f deals/s\n", nt, m / cs.seconds);
    }
    for (long long i = 0; i < m * 20; i++) {
      if (got[i] == 0xFF) continue;
      bad += got[i] != cpu[i];
      checked++;
    }
  }
  printf("%lld wrong of %lld checked\n", bad, checked);
  return bad != 0;
}
