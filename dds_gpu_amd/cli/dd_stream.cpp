// dd_stream: the streaming solver on the MI355X.
//   dd_stream                        deals from stdin, 32 bytes each (4 x little-endian uint64
//                                    hands N,E,S,W); one 28-byte record per deal on stdout as
//                                    soon as it is solved (any order): int64 0-based input
//                                    index + 20 bytes tricks[strain C,D,H,S,NT][declarer N,E,S,W]
//   dd_stream bench SECONDS [TT_LOG2] [CHECK_EVERY]
//                                    random full deals; deals/s every 10 s; every CHECK_EVERY-th
//                                    deal (default 50) re-solved on the CPU and compared
// Same argv, records, stderr lines and exit codes as the reference's gpu/stream_main.cu
// (0 = no mismatch, 1 = mismatch or GPU error, 2 = usage / TT_LOG2 < 9). Environment:
//   THREADS      solver lanes (default 32768, rounded up to 256)
//   PER_LEAD     0 (default, as the reference): one job per strain; 1: one job per lead,
//                which keeps four times as many lanes busy when the ring holds few deals
//   RING         ring slots (default 4096)
//   RATE         bench: Poisson arrivals at RATE deals/s (default 0 = as fast as slots free)
//   SEED         bench: deal stream seed (default 2026)
//   TEAM=16      sixteen lanes per search (THREADS / 16 searches)
//   WAVES=3      hold the kernel to 3 waves per SIMD
//   MODEL=FILE   a trained trick-count net: each deal's 20 predictions are computed on the GPU
//                when it arrives (inside its latency) and seed the search (DDX_SEED_CERT)
// Bench prints the latency from arrival to result, p50 and p99 in ms, over the deals that
// arrived in the first half of the run. A deal still unsolved at the end counts with the
// lower bound (end - arrival); a percentile that falls on such a deal is printed as ">=".
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "ddx.h"

using Clock = std::chrono::steady_clock;

static double since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

int main(int argc, char** argv) {
  const bool bench = argc > 1 && !strcmp(argv[1], "bench");
  if (argc > 1 && !bench) {
    fprintf(stderr, "usage: %s < deals.bin > results.bin | %s bench SECONDS [TT_LOG2] [CHECK_EVERY]\n",
            argv[0], argv[0]);
    return 2;
  }
  const double seconds = bench && argc > 2 ? atof(argv[2]) : 0;
  const int tt_log2 = bench && argc > 3 ? atoi(argv[3]) : 10;
  if (tt_log2 < 9) {
    fprintf(stderr, "TT_LOG2 must be >= %d\n", 9);
    return 2;
  }
  const long check_every = bench && argc > 4 ? atol(argv[4]) : 50;
  ddx_gpu_config cfg;
  ddx_gpu_default_config(&cfg);
  long long threads = getenv("THREADS") ? atoll(getenv("THREADS")) : 32768;
  cfg.lanes = (threads + 255) / 256 * 256;
  cfg.tt_log2 = tt_log2;
  cfg.jobs_per_deal = getenv("PER_LEAD") && atoi(getenv("PER_LEAD")) ? 20 : 5;
  cfg.team = getenv("TEAM") ? atoi(getenv("TEAM")) : 1;
  cfg.waves_per_simd = getenv("WAVES") ? atoi(getenv("WAVES")) : 0;
  ddx_mlp* net = nullptr;
  if (getenv("MODEL")) {
    uint64_t h0[4];
    uint8_t s0[20];
    ddx_random_deals(1, 0, 1, 13, h0);
    // Load, then predict once so the net's buffers exist before the persistent kernel starts.
    if (ddx_mlp_create(0, 1, &net) != DDX_OK || ddx_mlp_load(net, getenv("MODEL")) != DDX_OK ||
        ddx_mlp_predict_host(net, h0, 1, s0) != DDX_OK) {
      fprintf(stderr, "%s\n", ddx_last_error());
      return 1;
    }
    cfg.seed_mode = DDX_SEED_CERT;
  }
  const int ring = getenv("RING") ? atoi(getenv("RING")) : 4096;
  const double rate = getenv("RATE") ? atof(getenv("RATE")) : 0;
  const uint64_t seed = getenv("SEED") ? strtoull(getenv("SEED"), 0, 10) : 2026;
  fprintf(stderr, "%lld threads (%d per search), TT %.2f GB total, %d jobs per deal, ring %d deals%s\n",
          (long long)cfg.lanes, cfg.team, 320.0 * cfg.lanes / cfg.team * (1 << tt_log2) / 1e9,
          cfg.jobs_per_deal, ring, net ? ", learned seeds" : "");
  ddx_stream* s = nullptr;
  if (ddx_stream_open(&cfg, ring, &s) != DDX_OK) {
    fprintf(stderr, "%s\n", ddx_last_error());
    return 1;
  }
  const auto t0 = Clock::now();
  std::mutex mu;
  std::vector<double> arrival;  // by deal id
  std::vector<uint64_t> kept;   // bench: hands of the deals to check, by id / check_every
  std::atomic<long long> produced{0};
  std::atomic<bool> producing{true}, input_done{false};

  std::thread producer([&] {
    std::mt19937_64 rng(seed ^ 0x5DEECE66Dull);
    std::exponential_distribution<double> gap(rate > 0 ? rate : 1.0);
    double next_arrival = 0;
    uint64_t h[4];
    uint8_t seeds[20];
    for (long long p = 0; producing; p++) {
      if (bench) {
        ddx_random_deals(seed, p, 1, 13, h);
        if (rate > 0) {
          next_arrival += gap(rng);
          double wait = next_arrival - since(t0);
          if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        }
      } else if (fread(h, 8, 4, stdin) != 4) {
        break;
      }
      const double arr = bench && rate > 0 ? next_arrival : since(t0);
      {
        std::lock_guard<std::mutex> g(mu);
        arrival.push_back(arr);
        if (bench && p % check_every == 0) kept.insert(kept.end(), h, h + 4);
      }
      if (net && ddx_mlp_predict_host(net, h, 1, seeds) != DDX_OK) {
        fprintf(stderr, "%s\n", ddx_last_error());
        break;
      }
      int rc;
      while ((rc = ddx_stream_push(s, h, net ? seeds : nullptr, p)) == DDX_E_FULL) {
        if (!producing) break;
        std::this_thread::sleep_for(std::chrono::microseconds(100));
      }
      if (rc != DDX_OK) break;
      produced = p + 1;
    }
    input_done = true;
    if (!bench) ddx_stream_finish(s);
  });

  // Checker threads (bench): re-solve sampled deals on the CPU.
  std::vector<std::vector<uint8_t>> to_check_tab;
  std::vector<long long> to_check_id;
  std::atomic<long> checked{0}, wrong{0};
  std::atomic<bool> checking{true};
  std::mutex qm;
  std::vector<std::thread> checkers;
  if (bench) {
    int nc = std::max(1, (int)std::thread::hardware_concurrency() - 2);
    for (int t = 0; t < nc; t++)
      checkers.emplace_back([&] {
        ddx_cpu* c = nullptr;
        if (ddx_cpu_create(12, &c) != DDX_OK) return;
        for (;;) {
          long long id = -1;
          std::vector<uint8_t> got;
          {
            std::lock_guard<std::mutex> g(qm);
            if (!to_check_id.empty()) {
              id = to_check_id.back();
              got = to_check_tab.back();
              to_check_id.pop_back();
              to_check_tab.pop_back();
            }
          }
          if (id < 0) {
            if (!checking) break;
            std::this_thread::sleep_for(std::chrono::milliseconds(20));
            continue;
          }
          uint64_t h[4];
          {
            std::lock_guard<std::mutex> g(mu);
            memcpy(h, &kept[4 * (id / check_every)], 32);
          }
          bool bad = false;
          for (int st = 0; st < 5 && !bad; st++) {
            int o4[4];
            ddx_cpu_solve_strain(c, h, st, o4);
            for (int d = 0; d < 4; d++) bad = bad || o4[d] != got[st * 4 + d];
          }
          checked++;
          if (bad) wrong++;
        }
        ddx_cpu_destroy(c);
      });
  }

  std::vector<double> done_at;  // by deal id; < 0 while unsolved
  std::vector<int64_t> ids(4096);
  std::vector<uint8_t> tabs(4096 * 20);
  long long consumed = 0, last_consumed = 0;
  double last = 0;
  int rc = 0;
  for (;;) {
    int got = 0;
    if (ddx_stream_poll(s, 4096, ids.data(), tabs.data(), &got) != DDX_OK) {
      rc = 1;
      break;
    }
    const double t = since(t0);
    for (int i = 0; i < got; i++) {
      const long long id = ids[i];
      if ((long long)done_at.size() <= id) done_at.resize(id + 1, -1.0);
      done_at[id] = t;
      if (!bench) {
        unsigned char rec[28];
        memcpy(rec, &id, 8);
        memcpy(rec + 8, &tabs[20 * i], 20);
        fwrite(rec, 1, 28, stdout);
      } else if (id % check_every == 0) {
        std::lock_guard<std::mutex> g(qm);
        to_check_id.push_back(id);
        to_check_tab.emplace_back(tabs.begin() + 20 * i, tabs.begin() + 20 * i + 20);
      }
    }
    consumed += got;
    if (t - last >= 10) {
      fprintf(stderr, "  %.0fs: %lld deals solved, last 10 s: %.1f deals/s, checked %ld wrong %ld\n", t,
              consumed, (consumed - last_consumed) / (t - last), checked.load(), wrong.load());
      last = t;
      last_consumed = consumed;
    }
    if (bench && t > seconds) break;
    if (!bench && input_done && consumed >= produced) break;
    if (!got) std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
  producing = false;
  const double el = since(t0);
  producer.join();
  fflush(stdout);
  ddx_stats st;
  memset(&st, 0, sizeof st);
  if (ddx_stream_close(s, &st) != DDX_OK) {
    fprintf(stderr, "%s\n", ddx_last_error());
    rc = 1;
  }
  checking = false;
  for (auto& t : checkers) t.join();
  // Latency over the deals that arrived in the first half; unsolved ones are censored at `el`.
  std::vector<std::pair<double, bool>> lat;  // (ms, exact)
  const double horizon = bench ? el / 2 : el;
  for (size_t id = 0; id < arrival.size(); id++) {
    if (arrival[id] > horizon) continue;
    const bool fin = id < done_at.size() && done_at[id] >= 0;
    lat.push_back({1000 * ((fin ? done_at[id] : el) - arrival[id]), fin});
  }
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double q, char* buf) {
    if (lat.empty()) return (char*)"-";
    const auto& v = lat[(size_t)(q * (lat.size() - 1))];
    snprintf(buf, 32, "%s%.0f", v.second ? "" : ">=", v.first);
    return buf;
  };
  long long unsolved = 0;
  for (auto& v : lat) unsolved += !v.second;
  char b50[32], b99[32];
  fprintf(stderr, "%lld deals solved in %.1f s (%.1f deals/s), CPU check: %ld wrong of %ld\n", consumed,
          el, consumed / el, wrong.load(), checked.load());
  fprintf(stderr, "latency ms: p50 %s p99 %s (arrival to result, %zu deals arrived by %.0f s, %lld unsolved)\n",
          pct(0.5, b50), pct(0.99, b99), lat.size(), horizon, unsolved);
  if (net) ddx_mlp_destroy(net);
  return rc ? rc : wrong.load() != 0;
}
