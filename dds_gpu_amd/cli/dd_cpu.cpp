// dd_cpu: solve Pgx deals on the CPU and check them against the stored DDS tables.
//   dd_cpu FILE START COUNT [THREADS] [TT_LOG2]
// Same argv, environment (GUESS=exact|off1), output line and exit codes as the reference's
// cpu/cpu_main.cpp:12-88 (0 = all equal, 1 = mismatch or load failure, 2 = usage or
// TT_LOG2 < 9). Runs the ddx CPU solver through the C ABI.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <thread>
#include <vector>

#include "ddx.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s FILE START COUNT [THREADS] [TT_LOG2]\n", argv[0]);
    return 2;
  }
  long long start = atoll(argv[2]), count = atoll(argv[3]);
  int threads = argc > 4 ? atoi(argv[4]) : (int)std::thread::hardware_concurrency();
  int tt_log2 = argc > 5 ? atoi(argv[5]) : 18;
  if (tt_log2 < 9) {
    fprintf(stderr, "TT_LOG2 must be >= %d\n", 9);
    return 2;
  }
  if (threads < 1) threads = 1;
  const char* guess = getenv("GUESS") ? getenv("GUESS") : "";
  std::vector<uint64_t> hands(count * 4);
  std::vector<uint8_t> want(count * 20), got(count * 20), seeds(count * 20);
  int64_t n = 0;
  if (ddx_pgx_read(argv[1], start, count, hands.data(), want.data(), &n) != DDX_OK || n != count) {
    fprintf(stderr, "cannot load %s\n", argv[1]);
    return 1;
  }
  int mode = DDX_SEED_PLAIN;
  if (*guess) {  // exact: the true NS tricks; off1: +-1 alternating (cpu/cpu_main.cpp:57-65)
    mode = DDX_SEED_STEP;
    for (long long i = 0; i < count; i++)
      for (int st = 0; st < 5; st++)
        for (int d = 0; d < 4; d++) {
          int t = want[i * 20 + st * 4 + d];
          int ns = (d & 1) == 0 ? t : 13 - t;
          if (!strcmp(guess, "off1")) ns += ((start + i + d + st) & 1) ? 1 : -1;
          seeds[i * 20 + st * 4 + d] = (uint8_t)(ns < 0 ? 0 : ns > 13 ? 13 : ns);
        }
  }
  ddx_stats stats;
  if (ddx_cpu_solve_batch(hands.data(), count, threads, tt_log2, 5, mode, *guess ? seeds.data() : nullptr,
                          got.data(), nullptr, &stats) != DDX_OK) {
    fprintf(stderr, "%s\n", ddx_last_error());
    return 1;
  }
  long long bad = 0;
  for (long long i = 0; i < count; i++)
    for (int st = 0; st < 5; st++)
      for (int seat = 0; seat < 4; seat++) {
        int w = want[i * 20 + st * 4 + seat], h = got[i * 20 + st * 4 + seat];
        if (w != h && bad++ < 10)
          printf("deal %lld strain %d seat %d: want %d got %d\n", start + i, st, seat, w, h);
      }
  printf("%lld deals, %d threads: %.2fs, %.1f deals/s, %.0f nodes/deal, %lld wrong of %lld\n", count,
         threads, stats.seconds, count / stats.seconds, (double)stats.nodes / count, bad, count * 20);
  return bad != 0;
}
