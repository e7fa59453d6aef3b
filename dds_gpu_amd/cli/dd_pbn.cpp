// dd_pbn: double-dummy tricks for PBN deals read from stdin.
//   dd_pbn < deals.pbn
// Input lines "N:hand hand hand hand" (hands N,E,S,W; suits S.H.D.C; ranks AKQJT98765432; k
// cards each, 1 <= k <= 13); other lines are skipped. Output, one line per deal: 20 numbers,
// for strain C,D,H,S,NT and leader N,E,S,W, the tricks won by the side on lead -- the
// reference's tests/dd_pbn.cpp:1-35, which its DDS cross-check (tests/vs_dds.py) consumes.
// Environment: GPU=1 solves all deals in one batch on the GPU (16-lane teams) instead of on
// the CPU (one context, TT_LOG2 14, as the reference).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "ddx.h"

static bool parse_pbn(const char* line, uint64_t h[4]) {
  static const char* R = "23456789TJQKA";
  if (strncmp(line, "N:", 2)) return false;
  h[0] = h[1] = h[2] = h[3] = 0;
  int p = 0, s = 0;
  for (const char* q = line + 2; *q && *q != '\n' && *q != '\r'; q++) {
    if (*q == ' ') {
      if (++p > 3) return false;
      s = 0;
      continue;
    }
    if (*q == '.') {
      if (++s > 3) return false;
      continue;
    }
    const char* r = strchr(R, *q);
    if (!r || !*q) return false;
    h[p] |= 1ull << (s * 16 + (int)(r - R));
  }
  return p == 3;
}

int main() {
  std::vector<uint64_t> hands;
  char line[512];
  while (fgets(line, sizeof line, stdin)) {
    uint64_t h[4];
    if (!parse_pbn(line, h)) continue;
    hands.insert(hands.end(), h, h + 4);
  }
  const long long n = (long long)hands.size() / 4;
  std::vector<int> ns(n * 20);  // [deal][strain][leader]: N/S tricks
  if (getenv("GPU") && atoi(getenv("GPU"))) {
    ddx_gpu_config cfg;
    ddx_gpu_default_config(&cfg);
    cfg.team = 16;
    cfg.tt_log2 = 14;
    cfg.jobs_per_deal = 20;
    cfg.schedule = DDX_SCHED_SPLIT;
    cfg.lanes = ((n * 20 * 16 + 255) / 256) * 256;
    if (cfg.lanes > 65536) cfg.lanes = 65536;
    std::vector<uint8_t> tab(n * 20);
    ddx_gpu* g = nullptr;
    if (ddx_gpu_open(&cfg, &g) != DDX_OK || ddx_gpu_solve(g, hands.data(), n, nullptr, tab.data(), nullptr) != DDX_OK) {
      fprintf(stderr, "%s\n", ddx_last_error());
      return 1;
    }
    ddx_gpu_close(g);
    for (long long d = 0; d < n; d++)
      for (int st = 0; st < 5; st++)
        for (int leader = 0; leader < 4; leader++) {
          const int decl = (leader + 3) & 3, t = tab[d * 20 + st * 4 + decl];
          ns[d * 20 + st * 4 + leader] = (decl & 1) == 0 ? t : 13 - t;
        }
  } else {
    ddx_cpu* c = nullptr;
    if (ddx_cpu_create(14, &c) != DDX_OK) {
      fprintf(stderr, "%s\n", ddx_last_error());
      return 1;
    }
    for (long long d = 0; d < n; d++)
      for (int st = 0; st < 5; st++)
        for (int leader = 0; leader < 4; leader++)
          if (ddx_cpu_ns_tricks(c, &hands[4 * d], st, leader, 7, &ns[d * 20 + st * 4 + leader]) != DDX_OK) {
            fprintf(stderr, "%s\n", ddx_last_error());
            return 1;
          }
    ddx_cpu_destroy(c);
  }
  for (long long d = 0; d < n; d++) {
    int k = 0;
    for (uint64_t x = hands[4 * d]; x; x &= x - 1) k++;
    for (int i = 0; i < 20; i++) {
      const int leader = i & 3, v = ns[d * 20 + i];
      printf("%d ", (leader & 1) == 0 ? v : k - v);
    }
    printf("\n");
  }
  return 0;
}
