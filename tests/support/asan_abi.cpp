// The host side of the C ABI under the address and undefined-behaviour sanitizers (g++, host
// code only, built with dds_gpu_amd/csrc/ddx_host.cpp): CPU batch solves with one and with
// several threads, a context's strain solves, a Pgx write and read-back, and argument errors.
// Exit 0 and "ok" when every call behaves; the sanitizers abort on any bad access or leak.
//   asan_abi SCRATCH.npy MISSING.npy
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ddx.h"

int main(int argc, char** argv) {
  if (argc < 3) return 20;
  const char* path = argv[1];  // a scratch .npy to write and read back
  const char* missing = argv[2];  // a path that does not exist
  const int n = 16;  // 8-card endings for the solves
  uint64_t* hands = (uint64_t*)malloc(n * 32);
  if (ddx_random_deals(7, 0, n, 8, hands) != 0) return 1;
  uint8_t* out = (uint8_t*)malloc(n * 20);
  uint64_t* jn = (uint64_t*)malloc(n * 5 * 8);
  ddx_stats st;
  uint8_t* out1 = (uint8_t*)malloc(n * 20);
  if (ddx_cpu_solve_batch(hands, n, 1, 10, 5, 0, nullptr, out1, jn, &st) != 0) return 2;
  if (ddx_cpu_solve_batch(hands, n, 4, 10, 20, 0, nullptr, out, nullptr, &st) != 0) return 3;
  if (memcmp(out1, out, n * 20)) return 12;  // per-strain and per-lead jobs agree
  uint64_t* full = (uint64_t*)malloc(n * 32);  // Pgx files hold full deals
  if (ddx_random_deals(8, 0, n, 13, full) != 0) return 13;
  for (int i = 0; i < n * 20; i++) out1[i] = (uint8_t)((i * 7) % 14);
  if (ddx_pgx_write(path, full, out1, n) != 0) return 4;
  int64_t rows = 0;
  if (ddx_pgx_rows(path, &rows) != 0 || rows != n) return 5;
  uint64_t* h2 = (uint64_t*)malloc(n * 32);
  uint8_t* t2 = (uint8_t*)malloc(n * 20);
  int64_t got = 0;
  if (ddx_pgx_read(path, 0, n, h2, t2, &got) != 0 || got != n) return 6;
  if (memcmp(h2, full, n * 32) || memcmp(t2, out1, n * 20)) return 7;
  ddx_cpu* c;
  if (ddx_cpu_create(10, &c) != 0) return 8;
  int o4[4];
  for (int i = 0; i < n; i++) if (ddx_cpu_solve_strain(c, hands + 4 * i, i % 5, o4) != 0) return 9;
  ddx_cpu_destroy(c);
  // bad arguments fail cleanly
  if (ddx_cpu_solve_batch(hands, n, 0, 10, 5, 0, nullptr, out, nullptr, &st) == 0) return 10;
  if (ddx_pgx_read(missing, 0, 1, h2, t2, &got) == 0) return 11;
  printf("ok %llu nodes\n", (unsigned long long)st.nodes);
  free(hands);
  free(full);
  free(out);
  free(out1);
  free(jn);
  free(h2);
  free(t2);
  return 0;
}
