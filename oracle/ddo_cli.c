/* ddo_cli: the oracle from the command line (test infrastructure only).
 *   ddo_cli tables HANDS.bin N THREADS OUT.bin   tables of n deals (n x 4 uint64 in, n x 20 out)
 *   ddo_cli deals SEED FIRST N K OUT.bin         deals of the product's seeded generator
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ddo.h"

int main(int argc, char** argv) {
  if (argc >= 6 && !strcmp(argv[1], "tables")) {
    long n = atol(argv[3]);
    uint64_t* h = malloc(n * 32);
    uint8_t* out = malloc(n * 20);
    FILE* f = fopen(argv[2], "rb");
    if (!f || fread(h, 32, n, f) != (size_t)n) return 1;
    fclose(f);
    if (ddo_tables(h, n, atoi(argv[4]), out)) return 1;
    f = fopen(argv[5], "wb");
    if (!f || fwrite(out, 20, n, f) != (size_t)n) return 1;
    fclose(f);
    return 0;
  }
  if (argc >= 7 && !strcmp(argv[1], "deals")) {
    long first = atol(argv[3]), n = atol(argv[4]);
    int k = atoi(argv[5]);
    FILE* f = fopen(argv[6], "wb");
    if (!f) return 1;
    for (long i = 0; i < n; i++) {
      uint64_t h[4];
      ddo_random_deal(strtoull(argv[2], 0, 10), first + i, k, h);
      fwrite(h, 8, 4, f);
    }
    fclose(f);
    return 0;
  }
  fprintf(stderr, "usage: ddo_cli tables HANDS N THREADS OUT | deals SEED FIRST N K OUT\n");
  return 2;
}
