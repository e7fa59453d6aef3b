// pbn: print deal I of a Pgx .npy file as PBN (North first) and its stored table.
//   pbn FILE.npy I
// Output: "N:<N> <E> <S> <W>" (suits S.H.D.C, ranks high to low), then the 20 stored tricks,
// declarer seat N,E,S,W major and strain C,D,H,S,NT minor, comma-separated -- the reference's
// tests/pbn.cpp:1-21. Exit 1 if the file or deal cannot be read.
#include <stdio.h>
#include <stdlib.h>

#include "ddx.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s FILE.npy I\n", argv[0]);
    return 2;
  }
  uint64_t h[4];
  uint8_t t[20];
  int64_t got = 0;
  if (ddx_pgx_read(argv[1], atoll(argv[2]), 1, h, t, &got) != DDX_OK || got != 1) return 1;
  static const char* R = "23456789TJQKA";
  printf("N:");
  for (int p = 0; p < 4; p++) {
    for (int s = 0; s < 4; s++) {
      for (int r = 12; r >= 0; r--)
        if (h[p] >> (s * 16 + r) & 1) putchar(R[r]);
      if (s < 3) putchar('.');
    }
    putchar(p < 3 ? ' ' : '\n');
  }
  for (int seat = 0; seat < 4; seat++)
    for (int st = 0; st < 5; st++) printf("%d%c", t[st * 4 + seat], seat == 3 && st == 4 ? '\n' : ',');
  return 0;
}
