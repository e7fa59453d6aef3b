// dd_train: trains the trick-count net on solved deals, on the GPU.
//   dd_train DATA.bin MODEL.bin [EPOCHS] [BATCH_DEALS] [LR] [HOLDOUT]
// DATA.bin: records of 52 bytes, 4 x little-endian uint64 hands (N,E,S,W) + 20 uint8 tricks
// [strain C,D,H,S,NT][declarer N,E,S,W], 255 = unsolved (no label: skipped by the loss and the
// held-out score). The last HOLDOUT deals (default 2000) are held out; the
// held-out accuracy of the predicted NS tricks (exact, within one) is printed after each
// epoch. Defaults: 20 epochs, 512 deals per batch, learning rate 1e-3 (Adam). INIT=MODEL.bin
// starts from a saved net instead of a random one (e.g. a second pass at a lower rate).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "ddx.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s DATA.bin MODEL.bin [EPOCHS] [BATCH_DEALS] [LR] [HOLDOUT]\n", argv[0]);
    return 2;
  }
  const int epochs = argc > 3 ? atoi(argv[3]) : 20;
  const int batch = argc > 4 ? atoi(argv[4]) : 512;
  const float lr = argc > 5 ? (float)atof(argv[5]) : 1e-3f;
  long long holdout = argc > 6 ? atoll(argv[6]) : 2000;
  FILE* f = fopen(argv[1], "rb");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  fseek(f, 0, SEEK_END);
  const long long n = ftell(f) / 52;
  fseek(f, 0, SEEK_SET);
  std::vector<uint64_t> hands(n * 4);
  std::vector<uint8_t> tables(n * 20);
  for (long long i = 0; i < n; i++) {
    if (fread(&hands[i * 4], 8, 4, f) != 4 || fread(&tables[i * 20], 1, 20, f) != 20) {
      fprintf(stderr, "short read in %s\n", argv[1]);
      return 1;
    }
  }
  fclose(f);
  if (holdout >= n) holdout = n / 10;
  const long long ntrain = n - holdout;
  printf("%lld deals (%lld train, %lld held out), %d epochs, batch %d deals, lr %g\n", n, ntrain,
         holdout, epochs, batch, lr);
  ddx_mlp* net = nullptr;
  if (ddx_mlp_create(0, 12345, &net) != DDX_OK || (getenv("INIT") && ddx_mlp_load(net, getenv("INIT")) != DDX_OK)) {
    fprintf(stderr, "%s\n", ddx_last_error());
    return 1;
  }
  if (getenv("INIT")) printf("starting from %s\n", getenv("INIT"));
  std::vector<uint8_t> seeds(holdout * 20);
  for (int e = 0; e < epochs; e++) {
    float la[2];
    if (ddx_mlp_train(net, hands.data(), tables.data(), ntrain, 1, batch, lr, 1000 + e, la) != DDX_OK ||
        ddx_mlp_predict_host(net, &hands[ntrain * 4], holdout, seeds.data()) != DDX_OK) {
      fprintf(stderr, "%s\n", ddx_last_error());
      return 1;
    }
    long long exact = 0, near = 0, known = 0;
    for (long long i = 0; i < holdout; i++)
      for (int j = 0; j < 20; j++) {
        const int t = tables[(ntrain + i) * 20 + j];
        if (t == 255) continue;  // unsolved: no label
        const int ns = (j & 1) == 0 ? t : 13 - t;
        const int d = (int)seeds[i * 20 + j] - ns;
        exact += d == 0;
        near += d >= -1 && d <= 1;
        known++;
      }
    printf("epoch %2d: train loss %.4f acc %.3f | held out exact %.3f within one %.3f (%lld answers)\n", e + 1,
           la[0], la[1], exact / (double)(known ? known : 1), near / (double)(known ? known : 1), known);
    fflush(stdout);
  }
  if (ddx_mlp_save(net, argv[2]) != DDX_OK) {
    fprintf(stderr, "%s\n", ddx_last_error());
    return 1;
  }
  ddx_mlp_destroy(net);
  printf("saved %s\n", argv[2]);
  return 0;
}
