// The search machine under the memory sanitizer (clang++ -fsanitize=memory, host code only): the
// TT entries and the frame stack are left uninitialized on purpose (only the tag lines are
// cleared, as the GPU kernel clears them), so any read of a value the search never wrote is
// reported. Jobs run in index order on one lane; prints the nodes and the tables' checksum.
//   msan_machine HANDS.bin N
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "ddx/machine.hpp"

using namespace ddx;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const long n = atol(argv[2]);
  uint64_t* hands = (uint64_t*)malloc((size_t)n * 32);
  FILE* f = fopen(argv[1], "rb");
  if (!f || fread(hands, 32, (size_t)n, f) != (size_t)n) return 2;
  fclose(f);
  const size_t buckets = (size_t)1 << 10;
  u32* tags = (u32*)aligned_alloc(64, buckets * kWays * sizeof(u32));
  for (size_t i = 0; i < buckets * kWays; i++) tags[i] = 0;
  Entry* ents = (Entry*)aligned_alloc(64, buckets * kBucketEntryBytes);  // the build's entry format
  TTView v{tags, ents, (u32)(buckets - 1)};
  uint8_t* out = (uint8_t*)malloc((size_t)n * 20);
  Batch B{hands, out, nullptr, n * 5, 5, SEED_PLAIN};
  Lane* L = (Lane*)malloc(sizeof(Lane));
  lane_init(*L, v);
  ArrayStack* st = (ArrayStack*)malloc(sizeof(ArrayStack));
  long long next = 0;
  auto take = [&]() { return next++; };
  while (L->stage != ST_EXIT) lane_step(*L, B, *st, take);
  unsigned long long sum = 0;
  for (long i = 0; i < n * 20; i++) sum += out[i];
  printf("%llu %llu\n", (unsigned long long)L->c.nodes, sum);
  return 0;
}
