/* Runs the oracle's threaded raster schedule -- the raster bins
 * block-partitioned over host threads, as the reference CPU backend's process
 * kernels are (CPUBackend.cpp:522-571) -- under ThreadSanitizer, and checks
 * its frame against the single-threaded one.  TEST INFRASTRUCTURE: built and
 * run by `make -C oracle tsan` (tests/test_oracle.py). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "piko_oracle.h"

int piko_scene_tris_uniform(uint64_t seed, int n, int W, int H, float area_median, float sigma, float z_lo,
                            float z_hi, float *pos, uint32_t *col);

int main(void) {
  const int n = 20000, W = 256, H = 192, S = 8;
  float *pos = malloc(sizeof(float) * 9 * n);
  uint32_t *col = malloc(sizeof(uint32_t) * 3 * n);
  uint32_t *c1 = malloc(sizeof(uint32_t) * W * H), *c8 = malloc(sizeof(uint32_t) * W * H);
  float *d1 = malloc(sizeof(float) * W * H), *d8 = malloc(sizeof(float) * W * H);
  if (!pos || !col || !c1 || !c8 || !d1 || !d8) return 2;
  if (piko_scene_tris_uniform(17, n, W, H, 20.0f, 1.0f, 0.1f, 1.0f, pos, col) != 0) return 2;
  int bad = 0;
  for (int T = 8; T <= 64; T *= 2) {
    const long long p1 = oracle_frame_pretransformed(pos, col, n, W, H, S, T, 1, c1, d1);
    const long long p8 = oracle_frame_pretransformed(pos, col, n, W, H, S, T, 8, c8, d8);
    const int same = p1 == p8 && !memcmp(c1, c8, sizeof(uint32_t) * W * H) && !memcmp(d1, d8, sizeof(float) * W * H);
    printf("bins %d: %lld pairs, 8 threads %s the single-threaded frame\n", T, p1, same ? "equal" : "DIFFER from");
    bad |= !same;
  }
  free(pos); free(col); free(c1); free(c8); free(d1); free(d8);
  return bad;
}
