// Runs the edge-case pipe on `count` boxes (0 allowed) at 200x120 and prints
// one JSON line: covered pixels, the colour words' checksum, and each
// pixel's expected winner recomputed here (the nearest box, the later one on
// a tie).  A stage-code error aborts through PIKO_CHECK with its message.
//   edge_pipe <count>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "edgePipe.h"

int main(int argc, char **argv) {
  const int count = argc > 1 ? atoi(argv[1]) : 100;
  const int W = 200, H = 120;
  ConstantState cs;
  memset(&cs, 0, sizeof(cs));
  cs.screenSizeX = W;
  cs.screenSizeY = H;
  std::vector<edge_tri> tris(count);
  uint64_t s = 12345;
  auto next = [&]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(s >> 33);
  };
  for (int i = 0; i < count; ++i) {
    const float x = (float)(next() % (W + 40)) - 20.0f, y = (float)(next() % (H + 40)) - 20.0f;
    const float z = (float)(next() % 8) / 8.0f;  // few depths: ties happen
    tris[i].id = i;
    tris[i].vertPos[0] = {x, y, z};
    tris[i].vertPos[1] = {x + (float)(next() % 40), y, z};
    tris[i].vertPos[2] = {x, y + (float)(next() % 30), z};
    tris[i].color = 0xff000000u | next();
  }
  std::vector<float> depth((size_t)W * H);
  MutableState ms;
  ms.zBuffer = depth.data();
  EdgePipe pipe;
  pipe.allocate(cs, ms, tris.data(), count);
  pipe.prepare();
  pipe.run_single();
  const unsigned *got = pipe.pikoScreen.getData();
  // expected: boxes in order, a later box wins a tie (<=)
  std::vector<uint32_t> want((size_t)W * H, 0xff663313u);
  std::vector<float> wz((size_t)W * H, 1.0f);
  int covered = 0, wrong = 0;
  for (const edge_tri &t : tris) {
    int x0 = (int)fminf(fminf(t.vertPos[0].x, t.vertPos[1].x), t.vertPos[2].x);
    int y0 = (int)fminf(fminf(t.vertPos[0].y, t.vertPos[1].y), t.vertPos[2].y);
    int x1 = (int)fmaxf(fmaxf(t.vertPos[0].x, t.vertPos[1].x), t.vertPos[2].x) + 1;
    int y1 = (int)fmaxf(fmaxf(t.vertPos[0].y, t.vertPos[1].y), t.vertPos[2].y) + 1;
    for (int y = y0 < 0 ? 0 : y0; y < (y1 < H ? y1 : H); ++y)
      for (int x = x0 < 0 ? 0 : x0; x < (x1 < W ? x1 : W); ++x)
        if (t.vertPos[0].z <= wz[(size_t)y * W + x]) {
          wz[(size_t)y * W + x] = t.vertPos[0].z;
          want[(size_t)y * W + x] = t.color;
        }
  }
  for (size_t p = 0; p < want.size(); ++p) {
    covered += want[p] != 0xff663313u;
    wrong += got[p] != want[p] || depth[p] != wz[p];
  }
  pipe.destroy();
  printf("{\"count\": %d, \"covered\": %d, \"wrong\": %d}\n", count, covered, wrong);
  return 0;
}
