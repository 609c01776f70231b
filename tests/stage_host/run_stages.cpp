// run_stages.cpp -- executes the sample pipes' stage bodies (the C++ inside
// their #ifdef __PIKOC_DEVICE__ blocks) serially on the host, so a test can
// compare what those bodies compute with the oracle's frame.  Test
// infrastructure only.
//
// Schedule: every stage runs to completion before the next (the oracle's
// serial order).  Raster pipe: the vertex stage over the input in id order,
// the raster stage's assignBin over the setup triangles in emission order,
// then each bin's queue in order.  Reyes pipe: the split stage level by
// level (port 1 feeds the next level, port 0 the leaves), dice over the
// leaves in order, shade's assignBin over the micropolygons in order, then
// each bin's queue; the screen averages each pixel's samples.
//
// Build (one binary per setting of the sample macros):
//   g++ -std=c++17 -O1 -ffp-contract=off -D__PIKOC_DEVICE__ [-DRUN_REYES ...]
//       -I include -I piko-public_amd/include -I <sample dir> run_stages.cpp
// Usage: run_stages <in.bin> <out.bin>
//   in  (raster): int W, H, n; piko_constant_state; n*9 f32 positions;
//                 n*3 u32 colours (VTX_PRETRANSFORM) or n*9 f32 normals
//   out (raster): W*H u32 colours, W*H f32 depths
//   in  (Reyes):  int W, H, n; piko_constant_state; n*64 f32 control points
//   out (Reyes):  W*H*4 f32 RGBA, W*H*spp i32 sample depth bits
#include <stdio.h>
#include <stdlib.h>

#include <functional>
#include <vector>

#include "piko/piko.h"
#ifdef RUN_REYES
#include "stages.h"
#else
#include "raster.h"
#include "vertexShader.h"
#endif

namespace {
struct Grid {
  int bin = 0, nbx = 1, nby = 1;
};
Grid g_grid;
std::function<void(const void *, int)> g_emit;
std::function<void(int)> g_assign;

void grid_for(int binSize, int W, int H) {
  g_grid.nbx = binSize > 0 ? (W + binSize - 1) / binSize : 1;
  g_grid.nby = binSize > 0 ? (H + binSize - 1) / binSize : 1;
  g_grid.bin = 0;
}

template <class S>
struct Host : S {
  void bind(const ConstantState &cs, MutableState *ms) {
    this->constState = cs;
    this->mutableState = ms;
  }
};

FILE *open_or_die(const char *path, const char *mode) {
  FILE *f = fopen(path, mode);
  if (!f) {
    perror(path);
    exit(2);
  }
  return f;
}

template <class T>
void read_n(FILE *f, T *dst, size_t n) {
  if (fread(dst, sizeof(T), n, f) != n) {
    fprintf(stderr, "short input\n");
    exit(2);
  }
}
}  // namespace

template <int BX, int BY, int TT, class I, class O>
void Stage<BX, BY, TT, I, O>::assignToBin(const I &, int bin) {
  g_assign(bin);
}
template <int BX, int BY, int TT, class I, class O>
void Stage<BX, BY, TT, I, O>::assignToBin(const I &, AssignPolicy) {
  fprintf(stderr, "policy binning is not used by the samples\n");
  exit(3);
}
template <int BX, int BY, int TT, class I, class O>
int Stage<BX, BY, TT, I, O>::getBinID() const {
  return g_grid.bin;
}
template <int BX, int BY, int TT, class I, class O>
int Stage<BX, BY, TT, I, O>::getNumBins() const {
  return g_grid.nbx * g_grid.nby;
}
template <int BX, int BY, int TT, class I, class O>
int Stage<BX, BY, TT, I, O>::getNumBinsX() const {
  return g_grid.nbx;
}
template <int BX, int BY, int TT, class I, class O>
int Stage<BX, BY, TT, I, O>::getNumBinsY() const {
  return g_grid.nby;
}
template <int BX, int BY, int TT, class I, class O>
void Stage<BX, BY, TT, I, O>::emit(const O &p, int port) {
  g_emit(&p, port);
}

int piko::atomicMinInt(int *address, int value) {
  const int old = *address;
  if (value < old) *address = value;
  return old;
}

// Runs `stage` over `prims` bin by bin: assignBin in order, then each bin's
// queue in order, emitting through `sink`.
template <class S, class In>
void run_binned(S &stage, const std::vector<In> &prims, int binSize, int W, int H,
                const std::function<void(const void *, int)> &sink) {
  grid_for(binSize, W, H);
  std::vector<std::vector<int>> bins((size_t)g_grid.nbx * g_grid.nby);
  int cur = 0;
  g_assign = [&](int b) {
    if (b < 0 || b >= (int)bins.size()) {
      fprintf(stderr, "bin %d outside the grid\n", b);
      exit(3);
    }
    bins[b].push_back(cur);
  };
  for (cur = 0; cur < (int)prims.size(); ++cur) stage.assignBin(prims[cur]);
  g_emit = sink;
  for (int b = 0; b < (int)bins.size(); ++b) {
    g_grid.bin = b;
    for (int k : bins[b]) stage.process(prims[k]);
  }
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: run_stages in.bin out.bin\n");
    return 2;
  }
  FILE *in = open_or_die(argv[1], "rb");
  int hdr[3];
  read_n(in, hdr, 3);
  const int W = hdr[0], H = hdr[1], n = hdr[2];
  ConstantState cs;
  read_n(in, &cs, 1);
  const size_t npix = (size_t)W * H;
#ifndef RUN_REYES
  std::vector<raster_wtri> input(n);
  std::vector<float> pos((size_t)n * 9);
  read_n(in, pos.data(), pos.size());
#ifdef VTX_PRETRANSFORM
  std::vector<uint32_t> col((size_t)n * 3);
  read_n(in, col.data(), col.size());
#else
  std::vector<float> nrm((size_t)n * 9);
  read_n(in, nrm.data(), nrm.size());
#endif
  for (int i = 0; i < n; ++i) {
    input[i].id = i;
    for (int v = 0; v < 3; ++v) {
      input[i].vertPos[v] = {pos[9 * i + 3 * v], pos[9 * i + 3 * v + 1], pos[9 * i + 3 * v + 2]};
#ifdef VTX_PRETRANSFORM
      input[i].vertColor[v] = col[3 * i + v];
#else
      input[i].vertNormal[v] = {nrm[9 * i + 3 * v], nrm[9 * i + 3 * v + 1], nrm[9 * i + 3 * v + 2]};
#endif
    }
  }
  std::vector<float> depth(npix, 1.0f);
  std::vector<uint32_t> color(npix, 0xff663313u);
  MutableState ms;
  ms.zBuffer = depth.data();
  Host<VertexShaderStage> vs;
  Host<RasterStage> raster;
  vs.bind(cs, &ms);
  raster.bind(cs, &ms);
  // vertex stage in submission order; its bins only spread the work
  std::vector<raster_stri> setup;
  grid_for(VS_BINSIZE, W, H);
  g_assign = [](int) {};
  g_emit = [&](const void *p, int) { setup.push_back(*(const raster_stri *)p); };
  for (const raster_wtri &t : input) {
    vs.assignBin(t);
    vs.process(t);
  }
  run_binned(raster, setup, RASTER_BINSIZE, W, H, [&](const void *p, int) {
    const Pixel &px = *(const Pixel *)p;
    color[(size_t)px.pos.y * W + px.pos.x] = px.color;
  });
  FILE *out = open_or_die(argv[2], "wb");
  fwrite(color.data(), 4, npix, out);
  fwrite(depth.data(), 4, npix, out);
  fclose(out);
  printf("raster: %d triangles, %zu set up\n", n, setup.size());
#else
  std::vector<piko_patch> level(n);
  for (int i = 0; i < n; ++i) {
    read_n(in, level[i].CP, 16);
    level[i].numSplits = 0;
    level[i].nu = level[i].nv = 0;
    level[i].id = i;
  }
  const int spp = PIKO_REYES_SPP;
  std::vector<float> zbuf(npix * spp, 1.0f);
  MutableState ms;
  ms.zBuffer = zbuf.data();
  Host<SplitStage> split;
  Host<DiceStage> dice;
  Host<ShadeStage> shade;
  split.bind(cs, &ms);
  dice.bind(cs, &ms);
  shade.bind(cs, &ms);
  std::vector<piko_patch> leaves, next;
  g_emit = [&](const void *p, int port) { (port == 1 ? next : leaves).push_back(*(const piko_patch *)p); };
  g_assign = [](int) {};
  grid_for(SPLIT_BINSIZE, W, H);
  while (!level.empty()) {
    next.clear();
    for (const piko_patch &p : level) {
      split.assignBin(p);
      split.process(p);
    }
    level.swap(next);
  }
  std::vector<piko_upoly> upolys;
  grid_for(DICE_BINSIZE, W, H);
  g_emit = [&](const void *p, int) { upolys.push_back(*(const piko_upoly *)p); };
  for (const piko_patch &p : leaves) {
    dice.assignBin(p);
    dice.process(p);
  }
  float clear[3];
  const cvec3f c = piko::unpackABGR(0xff663313u);
  clear[0] = c.x, clear[1] = c.y, clear[2] = c.z;
  std::vector<float> samples(npix * spp * 3);
  for (size_t s = 0; s < npix * spp; ++s)
    for (int k = 0; k < 3; ++k) samples[3 * s + k] = clear[k];
  run_binned(shade, upolys, SHADE_BINSIZE, W, H, [&](const void *p, int) {
    const piko_sample &s = *(const piko_sample *)p;
    const size_t slot = ((size_t)s.pos.y * W + s.pos.x) * spp + s.index;
    for (int k = 0; k < 3; ++k) samples[3 * slot + k] = s.rgb[k];
  });
  std::vector<float> rgba(npix * 4);
  for (size_t px = 0; px < npix; ++px) {
    for (int k = 0; k < 3; ++k) {
      float a = samples[3 * (px * spp) + k];
      for (int s = 1; s < spp; ++s) a = a + samples[3 * (px * spp + s) + k];
      rgba[4 * px + k] = a * (1.0f / (float)spp);
    }
    rgba[4 * px + 3] = 1.0f;
  }
  FILE *out = open_or_die(argv[2], "wb");
  fwrite(rgba.data(), 4, rgba.size(), out);
  fwrite(zbuf.data(), 4, zbuf.size(), out);
  fclose(out);
  printf("reyes: %d patches, %zu leaves, %zu micropolygons\n", n, leaves.size(), upolys.size());
#endif
  fclose(in);
  return 0;
}
