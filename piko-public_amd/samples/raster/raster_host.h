// Host program of the raster samples, a template over the pipe class: builds
// a synthetic scene, runs the pipe the way the reference's host does
// (rasterPipelineFixPt/main.cpp:169-231: allocate once, then prepare +
// run_single per frame, pikoScreen.getData()), and prints one JSON line with
// checksums of the colour and depth buffers, the time per frame and the
// kernels one frame launches.  Included by each sample's main.cpp after its
// rasterPipe.h.
//   [--scene uniform|heightfield] [--runs N] [--dump DIR]
// (--dump writes the first frame's DIR/color.bin and DIR/depth.bin)
// uniform: 50,000 screen-space triangles at 512x512 from seed 1 (needs a
// pretransformed build); heightfield: the 2M-triangle 1080p height field
// (needs a world-space build).
#ifndef SAMPLE_RASTER_HOST_H
#define SAMPLE_RASTER_HOST_H

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "piko_scene.h"

inline bool write_file(const std::string &path, const void *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = fwrite(p, 1, n, f) == n;
  return fclose(f) == 0 && ok;
}

inline uint64_t fnv1a(const void *p, size_t n) {
  const unsigned char *b = (const unsigned char *)p;
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

template <class Pipe>
int raster_host_main(int argc, char **argv) {
  std::string scene = "uniform", dump;
  int runs = 20;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--scene") && i + 1 < argc) scene = argv[++i];
    else if (!strcmp(argv[i], "--runs") && i + 1 < argc) runs = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
  }
  ConstantState cs;
  memset(&cs, 0, sizeof(cs));
  const float light[3] = {0.5773f, 0.5773f, 0.5773f}, matcol[3] = {0.9f, 0.9f, 0.6f},
              ambient[3] = {0.15f, 0.15f, 0.30f};
  memcpy(cs.light, light, sizeof(light));
  memcpy(cs.matcol, matcol, sizeof(matcol));
  memcpy(cs.ambient, ambient, sizeof(ambient));
  std::vector<raster_wtri> tris;
#ifdef VTX_PRETRANSFORM
  if (scene != "uniform") {
    fprintf(stderr, "this build takes screen-space input: --scene uniform\n");
    return 2;
  }
  const int W = 512, H = 512, n = 50000;
  std::vector<float> pos((size_t)n * 9);
  std::vector<uint32_t> col((size_t)n * 3);
  if (piko_scene_tris_uniform(1, n, W, H, 20.0f, 1.0f, 0.1f, 1.0f, pos.data(), col.data()) != 0) return 1;
  for (int i = 0; i < 16; ++i) cs.viewProjMatrix[i] = (i % 5 == 0) ? 1.0f : 0.0f;
#else
  if (scene != "heightfield") {
    fprintf(stderr, "this build takes world-space input: --scene heightfield\n");
    return 2;
  }
  const int W = 1920, H = 1080, grid = 1000;
  const int n = (int)piko_scene_heightfield_count(grid);
  std::vector<float> pos((size_t)n * 9), nrm((size_t)n * 9);
  if (piko_scene_heightfield_freq(2, grid, 0.2f, 6, 16.0f, pos.data(), nrm.data()) != 0) return 1;
  const float eye[3] = {0.0f, 0.12f, 0.7f}, target[3] = {0.0f, -0.3f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
  if (piko_camera_view_proj(eye, target, up, 60.0f, (float)W / (float)H, 0.02f, 4.0f, cs.viewProjMatrix) != 0)
    return 1;
#endif
  cs.screenSizeX = W;
  cs.screenSizeY = H;
  cs.halfW = 0.5f * (float)W;  // initScreen (main.cpp:104-105)
  cs.halfH = 0.5f * (float)H;
  tris.resize(n);
  for (int i = 0; i < n; ++i) {
    raster_wtri &t = tris[i];
    t.id = i;
    for (int v = 0; v < 3; ++v) {
      t.vertPos[v] = {pos[9 * (size_t)i + 3 * v], pos[9 * (size_t)i + 3 * v + 1], pos[9 * (size_t)i + 3 * v + 2]};
#ifdef VTX_PRETRANSFORM
      t.vertColor[v] = col[3 * (size_t)i + v];
#else
      t.vertNormal[v] = {nrm[9 * (size_t)i + 3 * v], nrm[9 * (size_t)i + 3 * v + 1], nrm[9 * (size_t)i + 3 * v + 2]};
#endif
    }
  }
  std::vector<float> depth((size_t)W * H);
  MutableState ms;
  ms.zBuffer = depth.data();
  Pipe pipe;
  pipe.allocate(cs, ms, tris.data(), n);
  pipe.prepare();
  pipe.run_single();
  unsigned *data = pipe.pikoScreen.getData();
  const uint64_t hc = fnv1a(data, (size_t)W * H * 4), hd = fnv1a(depth.data(), (size_t)W * H * 4);
  if (!dump.empty() && !(write_file(dump + "/color.bin", data, (size_t)W * H * 4) &&
                         write_file(dump + "/depth.bin", depth.data(), (size_t)W * H * 4)))
    return 3;
  ms.zBuffer = nullptr;  // timing loop: no depth read-back
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < runs; ++r) {
    pipe.prepare();
    pipe.run_single();
  }
  const double ms_per = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                        (runs > 0 ? runs : 1);
  // the kernels one frame of the generated pipe launches, with their times
  float kms[64], frame_ms = 0.0f;
  const char *knames[64];
  int nk = 0;
  PIKO_CHECK(piko_profile_frame(pipe, kms, knames, 64, &nk, &frame_ms));
  std::string kernels;
  std::string kernel_ms;
  for (int k = 0; k < nk; ++k) {
    kernels += std::string(k ? ", \"" : "\"") + knames[k] + "\"";
    char t[32];
    snprintf(t, sizeof(t), "%s%.4f", k ? ", " : "", kms[k]);
    kernel_ms += t;
  }
  pipe.destroy();
  printf("{\"scene\": \"%s\", \"W\": %d, \"H\": %d, \"triangles\": %d, \"color_fnv1a\": \"%016llx\", "
         "\"depth_fnv1a\": \"%016llx\", \"ms_per_frame\": %.4f, \"kernels\": [%s], \"kernel_ms\": [%s]}\n",
         scene.c_str(), W, H, n, (unsigned long long)hc, (unsigned long long)hd, ms_per, kernels.c_str(), kernel_ms.c_str());
  return 0;
}

#endif
