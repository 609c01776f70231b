// Host program of the Reyes sample: the 4,096-patch noisy sphere at 1280x720
// (patches from libpiko_scene, patch set-up as reyesPipe/main.cpp:271-301
// does), allocate once, prepare + run_single per frame; prints one JSON line
// with checksums of the PikoScreen words and the depth, and the time per
// frame.   reyes_sample [--runs N] [--dump DIR]  (DIR/color.bin, DIR/depth.bin)
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "piko_scene.h"
#include "reyesPipe.h"

static bool write_file(const std::string &path, const void *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = fwrite(p, 1, n, f) == n;
  return fclose(f) == 0 && ok;
}

static uint64_t fnv1a(const void *p, size_t n) {
  const unsigned char *b = (const unsigned char *)p;
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

int main(int argc, char **argv) {
  int runs = 20;
  std::string dump;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--runs") && i + 1 < argc) runs = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
  }
  const int W = 1280, H = 720, nlat = 64, nlon = 64, n = nlat * nlon;
  std::vector<float> cp((size_t)n * 64);
  if (piko_scene_sphere_patches(3, nlat, nlon, 0.05f, cp.data()) != 0) return 1;
  std::vector<piko_patch> patches(n);
  for (int i = 0; i < n; ++i) {
    piko_patch &p = patches[i];
    for (int k = 0; k < 16; ++k)
      p.CP[k] = {cp[64 * (size_t)i + 4 * k], cp[64 * (size_t)i + 4 * k + 1], cp[64 * (size_t)i + 4 * k + 2],
                 cp[64 * (size_t)i + 4 * k + 3]};
    p.numSplits = 0;  // all patches begin with zero splits
    p.id = i;
    p.nu = p.nv = 0;
    p.bbmin = {99999.0f, 99999.0f};
    p.bbmax = {-99999.0f, -99999.0f};
  }
  ConstantState cs;
  memset(&cs, 0, sizeof(cs));
  cs.screenSizeX = W;
  cs.screenSizeY = H;
  cs.halfW = 0.5f * (float)W;
  cs.halfH = 0.5f * (float)H;
  const float eye[3] = {0.4f, 0.9f, 2.8f}, target[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
  if (piko_camera_view_proj(eye, target, up, 40.0f, (float)W / (float)H, 1.0f, 5.0f, cs.viewProjMatrix) != 0)
    return 1;
  std::vector<float> depth((size_t)W * H);
  MutableState ms;
  ms.zBuffer = depth.data();
  ReyesPipe pipe;
  pipe.allocate(cs, ms, patches.data(), n);
  pipe.prepare();
  pipe.run_single();
  unsigned *data = pipe.pikoScreen.getData();
  const uint64_t hc = fnv1a(data, (size_t)W * H * 4), hd = fnv1a(depth.data(), (size_t)W * H * 4);
  if (!dump.empty() && !(write_file(dump + "/color.bin", data, (size_t)W * H * 4) &&
                         write_file(dump + "/depth.bin", depth.data(), (size_t)W * H * 4)))
    return 3;
  ms.zBuffer = nullptr;
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
  printf("{\"scene\": \"sphere-4096\", \"W\": %d, \"H\": %d, \"patches\": %d, \"color_fnv1a\": \"%016llx\", "
         "\"depth_fnv1a\": \"%016llx\", \"ms_per_frame\": %.4f, \"kernels\": [%s], \"kernel_ms\": [%s]}\n",
         W, H, n, (unsigned long long)hc, (unsigned long long)hd, ms_per, kernels.c_str(), kernel_ms.c_str());
  return 0;
}
