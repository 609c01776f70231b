// Raster stage of the per-pixel-lit raster sample: the raster sample's bin
// walk, coverage and depth rule; the colour of a pixel that passes comes
// from the normal interpolated at the pixel, normalised and lit there
// (diffuse term clamped at zero, times the material, plus ambient).
#ifndef SAMPLE_RASTER_LIT_RASTER_H
#define SAMPLE_RASTER_LIT_RASTER_H

#include "rasterTypes.h"

#define RASTER_BINSIZE_LG2 3
#define RASTER_BINSIZE (1 << RASTER_BINSIZE_LG2)
#define RASTER_THREADCOUNT 64

class RasterStage : public Stage<RASTER_BINSIZE, RASTER_BINSIZE, RASTER_THREADCOUNT, raster_stri, Pixel> {
#ifdef __PIKOC_DEVICE__
 public:
  void assignBin(raster_stri p) {
    const int shift = PIKO_SUBPIXEL_BITS + RASTER_BINSIZE_LG2, size = 1 << shift;
    int lo[2], hi[2];
    pixelBox(p, lo, hi);
    const int bx0 = piko::imax(lo[0] >> shift, 0), bx1 = piko::imin(hi[0] >> shift, getNumBinsX() - 1);
    const int by0 = piko::imax(lo[1] >> shift, 0), by1 = piko::imin(hi[1] >> shift, getNumBinsY() - 1);
    for (int by = by0; by <= by1; ++by)
      for (int bx = bx0; bx <= bx1; ++bx) {
        bool keep = true;
        for (int e = 0; e < 3; ++e) {
          const long long dx = edgeDx(p, e), dy = edgeDy(p, e);
          const long long cx = ((long long)bx << shift) + (dy < 0 ? size : 0);
          const long long cy = ((long long)by << shift) + (dx >= 0 ? size : 0);
          keep = keep && edgeAt(p, e, cx, cy) >= 0;
        }
        if (keep) this->assignToBin(p, by * getNumBinsX() + bx);
      }
  }

  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }

  void process(raster_stri p) {
    const int S = PIKO_SUBPIXEL_BITS;
    const int W = constState.screenSizeX, H = constState.screenSizeY;
    const int bx = getBinID() % getNumBinsX(), by = getBinID() / getNumBinsX();
    int lo[2], hi[2];
    pixelBox(p, lo, hi);
    const int x0 = piko::imax(lo[0] >> S, bx * RASTER_BINSIZE);
    const int x1 = piko::imin(piko::imin(hi[0] >> S, (bx + 1) * RASTER_BINSIZE), W);
    const int y0 = piko::imax(lo[1] >> S, by * RASTER_BINSIZE);
    const int y1 = piko::imin(piko::imin(hi[1] >> S, (by + 1) * RASTER_BINSIZE), H);
    const long long area = -edgeDy(p, 2) * edgeDx(p, 0) + edgeDx(p, 2) * edgeDy(p, 0);
    const float invArea = 1.0f / (float)area;
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        const long long cx = ((long long)x << S) + (1 << (S - 1)), cy = ((long long)y << S) + (1 << (S - 1));
        const long long e0 = edgeAt(p, 0, cx, cy), e1 = edgeAt(p, 1, cx, cy), e2 = edgeAt(p, 2, cx, cy);
        if ((e0 | e1 | e2) < 0) continue;
        const float a = (float)e1 * invArea, b = (float)e2 * invArea;
        const float z = blend(p.z[0] - p.z[2], p.z[1] - p.z[2], p.z[2], a, b);
        int *stored = (int *)&mutableState->zBuffer[y * W + x];
        if (piko::atomicMinInt(stored, piko::floatBits(z)) < piko::floatBits(z)) continue;
        Pixel out;
        out.pos.x = x;
        out.pos.y = y;
        out.color = shade(blend(p.nrm[0].x - p.nrm[2].x, p.nrm[1].x - p.nrm[2].x, p.nrm[2].x, a, b),
                          blend(p.nrm[0].y - p.nrm[2].y, p.nrm[1].y - p.nrm[2].y, p.nrm[2].y, a, b),
                          blend(p.nrm[0].z - p.nrm[2].z, p.nrm[1].z - p.nrm[2].z, p.nrm[2].z, a, b));
        this->emit(out, 0);
      }
  }

 private:
  // the interpolated normal, normalised, lit by the frame's light
  uint32_t shade(float nx, float ny, float nz) const {
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    const float inv = len > 0.0f ? 1.0f / len : 0.0f;
    float d = (nx * inv) * constState.light[0] + (ny * inv) * constState.light[1] + (nz * inv) * constState.light[2];
    d = d < 0.0f ? 0.0f : d;
    return piko::packABGR(piko::saturate(d * constState.matcol[0] + constState.ambient[0]),
                          piko::saturate(d * constState.matcol[1] + constState.ambient[1]),
                          piko::saturate(d * constState.matcol[2] + constState.ambient[2]));
  }
  static long long edgeDx(const raster_stri &p, int e) { return (long long)p.xy[2 * ((e + 1) % 3)] - p.xy[2 * e]; }
  static long long edgeDy(const raster_stri &p, int e) {
    return (long long)p.xy[2 * ((e + 1) % 3) + 1] - p.xy[2 * e + 1];
  }
  static long long edgeAt(const raster_stri &p, int e, long long x, long long y) {
    return (y - p.xy[2 * e + 1]) * edgeDx(p, e) - (x - p.xy[2 * e]) * edgeDy(p, e);
  }
  static void pixelBox(const raster_stri &p, int lo[2], int hi[2]) {
    for (int k = 0; k < 2; ++k) {
      lo[k] = piko::imin(piko::imin(p.xy[k], p.xy[k + 2]), p.xy[k + 4]) & ~0xff;
      hi[k] = (piko::imax(piko::imax(p.xy[k], p.xy[k + 2]), p.xy[k + 4]) + 0xff) & ~0xff;
    }
  }
  static float blend(float d0, float d1, float base, float a, float b) { return base + a * d0 + b * d1; }
#endif
};

#endif
