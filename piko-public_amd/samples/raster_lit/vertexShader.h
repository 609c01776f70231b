// Vertex stage of the per-pixel-lit raster sample: the raster sample's
// transform, culls and chunk bins, with the vertex normals passed on unlit.
#ifndef SAMPLE_RASTER_LIT_VERTEX_SHADER_H
#define SAMPLE_RASTER_LIT_VERTEX_SHADER_H

#include "rasterTypes.h"

#define VS_BINSIZE 32

class VertexShaderStage : public Stage<VS_BINSIZE, VS_BINSIZE, 512, raster_wtri, raster_stri> {
#ifdef __PIKOC_DEVICE__
 public:
  void assignBin(raster_wtri p) { this->assignToBin(p, (p.id / 512) % getNumBins()); }

  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }

  void process(raster_wtri p) {
    const int S = PIKO_SUBPIXEL_BITS;
    raster_stri t;
    for (int v = 0; v < 3; ++v) {
      const cvec4f world = {p.vertPos[v].x, p.vertPos[v].y, p.vertPos[v].z, 1.0f};
      const cvec4f clip = piko::transform4(constState.viewProjMatrix, world);
      const float invW = 1.0f / clip.w;
      t.xy[2 * v] = piko::toFixed((clip.x * invW + 1.0f) * constState.halfW, S);
      t.xy[2 * v + 1] = piko::toFixed((clip.y * invW + 1.0f) * constState.halfH, S);
      t.z[v] = clip.z * invW;
      t.nrm[v] = p.vertNormal[v];
    }
    if (facesCamera(t) && insideGuardBand(t, S) && spansPixelRow(t, S)) this->emit(t, 0);
  }

 private:
  static bool facesCamera(const raster_stri &t) {
    const long long ax = t.xy[2] - t.xy[0], ay = t.xy[3] - t.xy[1];
    const long long bx = t.xy[4] - t.xy[0], by = t.xy[5] - t.xy[1];
    return by * ax - bx * ay > 0;
  }

  bool insideGuardBand(const raster_stri &t, int S) const {
    const int guard = 225 << S, xmax = constState.screenSizeX << S, ymax = constState.screenSizeY << S;
    for (int v = 0; v < 3; ++v) {
      const int x = t.xy[2 * v], y = t.xy[2 * v + 1];
      if (x < -guard || x >= xmax + guard || y < -guard || y >= ymax + guard) return false;
    }
    return true;
  }

  static bool spansPixelRow(const raster_stri &t, int S) {
    const int lo = piko::imin(piko::imin(t.xy[1], t.xy[3]), t.xy[5]);
    const int hi = piko::imax(piko::imax(t.xy[1], t.xy[3]), t.xy[5]);
    const int half = 1 << (S - 1);
    return ((lo - half) >> S) != ((hi - half) >> S);
  }
#endif
};

#endif
