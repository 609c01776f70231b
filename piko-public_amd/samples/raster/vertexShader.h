// Vertex stage of the raster sample (the role of rasterPipelineFixPt/
// vertexShader.pikostage): triangles go to chunk bins of 512 consecutive ids;
// process maps each vertex to fixed-point screen space (already there with
// VTX_PRETRANSFORM, else through the view-projection and the viewport),
// lights it (Gouraud: the vertex colour is computed here), and emits the
// setup triangle when it faces the camera, lies inside the guard band and
// spans a pixel-centre row.
#ifndef SAMPLE_VERTEX_SHADER_H
#define SAMPLE_VERTEX_SHADER_H

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
#ifdef VTX_PRETRANSFORM
      const cvec3f screen = p.vertPos[v];
      t.col[v] = p.vertColor[v];
#else
      const cvec4f world = {p.vertPos[v].x, p.vertPos[v].y, p.vertPos[v].z, 1.0f};
      const cvec4f clip = piko::transform4(constState.viewProjMatrix, world);
      const float invW = 1.0f / clip.w;
      const cvec3f screen = {(clip.x * invW + 1.0f) * constState.halfW, (clip.y * invW + 1.0f) * constState.halfH,
                             clip.z * invW};
      t.col[v] = lit(p.vertNormal[v]);
#endif
      t.xy[2 * v] = piko::toFixed(screen.x, S);
      t.xy[2 * v + 1] = piko::toFixed(screen.y, S);
      t.z[v] = screen.z;
    }
    if (facesCamera(t) && insideGuardBand(t, S) && spansPixelRow(t, S)) this->emit(t, 0);
  }

 private:
  // diffuse term clamped at zero, times the material, plus ambient, each
  // channel saturated
  uint32_t lit(cvec3f n) const {
    float d = n.x * constState.light[0] + n.y * constState.light[1] + n.z * constState.light[2];
    d = d < 0.0f ? 0.0f : d;
    return piko::packABGR(piko::saturate(d * constState.matcol[0] + constState.ambient[0]),
                          piko::saturate(d * constState.matcol[1] + constState.ambient[1]),
                          piko::saturate(d * constState.matcol[2] + constState.ambient[2]));
  }

  // counter-clockwise in screen space with a non-zero area
  static bool facesCamera(const raster_stri &t) {
    const long long ax = t.xy[2] - t.xy[0], ay = t.xy[3] - t.xy[1];
    const long long bx = t.xy[4] - t.xy[0], by = t.xy[5] - t.xy[1];
    return by * ax - bx * ay > 0;
  }

  // every vertex within 225 pixels of the screen
  bool insideGuardBand(const raster_stri &t, int S) const {
    const int guard = 225 << S, xmax = constState.screenSizeX << S, ymax = constState.screenSizeY << S;
    for (int v = 0; v < 3; ++v) {
      const int x = t.xy[2 * v], y = t.xy[2 * v + 1];
      if (x < -guard || x >= xmax + guard || y < -guard || y >= ymax + guard) return false;
    }
    return true;
  }

  // a pixel-centre row crosses the vertical extent (the horizontal extent is
  // not tested, as in the reference)
  static bool spansPixelRow(const raster_stri &t, int S) {
    const int lo = piko::imin(piko::imin(t.xy[1], t.xy[3]), t.xy[5]);
    const int hi = piko::imax(piko::imax(t.xy[1], t.xy[3]), t.xy[5]);
    const int half = 1 << (S - 1);
    return ((lo - half) >> S) != ((hi - half) >> S);
  }
#endif
};

#endif
