// raster_kernels.h -- device-side arithmetic of the Piko fixed-point raster
// path for gfx950.  Every helper restates one piece of the reference stage
// code (file:line given per function) with identical integer results and
// identical float operation order; compile with -ffp-contract=off so no
// multiply-add is fused (the reference CPU backend evaluates them unfused,
// api/include/piko/math.h:258-289).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace piko {

constexpr uint32_t kClearColor = 0xff663313u;  // PikoScreen clear, stage.h:83-85
constexpr int32_t kDepthClearBits = 0x3F800000; // 1.0f, main.cpp:402-409
// LDS visibility key: (signed z bits << 32) | (0xFFFFFFFE - tri).  The
// sentinel's low word exceeds every triangle's, so it loses every tie.
constexpr unsigned long long kVisSentinel =
    ((unsigned long long)(uint32_t)kDepthClearBits << 32) | 0xFFFFFFFFull;

// Per-frame uniforms passed by value (the reference keeps them in the
// module-global __constant__ ConstantState, stage.h:22-24; passing them as
// kernel arguments lets several pipes run concurrently).
// In a batch every job (a scene, or a request on a resident scene) carries
// its own copy, plus where its triangles and bins live:
//   vbase    first triangle of the job in the batch's triangle numbering
//   src_base first triangle of the job in the input buffer
//   bin_base first bin of the job in the batch's bin numbering
//   fb_base  first pixel of the job's framebuffer in the batch's buffers
// The tiled path also reads where the job's chunks and their offset table
// start (chunk_base, tab_base) and how many chunks it has (nchunks); the
// queue path whether it walks the columns of every int32-path triangle
// (qwalk = 1) or covers compact ones by spans (qwalk = 0).
struct FrameArgs {
  int W, H, S, lgT, nBx, nBy;
  float halfW, halfH;
  float vp[16];
  float light[3], matcol[3], ambient[3];
  uint32_t vbase, src_base, bin_base, nbins;
  uint32_t chunk_base, nchunks, tab_base, qwalk;
  unsigned long long fb_base;
};

__device__ __forceinline__ int imin3(int a, int b, int c) { return min(min(a, b), c); }
__device__ __forceinline__ int imax3(int a, int b, int c) { return max(max(a, b), c); }

// (int)(v * 2^S), vertexShader.pikostage:46-53; clamped to +-2^29 first so
// the conversion is defined for every input.
__device__ __forceinline__ int fixpt(float v, int S) {
  float s = v * (float)(1 << S);
  s = fminf(s, 536870912.0f);
  s = fmaxf(s, -536870912.0f);
  return (int)s;
}

// isFrontFacingFixPt (rasterMacros.h:23-30) && isInsideFrustumFixPt (32-46)
// && !isBBBetweenSamples (65-77; Y axis only, as the reference's typo does).
__device__ __forceinline__ bool vs_keep(const int p[6], int W, int H, int S) {
  const long long area = (long long)(p[5] - p[1]) * (long long)(p[2] - p[0]) -
                         (long long)(p[4] - p[0]) * (long long)(p[3] - p[1]);
  if (!(area > 0)) return false;
  const int guard = 225 << S, Wf = W << S, Hf = H << S;
  bool in = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    in &= !(p[2 * k] < -guard || p[2 * k] >= Wf + guard);
    in &= !(p[2 * k + 1] < -guard || p[2 * k + 1] >= Hf + guard);
  }
  if (!in) return false;
  const int half = 1 << (S - 1);
  const int lo = imin3(p[1], p[3], p[5]), hi = imax3(p[1], p[3], p[5]);
  return ((lo - half) >> S) != ((hi - half) >> S);
}

// computeLightingHost + saturatePixelHost + toABGRHost (main.cpp:310-334).
__device__ __forceinline__ uint32_t to_abgr(float r, float g, float b) {
  return (255u << 24) | ((unsigned)(b * 255.0f) << 16) | ((unsigned)(g * 255.0f) << 8) |
         (unsigned)(r * 255.0f);
}

__device__ __forceinline__ uint32_t light_vertex(float nx, float ny, float nz, const FrameArgs &fa) {
  float d = nx * fa.light[0] + ny * fa.light[1] + nz * fa.light[2];
  d = d < 0.0f ? 0.0f : d;
  float o[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float v = d * fa.matcol[c] + fa.ambient[c];
    o[c] = v > 1.0f ? 1.0f : (v < 0.0f ? 0.0f : v);
  }
  return to_abgr(o[0], o[1], o[2]);
}

// matmultfloat4 (math.h:365-373) + vtxTransform (rasterMacros.h:14-21) +
// fixed-point conversion.
__device__ __forceinline__ void transform_vertex(const float *v, const FrameArgs &fa, int &x, int &y,
                                                 float &z) {
  const float *m = fa.vp;
  const float cx = m[0] * v[0] + m[4] * v[1] + m[8] * v[2] + m[12];
  const float cy = m[1] * v[0] + m[5] * v[1] + m[9] * v[2] + m[13];
  const float cz = m[2] * v[0] + m[6] * v[1] + m[10] * v[2] + m[14];
  const float cw = m[3] * v[0] + m[7] * v[1] + m[11] * v[2] + m[15];
  const float onebyw = 1.0f / cw;
  const float sx = (cx * onebyw + 1.0f) * fa.halfW;
  const float sy = (cy * onebyw + 1.0f) * fa.halfH;
  z = (cz * onebyw);
  x = fixpt(sx, fa.S);
  y = fixpt(sy, fa.S);
}

// Edge deltas in the reference's naming (raster.pikostage:70-75, 166-171).
struct Edges {
  int x0, y0, x1, y1, x2, y2;
  int x1mx0, x2mx1, x0mx2, y1my0, y2my1, y0my2;
  __device__ __forceinline__ void set(const int p[6]) {
    x0 = p[0]; y0 = p[1]; x1 = p[2]; y1 = p[3]; x2 = p[4]; y2 = p[5];
    x1mx0 = x1 - x0; x2mx1 = x2 - x1; x0mx2 = x0 - x2;
    y1my0 = y1 - y0; y2my1 = y2 - y1; y0my2 = y0 - y2;
  }
  // barydenom, raster.pikostage:173
  __device__ __forceinline__ long long denom() const {
    return -(long long)y0my2 * (long long)x1mx0 + (long long)x0mx2 * (long long)y1my0;
  }
};

// Conservative bin range: computePixelBoundingBoxFixPt (rasterMacros.h:
// 48-49, 58-63) shifted by S+log2(T) and clamped (raster.pikostage:41-56).
__device__ __forceinline__ void bin_rect(const int p[6], int bsb, int nBx, int nBy, int &bx1, int &by1,
                                         int &bx2, int &by2) {
  const int lox = imin3(p[0], p[2], p[4]) & (int)0xffffff00;
  const int loy = imin3(p[1], p[3], p[5]) & (int)0xffffff00;
  const int hix = (imax3(p[0], p[2], p[4]) + 0xff) & (int)0xffffff00;
  const int hiy = (imax3(p[1], p[3], p[5]) + 0xff) & (int)0xffffff00;
  bx1 = max(0, lox >> bsb);
  by1 = max(0, loy >> bsb);
  bx2 = min(hix >> bsb, nBx - 1);
  by2 = min(hiy >> bsb, nBy - 1);
}

// Trivial-reject test of one bin (raster.pikostage:77-114 with
// GetTrivialRejectOffset, rasterMacros.h:152-160), evaluated directly at the
// bin corner that maximises each edge function; equal to the reference's
// incremental walk in exact arithmetic.
__device__ __forceinline__ bool tr_accept(const Edges &e, int bx, int by, int bsb) {
  const int B = 1 << bsb;
  const int bxs = bx << bsb, bys = by << bsb;
  const long long e0 = (long long)(bys + (e.x1mx0 >= 0 ? B : 0) - e.y0) * e.x1mx0 -
                       (long long)(bxs + (e.y1my0 < 0 ? B : 0) - e.x0) * e.y1my0;
  const long long e1 = (long long)(bys + (e.x2mx1 >= 0 ? B : 0) - e.y1) * e.x2mx1 -
                       (long long)(bxs + (e.y2my1 < 0 ? B : 0) - e.x1) * e.y2my1;
  const long long e2 = (long long)(bys + (e.x0mx2 >= 0 ? B : 0) - e.y2) * e.x0mx2 -
                       (long long)(bxs + (e.y0my2 < 0 ? B : 0) - e.x2) * e.y0my2;
  return (e0 | e1 | e2) >= 0;
}

// Edge functions at the centre of pixel (px, py) (raster.pikostage:176-180,
// 223-241).  Covered iff all three are >= 0 (no fill rule).
__device__ __forceinline__ void pixel_edges(const Edges &e, int px, int py, int S, long long &e0,
                                            long long &e1, long long &e2) {
  const int half = 1 << (S - 1);
  const int cx = (px << S) + half, cy = (py << S) + half;
  e0 = (long long)(cy - e.y0) * e.x1mx0 - (long long)(cx - e.x0) * e.y1my0;
  e1 = (long long)(cy - e.y1) * e.x2mx1 - (long long)(cx - e.x1) * e.y2my1;
  e2 = (long long)(cy - e.y2) * e.x0mx2 - (long long)(cx - e.x2) * e.y0my2;
}

// interpolate_alphabeta (raster.pikostage:28-32): f2 + alpha*f0mf2 + beta*f1mf2.
__device__ __forceinline__ float interp(float f0mf2, float f1mf2, float f2, float alpha, float beta) {
  return f2 + alpha * f0mf2 + beta * f1mf2;
}

__device__ __forceinline__ unsigned long long vis_key(float z, uint32_t tri) {
  return ((unsigned long long)(uint32_t)__float_as_int(z) << 32) | (unsigned long long)(0xFFFFFFFEu - tri);
}

}  // namespace piko
