/*
 * reyes_oracle.c -- CPU restatement of the Piko Reyes pipeline
 * (Split -> Dice -> Shade, pikoc/samples/reyesPipe/{split,dice,shade}.pikostage)
 * run single-threaded.  TEST INFRASTRUCTURE ONLY: the checker of the HIP
 * Reyes path, never part of it.
 *
 * Order of work (the reference CPU backend's, made deterministic): the split
 * stage runs level by level (each kernel launch of the split self-loop
 * processes the patches present when it starts, CPUBackend.cpp:459-515), so
 * leaves are numbered by (level, position in level); a leaf's micropolygons
 * are numbered in dice order (k outer, j inner, dice.pikostage:176-177).
 * Depth ties resolve to the later micropolygon in that numbering, as the
 * shade stage's `zBuffer == z` test after its atomicMin loop does
 * (shade.pikostage:226-241).
 *
 * Parameters beyond the reference (documented deviations):
 *   threshold  split threshold in pixels (reference THRESHOLD_PIXELS 4)
 *   dice       0: the reference's grid (nu,nv from pow2roundup, clamped to
 *              [2,32], quads k<=nu, j<=nv including one row/column past the
 *              patch, dice.pikostage:174-177); N>0: an N x N grid of quads
 *              k<N, j<N covering exactly the patch
 *   spp        1: the pixel centre (shade.pikostage:166-167); 4: the centres
 *              of the four 2x2 strata
 *   disp       displacement along the normal by disp*(Px*Px - Py*Py) before
 *              projection (0 = none, as the reference)
 *   output     FP32 RGBA, samples averaged (box filter); uncovered samples
 *              take the PikoScreen clear colour 0xff663313 as floats.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "piko_oracle.h"

typedef struct { float x, y, z, w; } v4;

static inline float fmn(float a, float b) { return (b < a) ? b : a; }  /* std::min, math.h:24-31 */
static inline float fmx(float a, float b) { return (a < b) ? b : a; }  /* std::max, math.h:44-51 */
static inline int32_t fbits(float f) { int32_t i; memcpy(&i, &f, 4); return i; }

/* matmultfloat4 for cvec4f (math.h:384-392) */
static v4 mat4(const float *m, v4 v) {
  v4 o;
  o.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  o.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  o.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  o.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return o;
}

static float dist2(float ax, float ay, float bx, float by) { /* piko::distance, math.h:334-338 */
  const float cx = ax - bx, cy = ay - by;
  return sqrtf(cx * cx + cy * cy);
}

/* getUVPixelExtents (split.pikostage:47-94) of the projected patch */
static void extents(const float *cl, float *u, float *v) {
  *u = 0.0f;
  for (int i = 0; i < 4; ++i) {
    const float d01 = dist2(cl[i * 16 + 0], cl[i * 16 + 1], cl[i * 16 + 4], cl[i * 16 + 5]);
    const float d12 = dist2(cl[i * 16 + 4], cl[i * 16 + 5], cl[i * 16 + 8], cl[i * 16 + 9]);
    const float d23 = dist2(cl[i * 16 + 8], cl[i * 16 + 9], cl[i * 16 + 12], cl[i * 16 + 13]);
    *u = fmx(*u, d01 + d12 + d23);
  }
  *v = 0.0f;
  for (int i = 0; i < 4; ++i) {
    const float d01 = dist2(cl[0 + i * 4], cl[1 + i * 4], cl[16 + i * 4], cl[17 + i * 4]);
    const float d12 = dist2(cl[16 + i * 4], cl[17 + i * 4], cl[32 + i * 4], cl[33 + i * 4]);
    const float d23 = dist2(cl[32 + i * 4], cl[33 + i * 4], cl[48 + i * 4], cl[49 + i * 4]);
    *v = fmx(*v, d01 + d12 + d23);
  }
}

/* SplitStage::process projection (split.pikostage:245-276): float divide by
 * w, then the viewport mapping in double. */
static void project_patch(const v4 *cp, const float *vp, int W, int H, float *cl) {
  for (int i = 0; i < 16; ++i) {
    v4 c = mat4(vp, cp[i]);
    if (c.w == 0.0f) c.w = 1.0f;
    c.x /= c.w;
    c.y /= c.w;
    c.z /= c.w;
    c.x = (float)(((double)c.x + 1.0) * 0.5 * (double)W);
    c.y = (float)(((double)c.y + 1.0) * 0.5 * (double)H);
    cl[i * 4 + 0] = c.x;
    cl[i * 4 + 1] = c.y;
    cl[i * 4 + 2] = c.z;
    cl[i * 4 + 3] = c.w;
  }
}

static v4 add4(v4 a, v4 b) { v4 o = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; return o; }
static v4 half4(v4 a) { v4 o = {a.x * 0.5f, a.y * 0.5f, a.z * 0.5f, a.w * 0.5f}; return o; }

/* splitU / splitV (split.pikostage:95-224): midpoint de Casteljau per row
 * (stride 1) or per column (stride 4). */
static void split_dir(const v4 *cp, v4 *n1, v4 *n2, int along_u) {
  for (int i = 0; i < 4; ++i) {
    const int s = along_u ? 1 : 4, b = along_u ? i * 4 : i;
    const int c0 = b, c1 = b + s, c2 = b + 2 * s, c3 = b + 3 * s;
    n1[c0] = cp[c0];
    n2[c3] = cp[c3];
    n1[c1] = half4(add4(cp[c0], cp[c1]));
    n2[c2] = half4(add4(cp[c2], cp[c3]));
    const v4 tmp = half4(add4(cp[c1], cp[c2]));
    n1[c2] = half4(add4(n1[c1], tmp));
    n2[c1] = half4(add4(n2[c2], tmp));
    n1[c3] = half4(add4(n1[c2], n2[c1]));
    n2[c0] = n1[c3];
  }
}

/* (int)ceil(ext / threshold) as split.pikostage:307-308 computes it, with
 * the conversion made total: values of 64 and above give the same clamped
 * grid (32) as any larger value, NaN gives 0 (the reference converts out of
 * range, which is undefined). */
static int ext_steps(float ext, float thr) {
  const double q = ceil((double)(ext / thr));
  return q >= 64.0 ? 64 : (q >= 0.0 ? (int)q : 0);
}

static int pow2roundup(int x) { /* split.pikostage:34-45 */
  if (x < 0) return 0;
  --x;
  x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16;
  return x + 1;
}

/* Bernstein basis and derivatives, dice.pikostage:85-93 */
static float B0(float u) { return ((1.0f - u) * (1.0f - u) * (1.0f - u)); }
static float B1(float u) { return ((3.0f * u) * (1.0f - u) * (1.0f - u)); }
static float B2(float u) { return ((3.0f * u) * (u - u * u)); }
static float B3(float u) { return ((u) * (u * u)); }
static float dB0(float u) { return (-3.0f * (1.0f - u) * (1.0f - u)); }
static float dB1(float u) { return (3.0f * (1.0f - u) * (1.0f - 3.0f * u)); }
static float dB2(float u) { return (3.0f * u * (2.0f - 3.0f * u)); }
static float dB3(float u) { return (3.0f * u * u); }

static v4 bez(float t, v4 a, v4 b, v4 c, v4 d) { /* bez2D, dice.pikostage:96-101 */
  v4 o;
  o.x = B0(t) * a.x + B1(t) * b.x + B2(t) * c.x + B3(t) * d.x;
  o.y = B0(t) * a.y + B1(t) * b.y + B2(t) * c.y + B3(t) * d.y;
  o.z = B0(t) * a.z + B1(t) * b.z + B2(t) * c.z + B3(t) * d.z;
  o.w = B0(t) * a.w + B1(t) * b.w + B2(t) * c.w + B3(t) * d.w;
  return o;
}
static v4 dbez(float t, v4 a, v4 b, v4 c, v4 d) { /* dbez2D, dice.pikostage:104-109 */
  v4 o;
  o.x = dB0(t) * a.x + dB1(t) * b.x + dB2(t) * c.x + dB3(t) * d.x;
  o.y = dB0(t) * a.y + dB1(t) * b.y + dB2(t) * c.y + dB3(t) * d.y;
  o.z = dB0(t) * a.z + dB1(t) * b.z + dB2(t) * c.z + dB3(t) * d.z;
  o.w = dB0(t) * a.w + dB1(t) * b.w + dB2(t) * c.w + dB3(t) * d.w;
  return o;
}
static v4 patch_eval(const v4 *P, float u, float v) { /* dice.pikostage:111-121 */
  const v4 r0 = bez(u, P[0], P[1], P[2], P[3]), r1 = bez(u, P[4], P[5], P[6], P[7]);
  const v4 r2 = bez(u, P[8], P[9], P[10], P[11]), r3 = bez(u, P[12], P[13], P[14], P[15]);
  return bez(v, r0, r1, r2, r3);
}
static v4 norm_eval(const v4 *P, float u, float v) { /* dice.pikostage:123-146 */
  v4 r0 = dbez(u, P[0], P[1], P[2], P[3]), r1 = dbez(u, P[4], P[5], P[6], P[7]);
  v4 r2 = dbez(u, P[8], P[9], P[10], P[11]), r3 = dbez(u, P[12], P[13], P[14], P[15]);
  const v4 a = bez(v, r0, r1, r2, r3);
  r0 = bez(u, P[0], P[1], P[2], P[3]);
  r1 = bez(u, P[4], P[5], P[6], P[7]);
  r2 = bez(u, P[8], P[9], P[10], P[11]);
  r3 = bez(u, P[12], P[13], P[14], P[15]);
  const v4 b = dbez(v, r0, r1, r2, r3);
  /* piko::cross (math.h:420-427) then normalize (math.h:344-353) */
  float x = a.y * b.z - a.z * b.y, y = a.z * b.x - a.x * b.z, z = a.x * b.y - a.y * b.x;
  const float mag = sqrtf(x * x + y * y + z * z);
  x /= mag;
  y /= mag;
  z /= mag;
  v4 o = {x, y, z, 1.0f};
  return o;
}

/* phongShadeReyes + computeLighting (shade.pikostage:22-30, 64-85).  The call
 * passes (normal, matcol, lightvec) into the macro's (_mynor, _lightvec,
 * _matcol), so the material colour acts as the light vector and vice versa. */
static void phong(float nx, float ny, float nz, float *o) {
  nx = -nx; ny = -ny; nz = -nz;
  const float matcol[3] = {0.10f, 0.35f, 0.85f};
  const float lightvec[3] = {0.1773f, 0.5773f, 0.973f};
  float d = nx * matcol[0] + ny * matcol[1] + nz * matcol[2];
  d = d < 0.0f ? 0.0f : d;
  for (int c = 0; c < 3; ++c) {
    float p = 0.2f + d * lightvec[c];
    o[c] = p > 1.0f ? 1.0f : (p < 0.0f ? 0.0f : p);
  }
}

typedef oracle_upoly upoly;

/* Dice one corner: patchEval, normEval, optional displacement, projection
 * (dice.pikostage:206-229). */
static void corner(const v4 *P, float u, float v, const float *vp, int W, int H, float disp, float *sx,
                   float *sy, float *sz, float *n) {
  v4 p = patch_eval(P, u, v);
  const v4 nn = norm_eval(P, u, v);
  if (disp != 0.0f) {
    const float d = disp * (p.x * p.x - p.y * p.y);
    p.x = p.x + d * nn.x;
    p.y = p.y + d * nn.y;
    p.z = p.z + d * nn.z;
  }
  const v4 s = mat4(vp, p);
  const float onebyw = 1.0f / s.w;
  *sx = ((s.x * onebyw + 1.0f) * (float)W) * 0.5f;
  *sy = ((s.y * onebyw + 1.0f) * (float)H) * 0.5f;
  *sz = (s.z * onebyw);
  n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
}

static void make_upoly(const v4 *P, int nu, int nv, int k, int j, const float *vp, int W, int H, float disp,
                       upoly *o) {
  const float du = 1.0f / ((float)nu), dv = 1.0f / ((float)nv);
  const float myU = (k) * du, myV = (j) * dv;
  const float us[4] = {myU, myU + du, myU + du, myU};
  const float vs[4] = {myV, myV, myV + dv, myV + dv};
  float n[4][3];
  for (int c = 0; c < 4; ++c) {
    float zc;
    corner(P, us[c], vs[c], vp, W, H, disp, &o->x[c], &o->y[c], &zc, n[c]);
    if (c == 0) o->z = zc; else o->z = o->z + zc;  /* ((z0+z1)+z2)+z3 */
  }
  o->z = o->z * 0.25f;                               /* shade.pikostage:98 */
  float col[3] = {0, 0, 0};
  for (int c = 0; c < 4; ++c) {                       /* shade.pikostage:108-138 */
    float ci[3];
    phong(n[c][0], n[c][1], n[c][2], ci);
    for (int q = 0; q < 3; ++q) col[q] = c == 0 ? ci[q] : col[q] + ci[q];
  }
  for (int q = 0; q < 3; ++q) o->rgb[q] = fmn(col[q], 4.0f) * 0.25f;
}

/* Two 4-edge fan tests at the sample point (shade.pikostage:176-218): hit if
 * either orientation has all edge functions > 0. */
static int quad_hit(const upoly *q, float px, float py) {
  const float *X = q->x, *Y = q->y;
  int r1 = 1, r2 = 1;
  static const int e1[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int e2[4][2] = {{0, 3}, {3, 2}, {2, 1}, {1, 0}};
  for (int e = 0; e < 4; ++e) {
    const int a = e1[e][0], b = e1[e][1];
    const float t = ((py - Y[a]) * (X[b] - X[a]) - (px - X[a]) * (Y[b] - Y[a]));
    if (t <= 0.0f) r1 = 0;
  }
  for (int e = 0; e < 4; ++e) {
    const int a = e2[e][0], b = e2[e][1];
    const float t = ((py - Y[a]) * (X[b] - X[a]) - (px - X[a]) * (Y[b] - Y[a]));
    if (t <= 0.0f) r2 = 0;
  }
  return r1 + r2;
}

static void sample_offsets(int spp, int s, float *ox, float *oy) {
  if (spp == 1) { *ox = 0.5f; *oy = 0.5f; return; }
  *ox = (s & 1) ? 0.75f : 0.25f;
  *oy = (s & 2) ? 0.75f : 0.25f;
}

long long oracle_reyes_split(const float *patches, int np, int W, int H, const float *vp, float threshold,
                             int max_splits, float *leaves, int *leaf_nu, int *leaf_nv, long long cap) {
  long long ncur = np, nleaf = 0;
  v4 *cur = (v4 *)malloc(sizeof(v4) * 16 * (size_t)(np > 0 ? np : 1));
  int *cur_splits = (int *)calloc((size_t)(np > 0 ? np : 1), sizeof(int));
  memcpy(cur, patches, sizeof(v4) * 16 * (size_t)np);
  float cl[64];
  while (ncur > 0) {
    v4 *nxt = (v4 *)malloc(sizeof(v4) * 16 * (size_t)(2 * ncur));
    int *nxt_splits = (int *)malloc(sizeof(int) * (size_t)(2 * ncur));
    long long nn = 0;
    for (long long i = 0; i < ncur; ++i) {
      const v4 *cp = cur + 16 * i;
      project_patch(cp, vp, W, H, cl);
      float us, vs;
      extents(cl, &us, &vs);
      int split = (us >= threshold) || (vs >= threshold);   /* split.pikostage:283-285 */
      if (cur_splits[i] >= max_splits) split = 0;
      if (split) {
        split_dir(cp, nxt + 16 * nn, nxt + 16 * (nn + 1), us >= threshold);
        nxt_splits[nn] = nxt_splits[nn + 1] = cur_splits[i] + 1;
        nn += 2;
      } else {
        if (nleaf < cap) {
          memcpy(leaves + 64 * nleaf, cp, sizeof(v4) * 16);
          int nu = pow2roundup(ext_steps(us, threshold) + 1);  /* split.pikostage:307-314 */
          int nv = pow2roundup(ext_steps(vs, threshold) + 1);
          nu = nu < 2 ? 2 : (nu > 32 ? 32 : nu);
          nv = nv < 2 ? 2 : (nv > 32 ? 32 : nv);
          if (leaf_nu) leaf_nu[nleaf] = nu;
          if (leaf_nv) leaf_nv[nleaf] = nv;
        }
        ++nleaf;
      }
    }
    free(cur);
    free(cur_splits);
    cur = nxt;
    cur_splits = nxt_splits;
    ncur = nn;
  }
  free(cur);
  free(cur_splits);
  return nleaf;
}

static int clampi(float f) { /* float -> int, saturating (the reference conversion is UB outside int) */
  if (!(f > -1073741824.0f)) return -1073741824;
  if (f > 1073741824.0f) return 1073741824;
  return (int)f;
}

/* Pixels a micropolygon is tested on: ShadeStage::assignBin puts it in the
 * 64-px bins overlapping the bounding box of its first THREE corners
 * (shade.pikostage:51-56 with computeBoundingBox/assignToBB,
 * dice.pikostage:18-43), and each bin tests the pixels of the four-corner
 * bounding box inside the bin (shade.pikostage:140-163).  Bins per axis are
 * ceil(W/64) with pixels clipped to the screen (the reference floors). */
static int upoly_pixels(const upoly *q, int W, int H, int *x0, int *x1, int *y0, int *y1) {
  const int nbx = (W + 63) / 64, nby = (H + 63) / 64;
  const float lo3x = floorf(fmn(q->x[0], fmn(q->x[1], q->x[2])));
  const float lo3y = floorf(fmn(q->y[0], fmn(q->y[1], q->y[2])));
  const float hi3x = ceilf(fmx(q->x[0], fmx(q->x[1], q->x[2])));
  const float hi3y = ceilf(fmx(q->y[0], fmx(q->y[1], q->y[2])));
  int bx1 = clampi(floorf(lo3x / 64.0f)), bx2 = clampi(floorf(hi3x / 64.0f));
  int by1 = clampi(floorf(lo3y / 64.0f)), by2 = clampi(floorf(hi3y / 64.0f));
  if (bx1 < 0) bx1 = 0;
  if (by1 < 0) by1 = 0;
  if (bx2 >= nbx) bx2 = nbx - 1;
  if (by2 >= nby) by2 = nby - 1;
  if (bx1 > bx2 || by1 > by2) return 0;
  const float b4x0 = floorf(fmn(fmn(q->x[0], q->x[1]), fmn(q->x[2], q->x[3])));
  const float b4y0 = floorf(fmn(fmn(q->y[0], q->y[1]), fmn(q->y[2], q->y[3])));
  const float b4x1 = ceilf(fmx(fmx(q->x[0], q->x[1]), fmx(q->x[2], q->x[3])));
  const float b4y1 = ceilf(fmx(fmx(q->y[0], q->y[1]), fmx(q->y[2], q->y[3])));
  int a = clampi(b4x0), b = clampi(b4x1), c = clampi(b4y0), d = clampi(b4y1);
  if (a < bx1 * 64) a = bx1 * 64;
  if (c < by1 * 64) c = by1 * 64;
  if (b > (bx2 + 1) * 64) b = (bx2 + 1) * 64;
  if (d > (by2 + 1) * 64) d = (by2 + 1) * 64;
  if (b > W) b = W;
  if (d > H) d = H;
  *x0 = a; *x1 = b; *y0 = c; *y1 = d;
  return a < b && c < d;
}

long long oracle_reyes_dice(const float *leaves, const int *leaf_nu, const int *leaf_nv, long long nleaf, int dice,
                            int W, int H, const float *vp, float disp, oracle_upoly *out, long long cap) {
  long long nup = 0;
  for (long long l = 0; l < nleaf; ++l) {
    const v4 *P = (const v4 *)(leaves + 64 * l);
    const int nu = dice > 0 ? dice : leaf_nu[l], nv = dice > 0 ? dice : leaf_nv[l];
    const int ku = dice > 0 ? nu - 1 : nu, kv = dice > 0 ? nv - 1 : nv;  /* reference: k <= nu */
    for (int k = 0; k <= ku; ++k)
      for (int j = 0; j <= kv; ++j, ++nup)
        if (nup < cap) make_upoly(P, nu, nv, k, j, vp, W, H, disp, out + nup);
  }
  return nup;
}

void oracle_reyes_shade(const oracle_upoly *up, long long nup, int W, int H, int spp, float *rgba,
                        int32_t *sample_z) {
  const size_t ns = (size_t)W * H * spp;
  int32_t *zb = (int32_t *)malloc(ns * 4 + 4);
  float *cb = (float *)malloc(ns * 3 * 4 + 4);
  float clear[3];
  oracle_from_abgr(0xff663313u, clear);           /* PikoScreen clear colour, stage.h:83-85 */
  for (size_t s = 0; s < ns; ++s) {
    zb[s] = fbits(1.0f);                           /* resetDepthBuffer, reyesPipe/main.cpp:324-329 */
    for (int c = 0; c < 3; ++c) cb[3 * s + c] = clear[c];
  }
  for (long long u = 0; u < nup; ++u) {
    const upoly *q = up + u;
    int x0, x1, y0, y1;
    if (!upoly_pixels(q, W, H, &x0, &x1, &y0, &y1)) continue;
    const int32_t zi = fbits(q->z);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x)
        for (int s = 0; s < spp; ++s) {
          float ox, oy;
          sample_offsets(spp, s, &ox, &oy);
          if (!quad_hit(q, (float)x + ox, (float)y + oy)) continue;
          const size_t si = ((size_t)y * W + x) * spp + s;
          if (zi <= zb[si]) {  /* atomicMin loop + (zBuffer == z), shade.pikostage:226-241 */
            zb[si] = zi;
            for (int c = 0; c < 3; ++c) cb[3 * si + c] = q->rgb[c];
          }
        }
  }
  for (size_t p = 0; p < (size_t)W * H; ++p) {
    for (int c = 0; c < 3; ++c) {
      float a = cb[3 * (p * spp) + c];
      for (int s = 1; s < spp; ++s) a = a + cb[3 * (p * spp + s) + c];
      rgba[4 * p + c] = a * (1.0f / (float)spp);
    }
    rgba[4 * p + 3] = 1.0f;
  }
  if (sample_z) memcpy(sample_z, zb, ns * 4);
  free(zb);
  free(cb);
}

long long oracle_reyes_frame(const float *patches, int np, int W, int H, const float *vp, float threshold,
                             int dice, int spp, float disp, float *rgba, int32_t *sample_z) {
  const long long nleaf = oracle_reyes_split(patches, np, W, H, vp, threshold, 20, NULL, NULL, NULL, 0);
  float *leaves = (float *)malloc(sizeof(float) * 64 * (size_t)(nleaf + 1));
  int *lnu = (int *)malloc(sizeof(int) * (size_t)(nleaf + 1));
  int *lnv = (int *)malloc(sizeof(int) * (size_t)(nleaf + 1));
  oracle_reyes_split(patches, np, W, H, vp, threshold, 20, leaves, lnu, lnv, nleaf);
  const long long nup = oracle_reyes_dice(leaves, lnu, lnv, nleaf, dice, W, H, vp, disp, NULL, 0);
  oracle_upoly *up = (oracle_upoly *)malloc(sizeof(oracle_upoly) * (size_t)(nup + 1));
  oracle_reyes_dice(leaves, lnu, lnv, nleaf, dice, W, H, vp, disp, up, nup);
  oracle_reyes_shade(up, nup, W, H, spp, rgba, sample_z);
  free(up);
  free(leaves);
  free(lnu);
  free(lnv);
  return nup;
}
