/*
 * piko_oracle.c -- CPU restatement of the Piko reference CPU backend for the
 * fixed-point raster pipeline.  TEST INFRASTRUCTURE ONLY (the checker); the
 * product never links or calls it.  See piko_oracle.h for the deviations
 * from the reference and their reasons.
 *
 * Build: gcc -O2 -ffp-contract=off -fwrapv -fPIC -shared (oracle/Makefile).
 * -ffp-contract=off keeps every a*b+c unfused, as the reference CPU branch
 * writes it (math.h:258-289); SSE float arithmetic (x86-64 default) rounds
 * every operation to binary32.
 */
#include "piko_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define CLEAR_COLOR 0xff663313u  /* PikoScreen::allocate, stage.h:83-85 */
#define DEPTH_CLEAR 1.0f         /* resetDepthBuffer, main.cpp:402-409 */

static inline int32_t f2i_bits(float f) { int32_t i; memcpy(&i, &f, 4); return i; }
static inline float i2f_bits(int32_t i) { float f; memcpy(&f, &i, 4); return f; }
static inline int imin(int a, int b) { return a < b ? a : b; }   /* math.h:215-216 */
static inline int imax(int a, int b) { return a > b ? a : b; }

/* (int)(v * 2^S) of vertexShader.pikostage:46-53, with the out-of-range
 * clamp described in the header. */
int32_t oracle_fixpt(float v, int S) {
  float s = v * (float)(1 << S);
  s = fminf(s, 536870912.0f);
  s = fmaxf(s, -536870912.0f);
  return (int32_t)s;
}

/* CPU toABGR, math.h:286-289: truncating conversion, alpha 255. */
uint32_t oracle_to_abgr(float r, float g, float b) {
  return (255u << 24) | ((unsigned)(b * 255.0f) << 16) |
         ((unsigned)(g * 255.0f) << 8) | (unsigned)(r * 255.0f);
}

/* fromABGR, math.h:323-330. */
void oracle_from_abgr(uint32_t c, float *rgb) {
  rgb[0] = (float)((c >> 0) & 255) / 255.0f;
  rgb[1] = (float)((c >> 8) & 255) / 255.0f;
  rgb[2] = (float)((c >> 16) & 255) / 255.0f;
}

/* isFrontFacingFixPt (rasterMacros.h:23-30), isInsideFrustumFixPt (32-46),
 * computeBoundingBoxFixPt (51-56), isBBBetweenSamples (65-77, which tests
 * the Y axis twice: the X-axis cull is computed and unused). */
static int vs_keep(const int32_t *p, int W, int H, int S) {
  const int64_t x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3], x2 = p[4], y2 = p[5];
  const int64_t area = (y2 - y0) * (x1 - x0) - (x2 - x0) * (y1 - y0);
  if (!(area > 0)) return 0;
  const int32_t guard = 225 << S;
  const int32_t Wf = W << S, Hf = H << S;
  for (int k = 0; k < 3; ++k) {
    const int32_t x = p[2 * k], y = p[2 * k + 1];
    if (x < -guard || x >= Wf + guard) return 0;
    if (y < -guard || y >= Hf + guard) return 0;
  }
  const int32_t lo_y = imin(imin(p[1], p[3]), p[5]);
  const int32_t hi_y = imax(imax(p[1], p[3]), p[5]);
  const int32_t half = 1 << (S - 1);
  const int cull_y = ((lo_y - half) >> S) == ((hi_y - half) >> S);
  return !cull_y;
}

/* vertexShader.pikostage:36-109 with VTX_PRETRANSFORM and GORAUD. */
int oracle_vs_pretransformed(const float *pos, const uint32_t *col, int n,
                             int W, int H, int S, const oracle_stri *outp) {
  const oracle_stri out = *outp;
  int emitted = 0;
  for (int i = 0; i < n; ++i) {
    int32_t *p = out.xy + 6 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      p[2 * k] = oracle_fixpt(pos[9 * (size_t)i + 3 * k + 0], S);
      p[2 * k + 1] = oracle_fixpt(pos[9 * (size_t)i + 3 * k + 1], S);
      out.z[3 * (size_t)i + k] = pos[9 * (size_t)i + 3 * k + 2];
      out.col[3 * (size_t)i + k] = col[3 * (size_t)i + k];
    }
    out.keep[i] = (uint8_t)vs_keep(p, W, H, S);
    emitted += out.keep[i];
  }
  return emitted;
}

/* computeLightingHost + saturatePixelHost + toABGRHost (main.cpp:310-334). */
static uint32_t light_vertex(const float *nrm, const float *l, const float *m,
                             const float *amb) {
  float d = nrm[0] * l[0] + nrm[1] * l[1] + nrm[2] * l[2];
  d = d < 0.0f ? 0.0f : d;
  float o[3];
  for (int c = 0; c < 3; ++c) {
    float v = d * m[c] + amb[c];
    o[c] = v > 1.0f ? 1.0f : (v < 0.0f ? 0.0f : v);
  }
  return oracle_to_abgr(o[0], o[1], o[2]);
}

int oracle_vs_world(const float *pos, const float *nrm, int n, int W, int H,
                    int S, const float *m, const float *light,
                    const float *matcol, const float *ambient,
                    const oracle_stri *outp) {
  const oracle_stri out = *outp;
  const float halfW = 0.5f * (float)W, halfH = 0.5f * (float)H; /* main.cpp:104-105 */
  int emitted = 0;
  for (int i = 0; i < n; ++i) {
    int32_t *p = out.xy + 6 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      const float *v = pos + 9 * (size_t)i + 3 * k;
      /* matmultfloat4, math.h:365-373 (left-to-right sums, unfused) */
      float cx = m[0] * v[0] + m[4] * v[1] + m[8] * v[2] + m[12];
      float cy = m[1] * v[0] + m[5] * v[1] + m[9] * v[2] + m[13];
      float cz = m[2] * v[0] + m[6] * v[1] + m[10] * v[2] + m[14];
      float cw = m[3] * v[0] + m[7] * v[1] + m[11] * v[2] + m[15];
      /* vtxTransform, rasterMacros.h:14-21 (rcp_approx = 1.0f/a, math.h:279) */
      float onebyw = 1.0f / cw;
      float sx = (cx * onebyw + 1.0f) * halfW;
      float sy = (cy * onebyw + 1.0f) * halfH;
      float sz = (cz * onebyw);
      p[2 * k] = oracle_fixpt(sx, S);
      p[2 * k + 1] = oracle_fixpt(sy, S);
      out.z[3 * (size_t)i + k] = sz;
      out.col[3 * (size_t)i + k] =
          light_vertex(nrm + 9 * (size_t)i + 3 * k, light, matcol, ambient);
    }
    out.keep[i] = (uint8_t)vs_keep(p, W, H, S);
    emitted += out.keep[i];
  }
  return emitted;
}

/* computePixelBoundingBoxFixPt with RoundDown/RoundUpFixPt
 * (rasterMacros.h:48-49, 58-63). */
static void pixel_bbox(const int32_t *p, int32_t *lo_x, int32_t *lo_y,
                       int32_t *hi_x, int32_t *hi_y) {
  *lo_x = imin(imin(p[0], p[2]), p[4]) & (int32_t)0xffffff00;
  *lo_y = imin(imin(p[1], p[3]), p[5]) & (int32_t)0xffffff00;
  *hi_x = (imax(imax(p[0], p[2]), p[4]) + 0xff) & (int32_t)0xffffff00;
  *hi_y = (imax(imax(p[1], p[3]), p[5]) + 0xff) & (int32_t)0xffffff00;
}

static int ilog2i(int t) { int l = 0; while ((1 << l) < t) ++l; return l; }

/* Enumerates the bins RasterStage::assignBin assigns (raster.pikostage:
 * 40-117), in the walk order (rows by, then columns bx), calling emit. */
typedef void (*bin_emit_fn)(void *ctx, int tri, int bin);
static void assign_bins(const int32_t *p, int tri, int nBx, int nBy, int S,
                        int T, bin_emit_fn emit, void *ctx) {
  int32_t lo_x, lo_y, hi_x, hi_y;
  pixel_bbox(p, &lo_x, &lo_y, &hi_x, &hi_y);
  const int bsb = S + ilog2i(T);
  int bx1 = lo_x >> bsb, bx2 = hi_x >> bsb;
  int by1 = lo_y >> bsb, by2 = hi_y >> bsb;
  bx1 = imax(0, bx1);
  by1 = imax(0, by1);
  bx2 = imin(bx2, nBx - 1);
  by2 = imin(by2, nBy - 1);
  const int64_t x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3], x2 = p[4], y2 = p[5];
  const int64_t x1mx0 = x1 - x0, x2mx1 = x2 - x1, x0mx2 = x0 - x2;
  const int64_t y1my0 = y1 - y0, y2my1 = y2 - y1, y0my2 = y0 - y2;
  const int64_t startx = (int64_t)bx1 << bsb, starty = (int64_t)by1 << bsb;
  const int64_t s0x = -(y1my0 << bsb), s1x = -(y2my1 << bsb), s2x = -(y0my2 << bsb);
  const int64_t s0y = x1mx0 << bsb, s1y = x2mx1 << bsb, s2y = x0mx2 << bsb;
  /* GetTrivialRejectOffset, rasterMacros.h:152-160 */
  const int64_t tr0 = (y1my0 < 0 ? s0x : 0) + (x1mx0 >= 0 ? s0y : 0);
  const int64_t tr1 = (y2my1 < 0 ? s1x : 0) + (x2mx1 >= 0 ? s1y : 0);
  const int64_t tr2 = (y0my2 < 0 ? s2x : 0) + (x0mx2 >= 0 ? s2y : 0);
  int64_t r0 = tr0 + (starty - y0) * x1mx0 - (startx - x0) * y1my0;
  int64_t r1 = tr1 + (starty - y1) * x2mx1 - (startx - x1) * y2my1;
  int64_t r2 = tr2 + (starty - y2) * x0mx2 - (startx - x2) * y0my2;
  for (int by = by1; by <= by2; ++by) {
    int64_t e0 = r0, e1 = r1, e2 = r2;
    for (int bx = bx1; bx <= bx2; ++bx) {
      if ((e0 | e1 | e2) >= 0) emit(ctx, tri, by * nBx + bx);
      e0 += s0x; e1 += s1x; e2 += s2x;
    }
    r0 += s0y; r1 += s1y; r2 += s2y;
  }
}

typedef struct { int64_t *cnt; int32_t *tris; int64_t *cursor; long long total; long long cap; } bin_ctx;
static void emit_count(void *c, int tri, int bin) {
  bin_ctx *b = (bin_ctx *)c; (void)tri; b->cnt[bin]++; b->total++;
}
static void emit_fill(void *c, int tri, int bin) {
  bin_ctx *b = (bin_ctx *)c; b->tris[b->cursor[bin]++] = tri;
}

long long oracle_bin(const oracle_stri *inp, int n, int W, int H, int S, int T,
                     int64_t *bin_start, int32_t *bin_tris, long long cap) {
  const oracle_stri in = *inp;
  const int nBx = (W + T - 1) / T, nBy = (H + T - 1) / T, nb = nBx * nBy;
  bin_ctx c;
  c.cnt = (int64_t *)calloc((size_t)nb, sizeof(int64_t));
  c.total = 0;
  c.cap = cap;
  /* kernel0 + VertexShader.process run serially in submission order, so each
   * raster bin's queue holds its triangles in increasing id order. */
  for (int i = 0; i < n; ++i)
    if (in.keep[i]) assign_bins(in.xy + 6 * (size_t)i, i, nBx, nBy, S, T, emit_count, &c);
  bin_start[0] = 0;
  for (int b = 0; b < nb; ++b) bin_start[b + 1] = bin_start[b] + c.cnt[b];
  long long total = c.total;
  if (total > cap) { free(c.cnt); return -1; }
  c.cursor = (int64_t *)malloc((size_t)nb * sizeof(int64_t));
  memcpy(c.cursor, bin_start, (size_t)nb * sizeof(int64_t));
  c.tris = bin_tris;
  for (int i = 0; i < n; ++i)
    if (in.keep[i]) assign_bins(in.xy + 6 * (size_t)i, i, nBx, nBy, S, T, emit_fill, &c);
  free(c.cursor);
  free(c.cnt);
  return total;
}

/* One fragment of RasterStage::process (raster.pikostage:257-319):
 * barycentrics, z interpolation, depth atomicMin loop (single-threaded, so
 * the loop is: pass iff z <= stored as signed int bits), Gouraud colour and
 * the PikoScreen store (stage.h:57-61). */
typedef struct {
  int64_t x0, y0, x1, y1, x2, y2;
  int64_t x1mx0, x2mx1, x0mx2, y1my0, y2my1, y0my2;
  float onebybary, z2, z0mz2, z1mz2;
  float vcol2[3], d0[3], d1[3];
} tri_setup;

static void setup_tri(const oracle_stri in, int i, tri_setup *t) {
  const int32_t *p = in.xy + 6 * (size_t)i;
  t->x0 = p[0]; t->y0 = p[1]; t->x1 = p[2]; t->y1 = p[3]; t->x2 = p[4]; t->y2 = p[5];
  t->x1mx0 = t->x1 - t->x0; t->x2mx1 = t->x2 - t->x1; t->x0mx2 = t->x0 - t->x2;
  t->y1my0 = t->y1 - t->y0; t->y2my1 = t->y2 - t->y1; t->y0my2 = t->y0 - t->y2;
  const int64_t barydenom = -t->y0my2 * t->x1mx0 + t->x0mx2 * t->y1my0; /* :173 */
  t->onebybary = 1.0f / (float)barydenom;                                  /* :174 */
  const float *z = in.z + 3 * (size_t)i;
  t->z2 = z[2];
  t->z0mz2 = z[0] - z[2];                                                  /* :251 */
  t->z1mz2 = z[1] - z[2];
  float c0[3], c1[3], c2[3];
  oracle_from_abgr(in.col[3 * (size_t)i + 0], c0);                        /* :183-185 */
  oracle_from_abgr(in.col[3 * (size_t)i + 1], c1);
  oracle_from_abgr(in.col[3 * (size_t)i + 2], c2);
  for (int c = 0; c < 3; ++c) {
    t->vcol2[c] = c2[c];
    t->d0[c] = c0[c] - c2[c];                                              /* :254-255 */
    t->d1[c] = c1[c] - c2[c];
  }
}

static inline void shade_pixel(const tri_setup *t, int px, int py, int S, int W,
                               uint32_t *color, int32_t *zbits) {
  const int64_t half = 1 << (S - 1);
  const int64_t cx = ((int64_t)px << S) + half, cy = ((int64_t)py << S) + half;
  const int64_t e0 = (cy - t->y0) * t->x1mx0 - (cx - t->x0) * t->y1my0;
  const int64_t e1 = (cy - t->y1) * t->x2mx1 - (cx - t->x1) * t->y2my1;
  const int64_t e2 = (cy - t->y2) * t->x0mx2 - (cx - t->x2) * t->y0my2;
  if ((e0 | e1 | e2) < 0) return;                                          /* :228 */
  const float alpha = (float)e1 * t->onebybary;                            /* :269 */
  const float beta = (float)e2 * t->onebybary;                             /* :270 */
  const float z = t->z2 + alpha * t->z0mz2 + beta * t->z1mz2;              /* :29-32, :272 */
  const int32_t zi = f2i_bits(z);
  const size_t pid = (size_t)py * (size_t)W + (size_t)px;                  /* :274 */
  const int32_t old = zbits[pid];                                          /* :280-285 */
  zbits[pid] = old < zi ? old : zi;
  if (!(old >= zi)) return;
  float c[3];
  for (int k = 0; k < 3; ++k) c[k] = t->vcol2[k] + alpha * t->d0[k] + beta * t->d1[k]; /* :291-293 */
  color[pid] = oracle_to_abgr(c[0], c[1], c[2]);                           /* :309, stage.h:57-61 */
}

/* RasterStage::process for one (triangle, bin): pixels of
 * bbox(rounded) ∩ bin ∩ screen, in raster order (the reference walks the
 * coverage mask low bit first, i.e. row-major; fragments of one triangle
 * never share a pixel so their order is immaterial). */
static void raster_tri_bin(const tri_setup *t, const int32_t *p, int bx, int by,
                           int W, int H, int S, int T, uint32_t *color,
                           int32_t *zbits) {
  int32_t lo_x, lo_y, hi_x, hi_y;
  pixel_bbox(p, &lo_x, &lo_y, &hi_x, &hi_y);
  int xb = imax(lo_x >> S, bx * T), xe = imin(imin(hi_x >> S, bx * T + T), W);
  int yb = imax(lo_y >> S, by * T), ye = imin(imin(hi_y >> S, by * T + T), H);
  for (int py = yb; py < ye; ++py)
    for (int px = xb; px < xe; ++px) shade_pixel(t, px, py, S, W, color, zbits);
}

typedef struct {
  const oracle_stri *in; int W, H, S, T, nBx; const int64_t *bs; const int32_t *bt;
  uint32_t *color; int32_t *zbits; int b0, b1;
} raster_job;

static void *raster_bins(void *arg) {
  raster_job *j = (raster_job *)arg;
  for (int b = j->b0; b < j->b1; ++b) {
    const int bx = b % j->nBx, by = b / j->nBx;
    for (int64_t k = j->bs[b]; k < j->bs[b + 1]; ++k) {
      const int i = j->bt[k];
      tri_setup t;
      setup_tri(*j->in, i, &t);
      raster_tri_bin(&t, j->in->xy + 6 * (size_t)i, bx, by, j->W, j->H, j->S,
                     j->T, j->color, j->zbits);
    }
  }
  return NULL;
}

static void clear_fb(uint32_t *color, float *depth, size_t np) {
  for (size_t k = 0; k < np; ++k) { color[k] = CLEAR_COLOR; depth[k] = DEPTH_CLEAR; }
}

/* Process kernel over bins.  threads>1 block-partitions the bins over host
 * threads as CPUBackend.cpp:534-557 does; bins own disjoint pixels so the
 * result does not depend on the thread count. */
static void raster_threads(const oracle_stri in, int W, int H, int S, int T,
                           const int64_t *bin_start, const int32_t *bin_tris,
                           uint32_t *color, float *depth, int threads) {
  const int nBx = (W + T - 1) / T, nBy = (H + T - 1) / T, nb = nBx * nBy;
  clear_fb(color, depth, (size_t)W * H);
  int32_t *zbits = (int32_t *)depth; /* the reference's int view of zBuffer, :277 */
  if (threads < 1) threads = 1;
  raster_job *jobs = (raster_job *)calloc((size_t)threads, sizeof(raster_job));
  pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  const int per = (nb + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    raster_job j = {&in, W, H, S, T, nBx, bin_start, bin_tris, color, zbits,
                    imin(nb, t * per), imin(nb, (t + 1) * per)};
    jobs[t] = j;
  }
  if (threads == 1) raster_bins(&jobs[0]);
  else {
    for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, raster_bins, &jobs[t]);
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  }
  free(th);
  free(jobs);
}

void oracle_raster(const oracle_stri *in, int W, int H, int S, int T,
                   const int64_t *bin_start, const int32_t *bin_tris,
                   uint32_t *color, float *depth) {
  raster_threads(*in, W, H, S, T, bin_start, bin_tris, color, depth, 1);
}

void oracle_raster_bruteforce(const oracle_stri *inp, int n, int W, int H, int S,
                              uint32_t *color, float *depth) {
  const oracle_stri in = *inp;
  clear_fb(color, depth, (size_t)W * H);
  int32_t *zbits = (int32_t *)depth;
  for (int i = 0; i < n; ++i) {
    if (!in.keep[i]) continue;
    tri_setup t;
    setup_tri(in, i, &t);
    for (int py = 0; py < H; ++py)
      for (int px = 0; px < W; ++px) shade_pixel(&t, px, py, S, W, color, zbits);
  }
}

void oracle_fragment_stats(const oracle_stri *inp, int n, int W, int H, int S,
                           long long *fragments, long long *pixels_covered) {
  const oracle_stri in = *inp;
  uint8_t *hit = (uint8_t *)calloc((size_t)W * H, 1);
  long long frags = 0;
  const int64_t half = 1 << (S - 1);
  for (int i = 0; i < n; ++i) {
    if (!in.keep[i]) continue;
    tri_setup t;
    setup_tri(in, i, &t);
    int32_t lo_x, lo_y, hi_x, hi_y;
    pixel_bbox(in.xy + 6 * (size_t)i, &lo_x, &lo_y, &hi_x, &hi_y);
    int xb = imax(lo_x >> S, 0), xe = imin(hi_x >> S, W);
    int yb = imax(lo_y >> S, 0), ye = imin(hi_y >> S, H);
    for (int py = yb; py < ye; ++py)
      for (int px = xb; px < xe; ++px) {
        const int64_t cx = ((int64_t)px << S) + half, cy = ((int64_t)py << S) + half;
        const int64_t e0 = (cy - t.y0) * t.x1mx0 - (cx - t.x0) * t.y1my0;
        const int64_t e1 = (cy - t.y1) * t.x2mx1 - (cx - t.x1) * t.y2my1;
        const int64_t e2 = (cy - t.y2) * t.x0mx2 - (cx - t.x2) * t.y0my2;
        if ((e0 | e1 | e2) < 0) continue;
        ++frags;
        hit[(size_t)py * W + px] = 1;
      }
  }
  long long cov = 0;
  for (size_t k = 0; k < (size_t)W * H; ++k) cov += hit[k];
  free(hit);
  *fragments = frags;
  *pixels_covered = cov;
}

static long long frame_common(oracle_stri s, int n, int W, int H, int S, int T,
                              int threads, uint32_t *color, float *depth) {
  const int nBx = (W + T - 1) / T, nBy = (H + T - 1) / T, nb = nBx * nBy;
  int64_t *bs = (int64_t *)malloc(((size_t)nb + 1) * sizeof(int64_t));
  /* sizing call: fills bin_start, writes no list entries */
  oracle_bin(&s, n, W, H, S, T, bs, NULL, 0);
  const long long pairs = bs[nb];
  int32_t *bt = (int32_t *)malloc(((size_t)pairs + 1) * sizeof(int32_t));
  long long total = oracle_bin(&s, n, W, H, S, T, bs, bt, pairs);
  if (total >= 0) raster_threads(s, W, H, S, T, bs, bt, color, depth, threads);
  free(bt);
  free(bs);
  return total;
}

static oracle_stri alloc_stri(int n) {
  oracle_stri s;
  s.xy = (int32_t *)malloc((size_t)n * 6 * sizeof(int32_t) + 4);
  s.z = (float *)malloc((size_t)n * 3 * sizeof(float) + 4);
  s.col = (uint32_t *)malloc((size_t)n * 3 * sizeof(uint32_t) + 4);
  s.keep = (uint8_t *)malloc((size_t)n + 1);
  return s;
}
static void free_stri(oracle_stri s) { free(s.xy); free(s.z); free(s.col); free(s.keep); }

long long oracle_frame_pretransformed(const float *pos, const uint32_t *col,
                                      int n, int W, int H, int S, int T,
                                      int threads, uint32_t *color,
                                      float *depth) {
  oracle_stri s = alloc_stri(n);
  oracle_vs_pretransformed(pos, col, n, W, H, S, &s);
  long long r = frame_common(s, n, W, H, S, T, threads, color, depth);
  free_stri(s);
  return r;
}

long long oracle_frame_world(const float *pos, const float *nrm, int n, int W,
                             int H, int S, int T, const float *vp,
                             const float *light, const float *matcol,
                             const float *ambient, int threads,
                             uint32_t *color, float *depth) {
  oracle_stri s = alloc_stri(n);
  oracle_vs_world(pos, nrm, n, W, H, S, vp, light, matcol, ambient, &s);
  long long r = frame_common(s, n, W, H, S, T, threads, color, depth);
  free_stri(s);
  return r;
}
