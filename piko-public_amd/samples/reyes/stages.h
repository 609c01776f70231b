// Stages of the Reyes sample (the roles of reyesPipe/{split,dice,shade}
// .pikostage): Split loops on itself through out port 1 until a patch is
// small on screen, Dice turns each leaf into a grid of micropolygons, Shade
// tests the samples under each micropolygon and keeps the nearest.
//
// Settings (macros, read by pikoc_hip as the pipe's options):
//   THRESHOLD_PIXELS         split while a projected control-point row or
//                            column is at least this long (pixels)
//   PIKO_REYES_MAX_SPLITS    a patch split this many times is a leaf
//   PIKO_REYES_DICE          0: the leaf's own grid, (nu+1) x (nv+1) quads
//                            with nu, nv from its screen extents; N: N x N
//   PIKO_REYES_SPP           samples per pixel, 1 (centre) or 4 (2x2 strata)
//   PIKO_REYES_DISPLACEMENT  d: each diced point moves along its normal by
//                            d * (x*x - y*y)
#ifndef SAMPLE_REYES_STAGES_H
#define SAMPLE_REYES_STAGES_H

#include "reyesTypes.h"

#define SPLIT_BINSIZE 0
#define DICE_BINSIZE 64
#define SHADE_BINSIZE 64

#ifndef THRESHOLD_PIXELS
#define THRESHOLD_PIXELS 4.0f
#endif
#ifndef PIKO_REYES_MAX_SPLITS
#define PIKO_REYES_MAX_SPLITS 20
#endif
#ifndef PIKO_REYES_DICE
#define PIKO_REYES_DICE 0
#endif
#ifndef PIKO_REYES_SPP
#define PIKO_REYES_SPP 1
#endif
#ifndef PIKO_REYES_DISPLACEMENT
#define PIKO_REYES_DISPLACEMENT 0.0f
#endif

class SplitStage : public Stage<SPLIT_BINSIZE, SPLIT_BINSIZE, 256, piko_patch, piko_patch> {
#ifdef __PIKOC_DEVICE__
 public:
  void assignBin(piko_patch p) { this->assignToBin(p, 0); }

  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }

  void process(piko_patch p) {
    float sx[16], sy[16];
    for (int i = 0; i < 16; ++i) {
      cvec4f c = piko::transform4(constState.viewProjMatrix, p.CP[i]);
      if (c.w == 0.0f) c.w = 1.0f;
      sx[i] = (float)(((double)(c.x / c.w) + 1.0) * 0.5 * (double)constState.screenSizeX);
      sy[i] = (float)(((double)(c.y / c.w) + 1.0) * 0.5 * (double)constState.screenSizeY);
    }
    // longest projected control polyline along u (a row) and along v (a column)
    float extU = 0.0f, extV = 0.0f;
    for (int i = 0; i < 4; ++i) {
      extU = piko::fmax2(extU, polyline(sx, sy, 4 * i, 1));
      extV = piko::fmax2(extV, polyline(sx, sy, i, 4));
    }
    const bool big = extU >= THRESHOLD_PIXELS || extV >= THRESHOLD_PIXELS;
    if (big && p.numSplits < PIKO_REYES_MAX_SPLITS) {
      piko_patch lo = p, hi = p;
      halve(p, lo, hi, extU >= THRESHOLD_PIXELS);
      lo.numSplits = hi.numSplits = p.numSplits + 1;
      this->emit(lo, 1);
      this->emit(hi, 1);
    } else {
      p.nu = gridSize(extU);
      p.nv = gridSize(extV);
      this->emit(p, 0);
    }
  }

 private:
  static float polyline(const float *x, const float *y, int first, int stride) {
    float len = 0.0f;
    for (int k = 0; k < 3; ++k) {
      const int a = first + k * stride, b = a + stride;
      const float dx = x[a] - x[b], dy = y[a] - y[b];
      len = len + sqrtf(dx * dx + dy * dy);
    }
    return len;
  }

  static cvec4f mid(cvec4f a, cvec4f b) {
    const cvec4f m = {(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, (a.w + b.w) * 0.5f};
    return m;
  }

  // de Casteljau at t = 1/2 of every row (along u) or every column (along v)
  static void halve(const piko_patch &p, piko_patch &lo, piko_patch &hi, bool alongU) {
    const int step = alongU ? 1 : 4;
    for (int i = 0; i < 4; ++i) {
      const int c0 = alongU ? 4 * i : i, c1 = c0 + step, c2 = c1 + step, c3 = c2 + step;
      const cvec4f centre = mid(p.CP[c1], p.CP[c2]);
      lo.CP[c0] = p.CP[c0];
      lo.CP[c1] = mid(p.CP[c0], p.CP[c1]);
      lo.CP[c2] = mid(lo.CP[c1], centre);
      hi.CP[c3] = p.CP[c3];
      hi.CP[c2] = mid(p.CP[c2], p.CP[c3]);
      hi.CP[c1] = mid(hi.CP[c2], centre);
      lo.CP[c3] = mid(lo.CP[c2], hi.CP[c1]);
      hi.CP[c0] = lo.CP[c3];
    }
  }

  // the power of two at or above ceil(extent / threshold) + 1, in [2, 32]
  static int gridSize(float extent) {
    const double steps = ceil((double)(extent / THRESHOLD_PIXELS));
    int n = (steps >= 64.0 ? 64 : (steps >= 0.0 ? (int)steps : 0)) + 1;
    int pow2 = 1;
    while (pow2 < n) pow2 *= 2;
    return pow2 < 2 ? 2 : (pow2 > 32 ? 32 : pow2);
  }
#endif
};

class DiceStage : public Stage<DICE_BINSIZE, DICE_BINSIZE, 128, piko_patch, piko_upoly> {
#ifdef __PIKOC_DEVICE__
 public:
  void assignBin(piko_patch p) { this->assignToBin(p, (p.id / 128) % getNumBins()); }

  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }

  void process(piko_patch p) {
    const int nu = PIKO_REYES_DICE > 0 ? PIKO_REYES_DICE : p.nu;
    const int nv = PIKO_REYES_DICE > 0 ? PIKO_REYES_DICE : p.nv;
    // the leaf's own grid takes one row and one column of quads past the patch
    const int lastU = PIKO_REYES_DICE > 0 ? nu - 1 : nu, lastV = PIKO_REYES_DICE > 0 ? nv - 1 : nv;
    const float du = 1.0f / (float)nu, dv = 1.0f / (float)nv;
    for (int k = 0; k <= lastU; ++k)
      for (int j = 0; j <= lastV; ++j) {
        const float u = k * du, v = j * dv;
        const float cu[4] = {u, u + du, u + du, u}, cv[4] = {v, v, v + dv, v + dv};
        piko_upoly q;
        for (int c = 0; c < 4; ++c) corner(p, cu[c], cv[c], q.screenPos[c], q.normal[c]);
        this->emit(q, 0);
      }
  }

 private:
  // cubic Bernstein weights and their derivatives
  static void basis(float t, float b[4], float db[4]) {
    const float s = 1.0f - t;
    b[0] = s * s * s;
    b[1] = (3.0f * t) * s * s;
    b[2] = (3.0f * t) * (t - t * t);
    b[3] = t * (t * t);
    db[0] = -3.0f * s * s;
    db[1] = 3.0f * s * (1.0f - 3.0f * t);
    db[2] = 3.0f * t * (2.0f - 3.0f * t);
    db[3] = 3.0f * t * t;
  }

  static cvec4f combine(const float w[4], cvec4f a, cvec4f b, cvec4f c, cvec4f d) {
    const cvec4f o = {w[0] * a.x + w[1] * b.x + w[2] * c.x + w[3] * d.x, w[0] * a.y + w[1] * b.y + w[2] * c.y + w[3] * d.y,
                      w[0] * a.z + w[1] * b.z + w[2] * c.z + w[3] * d.z, w[0] * a.w + w[1] * b.w + w[2] * c.w + w[3] * d.w};
    return o;
  }

  // position, unit normal (dP/du x dP/dv) and displaced, projected corner
  void corner(const piko_patch &p, float u, float v, cvec4f &screen, cvec3f &normal) const {
    float bu[4], dbu[4], bv[4], dbv[4];
    basis(u, bu, dbu);
    basis(v, bv, dbv);
    cvec4f rows[4], drows[4];
    for (int r = 0; r < 4; ++r) {
      rows[r] = combine(bu, p.CP[4 * r], p.CP[4 * r + 1], p.CP[4 * r + 2], p.CP[4 * r + 3]);
      drows[r] = combine(dbu, p.CP[4 * r], p.CP[4 * r + 1], p.CP[4 * r + 2], p.CP[4 * r + 3]);
    }
    cvec4f pos = combine(bv, rows[0], rows[1], rows[2], rows[3]);
    const cvec4f tu = combine(bv, drows[0], drows[1], drows[2], drows[3]);
    const cvec4f tv = combine(dbv, rows[0], rows[1], rows[2], rows[3]);
    float nx = tu.y * tv.z - tu.z * tv.y, ny = tu.z * tv.x - tu.x * tv.z, nz = tu.x * tv.y - tu.y * tv.x;
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    normal.x = nx / len;
    normal.y = ny / len;
    normal.z = nz / len;
    if (PIKO_REYES_DISPLACEMENT != 0.0f) {
      const float d = PIKO_REYES_DISPLACEMENT * (pos.x * pos.x - pos.y * pos.y);
      pos.x = pos.x + d * normal.x;
      pos.y = pos.y + d * normal.y;
      pos.z = pos.z + d * normal.z;
    }
    const cvec4f clip = piko::transform4(constState.viewProjMatrix, pos);
    const float invW = 1.0f / clip.w;
    screen.x = ((clip.x * invW + 1.0f) * (float)constState.screenSizeX) * 0.5f;
    screen.y = ((clip.y * invW + 1.0f) * (float)constState.screenSizeY) * 0.5f;
    screen.z = clip.z * invW;
    screen.w = 1.0f;
  }
#endif
};

class ShadeStage : public Stage<SHADE_BINSIZE, SHADE_BINSIZE, 512, piko_upoly, piko_sample> {
#ifdef __PIKOC_DEVICE__
 public:
  // the bins under the bounding box of the first three corners
  void assignBin(piko_upoly q) {
    const float x0 = floorf(piko::fmin2(q.screenPos[0].x, piko::fmin2(q.screenPos[1].x, q.screenPos[2].x)));
    const float y0 = floorf(piko::fmin2(q.screenPos[0].y, piko::fmin2(q.screenPos[1].y, q.screenPos[2].y)));
    const float x1 = ceilf(piko::fmax2(q.screenPos[0].x, piko::fmax2(q.screenPos[1].x, q.screenPos[2].x)));
    const float y1 = ceilf(piko::fmax2(q.screenPos[0].y, piko::fmax2(q.screenPos[1].y, q.screenPos[2].y)));
    const int bx0 = piko::imax(toInt(floorf(x0 / SHADE_BINSIZE)), 0);
    const int by0 = piko::imax(toInt(floorf(y0 / SHADE_BINSIZE)), 0);
    const int bx1 = piko::imin(toInt(floorf(x1 / SHADE_BINSIZE)), getNumBinsX() - 1);
    const int by1 = piko::imin(toInt(floorf(y1 / SHADE_BINSIZE)), getNumBinsY() - 1);
    for (int by = by0; by <= by1; ++by)
      for (int bx = bx0; bx <= bx1; ++bx) this->assignToBin(q, by * getNumBinsX() + bx);
  }

  void schedule(int binID) { specifySchedule(LOAD_BALANCE); }

  void process(piko_upoly q) {
    const int bx = getBinID() % getNumBinsX(), by = getBinID() / getNumBinsX();
    float loX = q.screenPos[0].x, loY = q.screenPos[0].y, hiX = loX, hiY = loY;
    float z = q.screenPos[0].z;
    for (int c = 1; c < 4; ++c) {
      loX = piko::fmin2(loX, q.screenPos[c].x);
      loY = piko::fmin2(loY, q.screenPos[c].y);
      hiX = piko::fmax2(hiX, q.screenPos[c].x);
      hiY = piko::fmax2(hiY, q.screenPos[c].y);
      z = z + q.screenPos[c].z;
    }
    z = z * 0.25f;
    float rgb[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < 4; ++c) {
      float lit[3];
      phong(q.normal[c], lit);
      for (int k = 0; k < 3; ++k) rgb[k] = c == 0 ? lit[k] : rgb[k] + lit[k];
    }
    const int x0 = piko::imax(toInt(floorf(loX)), bx * SHADE_BINSIZE);
    const int y0 = piko::imax(toInt(floorf(loY)), by * SHADE_BINSIZE);
    const int x1 = piko::imin(piko::imin(toInt(ceilf(hiX)), (bx + 1) * SHADE_BINSIZE), constState.screenSizeX);
    const int y1 = piko::imin(piko::imin(toInt(ceilf(hiY)), (by + 1) * SHADE_BINSIZE), constState.screenSizeY);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x)
        for (int s = 0; s < PIKO_REYES_SPP; ++s) {
          const float ox = PIKO_REYES_SPP == 1 ? 0.5f : ((s & 1) ? 0.75f : 0.25f);
          const float oy = PIKO_REYES_SPP == 1 ? 0.5f : ((s & 2) ? 0.75f : 0.25f);
          if (!covers(q, (float)x + ox, (float)y + oy)) continue;
          // the nearest sample wins; a later micropolygon wins a tie
          const int slot = (y * constState.screenSizeX + x) * PIKO_REYES_SPP + s;
          int *stored = (int *)&mutableState->zBuffer[slot];
          if (piko::atomicMinInt(stored, piko::floatBits(z)) < piko::floatBits(z)) continue;
          piko_sample out;
          out.pos.x = x;
          out.pos.y = y;
          out.index = s;
          for (int k = 0; k < 3; ++k) out.rgb[k] = piko::fmin2(rgb[k], 4.0f) * 0.25f;
          this->emit(out, 0);
        }
  }

 private:
  static int toInt(float f) {  // saturating float -> int
    if (!(f > -1073741824.0f)) return -1073741824;
    return f > 1073741824.0f ? 1073741824 : (int)f;
  }

  // diffuse light from direction (0.10, 0.35, 0.85) on the flipped normal,
  // coloured (0.1773, 0.5773, 0.973), plus 0.2 ambient, saturated
  static void phong(cvec3f n, float out[3]) {
    float d = -n.x * 0.10f + -n.y * 0.35f + -n.z * 0.85f;
    d = d < 0.0f ? 0.0f : d;
    out[0] = piko::saturate(0.2f + d * 0.1773f);
    out[1] = piko::saturate(0.2f + d * 0.5773f);
    out[2] = piko::saturate(0.2f + d * 0.973f);
  }

  // strictly inside the quad traversed either way round
  static bool covers(const piko_upoly &q, float px, float py) {
    bool fwd = true, back = true;
    for (int e = 0; e < 4; ++e) {
      const cvec4f a = q.screenPos[e], b = q.screenPos[(e + 1) % 4];
      fwd = fwd && (py - a.y) * (b.x - a.x) - (px - a.x) * (b.y - a.y) > 0.0f;
      const cvec4f c = q.screenPos[(4 - e) % 4], d = q.screenPos[3 - e];
      back = back && (py - c.y) * (d.x - c.x) - (px - c.x) * (d.y - c.y) > 0.0f;
    }
    return fwd || back;
  }
#endif
};

#endif
