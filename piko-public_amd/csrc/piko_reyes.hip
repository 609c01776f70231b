// piko_reyes.hip -- the Reyes pipeline (Split -> Dice -> Shade -> PikoScreen)
// on gfx950 (reyesPipe/reyesPipe.h:12-40, split/dice/shade.pikostage).
//
// Frame = five launches on the pipe's stream, no host round trip:
//   k_reyes_split   one wave walks the split trees of a contiguous run of
//                   input patches level by level (16 lanes per patch, one per
//                   control point, 4 patches per pass).  Its frontier lives
//                   in a per-wave region; leaves go to a per-wave region in
//                   (level, position) order with per-level counts.
//   k_reyes_levels, k_reyes_rank
//                   one block per level: level totals, then the exclusive scan
//                   of the per-(level, wave) leaf and quad counts in
//                   level-major order.  Because
//                   each wave owns a contiguous run of roots, a leaf's index in
//                   the split stage's emission order (level by level, in order
//                   within a level: the reference's split self-loop, CPU
//                   backend one launch per level) is
//                   offset[level][wave] + its position in the wave's list; the
//                   kernel writes that index densely (rank -> region slot).
//   k_reyes_dice    one wave per leaf (persistent grid): per-u Bernstein rows,
//                   the (g+1)^2 grid points once each (position, normal,
//                   displacement, projection, Phong colour) into LDS, then one
//                   lane per micropolygon for flat depth, colour and pixel
//                   box, and the (micropolygon, pixel) pairs spread evenly over
//                   the lanes for the sample tests, each hit a 64-bit atomicMin
//                   of (depth, reversed emission order), so the nearest micropolygon
//                   wins and on equal depth the later one -- the outcome of
//                   the reference's atomicMin loop + (zBuffer == z) test run
//                   in emission order (shade.pikostage:226-241).  A hit that
//                   already loses to the sample's current key takes no atomic.
//   k_reyes_resolve one thread per pixel: box filter of the samples' colours
//                   to FP32 RGBA, nearest depth, and the visibility buffer
//                   reset to empty for the next frame.
// All floating-point expressions keep the reference's operation order and
// are compiled without contraction (-ffp-contract=off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "piko_hip.h"

namespace {

constexpr int kLevels = 21;  // levels 0..20: a patch splits while numSplits < 20 (split.pikostage:284)
constexpr int kSplitCapMax = 20;
constexpr unsigned long long kVisEmpty = ((unsigned long long)(0x3F800000u ^ 0x80000000u) << 32) | 0xFFFFFFFFull;
constexpr unsigned long long kMaxOrder = 0xFFFFFFFEull;  // quads per frame must stay below this
constexpr unsigned long long kMaxPairs = 0xFFFFFFFFull;  // (tile, leaf) pairs of tiled sampling

struct ReyesState {
  unsigned long long nleaf;
  unsigned long long nquad;
  unsigned int overflow;  // bit 0: a split frontier, bit 1: a leaf region, bit 2: quad colours,
                          // bit 3: the (tile, leaf) lists of tiled sampling
  unsigned int levels;
  unsigned int npairs;  // (tile, leaf) pairs of tiled sampling, saturated at 2^32 - 1
  unsigned int pad;
};

struct ReyesArgs {
  int W, H;
  float vp[16];
  float threshold;
  int max_splits, dice, spp;
  float disp;
  int roots_per_wave, nroots, G;  // G: split waves
  int front_cap, leaf_cap;        // per split wave: nodes per level buffer, leaves
  unsigned long long qcap;        // quad colour slots
  int keep_samples;
  // tiled sampling: tiles of 2^tsb pixels, tnx x tny of them; the grid
  // points of a leaf start at pt_num / 4 times its first order index
  int tsb, tnx, tny;
  uint32_t pair_cap, pt_num;
};

// piko::fminf / fmaxf of the CPU backend are std::min / std::max
// (math.h:24-51): the first argument unless the second compares smaller /
// larger.
__device__ __forceinline__ float fmn(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmx(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ float4 f4add(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 f4half(float4 a) {
  return make_float4(a.x * 0.5f, a.y * 0.5f, a.z * 0.5f, a.w * 0.5f);
}
__device__ __forceinline__ float4 shfl4(float4 v, int src) {
  return make_float4(__shfl(v.x, src, 16), __shfl(v.y, src, 16), __shfl(v.z, src, 16), __shfl(v.w, src, 16));
}

// Synchronises the lanes of a one-wave workgroup around LDS traffic.  The
// fences are wavefront-scoped, so outstanding global stores and atomics are
// not waited for (a __syncthreads would wait for them).  LDS accesses of one
// wave execute in order, so this orders them.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (int)ceil(ext / threshold) of split.pikostage:307-308 made total: 64 and
// above give the same clamped grid as any larger value, NaN gives 0.
__device__ __forceinline__ int ext_steps(float ext, float thr) {
  const double q = ceil((double)(ext / thr));
  return q >= 64.0 ? 64 : (q >= 0.0 ? (int)q : 0);
}

__device__ __forceinline__ int pow2roundup(int x) {  // split.pikostage:34-45
  if (x < 0) return 0;
  --x;
  x |= x >> 1;
  x |= x >> 2;
  x |= x >> 4;
  x |= x >> 8;
  x |= x >> 16;
  return x + 1;
}

__device__ __forceinline__ int grid_steps(float ext, float thr) {  // split.pikostage:307-314
  int n = pow2roundup(ext_steps(ext, thr) + 1);
  return n < 2 ? 2 : (n > 32 ? 32 : n);
}

// ---------------------------------------------------------------------------
// Split.  Lane l of the wave holds control point (l & 15) of node slot l >> 4.

__global__ __launch_bounds__(64) void k_reyes_split(const float4 *__restrict__ roots, ReyesArgs a,
                                                    float4 *__restrict__ front, float4 *__restrict__ lcp,
                                                    uint32_t *__restrict__ lmeta, uint32_t *__restrict__ lqloc,
                                                    uint32_t *__restrict__ cnt, uint32_t *__restrict__ qcnt,
                                                    uint32_t *__restrict__ lvlpre, uint32_t *__restrict__ wstat) {
  const int w = blockIdx.x, lane = threadIdx.x;
  const int slot = lane >> 4, i = lane & 15, r = i >> 2, c = i & 3;
  const int G = a.G;
  const int r0 = w * a.roots_per_wave;
  int n = min(r0 + a.roots_per_wave, a.nroots) - r0;
  float4 *fb0 = front + (size_t)w * 2 * a.front_cap * 16;
  float4 *fb1 = fb0 + (size_t)a.front_cap * 16;
  const float4 *fcur = roots + (size_t)r0 * 16;
  float4 *fnext = fb0;
  float4 *leaf = lcp + (size_t)w * a.leaf_cap * 16;
  uint32_t *meta = lmeta + (size_t)w * a.leaf_cap;
  uint32_t *qloc = lqloc + (size_t)w * a.leaf_cap;
  const float *m = a.vp;
  uint32_t nleaves = 0, status = 0;
  int L = 0;
  for (; L < kLevels; ++L) {
    const uint32_t lvl0 = nleaves;
    uint32_t qrun = 0;
    int nn = 0;
    for (int base = 0; base < n && !status; base += 4) {
      const int node = base + slot;
      const bool valid = node < n;
      const float4 p = valid ? fcur[(size_t)node * 16 + i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
      // SplitStage::process projection (split.pikostage:245-268)
      float cx = ((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12] * p.w;
      float cy = ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13] * p.w;
      float cw = ((m[3] * p.x + m[7] * p.y) + m[11] * p.z) + m[15] * p.w;
      if (cw == 0.0f) cw = 1.0f;
      cx = cx / cw;
      cy = cy / cw;
      const float X = (float)(((double)cx + 1.0) * 0.5 * (double)a.W);
      const float Y = (float)(((double)cy + 1.0) * 0.5 * (double)a.H);
      // getUVPixelExtents (split.pikostage:47-94): longest row / column
      // control polyline on screen.
      float nx = __shfl(X, (i + 1) & 15, 16), ny = __shfl(Y, (i + 1) & 15, 16);
      float dx = X - nx, dy = Y - ny;
      const float du = sqrtf(dx * dx + dy * dy);
      const float rows = (__shfl(du, 4 * r, 16) + __shfl(du, 4 * r + 1, 16)) + __shfl(du, 4 * r + 2, 16);
      nx = __shfl(X, (i + 4) & 15, 16);
      ny = __shfl(Y, (i + 4) & 15, 16);
      dx = X - nx;
      dy = Y - ny;
      const float dv = sqrtf(dx * dx + dy * dy);
      const float cols = (__shfl(dv, c, 16) + __shfl(dv, c + 4, 16)) + __shfl(dv, c + 8, 16);
      float ue = 0.0f, ve = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) ue = fmx(ue, __shfl(rows, 4 * k, 16));
#pragma unroll
      for (int k = 0; k < 4; ++k) ve = fmx(ve, __shfl(cols, k, 16));
      const bool dirU = ue >= a.threshold;
      const bool split = valid && (dirU || ve >= a.threshold) && L < a.max_splits;
      const bool isleaf = valid && !split;
      const int gu = grid_steps(ue, a.threshold), gv = grid_steps(ve, a.threshold);
      const uint32_t nq = !isleaf ? 0u : (a.dice > 0 ? (uint32_t)(a.dice * a.dice) : (uint32_t)((gu + 1) * (gv + 1)));
      // positions: children and leaves of the four slots in slot order
      int pch = 0, plf = 0, tch = 0, tlf = 0;
      uint32_t pq = 0, tq = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int ck = __shfl(split ? 2 : 0, 16 * k), lk = __shfl(isleaf ? 1 : 0, 16 * k);
        const uint32_t qk = __shfl(nq, 16 * k);
        if (k < slot) {
          pch += ck;
          plf += lk;
          pq += qk;
        }
        tch += ck;
        tlf += lk;
        tq += qk;
      }
      if (nn + tch > a.front_cap) status |= 1u;
      if ((int)(nleaves + tlf) > a.leaf_cap) status |= 2u;
      if (status) break;  // uniform across the wave
      // splitU / splitV (split.pikostage:95-224): midpoint de Casteljau on
      // the row (U) or column (V) through this lane's control point.
      const int s0 = dirU ? 4 * r : c, st = dirU ? 1 : 4, pos = dirU ? c : r;
      const float4 a0 = shfl4(p, s0), a1 = shfl4(p, s0 + st), a2 = shfl4(p, s0 + 2 * st), a3 = shfl4(p, s0 + 3 * st);
      if (split) {
        const float4 n11 = f4half(f4add(a0, a1)), n22 = f4half(f4add(a2, a3)), tmp = f4half(f4add(a1, a2));
        const float4 n12 = f4half(f4add(n11, tmp)), n21 = f4half(f4add(n22, tmp));
        const float4 n13 = f4half(f4add(n12, n21));
        const float4 c1 = pos == 0 ? a0 : (pos == 1 ? n11 : (pos == 2 ? n12 : n13));
        const float4 c2 = pos == 0 ? n13 : (pos == 1 ? n21 : (pos == 2 ? n22 : a3));
        fnext[(size_t)(nn + pch) * 16 + i] = c1;
        fnext[(size_t)(nn + pch + 1) * 16 + i] = c2;
      }
      if (isleaf) {
        const uint32_t at = nleaves + plf;
        leaf[(size_t)at * 16 + i] = p;
        if (i == 0) {
          meta[at] = (uint32_t)gu | ((uint32_t)gv << 8) | ((uint32_t)L << 16);
          qloc[at] = qrun + pq;
        }
      }
      nn += tch;
      nleaves += tlf;
      qrun += tq;
    }
    if (lane == 0) {
      cnt[L * G + w] = nleaves - lvl0;
      qcnt[L * G + w] = qrun;
      lvlpre[L * G + w] = lvl0;
    }
    if (status || nn == 0) break;
    __syncthreads();  // the next level list is read by other lanes
    fcur = fnext;
    fnext = (fnext == fb0) ? fb1 : fb0;
    n = nn;
  }
  if (lane == 0) {
    for (int l = L + 1; l < kLevels; ++l) {
      cnt[l * G + w] = 0;
      qcnt[l * G + w] = 0;
      lvlpre[l * G + w] = nleaves;
    }
    wstat[w] = status | ((uint32_t)(L + 1) << 8);
  }
}

// ---------------------------------------------------------------------------
// Ranks.  k_reyes_levels: one block per level sums the level's leaf and quad
// counts over the waves.  k_reyes_rank: one block per level adds the levels
// before it and scans the level's counts over the waves (each thread a
// contiguous run of waves), writing the dense emission-order index
// leaf_at[rank] = region slot and qoff[level][wave], the order index of the
// group's first quad.  The last level's block writes the frame totals and
// status.

constexpr int kRankThreads = 1024;

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(v, d);
    if (lane >= d) v += y;
  }
  return v;
}

// Block-wide sums of two values (all threads receive them).
__device__ __forceinline__ void block_sum2(unsigned long long &a, unsigned long long &b,
                                           unsigned long long (*sh)[kRankThreads / 64]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  a = wave_sum_u64(a);
  b = wave_sum_u64(b);
  if (lane == 0) {
    sh[0][wv] = a;
    sh[1][wv] = b;
  }
  __syncthreads();
  a = b = 0;
#pragma unroll
  for (int k = 0; k < kRankThreads / 64; ++k) {
    a += sh[0][k];
    b += sh[1][k];
  }
  __syncthreads();
}

__global__ __launch_bounds__(kRankThreads) void k_reyes_levels(const uint32_t *__restrict__ cnt,
                                                               const uint32_t *__restrict__ qcnt, int G,
                                                               unsigned long long *__restrict__ lvltot) {
  __shared__ unsigned long long sh[2][kRankThreads / 64];
  const int L = blockIdx.x;
  unsigned long long a = 0, b = 0;
  for (int w = threadIdx.x; w < G; w += kRankThreads) {
    a += cnt[L * G + w];
    b += qcnt[L * G + w];
  }
  block_sum2(a, b, sh);
  if (threadIdx.x == 0) {
    lvltot[2 * L] = a;
    lvltot[2 * L + 1] = b;
  }
}

__global__ __launch_bounds__(kRankThreads) void k_reyes_rank(
    const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ qcnt, const uint32_t *__restrict__ lvlpre,
    const uint32_t *__restrict__ wstat, const unsigned long long *__restrict__ lvltot, int G, int leaf_cap,
    unsigned long long qcap, uint32_t *__restrict__ off, uint32_t *__restrict__ qoff, uint32_t *__restrict__ leaf_at,
    ReyesState *__restrict__ st) {
  __shared__ unsigned long long sh[2][kRankThreads / 64];
  __shared__ uint32_t s_or, s_lv;
  const int L = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (lvltot[2 * L] == 0 && L != kLevels - 1) return;  // no leaves at this level
  // leaves and quads of the levels before this one, one level per lane: a
  // loop over them waited for each level's load in turn, and the last
  // level's block (which always runs) waited twenty times
  static_assert(kLevels <= 64, "one lane per level");
  const unsigned long long base_l = wave_sum_u64(lane < L ? lvltot[2 * lane] : 0ull);
  const unsigned long long base_q = wave_sum_u64(lane < L ? lvltot[2 * lane + 1] : 0ull);
  const int run = (G + kRankThreads - 1) / kRankThreads;
  const int w0 = min(t * run, G), w1 = min(w0 + run, G);
  // A run of up to kRunRegs waves is read into registers with every load in
  // flight at once; a loop over the run waited for each wave's loads in turn.
  // Longer runs (more than kRunRegs * kRankThreads split waves) loop.
  constexpr int kRunRegs = 4;
  const bool regs = run <= kRunRegs;
  uint32_t rc[kRunRegs] = {}, rq[kRunRegs] = {}, rp[kRunRegs] = {};
  unsigned long long sl = 0, sq = 0;
  if (lvltot[2 * L] != 0) {
    if (regs) {
      if (w0 < w1) {
#pragma unroll
        for (int i = 0; i < kRunRegs; ++i) {
          const int w = min(w0 + i, w1 - 1);  // in the run: no branch around the loads
          rc[i] = cnt[L * G + w];
          rq[i] = qcnt[L * G + w];
          rp[i] = lvlpre[L * G + w];
        }
#pragma unroll
        for (int i = 0; i < kRunRegs; ++i)
          if (w0 + i < w1) {
            sl += rc[i];
            sq += rq[i];
          }
      }
    } else {
      for (int w = w0; w < w1; ++w) {
        sl += cnt[L * G + w];
        sq += qcnt[L * G + w];
      }
    }
  }
  // exclusive block scan of the per-thread run totals
  const unsigned long long il = wave_incl_scan<unsigned long long>(sl, lane);
  const unsigned long long iq = wave_incl_scan<unsigned long long>(sq, lane);
  if (lane == 63) {
    sh[0][wv] = il;
    sh[1][wv] = iq;
  }
  __syncthreads();
  unsigned long long o = base_l + il - sl, oq = base_q + iq - sq;
#pragma unroll
  for (int k = 0; k < kRankThreads / 64; ++k)
    if (k < wv) {
      o += sh[0][k];
      oq += sh[1][k];
    }
  // (the last level's block runs for the frame totals below; with no leaves
  // at that level it has no ranks to write)
  if (lvltot[2 * L] != 0) {
    if (regs) {
#pragma unroll
      for (int i = 0; i < kRunRegs; ++i) {
        const int w = w0 + i;
        if (w < w1) {
          off[L * G + w] = (uint32_t)o;
          qoff[L * G + w] = (uint32_t)oq;  // exact whenever the frame's quads fit (status bit 2 clear)
          const uint32_t slot0 = (uint32_t)w * (uint32_t)leaf_cap + rp[i];
          for (uint32_t k = 0; k < rc[i]; ++k) leaf_at[o + k] = slot0 + k;
          o += rc[i];
          oq += rq[i];
        }
      }
    } else {
      for (int w = w0; w < w1; ++w) {
        const uint32_t c = cnt[L * G + w];
        off[L * G + w] = (uint32_t)o;
        qoff[L * G + w] = (uint32_t)oq;
        const uint32_t slot0 = (uint32_t)w * (uint32_t)leaf_cap + lvlpre[L * G + w];
        for (uint32_t k = 0; k < c; ++k) leaf_at[o + k] = slot0 + k;
        o += c;
        oq += qcnt[L * G + w];
      }
    }
  }
  if (L != kLevels - 1) return;
  if (t == 0) {
    s_or = 0;
    s_lv = 0;
  }
  __syncthreads();
  uint32_t ovf = 0, lv = 0;
  if (regs) {
    uint32_t ws[kRunRegs];
#pragma unroll
    for (int i = 0; i < kRunRegs; ++i) ws[i] = wstat[min(t + i * kRankThreads, G - 1)];
#pragma unroll
    for (int i = 0; i < kRunRegs; ++i)
      if (t + i * kRankThreads < G) {
        ovf |= ws[i] & 3u;
        lv = max(lv, ws[i] >> 8);
      }
  } else {
    for (int w = t; w < G; w += kRankThreads) {
      ovf |= wstat[w] & 3u;
      lv = max(lv, wstat[w] >> 8);
    }
  }
  atomicOr(&s_or, ovf);
  atomicMax(&s_lv, lv);
  __syncthreads();
  if (t == 0) {
    const unsigned long long nleaf = base_l + lvltot[2 * L], nquad = base_q + lvltot[2 * L + 1];
    st->nleaf = nleaf;
    st->nquad = nquad;
    st->overflow = s_or | ((nquad > qcap || nquad >= kMaxOrder) ? 4u : 0u);
    st->levels = s_lv;
  }
}

// ---------------------------------------------------------------------------
// Dice + sample.

// Bernstein basis and derivatives (dice.pikostage:85-93).
__device__ __forceinline__ void bern(float u, float *b, float *d) {
  b[0] = ((1.0f - u) * (1.0f - u) * (1.0f - u));
  b[1] = ((3.0f * u) * (1.0f - u) * (1.0f - u));
  b[2] = ((3.0f * u) * (u - u * u));
  b[3] = ((u) * (u * u));
  d[0] = (-3.0f * (1.0f - u) * (1.0f - u));
  d[1] = (3.0f * (1.0f - u) * (1.0f - 3.0f * u));
  d[2] = (3.0f * u * (2.0f - 3.0f * u));
  d[3] = (3.0f * u * u);
}

// bez2D / dbez2D (dice.pikostage:96-109) with the basis values given.
__device__ __forceinline__ float4 comb(const float *B, float4 v0, float4 v1, float4 v2, float4 v3) {
  return make_float4(B[0] * v0.x + B[1] * v1.x + B[2] * v2.x + B[3] * v3.x,
                     B[0] * v0.y + B[1] * v1.y + B[2] * v2.y + B[3] * v3.y,
                     B[0] * v0.z + B[1] * v1.z + B[2] * v2.z + B[3] * v3.z,
                     B[0] * v0.w + B[1] * v1.w + B[2] * v2.w + B[3] * v3.w);
}

// phongShadeReyes + computeLighting (shade.pikostage:22-30, 64-85): the call
// passes (normal, matcol, lightvec) into the macro's (_mynor, _lightvec,
// _matcol), so the material colour acts as the light vector and vice versa.
__device__ __forceinline__ float sat(float p) { return p > 1.0f ? 1.0f : (p < 0.0f ? 0.0f : p); }

// The sample position of sample s (pixel centre, or the centres of the 2x2
// strata in x-fastest order).
__device__ __forceinline__ float sample_off(int spp, int s, int axis) {
  if (spp == 1) return 0.5f;
  return ((s >> axis) & 1) ? 0.75f : 0.25f;
}

__device__ __forceinline__ int clampi(float f) {
  if (!(f > -1073741824.0f)) return -1073741824;
  if (f > 1073741824.0f) return 1073741824;
  return (int)f;
}

// Pixels a micropolygon is tested on: ShadeStage bins it by the box of its
// first three corners (computeBoundingBox, dice.pikostage:18-43) into 64-px
// bins and tests the four-corner box inside each bin (shade.pikostage:
// 140-163).  Bins per axis are ceil(W/64), pixels clipped to the screen.
__device__ __forceinline__ void quad_pixels(const float *X, const float *Y, int W, int H, int *x0, int *y0,
                                            int *x1, int *y1) {
  const int nbx = (W + 63) / 64, nby = (H + 63) / 64;
  const float lo3x = floorf(fmn(X[0], fmn(X[1], X[2]))), lo3y = floorf(fmn(Y[0], fmn(Y[1], Y[2])));
  const float hi3x = ceilf(fmx(X[0], fmx(X[1], X[2]))), hi3y = ceilf(fmx(Y[0], fmx(Y[1], Y[2])));
  const int bx1 = max(clampi(floorf(lo3x / 64.0f)), 0), bx2 = min(clampi(floorf(hi3x / 64.0f)), nbx - 1);
  const int by1 = max(clampi(floorf(lo3y / 64.0f)), 0), by2 = min(clampi(floorf(hi3y / 64.0f)), nby - 1);
  *x0 = max(clampi(floorf(fmn(fmn(X[0], X[1]), fmn(X[2], X[3])))), bx1 * 64);
  *y0 = max(clampi(floorf(fmn(fmn(Y[0], Y[1]), fmn(Y[2], Y[3])))), by1 * 64);
  *x1 = min(min(clampi(ceilf(fmx(fmx(X[0], X[1]), fmx(X[2], X[3])))), (bx2 + 1) * 64), W);
  *y1 = min(min(clampi(ceilf(fmx(fmx(Y[0], Y[1]), fmx(Y[2], Y[3])))), (by2 + 1) * 64), H);
  if (bx1 > bx2 || by1 > by2 || *x0 >= *x1 || *y0 >= *y1) {
    *x0 = *y0 = 0;
    *x1 = *y1 = 0;
  }
}

// Two 4-edge fan tests at a sample (shade.pikostage:176-218): hit if either
// orientation has all edge functions > 0.
__device__ __forceinline__ bool quad_hit(const float *X, const float *Y, float px, float py) {
  bool r1 = true, r2 = true;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int ea = e, eb = (e + 1) & 3;
    if (((py - Y[ea]) * (X[eb] - X[ea]) - (px - X[ea]) * (Y[eb] - Y[ea])) <= 0.0f) r1 = false;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int ea = (4 - e) & 3, eb = (3 - e) & 3;
    if (((py - Y[ea]) * (X[eb] - X[ea]) - (px - X[ea]) * (Y[eb] - Y[ea])) <= 0.0f) r2 = false;
  }
  return r1 || r2;
}

// One leaf's dice set-up, by the one wave of a workgroup: control points,
// grid size, the order index of its first micropolygon, then every grid
// point's screen position, depth and lit colour in s_pt.
struct LeafGrid {
  uint32_t qbase;  // order index of the leaf's first micropolygon
  int gu, gv;      // micropolygons along u and v
};

template <int MAXP>
__device__ __forceinline__ LeafGrid dice_leaf(const ReyesArgs &a, const float4 *__restrict__ lcp,
                                              const uint32_t *__restrict__ lmeta,
                                              const uint32_t *__restrict__ lqloc, const uint32_t *__restrict__ qoff,
                                              size_t at, float4 *s_cp, float4 (*s_row)[8], float (*s_pt)[6],
                                              int lane) {
  const float *m = a.vp;
  const int W = a.W, H = a.H;
  if (lane < 16) s_cp[lane] = lcp[at * 16 + lane];
  const uint32_t mt = lmeta[at];
  const uint32_t qbase = qoff[(mt >> 16) * (uint32_t)a.G + (uint32_t)(at / (size_t)a.leaf_cap)] + lqloc[at];
  int nu, nv, gu, gv;  // dice steps; micropolygons along u and v
  if (a.dice > 0) {
    nu = nv = gu = gv = a.dice;
  } else {
    nu = (int)(mt & 0xffu);
    nv = (int)((mt >> 8) & 0xffu);
    gu = nu + 1;  // k <= nu (dice.pikostage:176-177)
    gv = nv + 1;
  }
  const float du = 1.0f / ((float)nu), dv = 1.0f / ((float)nv);
  wave_sync();
  // grid u values: myU = k*du for k < gu, and the last quad's myU + du
  // (dice.pikostage:180, 233); with power-of-two steps these coincide with
  // the neighbouring quads' corners bit for bit.
  for (int e = lane; e < 8 * (gu + 1); e += 64) {  // (grid u, row) pairs
    const int iu = e >> 3, row = e & 7, r4 = 4 * (row & 3);
    const float u = iu < gu ? (float)iu * du : (float)(gu - 1) * du + du;
    float B[4], D[4];
    bern(u, B, D);
    s_row[iu][row] = comb(row < 4 ? B : D, s_cp[r4], s_cp[r4 + 1], s_cp[r4 + 2], s_cp[r4 + 3]);
  }
  wave_sync();
  const int np = (gu + 1) * (gv + 1);
  for (int pt = lane; pt < np; pt += 64) {
    const int iu = pt / (gv + 1), jv = pt - iu * (gv + 1);
    const float v = jv < gv ? (float)jv * dv : (float)(gv - 1) * dv + dv;
    float B[4], D[4];
    bern(v, B, D);
    // patchEval (dice.pikostage:111-121), normEval (123-146)
    float4 P = comb(B, s_row[iu][0], s_row[iu][1], s_row[iu][2], s_row[iu][3]);
    const float4 A = comb(B, s_row[iu][4], s_row[iu][5], s_row[iu][6], s_row[iu][7]);
    const float4 Bv = comb(D, s_row[iu][0], s_row[iu][1], s_row[iu][2], s_row[iu][3]);
    float nx = A.y * Bv.z - A.z * Bv.y, ny = A.z * Bv.x - A.x * Bv.z, nz = A.x * Bv.y - A.y * Bv.x;
    const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
    nx /= mag;
    ny /= mag;
    nz /= mag;
    if (a.disp != 0.0f) {
      const float d = a.disp * (P.x * P.x - P.y * P.y);
      P.x = P.x + d * nx;
      P.y = P.y + d * ny;
      P.z = P.z + d * nz;
    }
    // projection (dice.pikostage:209-216)
    const float sx = ((m[0] * P.x + m[4] * P.y) + m[8] * P.z) + m[12] * P.w;
    const float sy = ((m[1] * P.x + m[5] * P.y) + m[9] * P.z) + m[13] * P.w;
    const float sz = ((m[2] * P.x + m[6] * P.y) + m[10] * P.z) + m[14] * P.w;
    const float sw = ((m[3] * P.x + m[7] * P.y) + m[11] * P.z) + m[15] * P.w;
    const float onebyw = 1.0f / sw;
    float *o = s_pt[pt];
    o[0] = ((sx * onebyw + 1.0f) * (float)W) * 0.5f;
    o[1] = ((sy * onebyw + 1.0f) * (float)H) * 0.5f;
    o[2] = (sz * onebyw);
    const float ix = -nx, iy = -ny, iz = -nz;
    float dif = ix * 0.10f + iy * 0.35f + iz * 0.85f;
    dif = dif < 0.0f ? 0.0f : dif;
    o[3] = sat(0.2f + dif * 0.1773f);
    o[4] = sat(0.2f + dif * 0.5773f);
    o[5] = sat(0.2f + dif * 0.973f);
  }
  wave_sync();
  return LeafGrid{qbase, gu, gv};
}

template <int MAXP, int SPP>
__global__ __launch_bounds__(64) void k_reyes_dice(ReyesArgs a, const float4 *__restrict__ lcp,
                                                   const uint32_t *__restrict__ lmeta,
                                                   const uint32_t *__restrict__ lqloc,
                                                   const uint32_t *__restrict__ leaf_at,
                                                   const uint32_t *__restrict__ qoff,
                                                   const ReyesState *__restrict__ st, float *__restrict__ qcol,
                                                   unsigned long long *__restrict__ vis) {
  __shared__ float4 s_cp[16];
  __shared__ float4 s_row[MAXP][8];  // per grid u: bez2D rows 0..3, dbez2D rows 4..7
  __shared__ float s_pt[MAXP * MAXP][6];
  __shared__ float s_qx[64][4], s_qy[64][4];
  __shared__ unsigned long long s_key[64];
  __shared__ int s_x0[64], s_y0[64], s_wid[64], s_pre[64], s_own[64];
  const int lane = threadIdx.x;
  if (st->overflow & 4u) return;  // quad colours would not fit: the frame is re-run with room
  const unsigned long long nleaf = st->nleaf;
  const int W = a.W, H = a.H;
  for (unsigned long long g = blockIdx.x; g < nleaf; g += gridDim.x) {
    const LeafGrid lg = dice_leaf<MAXP>(a, lcp, lmeta, lqloc, qoff, leaf_at[g], s_cp, s_row, s_pt, lane);
    const uint32_t qbase = lg.qbase;
    const int gu = lg.gu, gv = lg.gv;
    const int nq = gu * gv;
    for (int q0 = 0; q0 < nq; q0 += 64) {
      // one lane per micropolygon: flat depth and colour, pixel box
      const int qd = q0 + lane;
      int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
      if (qd < nq) {
        const int k = qd / gv, j = qd - k * gv;
        const float *c0 = s_pt[k * (gv + 1) + j], *c1 = s_pt[(k + 1) * (gv + 1) + j];
        const float *c2 = s_pt[(k + 1) * (gv + 1) + j + 1], *c3 = s_pt[k * (gv + 1) + j + 1];
        const float X[4] = {c0[0], c1[0], c2[0], c3[0]}, Y[4] = {c0[1], c1[1], c2[1], c3[1]};
        const float z = (((c0[2] + c1[2]) + c2[2]) + c3[2]) * 0.25f;  // shade.pikostage:98
        const uint32_t order = qbase + (uint32_t)qd;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)  // shade.pikostage:130-138
          qcol[(size_t)order * 3 + ch] = fmn(((c0[3 + ch] + c1[3 + ch]) + c2[3 + ch]) + c3[3 + ch], 4.0f) * 0.25f;
        quad_pixels(X, Y, W, H, &x0, &y0, &x1, &y1);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          s_qx[lane][e] = X[e];
          s_qy[lane][e] = Y[e];
        }
        s_key[lane] = ((unsigned long long)(__float_as_uint(z) ^ 0x80000000u) << 32) |
                      (unsigned long long)(0xFFFFFFFEu - order);
      }
      // spread the (micropolygon, pixel) pairs of these 64 quads evenly
      // over the lanes
      const int wid = x1 - x0, npx = wid * (y1 - y0);
      const int inc = wave_incl_scan<int>(npx, lane);
      const int total = __shfl(inc, 63);
      s_x0[lane] = x0;
      s_y0[lane] = y0;
      s_wid[lane] = wid;
      s_pre[lane] = inc - npx;
      wave_sync();
      // Pair slot t belongs to the last quad whose first pair is at or before
      // t; quads with pairs start at distinct slots.  Per chunk of 64 slots
      // each such quad marks its start, the ballot of the marks gives every
      // slot the nearest mark at or below it, and a slot with none keeps the
      // previous chunk's last owner (a binary search over the starts took six
      // dependent LDS reads per slot).  The quotient local / w uses the float
      // reciprocal, exact for local < 2^21 as shown for the raster's span
      // index; a chunk with more pairs divides.
      const int pre = inc - npx;
      const bool fquot = total < (1 << 21);
      int carry = 0;
      for (int t0 = 0; t0 < total; t0 += 64) {
        s_own[lane] = -1;
        wave_sync();
        const int rel = pre - t0;
        if (npx > 0 && rel >= 0 && rel < 64) s_own[rel] = lane;
        wave_sync();
        const int mark = s_own[lane];
        const unsigned long long m = __ballot(mark >= 0) & ((2ull << lane) - 1ull);
        const int got = __shfl(mark, m ? 63 - __clzll((long long)m) : 0);
        const int q = m ? got : carry;
        carry = __shfl(q, 63);
        wave_sync();
        const int t = t0 + lane;
        if (t >= total) break;  // the last chunk
        const int local = t - s_pre[q], w = s_wid[q];
        const int dy = fquot ? (int)(((float)local + 0.5f) * __builtin_amdgcn_rcpf((float)w)) : local / w;
        const int dx = local - dy * w;
        const int x = s_x0[q] + dx, y = s_y0[q] + dy;
        const float X[4] = {s_qx[q][0], s_qx[q][1], s_qx[q][2], s_qx[q][3]};
        const float Y[4] = {s_qy[q][0], s_qy[q][1], s_qy[q][2], s_qy[q][3]};
        const unsigned long long key = s_key[q];
        unsigned long long *dst = vis + ((size_t)y * W + x) * SPP;
        bool hit[SPP];
        bool any = false;
#pragma unroll
        for (int smp = 0; smp < SPP; ++smp) {
          hit[smp] = quad_hit(X, Y, (float)x + sample_off(SPP, smp, 0), (float)y + sample_off(SPP, smp, 1));
          any |= hit[smp];
        }
        if (any) {
          // A hit whose key cannot beat the sample's current key is dropped
          // before its atomic.  The keys only decrease within the frame, and
          // the reset that starts the frame is the previous resolve of this
          // slot's visibility buffer: in run_frames that resolve runs on
          // another stream, and this dice is ordered after it only by the
          // wait on the slot's resolve event (e_res[slot]) made before the
          // launch.  With that wait in place whatever value this load
          // returns, however stale, is at least the sample's minimum so far,
          // so a dropped key would not have changed it; without it a stale
          // key of the previous frame could drop a winning hit.  The
          // multi-frame parity tests (frames in one call with regrowth,
          // repeated frames across camera changes) cover the ordering.  On
          // the 4,096-patch sphere (back faces drawn too) dice ran 30% faster.
          unsigned long long cur[SPP];
#pragma unroll
          for (int smp = 0; smp < SPP; ++smp) cur[smp] = dst[smp];
#pragma unroll
          for (int smp = 0; smp < SPP; ++smp)
            if (hit[smp] && key < cur[smp]) atomicMin(dst + smp, key);
        }
      }
      wave_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled sampling.  The dice kernel above composites every covered sample
// with a 64-bit atomicMin in a W x H x spp buffer; on this GPU those atomics
// are performed at the memory side.  Here k_reyes_dice_bin instead keeps
// each micropolygon (grid points, flat depth, colour, pixel box) and the
// leaf's pixel box; the leaves are listed per screen tile, and workgroups take the
// same minimum of the same keys over the same (micropolygon, sample) pairs in
// LDS, then resolve the tile's pixels.  A tile with more than kPartLeaves
// leaves is split into parts run by separate workgroups; their LDS keys meet
// in the global buffer (one atomicMin per covered sample and part) and the
// last part to finish resolves the tile.  The minimum does not depend on the
// order of the leaves or of the parts, so the frame is the same bit for bit.

constexpr int kMaxSampleTiles = 8192;  // tiles per screen (LDS histograms of the binning)
constexpr int kLeafChunk = 4096;       // leaves per binning block
constexpr int kBinThreads = 1024;
constexpr int kTileWaves = 4;          // waves per tile workgroup
constexpr int kPartLeaves = 64;        // leaves per tile part (at most 64 per wave); 32 and 128 were slower
static_assert(kPartLeaves <= 64 * kTileWaves, "one batch of leaf records per wave");

// Tiles covered by a leaf record's pixel box [x0, x1) x [y0, y1) (none when
// the box is empty).
struct TileRange {
  int tx0, ty0, ntx, nty;
};
__device__ __forceinline__ TileRange tile_range(uint4 info, int tsb) {
  const int x0 = (int)(info.z & 0xffffu), y0 = (int)(info.z >> 16);
  const int x1 = (int)(info.w & 0xffffu), y1 = (int)(info.w >> 16);
  TileRange r;
  if (x0 >= x1) {
    r.tx0 = r.ty0 = r.ntx = r.nty = 0;
    return r;
  }
  r.tx0 = x0 >> tsb;
  r.ty0 = y0 >> tsb;
  r.ntx = ((x1 - 1) >> tsb) - r.tx0 + 1;
  r.nty = ((y1 - 1) >> tsb) - r.ty0 + 1;
  return r;
}

// Grid point (iu, jv) of a leaf whose first micropolygon has order index
// qbase sits at qbase * pt_num / 4 + iu * (gv + 1) + jv.  A leaf of
// gu x gv micropolygons has (gu + 1)(gv + 1) grid points, at most 9/4 of
// gu * gv for gu, gv >= 2 and 4 times it for gu = gv = 1 (pt_num 9 or 16),
// and its successor's first order index is qbase + gu * gv, so the leaves'
// point runs do not overlap.
__device__ __forceinline__ size_t point_base(uint32_t qbase, uint32_t pt_num) {
  return (size_t)(((unsigned long long)qbase * pt_num) >> 2);
}

template <int MAXP>
__global__ __launch_bounds__(64) void k_reyes_dice_bin(
    ReyesArgs a, const float4 *__restrict__ lcp, const uint32_t *__restrict__ lmeta,
    const uint32_t *__restrict__ lqloc, const uint32_t *__restrict__ leaf_at, const uint32_t *__restrict__ qoff,
    const ReyesState *__restrict__ st, float *__restrict__ qcol, float2 *__restrict__ pxy, float *__restrict__ qz,
    ushort4 *__restrict__ qbox, uint4 *__restrict__ linfo) {
  __shared__ float4 s_cp[16];
  __shared__ float4 s_row[MAXP][8];
  __shared__ float s_pt[MAXP * MAXP][6];
  const int lane = threadIdx.x;
  if (st->overflow & 4u) return;  // quad colours would not fit: the frame is re-run with room
  const unsigned long long nleaf = st->nleaf;
  const int W = a.W, H = a.H;
  for (unsigned long long g = blockIdx.x; g < nleaf; g += gridDim.x) {
    const LeafGrid lg = dice_leaf<MAXP>(a, lcp, lmeta, lqloc, qoff, leaf_at[g], s_cp, s_row, s_pt, lane);
    const uint32_t qbase = lg.qbase;
    const int gu = lg.gu, gv = lg.gv;
    const size_t pbase = point_base(qbase, a.pt_num);
    const int np = (gu + 1) * (gv + 1);
    for (int pt = lane; pt < np; pt += 64) pxy[pbase + pt] = make_float2(s_pt[pt][0], s_pt[pt][1]);
    int bx0 = 0x7fffffff, by0 = 0x7fffffff, bx1 = 0, by1 = 0;  // union of the pixel boxes
    const int nq = gu * gv;
    for (int qd = lane; qd < nq; qd += 64) {
      const int k = qd / gv, j = qd - k * gv;
      const float *c0 = s_pt[k * (gv + 1) + j], *c1 = s_pt[(k + 1) * (gv + 1) + j];
      const float *c2 = s_pt[(k + 1) * (gv + 1) + j + 1], *c3 = s_pt[k * (gv + 1) + j + 1];
      const float X[4] = {c0[0], c1[0], c2[0], c3[0]}, Y[4] = {c0[1], c1[1], c2[1], c3[1]};
      const uint32_t order = qbase + (uint32_t)qd;
#pragma unroll
      for (int ch = 0; ch < 3; ++ch)  // shade.pikostage:130-138
        qcol[(size_t)order * 3 + ch] = fmn(((c0[3 + ch] + c1[3 + ch]) + c2[3 + ch]) + c3[3 + ch], 4.0f) * 0.25f;
      qz[order] = (((c0[2] + c1[2]) + c2[2]) + c3[2]) * 0.25f;  // shade.pikostage:98
      int x0, y0, x1, y1;
      quad_pixels(X, Y, W, H, &x0, &y0, &x1, &y1);
      qbox[order] = make_ushort4((unsigned short)x0, (unsigned short)y0, (unsigned short)x1, (unsigned short)y1);
      if (x0 < x1) {
        bx0 = min(bx0, x0);
        by0 = min(by0, y0);
        bx1 = max(bx1, x1);
        by1 = max(by1, y1);
      }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      bx0 = min(bx0, __shfl_xor(bx0, d));
      by0 = min(by0, __shfl_xor(by0, d));
      bx1 = max(bx1, __shfl_xor(bx1, d));
      by1 = max(by1, __shfl_xor(by1, d));
    }
    if (bx0 >= bx1) bx0 = by0 = bx1 = by1 = 0;  // no micropolygon covers a pixel
    if (lane == 0)
      linfo[g] = make_uint4(qbase, (uint32_t)gu | ((uint32_t)gv << 8), (uint32_t)bx0 | ((uint32_t)by0 << 16),
                            (uint32_t)bx1 | ((uint32_t)by1 << 16));
    wave_sync();
  }
}

// Tile histogram of one chunk of consecutive leaves in LDS.  Returns the
// chunk's first leaf and its leaf count (0 past the last leaf).
__device__ __forceinline__ int chunk_histogram(const ReyesArgs &a, unsigned long long nleaf,
                                               const uint4 *__restrict__ linfo, uint32_t *s_h,
                                               unsigned long long *g0) {
  const int nt = a.tnx * a.tny;
  *g0 = (unsigned long long)blockIdx.x * kLeafChunk;
  const int n = *g0 < nleaf ? (int)min((unsigned long long)kLeafChunk, nleaf - *g0) : 0;
  for (int k = threadIdx.x; k < nt; k += blockDim.x) s_h[k] = 0;
  __syncthreads();
  for (int l = threadIdx.x; l < n; l += blockDim.x) {
    const TileRange r = tile_range(linfo[*g0 + l], a.tsb);
    for (int ty = r.ty0; ty < r.ty0 + r.nty; ++ty)
      for (int tx = r.tx0; tx < r.tx0 + r.ntx; ++tx) atomicAdd(&s_h[ty * a.tnx + tx], 1u);
  }
  __syncthreads();
  return n;
}

// Leaves per tile: one global add per (chunk, tile) with leaves.
__global__ __launch_bounds__(kBinThreads) void k_reyes_tile_count(ReyesArgs a, const ReyesState *__restrict__ st,
                                                                  const uint4 *__restrict__ linfo,
                                                                  uint32_t *__restrict__ tcnt) {
  __shared__ uint32_t s_h[kMaxSampleTiles];
  if (st->overflow & 4u) return;
  unsigned long long g0;
  if (chunk_histogram(a, st->nleaf, linfo, s_h, &g0) == 0) return;
  for (int k = threadIdx.x; k < a.tnx * a.tny; k += blockDim.x)
    if (s_h[k]) atomicAdd(&tcnt[k], s_h[k]);
}

// Exclusive scans over the tiles (one block) of the leaf counts (list
// offsets and scatter cursors) and of the parts (at least one per tile, so
// every pixel is written), and the tile of every part; the counts are
// cleared for the next frame.
__global__ __launch_bounds__(kBinThreads) void k_reyes_tile_scan(uint32_t *__restrict__ tcnt,
                                                                 uint32_t *__restrict__ toff,
                                                                 uint32_t *__restrict__ tcur,
                                                                 uint32_t *__restrict__ ioff,
                                                                 uint32_t *__restrict__ ptile, int nt,
                                                                 uint32_t pair_cap, ReyesState *__restrict__ st) {
  __shared__ unsigned long long sh[2][kBinThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int run = (nt + kBinThreads - 1) / kBinThreads;
  const int w0 = min(t * run, nt), w1 = min(w0 + run, nt);
  unsigned long long sum = 0, parts = 0;
  for (int w = w0; w < w1; ++w) {
    const uint32_t c = tcnt[w];
    sum += c;
    parts += c > (uint32_t)kPartLeaves ? (c + kPartLeaves - 1) / kPartLeaves : 1u;
  }
  const unsigned long long inc = wave_incl_scan<unsigned long long>(sum, lane);
  const unsigned long long pinc = wave_incl_scan<unsigned long long>(parts, lane);
  if (lane == 63) {
    sh[0][wv] = inc;
    sh[1][wv] = pinc;
  }
  __syncthreads();
  unsigned long long o = inc - sum, po = pinc - parts, total = 0, ptotal = 0;
#pragma unroll
  for (int k = 0; k < kBinThreads / 64; ++k) {
    if (k < wv) {
      o += sh[0][k];
      po += sh[1][k];
    }
    total += sh[0][k];
    ptotal += sh[1][k];
  }
  const bool fits = total <= (unsigned long long)pair_cap;
  for (int w = w0; w < w1; ++w) {
    const uint32_t c = tcnt[w];
    const uint32_t np = c > (uint32_t)kPartLeaves ? (c + kPartLeaves - 1) / kPartLeaves : 1u;
    toff[w] = (uint32_t)o;
    tcur[w] = (uint32_t)o;
    ioff[w] = (uint32_t)po;
    if (fits)  // then the parts fit their table (nt + pair_cap / kPartLeaves entries)
      for (uint32_t q = 0; q < np; ++q) ptile[po + q] = (uint32_t)w;
    tcnt[w] = 0;
    o += c;
    po += np;
  }
  if (t == 0) {
    toff[nt] = (uint32_t)total;
    ioff[nt] = fits ? (uint32_t)ptotal : 0u;
    st->npairs = total > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total;
    if (!fits) atomicOr(&st->overflow, 8u);
  }
}

// Every leaf into the list of each tile its pixel box touches: one global
// add per (chunk, tile) reserves the chunk's run of the tile's list, LDS
// cursors fill it.  The order within a list is immaterial.
__global__ __launch_bounds__(kBinThreads) void k_reyes_tile_scatter(ReyesArgs a, const ReyesState *__restrict__ st,
                                                                    const uint4 *__restrict__ linfo,
                                                                    uint32_t *__restrict__ tcur,
                                                                    uint32_t *__restrict__ tlist) {
  __shared__ uint32_t s_h[kMaxSampleTiles];
  if (st->overflow & 12u) return;
  unsigned long long g0;
  const int n = chunk_histogram(a, st->nleaf, linfo, s_h, &g0);
  if (n == 0) return;
  for (int k = threadIdx.x; k < a.tnx * a.tny; k += blockDim.x) {
    const uint32_t c = s_h[k];
    if (c) s_h[k] = atomicAdd(&tcur[k], c);
  }
  __syncthreads();
  for (int l = threadIdx.x; l < n; l += blockDim.x) {
    const TileRange r = tile_range(linfo[g0 + l], a.tsb);
    for (int ty = r.ty0; ty < r.ty0 + r.nty; ++ty)
      for (int tx = r.tx0; tx < r.tx0 + r.ntx; ++tx)
        tlist[atomicAdd(&s_h[ty * a.tnx + tx], 1u)] = (uint32_t)(g0 + l);
  }
}

// One workgroup per tile part.  Consecutive parts go to different XCDs (the
// hardware deals workgroups out in turn), which spreads the dense tiles of a
// screen region over the whole GPU; giving each XCD a contiguous run of
// tiles instead left some XCDs with the empty part of the screen and took
// 40% longer.  Each wave takes every kTileWaves-th leaf of the part and
// reads their records at once.  In each round it loads the pixel boxes of up
// to kTileBatch chunks of 64 micropolygons together and queues in LDS those
// that touch the tile (order index, leaf, box clipped to the tile).  Every 64
// queued micropolygons are rebuilt from the stored grid points and their
// (micropolygon, pixel) pairs run with the dice kernel's fan tests, keeping
// the minimum key per sample in LDS.  The tile's pixels are resolved as
// k_reyes_resolve does, by the only part or by the last of several.
constexpr int kTileBatch = 4;
constexpr int kTileQueue = 64 * (kTileBatch + 1);

template <int SPP>
__global__ __launch_bounds__(64 * kTileWaves) void k_reyes_tile_sample(
    ReyesArgs a, const ReyesState *__restrict__ st, const uint32_t *__restrict__ toff,
    const uint32_t *__restrict__ ioff, const uint32_t *__restrict__ ptile, const uint32_t *__restrict__ tlist,
    const uint4 *__restrict__ linfo, const float2 *__restrict__ pxy, const float *__restrict__ qz,
    const ushort4 *__restrict__ qbox, const float *__restrict__ qcol, unsigned long long *__restrict__ vis,
    uint32_t *__restrict__ tdone, float4 *__restrict__ rgba, float *__restrict__ depth, int32_t *__restrict__ sz) {
  // keys, sample-major: sample s of tile pixel k at s * 2^(2 tsb) + k (neighbouring
  // pixels in neighbouring banks)
  extern __shared__ unsigned long long s_vis[];
  __shared__ float s_qx[kTileWaves][64][4], s_qy[kTileWaves][64][4];
  __shared__ unsigned long long s_key[kTileWaves][64];
  __shared__ uint32_t s_box[kTileWaves][64];  // x0 | y0 << 13 | width << 26
  __shared__ int s_pre[kTileWaves][64];
  __shared__ int s_own[kTileWaves][64];  // pair owners: start marks per chunk of 64 pairs
  // queue: order index; leaf (in this wave's batch) | clipped box << 6
  // (x0, y0, x1, y1 relative to the tile, 6 bits each)
  __shared__ uint32_t s_qord[kTileWaves][kTileQueue], s_qlr[kTileWaves][kTileQueue];
  __shared__ int s_last;
  if (st->overflow & 12u) return;  // the frame is re-run with room
  const int nt = a.tnx * a.tny;
  const uint32_t item = blockIdx.x;
  if (item >= ioff[nt]) return;
  const int t = (int)ptile[item];
  const int W = a.W, H = a.H, tsb = a.tsb, TS = 1 << tsb;
  for (int k = threadIdx.x; k < (TS * TS) * SPP; k += blockDim.x) s_vis[k] = kVisEmpty;
  const uint32_t part = item - ioff[t], nparts = ioff[t + 1] - ioff[t];
  const int ty = t / a.tnx, tx = t - ty * a.tnx;
  const int rx0 = tx << tsb, ry0 = ty << tsb, rx1 = min(rx0 + TS, W), ry1 = min(ry0 + TS, H);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t lbeg = toff[t] + part * (uint32_t)kPartLeaves;
  const uint32_t lend = min(lbeg + (uint32_t)kPartLeaves, toff[t + 1]);
  // this wave's leaves: lbeg + wv + kTileWaves * m, m < nmine (at most 64)
  const int nmine = lbeg + (uint32_t)wv < lend ? (int)((lend - lbeg - (uint32_t)wv + kTileWaves - 1) / kTileWaves) : 0;
  uint4 mine = make_uint4(0u, 0u, 0u, 0u);
  if (lane < nmine) mine = linfo[tlist[lbeg + (uint32_t)wv + (uint32_t)(kTileWaves * lane)]];
  __syncthreads();  // s_vis is cleared
  uint32_t *qord = s_qord[wv], *qlr = s_qlr[wv];
  int qn = 0;         // queued micropolygons
  int m = 0, q0 = 0;  // the next chunk: micropolygons q0.. of leaf m
  while (m < nmine || qn > 0) {
    ushort4 b[kTileBatch];
    uint32_t ord[kTileBatch], lf[kTileBatch];
    bool has[kTileBatch];
#pragma unroll
    for (int r = 0; r < kTileBatch; ++r) {
      has[r] = false;
      if (m < nmine) {
        const uint32_t qbase = __shfl(mine.x, m), grid = __shfl(mine.y, m);
        const int nq = (int)(grid & 0xffu) * (int)((grid >> 8) & 0xffu);
        const int qd = q0 + lane;
        if (qd < nq) {
          ord[r] = qbase + (uint32_t)qd;
          b[r] = qbox[ord[r]];
          lf[r] = (uint32_t)m;
          has[r] = true;
        }
        q0 += 64;
        if (q0 >= nq) {
          ++m;
          q0 = 0;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kTileBatch; ++r) {
      bool hit = false;
      uint32_t lr = 0;
      if (has[r]) {
        const int x0 = max((int)b[r].x, rx0), y0 = max((int)b[r].y, ry0);
        const int x1 = min((int)b[r].z, rx1), y1 = min((int)b[r].w, ry1);
        if (x0 < x1 && y0 < y1) {
          hit = true;
          lr = lf[r] | ((uint32_t)(x0 - rx0) << 6) | ((uint32_t)(y0 - ry0) << 12) | ((uint32_t)(x1 - rx0) << 18) |
               ((uint32_t)(y1 - ry0) << 24);
        }
      }
      const unsigned long long mask = __ballot(hit);
      if (hit) {
        const int at = qn + __popcll(mask & ((1ull << lane) - 1ull));
        qord[at] = ord[r];
        qlr[at] = lr;
      }
      qn += __popcll(mask);
    }
    wave_sync();
    while (qn >= 64 || (m >= nmine && qn > 0)) {  // 64 queued at a time; the rest once no leaf is left
      const int n = min(qn, 64);
      const uint32_t order = lane < n ? qord[lane] : 0u, lr = lane < n ? qlr[lane] : 0u;
      const int mi = (int)(lr & 63u);
      const uint32_t qbase = __shfl(mine.x, mi), grid = __shfl(mine.y, mi);
      int x0 = 0, y0 = 0, wid = 0, npx = 0;
      if (lane < n) {
        const int gv = (int)((grid >> 8) & 0xffu), qd = (int)(order - qbase);
        const size_t pbase = point_base(qbase, a.pt_num);
        const int k = qd / gv, j = qd - k * gv;
        const float2 c0 = pxy[pbase + k * (gv + 1) + j], c1 = pxy[pbase + (k + 1) * (gv + 1) + j];
        const float2 c2 = pxy[pbase + (k + 1) * (gv + 1) + j + 1], c3 = pxy[pbase + k * (gv + 1) + j + 1];
        const float z = qz[order];
        s_qx[wv][lane][0] = c0.x;
        s_qx[wv][lane][1] = c1.x;
        s_qx[wv][lane][2] = c2.x;
        s_qx[wv][lane][3] = c3.x;
        s_qy[wv][lane][0] = c0.y;
        s_qy[wv][lane][1] = c1.y;
        s_qy[wv][lane][2] = c2.y;
        s_qy[wv][lane][3] = c3.y;
        s_key[wv][lane] = ((unsigned long long)(__float_as_uint(z) ^ 0x80000000u) << 32) |
                          (unsigned long long)(0xFFFFFFFEu - order);
        x0 = rx0 + (int)((lr >> 6) & 63u);
        y0 = ry0 + (int)((lr >> 12) & 63u);
        wid = rx0 + (int)((lr >> 18) & 63u) - x0;
        npx = wid * (ry0 + (int)((lr >> 24) & 63u) - y0);
      }
      const int inc = wave_incl_scan<int>(npx, lane);
      const int total = __shfl(inc, 63);
      s_box[wv][lane] = (uint32_t)x0 | ((uint32_t)y0 << 13) | ((uint32_t)wid << 26);
      s_pre[wv][lane] = inc - npx;
      wave_sync();
      // pair owners and rows as in k_reyes_dice
      const int pre = inc - npx;
      const bool fquot = total < (1 << 21);
      int carry = 0;
      for (int p0 = 0; p0 < total; p0 += 64) {
        s_own[wv][lane] = -1;
        wave_sync();
        const int rel = pre - p0;
        if (npx > 0 && rel >= 0 && rel < 64) s_own[wv][rel] = lane;
        wave_sync();
        const int mark = s_own[wv][lane];
        const unsigned long long mk = __ballot(mark >= 0) & ((2ull << lane) - 1ull);
        const int got = __shfl(mark, mk ? 63 - __clzll((long long)mk) : 0);
        const int q = mk ? got : carry;
        carry = __shfl(q, 63);
        wave_sync();
        const int p = p0 + lane;
        if (p >= total) break;  // the last chunk
        const uint32_t bx = s_box[wv][q];
        const int local = p - s_pre[wv][q], w = (int)(bx >> 26);
        const int dy = fquot ? (int)(((float)local + 0.5f) * __builtin_amdgcn_rcpf((float)w)) : local / w;
        const int dx = local - dy * w;
        const int x = (int)(bx & 0x1fffu) + dx, y = (int)((bx >> 13) & 0x1fffu) + dy;
        const float X[4] = {s_qx[wv][q][0], s_qx[wv][q][1], s_qx[wv][q][2], s_qx[wv][q][3]};
        const float Y[4] = {s_qy[wv][q][0], s_qy[wv][q][1], s_qy[wv][q][2], s_qy[wv][q][3]};
        const unsigned long long key = s_key[wv][q];
        unsigned long long *dst = s_vis + (((y - ry0) << tsb) + (x - rx0));
#pragma unroll
        for (int smp = 0; smp < SPP; ++smp)
          if (quad_hit(X, Y, (float)x + sample_off(SPP, smp, 0), (float)y + sample_off(SPP, smp, 1)))
            atomicMin(dst + (smp << (2 * tsb)), key);
      }
      // the queue's remainder to its front (each pass reads 64 entries past
      // the ones it writes)
      for (int e = lane; e < qn - n; e += 64) {
        const uint32_t o = qord[e + n], l = qlr[e + n];
        qord[e] = o;
        qlr[e] = l;
      }
      qn -= n;
      wave_sync();
    }
  }
  __syncthreads();
  if (nparts > 1) {
    // meet the other parts in the global buffer; the last part resolves
    for (int k = threadIdx.x; k < (TS * TS) * SPP; k += blockDim.x) {
      const unsigned long long key = s_vis[k];
      if (key != kVisEmpty) {
        const int s = k >> (2 * tsb), px = k & (TS * TS - 1);
        const size_t p = (size_t)(ry0 + (px >> tsb)) * W + (size_t)(rx0 + (px & (TS - 1)));
        atomicMin(vis + p * SPP + s, key);
      }
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&tdone[t], 1u) == nparts - 1;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // the merged keys, and the words reset for the next frame
    for (int k = threadIdx.x; k < (TS * TS) * SPP; k += blockDim.x) {
      const int s = k >> (2 * tsb), px = k & (TS * TS - 1);
      const int x = rx0 + (px & (TS - 1)), y = ry0 + (px >> tsb);
      if (x < rx1 && y < ry1) s_vis[k] = atomicExch(vis + ((size_t)y * W + x) * SPP + s, kVisEmpty);
    }
    if (threadIdx.x == 0) atomicExch(&tdone[t], 0u);
    __syncthreads();
  }
  // PikoScreen clear colour 0xff663313 through fromABGR (math.h:323-330)
  const float clr[3] = {(float)0x13 / 255.0f, (float)0x33 / 255.0f, (float)0x66 / 255.0f};
  for (int k = threadIdx.x; k < TS * TS; k += blockDim.x) {
    const int x = rx0 + (k & (TS - 1)), y = ry0 + (k >> tsb);
    if (x >= rx1 || y >= ry1) continue;
    const size_t p = (size_t)y * W + x;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    int32_t zmin = 0;
#pragma unroll
    for (int s = 0; s < SPP; ++s) {
      const unsigned long long key = s_vis[(s << (2 * tsb)) + k];
      const uint32_t low = (uint32_t)key;
      float c[3];
      int32_t zi;
      if (low == 0xFFFFFFFFu) {
        c[0] = clr[0];
        c[1] = clr[1];
        c[2] = clr[2];
        zi = 0x3F800000;
      } else {
        const unsigned long long order = 0xFFFFFFFEull - low;
        c[0] = qcol[order * 3 + 0];
        c[1] = qcol[order * 3 + 1];
        c[2] = qcol[order * 3 + 2];
        zi = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
      }
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) acc[ch] = s == 0 ? c[ch] : acc[ch] + c[ch];
      zmin = (s == 0 || zi < zmin) ? zi : zmin;
      if (a.keep_samples) sz[p * SPP + s] = zi;
    }
    const float inv = 1.0f / (float)SPP;
    rgba[p] = make_float4(acc[0] * inv, acc[1] * inv, acc[2] * inv, 1.0f);
    depth[p] = __int_as_float(zmin);
  }
}

// ---------------------------------------------------------------------------
// Resolve: box filter, nearest depth, visibility reset.

__global__ __launch_bounds__(256) void k_reyes_resolve(ReyesArgs a, unsigned long long *__restrict__ vis,
                                                       const float *__restrict__ qcol, float4 *__restrict__ rgba,
                                                       float *__restrict__ depth, int32_t *__restrict__ sz) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long long)a.W * a.H) return;
  const int spp = a.spp;
  // PikoScreen clear colour 0xff663313 through fromABGR (math.h:323-330)
  const float clr[3] = {(float)0x13 / 255.0f, (float)0x33 / 255.0f, (float)0x66 / 255.0f};
  float acc[3] = {0.0f, 0.0f, 0.0f};
  int32_t zmin = 0;
  for (int s = 0; s < spp; ++s) {
    const size_t si = (size_t)p * spp + s;
    const unsigned long long key = vis[si];
    const uint32_t low = (uint32_t)key;
    float c[3];
    int32_t zi;
    if (low == 0xFFFFFFFFu) {
      c[0] = clr[0];
      c[1] = clr[1];
      c[2] = clr[2];
      zi = 0x3F800000;
    } else {
      const unsigned long long order = 0xFFFFFFFEull - low;
      c[0] = qcol[order * 3 + 0];
      c[1] = qcol[order * 3 + 1];
      c[2] = qcol[order * 3 + 2];
      zi = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[ch] = s == 0 ? c[ch] : acc[ch] + c[ch];
    zmin = (s == 0 || zi < zmin) ? zi : zmin;
    if (a.keep_samples) sz[si] = zi;
    if (low != 0xFFFFFFFFu) vis[si] = kVisEmpty;
  }
  const float inv = 1.0f / (float)spp;
  rgba[p] = make_float4(acc[0] * inv, acc[1] * inv, acc[2] * inv, 1.0f);
  depth[p] = __int_as_float(zmin);
}

__global__ void k_fill_u64(unsigned long long *p, long long n, unsigned long long v) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}

inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace

struct piko_reyes_pipe {
  int W = 0, H = 0, P = 0;
  piko_reyes_options opt{};
  ReyesArgs ra{};
  hipStream_t stream = nullptr;
  // Global sampling in run_frames: split and ranks run on stream, dice on
  // rstream, resolve on vstream, so a frame's split and ranks and the
  // previous frame's resolve run beside a frame's dice.  Consecutive frames
  // alternate between two slots of what split and ranks write and dice
  // reads, and of the visibility buffer and quad colours that dice writes
  // and resolve reads and clears (the d_* pointers below are the current
  // slot, a_* the other one).  A split waits for the dice that last read its
  // slot (e_slot), a dice for its frame's ranks (e_rank) and for the resolve
  // that last cleared its slot (e_res), a resolve for its frame's dice.
  hipStream_t rstream = nullptr, vstream = nullptr;
  hipEvent_t e_rank = nullptr, e_slot[2] = {nullptr, nullptr}, e_res[2] = {nullptr, nullptr};
  hipEvent_t e_last = nullptr;  // the last resolve enqueued on vstream
  bool res_pending = false;     // stream has not yet waited for e_last
  int slot = 0;
  bool alt_ready = false;
  float4 *a_lcp = nullptr;
  uint32_t *a_lmeta = nullptr, *a_lqloc = nullptr, *a_leaf_at = nullptr, *a_off = nullptr, *a_qoff = nullptr;
  ReyesState *a_st = nullptr;
  unsigned long long *a_vis = nullptr;
  float *a_qcol = nullptr;
  long long alt_bytes = 0;
  const float4 *d_cp = nullptr;
  bool owns_input = false;
  float4 *d_front = nullptr, *d_lcp = nullptr;
  uint32_t *d_lmeta = nullptr, *d_lqloc = nullptr;
  uint32_t *d_cnt = nullptr, *d_qcnt = nullptr, *d_lvlpre = nullptr, *d_off = nullptr, *d_wstat = nullptr;
  uint32_t *d_leaf_at = nullptr;  // emission rank -> region slot
  uint32_t *d_qoff = nullptr;     // [level][wave] order index of the group's first quad
  unsigned long long *d_lvltot = nullptr;  // [level][leaves, quads]
  ReyesState *d_st = nullptr;
  float *d_qcol = nullptr;
  unsigned long long *d_vis = nullptr;
  float4 *d_rgba = nullptr;
  float *d_depth = nullptr;
  int32_t *d_sz = nullptr;
  // tiled sampling
  bool tiled = false;
  float2 *d_pxy = nullptr;     // grid point screen positions
  float *d_qz = nullptr;       // flat depth per micropolygon (order index)
  ushort4 *d_qbox = nullptr;   // pixel box per micropolygon
  uint4 *d_linfo = nullptr;    // per leaf (emission rank): first order index, grid, pixel box
  uint32_t *d_tcnt = nullptr, *d_toff = nullptr, *d_tcur = nullptr, *d_tlist = nullptr;
  uint32_t *d_ioff = nullptr, *d_tdone = nullptr;  // parts per tile (scanned), finished parts
  uint32_t *d_ptile = nullptr;                      // tile of every part
  long long tlist_bytes = 0;
  int cus = 256, dice_grid = 0;
  int reruns = 0;  // capacity re-runs of the last run call
  long long device_bytes = 0;
  long long front_bytes = 0, leaf_bytes = 0, qcol_bytes = 0;
};

namespace {

#define HIPOK(x)                                \
  do {                                          \
    if ((x) != hipSuccess) return PIKO_ERR_HIP; \
  } while (0)

template <typename T>
int ralloc(piko_reyes_pipe *p, T **ptr, long long elems, long long *acct = nullptr) {
  const size_t bytes = (size_t)std::max<long long>(elems, 1) * sizeof(T);
  if (hipMalloc((void **)ptr, bytes) != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();  // reported as PIKO_ERR_OOM: not left for a later launch check to find
    return PIKO_ERR_OOM;
  }
  p->device_bytes += (long long)bytes;
  if (acct) *acct += (long long)bytes;
  return PIKO_OK;
}

template <typename T>
void rfree(T *&ptr) {
  if (ptr) (void)hipFree((void *)ptr);
  ptr = nullptr;
}

void free_front(piko_reyes_pipe *p) {
  rfree(p->d_front);
  p->device_bytes -= p->front_bytes;
  p->front_bytes = 0;
}

int alloc_front(piko_reyes_pipe *p) {
  return ralloc(p, &p->d_front, (long long)p->ra.G * 2 * p->ra.front_cap * 16, &p->front_bytes);
}

// The second slot of the split and rank outputs (see piko_reyes_pipe).
void free_alt(piko_reyes_pipe *p) {
  rfree(p->a_lcp);
  rfree(p->a_lmeta);
  rfree(p->a_lqloc);
  rfree(p->a_leaf_at);
  rfree(p->a_off);
  rfree(p->a_qoff);
  rfree(p->a_st);
  rfree(p->a_vis);
  rfree(p->a_qcol);
  p->device_bytes -= p->alt_bytes;
  p->alt_bytes = 0;
  p->alt_ready = false;
}

int alloc_alt(piko_reyes_pipe *p) {
  const long long n = (long long)p->ra.G * p->ra.leaf_cap, NL = (long long)kLevels * p->ra.G;
  int rc = ralloc(p, &p->a_lcp, n * 16, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_lmeta, n, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_lqloc, n, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_leaf_at, n, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_off, NL, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_qoff, NL, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_st, 1, &p->alt_bytes);
  const long long ns = (long long)p->W * p->H * p->ra.spp;
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_vis, ns, &p->alt_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->a_qcol, (long long)p->ra.qcap * 3, &p->alt_bytes);
  if (rc == PIKO_OK) {
    // empty keys before any frame uses the slot: the fill runs on stream,
    // ahead of every split, so ahead of every dice and resolve
    k_fill_u64<<<(unsigned)std::min<long long>(cdiv(ns, 256), 4096), 256, 0, p->stream>>>(p->a_vis, ns, kVisEmpty);
    if (hipGetLastError() != hipSuccess) rc = PIKO_ERR_HIP;
  }
  if (rc != PIKO_OK) {
    free_alt(p);
    return rc;
  }
  p->alt_ready = true;
  return PIKO_OK;
}

void swap_slot(piko_reyes_pipe *p) {
  std::swap(p->d_lcp, p->a_lcp);
  std::swap(p->d_lmeta, p->a_lmeta);
  std::swap(p->d_lqloc, p->a_lqloc);
  std::swap(p->d_leaf_at, p->a_leaf_at);
  std::swap(p->d_off, p->a_off);
  std::swap(p->d_qoff, p->a_qoff);
  std::swap(p->d_st, p->a_st);
  std::swap(p->d_vis, p->a_vis);
  std::swap(p->d_qcol, p->a_qcol);
  p->slot ^= 1;
}

void free_leaves(piko_reyes_pipe *p) {
  free_alt(p);  // sized by the leaf capacity too; made again when next needed
  rfree(p->d_lcp);
  rfree(p->d_lmeta);
  rfree(p->d_lqloc);
  rfree(p->d_leaf_at);
  rfree(p->d_linfo);
  p->device_bytes -= p->leaf_bytes;
  p->leaf_bytes = 0;
}

int alloc_leaves(piko_reyes_pipe *p) {
  const long long n = (long long)p->ra.G * p->ra.leaf_cap;
  int rc = ralloc(p, &p->d_lcp, n * 16, &p->leaf_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_lmeta, n, &p->leaf_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_lqloc, n, &p->leaf_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_leaf_at, n, &p->leaf_bytes);
  if (rc == PIKO_OK && p->tiled) rc = ralloc(p, &p->d_linfo, n, &p->leaf_bytes);
  return rc;
}

void free_qcol(piko_reyes_pipe *p) {
  free_alt(p);  // its quad colours are sized by the same capacity
  rfree(p->d_qcol);
  rfree(p->d_pxy);
  rfree(p->d_qz);
  rfree(p->d_qbox);
  p->device_bytes -= p->qcol_bytes;
  p->qcol_bytes = 0;
}

int alloc_qcol(piko_reyes_pipe *p) {
  const long long q = (long long)p->ra.qcap;
  int rc = ralloc(p, &p->d_qcol, q * 3, &p->qcol_bytes);
  if (rc == PIKO_OK && p->tiled) rc = ralloc(p, &p->d_pxy, ((q * p->ra.pt_num) >> 2) + 1, &p->qcol_bytes);
  if (rc == PIKO_OK && p->tiled) rc = ralloc(p, &p->d_qz, q, &p->qcol_bytes);
  if (rc == PIKO_OK && p->tiled) rc = ralloc(p, &p->d_qbox, q, &p->qcol_bytes);
  return rc;
}

// Parts of tiled sampling: at most one per tile plus one per kPartLeaves
// list entries.
long long max_parts(const ReyesArgs &a) {
  return (long long)a.tnx * a.tny + ((long long)a.pair_cap + kPartLeaves - 1) / kPartLeaves;
}

void free_tlist(piko_reyes_pipe *p) {
  rfree(p->d_tlist);
  rfree(p->d_ptile);
  p->device_bytes -= p->tlist_bytes;
  p->tlist_bytes = 0;
}

int alloc_tlist(piko_reyes_pipe *p) {
  int rc = ralloc(p, &p->d_tlist, (long long)p->ra.pair_cap, &p->tlist_bytes);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_ptile, max_parts(p->ra), &p->tlist_bytes);
  return rc;
}

const char *kReyesKernels[] = {"k_reyes_split", "k_reyes_levels+k_reyes_rank", "k_reyes_dice", "k_reyes_resolve"};
const char *kReyesTiledKernels[] = {"k_reyes_split", "k_reyes_levels+k_reyes_rank",
                                    "k_reyes_dice_bin+k_reyes_tile_count+k_reyes_tile_scan+k_reyes_tile_scatter",
                                    "k_reyes_tile_sample"};

// overlap: global sampling's split and ranks, dice and resolve on three
// streams, consecutive frames in alternate slots (see piko_reyes_pipe)
int enqueue_frame(piko_reyes_pipe *p, hipEvent_t *ev, bool overlap = false) {
  const ReyesArgs &a = p->ra;
  const int NL = kLevels * a.G;
  (void)NL;
  if (overlap) HIPOK(hipStreamWaitEvent(p->stream, p->e_slot[p->slot], 0));  // the last dice to read this slot
  k_reyes_split<<<a.G, 64, 0, p->stream>>>(p->d_cp, a, p->d_front, p->d_lcp, p->d_lmeta, p->d_lqloc, p->d_cnt,
                                           p->d_qcnt, p->d_lvlpre, p->d_wstat);
  if (ev) (void)hipEventRecord(ev[0], p->stream);
  k_reyes_levels<<<kLevels, kRankThreads, 0, p->stream>>>(p->d_cnt, p->d_qcnt, a.G, p->d_lvltot);
  k_reyes_rank<<<kLevels, kRankThreads, 0, p->stream>>>(p->d_cnt, p->d_qcnt, p->d_lvlpre, p->d_wstat, p->d_lvltot,
                                                        a.G, a.leaf_cap, a.qcap, p->d_off, p->d_qoff, p->d_leaf_at,
                                                        p->d_st);
  if (ev) (void)hipEventRecord(ev[1], p->stream);
  const bool small = a.dice > 0 && a.dice <= 8;
  if (p->tiled) {
    const int nt = a.tnx * a.tny;
#define PIKO_DICE_BIN(MAXP)                                                                                   \
  k_reyes_dice_bin<MAXP><<<p->dice_grid, 64, 0, p->stream>>>(a, p->d_lcp, p->d_lmeta, p->d_lqloc, p->d_leaf_at, \
                                                             p->d_qoff, p->d_st, p->d_qcol, p->d_pxy, p->d_qz,   \
                                                             p->d_qbox, p->d_linfo)
    if (small) PIKO_DICE_BIN(9);
    else PIKO_DICE_BIN(34);
#undef PIKO_DICE_BIN
    const unsigned chunks = (unsigned)cdiv((long long)a.G * a.leaf_cap, kLeafChunk);
    k_reyes_tile_count<<<chunks, kBinThreads, 0, p->stream>>>(a, p->d_st, p->d_linfo, p->d_tcnt);
    k_reyes_tile_scan<<<1, kBinThreads, 0, p->stream>>>(p->d_tcnt, p->d_toff, p->d_tcur, p->d_ioff, p->d_ptile, nt,
                                                       a.pair_cap, p->d_st);
    k_reyes_tile_scatter<<<chunks, kBinThreads, 0, p->stream>>>(a, p->d_st, p->d_linfo, p->d_tcur, p->d_tlist);
    if (ev) (void)hipEventRecord(ev[2], p->stream);
    const size_t lds = ((size_t)1 << (2 * a.tsb)) * (size_t)a.spp * 8;
    const unsigned grid = (unsigned)max_parts(a);
#define PIKO_TILE_SAMPLE(SPP)                                                                                    \
  k_reyes_tile_sample<SPP><<<grid, 64 * kTileWaves, lds, p->stream>>>(a, p->d_st, p->d_toff, p->d_ioff, p->d_ptile, \
                                                                      p->d_tlist,                                   \
                                                                      p->d_linfo, p->d_pxy, p->d_qz, p->d_qbox,    \
                                                                      p->d_qcol, p->d_vis, p->d_tdone, p->d_rgba,  \
                                                                      p->d_depth, p->d_sz)
    if (a.spp == 4) PIKO_TILE_SAMPLE(4);
    else PIKO_TILE_SAMPLE(1);
#undef PIKO_TILE_SAMPLE
    if (ev) (void)hipEventRecord(ev[3], p->stream);
    return hipGetLastError() == hipSuccess ? PIKO_OK : PIKO_ERR_HIP;
  }
  hipStream_t ds = p->stream, vs = p->stream;
  if (overlap) {
    HIPOK(hipEventRecord(p->e_rank, p->stream));
    HIPOK(hipStreamWaitEvent(p->rstream, p->e_rank, 0));
    HIPOK(hipStreamWaitEvent(p->rstream, p->e_res[p->slot], 0));  // this slot's buffer cleared
    ds = p->rstream;
    vs = p->vstream;
  } else if (p->res_pending) {
    HIPOK(hipStreamWaitEvent(p->stream, p->e_last, 0));
    p->res_pending = false;
  }
#define PIKO_DICE(MAXP, SPP)                                                                           \
  k_reyes_dice<MAXP, SPP><<<p->dice_grid, 64, 0, ds>>>(a, p->d_lcp, p->d_lmeta, p->d_lqloc, p->d_leaf_at, \
                                                       p->d_qoff, p->d_st, p->d_qcol, p->d_vis)
  if (a.spp == 4) {
    if (small) PIKO_DICE(9, 4);
    else PIKO_DICE(34, 4);
  } else {
    if (small) PIKO_DICE(9, 1);
    else PIKO_DICE(34, 1);
  }
#undef PIKO_DICE
  if (ev) (void)hipEventRecord(ev[2], p->stream);
  if (overlap) {
    HIPOK(hipEventRecord(p->e_slot[p->slot], p->rstream));
    HIPOK(hipStreamWaitEvent(p->vstream, p->e_slot[p->slot], 0));
  }
  k_reyes_resolve<<<(unsigned)cdiv((long long)a.W * a.H, 256), 256, 0, vs>>>(a, p->d_vis, p->d_qcol, p->d_rgba,
                                                                             p->d_depth, p->d_sz);
  if (overlap) {
    HIPOK(hipEventRecord(p->e_res[p->slot], p->vstream));
    p->e_last = p->e_res[p->slot];
    p->res_pending = true;
  }
  if (ev) (void)hipEventRecord(ev[3], p->stream);
  return hipGetLastError() == hipSuccess ? PIKO_OK : PIKO_ERR_HIP;
}

int read_state(piko_reyes_pipe *p, ReyesState *s) {
  HIPOK(hipMemcpyAsync(s, p->d_st, sizeof(ReyesState), hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

// Grows whatever overflowed in the last frame; *rerun = 1 if it did.
int check_capacity(piko_reyes_pipe *p, int *rerun) {
  ReyesState s;
  int rc = read_state(p, &s);
  if (rc != PIKO_OK) return rc;
  *rerun = 0;
  if (!s.overflow) return PIKO_OK;
  if (s.overflow & 1u) {
    free_front(p);
    p->ra.front_cap *= 2;
    if ((rc = alloc_front(p)) != PIKO_OK) return rc;
  }
  if (s.overflow & 2u) {
    free_leaves(p);
    p->ra.leaf_cap *= 2;
    if ((rc = alloc_leaves(p)) != PIKO_OK) return rc;
  }
  if ((s.overflow & 4u) && !(s.overflow & 3u)) {
    if (s.nquad >= kMaxOrder) return PIKO_ERR_CAPACITY;
    free_qcol(p);
    p->ra.qcap = std::min<unsigned long long>(s.nquad + s.nquad / 8 + 4096, kMaxOrder);
    if ((rc = alloc_qcol(p)) != PIKO_OK) return rc;
  }
  if ((s.overflow & 8u) && !(s.overflow & 7u)) {  // the pair count is exact only for a complete dice
    const unsigned long long want = (unsigned long long)s.npairs + s.npairs / 4 + 4096;
    if (s.npairs == 0xFFFFFFFFu || want > kMaxPairs) return PIKO_ERR_CAPACITY;
    free_tlist(p);
    p->ra.pair_cap = (uint32_t)want;
    if ((rc = alloc_tlist(p)) != PIKO_OK) return rc;
  }
  *rerun = 1;
  return PIKO_OK;
}

bool valid_dice(int d) { return d == 0 || d == 1 || d == 2 || d == 4 || d == 8 || d == 16 || d == 32; }

void set_camera(ReyesArgs *a, const piko_constant_state *cs) {
  std::memcpy(a->vp, cs->viewProjMatrix, sizeof(a->vp));
}

int make_pipe(piko_reyes_pipe **out, const piko_constant_state *cs, const float *patches, int count,
              const piko_reyes_options *opt, bool on_device) {
  if (!out) return PIKO_ERR_INVALID;
  *out = nullptr;
  if (!cs || !opt || count < 0 || (count > 0 && !patches)) return PIKO_ERR_INVALID;
  const int W = cs->screenSizeX, H = cs->screenSizeY;
  if (W <= 0 || H <= 0 || W > 8192 || H > 8192) return PIKO_ERR_INVALID;
  if (!(opt->threshold > 0.0f) || opt->max_splits < 0 || opt->max_splits > kSplitCapMax) return PIKO_ERR_INVALID;
  if (!valid_dice(opt->dice) || (opt->spp != 1 && opt->spp != 4)) return PIKO_ERR_INVALID;
  if (!(opt->displacement == opt->displacement) || opt->roots_per_wave < 0 || opt->roots_per_wave > 256)
    return PIKO_ERR_INVALID;
  if (opt->sampling < PIKO_SAMPLE_AUTO || opt->sampling > PIKO_SAMPLE_TILED) return PIKO_ERR_INVALID;
  if (opt->tile != 0 && opt->tile != 16 && opt->tile != 32) return PIKO_ERR_INVALID;
  // tiled sampling bins into at most kMaxSampleTiles tiles: 16 pixels (or
  // the tile asked for), 32 where a screen would have more; a TILED request
  // for a larger screen is refused
  int tsb = opt->tile == 32 ? 5 : 4;
  if (cdiv(W, 1 << tsb) * cdiv(H, 1 << tsb) > kMaxSampleTiles) tsb = 5;
  const bool tiles_fit = cdiv(W, 1 << tsb) * cdiv(H, 1 << tsb) <= kMaxSampleTiles;
  if (opt->sampling == PIKO_SAMPLE_TILED && !tiles_fit) return PIKO_ERR_INVALID;
  // global sampling scans the (micropolygon, pixel) pairs of a wave's 64
  // quads in int32: a quad's clipped box holds at most W*H pixels, so the
  // scan cannot overflow while 64*W*H < 2^31; larger screens take tiled
  // sampling under AUTO and are refused for GLOBAL
  const bool global_fits = 64ll * W * H < (1ll << 31);
  if (opt->sampling == PIKO_SAMPLE_GLOBAL && !global_fits) return PIKO_ERR_INVALID;
  if (opt->sampling == PIKO_SAMPLE_AUTO && !global_fits && !tiles_fit) return PIKO_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PIKO_ERR_NO_DEVICE;
  piko_reyes_pipe *p = new (std::nothrow) piko_reyes_pipe();
  if (!p) return PIKO_ERR_OOM;
  p->W = W;
  p->H = H;
  p->P = count;
  p->opt = *opt;
  ReyesArgs &a = p->ra;
  a.W = W;
  a.H = H;
  set_camera(&a, cs);
  a.threshold = opt->threshold;
  a.max_splits = opt->max_splits;
  a.dice = opt->dice;
  a.spp = opt->spp;
  a.disp = opt->displacement;
  a.roots_per_wave = opt->roots_per_wave > 0 ? opt->roots_per_wave : 1;
  a.nroots = count;
  a.G = (int)std::max<long long>(cdiv(count, a.roots_per_wave), 1);
  a.front_cap = 64 * a.roots_per_wave;
  a.leaf_cap = 64 * a.roots_per_wave;
  const unsigned long long per_leaf = a.dice > 0 ? (unsigned long long)a.dice * a.dice : 33ull * 33ull;
  a.qcap = std::min<unsigned long long>((unsigned long long)a.G * a.leaf_cap * per_leaf,
                                        std::max<unsigned long long>(1ull << 24, (unsigned long long)count * 64));
  a.keep_samples = opt->keep_samples ? 1 : 0;
  a.tsb = tsb;
  a.tnx = (int)cdiv(W, 1 << tsb);
  a.tny = (int)cdiv(H, 1 << tsb);
  p->tiled = opt->sampling == PIKO_SAMPLE_TILED || !global_fits;
  a.pt_num = a.dice == 1 ? 16u : 9u;
  a.pair_cap = (uint32_t)std::min<long long>(std::max<long long>(4096, (long long)count * 64), 1ll << 30);
  int rc = PIKO_OK;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0) p->cus = prop.multiProcessorCount;
  // one-wave dice workgroups, 128 per CU: on the sphere frame that is about
  // one leaf each, and the hardware hands the next workgroup to whichever
  // SIMD frees first.  16 per CU, each walking eight leaves in turn, took 15%
  // longer.
  p->dice_grid = p->cus * 128;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&p->rstream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&p->vstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->e_rank, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->e_slot[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->e_slot[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->e_res[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->e_res[1], hipEventDisableTiming) != hipSuccess)
    rc = PIKO_ERR_HIP;
  if (rc == PIKO_OK) {
    if (on_device) {
      p->d_cp = (const float4 *)patches;
    } else {
      float4 *d = nullptr;
      rc = ralloc(p, &d, (long long)count * 16);
      if (rc == PIKO_OK && count > 0 &&
          hipMemcpy(d, patches, (size_t)count * 256, hipMemcpyHostToDevice) != hipSuccess)
        rc = PIKO_ERR_HIP;
      p->d_cp = d;
      p->owns_input = true;
    }
  }
  const long long NL = (long long)kLevels * a.G;
  const long long npix = (long long)W * H, ns = npix * a.spp;
  if (rc == PIKO_OK) rc = alloc_front(p);
  if (rc == PIKO_OK) rc = alloc_leaves(p);
  if (rc == PIKO_OK) rc = alloc_qcol(p);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_cnt, NL);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_qcnt, NL);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_lvlpre, NL);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_off, NL);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_qoff, NL);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_lvltot, 2 * kLevels);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_wstat, a.G);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_st, 1);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_vis, ns);  // tiled: where split tiles meet
  if (rc == PIKO_OK && p->tiled) {
    const long long nt = (long long)a.tnx * a.tny;
    rc = ralloc(p, &p->d_tcnt, nt);
    if (rc == PIKO_OK) rc = ralloc(p, &p->d_toff, nt + 1);
    if (rc == PIKO_OK) rc = ralloc(p, &p->d_tcur, nt);
    if (rc == PIKO_OK) rc = ralloc(p, &p->d_ioff, nt + 1);
    if (rc == PIKO_OK) rc = ralloc(p, &p->d_tdone, nt);
    if (rc == PIKO_OK) rc = alloc_tlist(p);
    if (rc == PIKO_OK && (hipMemsetAsync(p->d_tcnt, 0, (size_t)nt * 4, p->stream) != hipSuccess ||
                          hipMemsetAsync(p->d_tdone, 0, (size_t)nt * 4, p->stream) != hipSuccess))
      rc = PIKO_ERR_HIP;
  }
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_rgba, npix);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_depth, npix);
  if (rc == PIKO_OK) rc = ralloc(p, &p->d_sz, a.keep_samples ? ns : 1);
  if (rc == PIKO_OK)
    k_fill_u64<<<(unsigned)std::min<long long>(cdiv(ns, 256), 4096), 256, 0, p->stream>>>(p->d_vis, ns, kVisEmpty);
  if (rc == PIKO_OK && (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess))
    rc = PIKO_ERR_HIP;
  if (rc != PIKO_OK) {
    piko_reyes_destroy(p);
    return rc;
  }
  *out = p;
  return PIKO_OK;
}

}  // namespace

extern "C" {

int piko_reyes_allocate(piko_reyes_pipe **out, const piko_constant_state *cs, const float *patches, int count,
                        const piko_reyes_options *opt) {
  return make_pipe(out, cs, patches, count, opt, false);
}

int piko_reyes_allocate_device(piko_reyes_pipe **out, const piko_constant_state *cs, const float *d_patches,
                               int count, const piko_reyes_options *opt) {
  return make_pipe(out, cs, d_patches, count, opt, true);
}

int piko_reyes_prepare(piko_reyes_pipe *p, const piko_constant_state *cs) {
  if (!p || !cs) return PIKO_ERR_INVALID;
  if (cs->screenSizeX != p->W || cs->screenSizeY != p->H) return PIKO_ERR_INVALID;
  set_camera(&p->ra, cs);
  return PIKO_OK;
}

int piko_reyes_run_frames(piko_reyes_pipe *p, int nframes) {
  if (!p || nframes < 0) return PIKO_ERR_INVALID;
  p->reruns = 0;
  // with more than one frame, each frame's dice and resolve overlap the next
  // frame's split and ranks, the frames alternating between two slots
  bool overlap = !p->tiled && nframes > 1;
  if (overlap && !p->alt_ready && alloc_alt(p) != PIKO_OK) overlap = false;  // no room: one frame at a time
  for (int attempt = 0; attempt < 12; ++attempt) {
    if (overlap && !p->alt_ready && alloc_alt(p) != PIKO_OK) overlap = false;  // regrown leaves freed it
    for (int f = 0; f < nframes; ++f) {
      if (overlap && f > 0) swap_slot(p);
      const int rc = enqueue_frame(p, nullptr, overlap);
      if (rc != PIKO_OK) {
        if (p->res_pending) (void)hipStreamWaitEvent(p->stream, p->e_last, 0);
        p->res_pending = false;
        return rc;
      }
    }
    if (p->res_pending) {  // the caller's later work on the pipe's stream follows the last resolve
      HIPOK(hipStreamWaitEvent(p->stream, p->e_last, 0));
      p->res_pending = false;
    }
    int rerun = 0;
    const int rc = check_capacity(p, &rerun);
    if (rc != PIKO_OK) return rc;
    if (!rerun) return PIKO_OK;
    if (nframes == 0) return PIKO_OK;
    ++p->reruns;
  }
  return PIKO_ERR_CAPACITY;
}

int piko_reyes_run_single(piko_reyes_pipe *p) { return piko_reyes_run_frames(p, 1); }

int piko_reyes_profile_frames(piko_reyes_pipe *p, int nframes, float *kernel_ms, const char **names, int cap,
                              int *nk, float *frame_ms) {
  if (!p || nframes <= 0 || !kernel_ms || !nk) return PIKO_ERR_INVALID;
  int rc = piko_reyes_run_frames(p, 1);  // settles the capacities
  if (rc != PIKO_OK) return rc;
  hipEvent_t ev0, ev[4];
  HIPOK(hipEventCreate(&ev0));
  for (int k = 0; k < 4; ++k) HIPOK(hipEventCreate(&ev[k]));
  double acc[4] = {0, 0, 0, 0}, facc = 0;
  for (int f = 0; f < nframes; ++f) {
    HIPOK(hipEventRecord(ev0, p->stream));
    if ((rc = enqueue_frame(p, ev)) != PIKO_OK) return rc;
    HIPOK(hipEventSynchronize(ev[3]));
    float ms;
    HIPOK(hipEventElapsedTime(&ms, ev0, ev[0]));
    acc[0] += ms;
    for (int k = 1; k < 4; ++k) {
      HIPOK(hipEventElapsedTime(&ms, ev[k - 1], ev[k]));
      acc[k] += ms;
    }
    HIPOK(hipEventElapsedTime(&ms, ev0, ev[3]));
    facc += ms;
  }
  const int n = std::min(4, cap);
  for (int k = 0; k < n; ++k) {
    kernel_ms[k] = (float)(acc[k] / nframes);
    if (names) names[k] = p->tiled ? kReyesTiledKernels[k] : kReyesKernels[k];
  }
  *nk = n;
  if (frame_ms) *frame_ms = (float)(facc / nframes);
  (void)hipEventDestroy(ev0);
  for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ev[k]);
  int rerun = 0;
  rc = check_capacity(p, &rerun);
  return rc != PIKO_OK ? rc : (rerun ? PIKO_ERR_CAPACITY : PIKO_OK);
}

int piko_reyes_get_color(piko_reyes_pipe *p, float *rgba) {
  if (!p || !rgba) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(rgba, p->d_rgba, (size_t)p->W * p->H * 16, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

int piko_reyes_get_depth(piko_reyes_pipe *p, float *depth) {
  if (!p || !depth) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(depth, p->d_depth, (size_t)p->W * p->H * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

int piko_reyes_get_samples(piko_reyes_pipe *p, int32_t *z) {
  if (!p || !z || !p->ra.keep_samples) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(z, p->d_sz, (size_t)p->W * p->H * p->ra.spp * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

long long piko_reyes_get_leaves(piko_reyes_pipe *p, float *cp, int *nu, int *nv, long long cap) {
  if (!p || cap < 0 || (cap > 0 && !cp)) return PIKO_ERR_INVALID;
  ReyesState s;
  int rc = read_state(p, &s);
  if (rc != PIKO_OK) return rc;
  const ReyesArgs &a = p->ra;
  const long long NL = (long long)kLevels * a.G, region = (long long)a.G * a.leaf_cap;
  std::vector<uint32_t> cnt(NL), off(NL), pre(NL), meta(region);
  std::vector<float4> leaves(cap > 0 ? region * 16 : 0);
  HIPOK(hipMemcpy(cnt.data(), p->d_cnt, NL * 4, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(off.data(), p->d_off, NL * 4, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(pre.data(), p->d_lvlpre, NL * 4, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(meta.data(), p->d_lmeta, region * 4, hipMemcpyDeviceToHost));
  if (cap > 0) HIPOK(hipMemcpy(leaves.data(), p->d_lcp, region * 256, hipMemcpyDeviceToHost));
  for (long long grp = 0; grp < NL; ++grp) {
    const long long w = grp % a.G;
    for (uint32_t k = 0; k < cnt[grp]; ++k) {
      const long long rank = (long long)off[grp] + k;
      if (rank >= cap) continue;
      const long long at = w * a.leaf_cap + pre[grp] + k;
      std::memcpy(cp + rank * 64, &leaves[at * 16], 256);
      if (nu) nu[rank] = (int)(meta[at] & 0xffu);
      if (nv) nv[rank] = (int)((meta[at] >> 8) & 0xffu);
    }
  }
  return (long long)s.nleaf;
}

int piko_reyes_get_stats(piko_reyes_pipe *p, piko_reyes_stats *st) {
  if (!p || !st) return PIKO_ERR_INVALID;
  ReyesState s;
  const int rc = read_state(p, &s);
  if (rc != PIKO_OK) return rc;
  st->patches = p->P;
  st->leaves = (long long)s.nleaf;
  st->micropolygons = (long long)s.nquad;
  st->levels = (int)s.levels;
  st->device_bytes = p->device_bytes;
  st->capacity_reruns = p->reruns;
  st->tile_pairs = p->tiled ? (long long)s.npairs : 0;
  return PIKO_OK;
}

int piko_reyes_destroy(piko_reyes_pipe *p) {
  if (!p) return PIKO_ERR_INVALID;
  if (p->rstream) (void)hipStreamSynchronize(p->rstream);
  if (p->vstream) (void)hipStreamSynchronize(p->vstream);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  free_front(p);
  free_leaves(p);
  free_qcol(p);
  rfree(p->d_cnt);
  rfree(p->d_qcnt);
  rfree(p->d_lvlpre);
  rfree(p->d_off);
  rfree(p->d_qoff);
  rfree(p->d_lvltot);
  rfree(p->d_wstat);
  rfree(p->d_st);
  rfree(p->d_vis);
  rfree(p->d_tcnt);
  rfree(p->d_toff);
  rfree(p->d_tcur);
  rfree(p->d_ioff);
  rfree(p->d_tdone);
  free_tlist(p);
  rfree(p->d_rgba);
  rfree(p->d_depth);
  rfree(p->d_sz);
  if (p->owns_input) {
    float4 *d = const_cast<float4 *>(p->d_cp);
    rfree(d);
  }
  if (p->e_rank) (void)hipEventDestroy(p->e_rank);
  for (hipEvent_t e : p->e_slot)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->e_res)
    if (e) (void)hipEventDestroy(e);
  if (p->vstream) (void)hipStreamDestroy(p->vstream);
  if (p->rstream) (void)hipStreamDestroy(p->rstream);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return PIKO_OK;
}

}  // extern "C"
