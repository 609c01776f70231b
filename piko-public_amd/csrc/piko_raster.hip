// piko_raster.hip -- the Piko fixed-point raster pipeline on MI355X (gfx950).
//
// Frame = the kernel sequence pikoc plans for RasterPipe with --opt
// (pikoc.cpp:206-318, PTXBackend.cpp:875-984), re-designed for CDNA4:
//
//   k_vs_count      VertexShaderStage::process fused with RasterStage::assignBin's
//                   bin count (vertexShader.pikostage:36-109, raster.pikostage:40-117)
//   scan            exclusive scan of per-triangle bin counts
//   k_emit_pairs    (bin, triangle) pairs in submission order
//   radix passes    stable count / scan / scatter of the pairs by bin: the per-bin
//                   queues, built with LDS histograms and wave ballots, no global
//                   atomics on bins (replaces Bin::insert, datatypes.h:88-93)
//   k_bin_bounds    bin -> [start, end) of its queue
//   k_raster<T>     RasterStage::process + depth test + Gouraud + PikoScreen store
//                   (raster.pikostage:124-321, stage.h:57-61) per bin, with
//                   the bin's depth/colour resolved in LDS and written once.
//
// Compositing: the reference's depth loop (raster.pikostage:280-287) passes a
// fragment iff its z bits (as signed int) are <= the stored value, and the last
// passing fragment's colour stays.  Over a submission-ordered stream that is:
// depth = min z, colour = the last fragment that attains the min.  k_raster
// computes exactly that with a 64-bit LDS atomicMin on (z bits, 0xFFFFFFFE-tri),
// then shades only the winner, so the result does not depend on the order in
// which lanes process the bin's queue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "piko_hip.h"
#include "raster_kernels.h"

using namespace piko;

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // elements per scan / sort block

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan over the block (NT threads, NT a multiple of 64).
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) {
      const uint32_t t = sh[k];
      sh[k] = run;
      run += t;
    }
    sh[NT / 64] = run;
  }
  __syncthreads();
  const uint32_t r = sh[w] + inc - v;
  total = sh[NT / 64];
  __syncthreads();
  return r;
}

// ---- generic exclusive scan (reduce, top, down) -------------------------

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t *__restrict__ in, long long n,
                                                           uint32_t *__restrict__ blk) {
  __shared__ uint32_t sh[kThreads / 64 + 1];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t total;
  block_excl_scan<kThreads>(s, sh, total);
  if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// One block scans the block sums in place (exclusive) and stores the total
// at out[n].
__global__ void __launch_bounds__(1024) k_scan_top(uint32_t *__restrict__ blk, int nblk,
                                                   uint32_t *__restrict__ out, long long n) {
  __shared__ uint32_t sh[1024 / 64 + 1];
  uint32_t carry = 0;
  for (int base = 0; base < nblk; base += 1024) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < nblk ? blk[i] : 0u;
    uint32_t total;
    const uint32_t e = block_excl_scan<1024>(v, sh, total);
    if (i < nblk) blk[i] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ void __launch_bounds__(kThreads) k_scan_down(const uint32_t *__restrict__ in, long long n,
                                                         const uint32_t *__restrict__ blk,
                                                         uint32_t *__restrict__ out) {
  __shared__ uint32_t sh[kThreads / 64 + 1];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_excl_scan<kThreads>(s, sh, total) + blk[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (base + k < n) {
      out[base + k] = run;
      run += v[k];
    }
}

// ---- vertex stage + bin count -------------------------------------------

template <int KIND>
__global__ void __launch_bounds__(kThreads) k_vs_count(const float *__restrict__ pos,
                                                        const void *__restrict__ attr, int n, FrameArgs fa,
                                                        int4 *__restrict__ geo, uint4 *__restrict__ colr,
                                                        uint32_t *__restrict__ cnt,
                                                        uint32_t *__restrict__ keep_blk) {
  __shared__ uint32_t sh[kThreads / 64 + 1];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  uint32_t count = 0, keep = 0;
  if (i < n) {
    int p[6];
    float z[3];
    uint32_t c[3];
    const float *v = pos + 9 * (size_t)i;
    if (KIND == PIKO_INPUT_PRETRANSFORMED) {
      const uint32_t *cc = (const uint32_t *)attr + 3 * (size_t)i;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        p[2 * k] = fixpt(v[3 * k], fa.S);
        p[2 * k + 1] = fixpt(v[3 * k + 1], fa.S);
        z[k] = v[3 * k + 2];
        c[k] = cc[k];
      }
    } else {
      const float *nn = (const float *)attr + 9 * (size_t)i;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        transform_vertex(v + 3 * k, fa, p[2 * k], p[2 * k + 1], z[k]);
        c[k] = light_vertex(nn[3 * k], nn[3 * k + 1], nn[3 * k + 2], fa);
      }
    }
    if (vs_keep(p, fa.W, fa.H, fa.S)) {
      keep = 1;
      Edges e;
      e.set(p);
      const float rcp = 1.0f / (float)e.denom();  // raster.pikostage:174
      const float z0mz2 = z[0] - z[2], z1mz2 = z[1] - z[2];
      geo[3 * (size_t)i + 0] = make_int4(p[0], p[1], p[2], p[3]);
      geo[3 * (size_t)i + 1] = make_int4(p[4], p[5], __float_as_int(z[2]), __float_as_int(z0mz2));
      geo[3 * (size_t)i + 2] = make_int4(__float_as_int(z1mz2), __float_as_int(rcp), 0, 0);
      colr[i] = make_uint4(c[0], c[1], c[2], 0u);
      const int bsb = fa.S + fa.lgT;
      int bx1, by1, bx2, by2;
      bin_rect(p, bsb, fa.nBx, fa.nBy, bx1, by1, bx2, by2);
      for (int by = by1; by <= by2; ++by)
        for (int bx = bx1; bx <= bx2; ++bx) count += tr_accept(e, bx, by, bsb) ? 1u : 0u;
    }
    cnt[i] = count;
  }
  uint32_t total;
  block_excl_scan<kThreads>(keep, sh, total);
  if (threadIdx.x == 0) keep_blk[blockIdx.x] = total;
}

// (bin, triangle) pairs at the scanned offsets, walk order within a triangle.
__global__ void __launch_bounds__(kThreads) k_emit_pairs(const int4 *__restrict__ geo,
                                                          const uint32_t *__restrict__ off, int n, FrameArgs fa,
                                                          long long cap, uint32_t *__restrict__ keys,
                                                          uint32_t *__restrict__ vals,
                                                          uint32_t *__restrict__ flags) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i == 0 && (long long)off[n] > cap) flags[0] = 1u;  // queue overflow: reported
  if (i >= n) return;
  uint32_t o = off[i];
  const uint32_t e_ = off[i + 1];
  if (o == e_) return;
  const int4 a = geo[3 * (size_t)i + 0], b = geo[3 * (size_t)i + 1];
  const int p[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
  Edges e;
  e.set(p);
  const int bsb = fa.S + fa.lgT;
  int bx1, by1, bx2, by2;
  bin_rect(p, bsb, fa.nBx, fa.nBy, bx1, by1, bx2, by2);
  for (int by = by1; by <= by2; ++by)
    for (int bx = bx1; bx <= bx2; ++bx)
      if (tr_accept(e, bx, by, bsb)) {
        if ((long long)o < cap) {
          keys[o] = (uint32_t)(by * fa.nBx + bx);
          vals[o] = (uint32_t)i;
        }
        ++o;
      }
}

// ---- stable LSD radix sort of the pairs by bin ---------------------------

__global__ void __launch_bounds__(kThreads) k_radix_hist(const uint32_t *__restrict__ keys,
                                                          const uint32_t *__restrict__ total, long long cap,
                                                          uint32_t *__restrict__ hist, int nblk, int shift) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long P = min((long long)*total, cap);
  const long long base = (long long)blockIdx.x * kTile;
#pragma unroll 4
  for (int r = 0; r < kItems; ++r) {
    const long long idx = base + r * kThreads + threadIdx.x;
    if (idx < P) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: rounds of 256 consecutive keys; within a wave the rank of a
// key among equal digits comes from 9 ballots (8 digit bits + validity),
// across the block's 4 waves from per-wave digit counts in LDS.
__global__ void __launch_bounds__(kThreads) k_radix_scatter(const uint32_t *__restrict__ kin,
                                                             const uint32_t *__restrict__ vin,
                                                             uint32_t *__restrict__ kout,
                                                             uint32_t *__restrict__ vout,
                                                             const uint32_t *__restrict__ total, long long cap,
                                                             const uint32_t *__restrict__ hist, int nblk,
                                                             int shift) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kThreads / 64][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long P = min((long long)*total, cap);
  const long long base = (long long)blockIdx.x * kTile;
  if (base >= P) return;
  run[t] = hist[(size_t)t * nblk + blockIdx.x];
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int r = 0; r < kItems; ++r) {
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) wcnt[k][t] = 0;
    __syncthreads();
    const long long idx = base + r * kThreads + t;
    const bool valid = idx < P;
    uint32_t key = 0, val = 0;
    if (valid) {
      key = kin[idx];
      val = vin[idx];
    }
    const uint32_t d = valid ? ((key >> shift) & 255u) : 256u;
    unsigned long long peers = ~0ull;
#pragma unroll
    for (int b = 0; b < 9; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = __popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t o = run[d] + rank;
      for (int k = 0; k < w; ++k) o += wcnt[k][d];
      kout[o] = key;
      vout[o] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) add += wcnt[k][t];
    run[t] += add;
  }
}

// bin_start[b] = first queue index with key >= b; bin_start[nbins] = P.
__global__ void __launch_bounds__(kThreads) k_bin_bounds(const uint32_t *__restrict__ keys,
                                                          const uint32_t *__restrict__ total, long long cap,
                                                          uint32_t *__restrict__ start, int nbins) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long P = min((long long)*total, cap);
  if (P == 0) {
    if (i <= nbins) start[i] = 0;
    return;
  }
  if (i >= P) return;
  const int k = (int)keys[i];
  const int kp = i == 0 ? -1 : (int)keys[i - 1];
  for (int b = kp + 1; b <= k; ++b) start[b] = (uint32_t)i;
  if (i == P - 1)
    for (int b = k + 1; b <= nbins; ++b) start[b] = (uint32_t)P;
}

// ---- tile raster ----------------------------------------------------------

template <int T, int NT>
__global__ void __launch_bounds__(NT) k_raster(const int4 *__restrict__ geo, const uint4 *__restrict__ colr,
                                               const uint32_t *__restrict__ bin_start,
                                               const uint32_t *__restrict__ bin_tris, FrameArgs fa,
                                               uint32_t *__restrict__ color, float *__restrict__ depth) {
  __shared__ unsigned long long vis[T * T];
  __shared__ uint32_t coff[NT + 1];
  __shared__ uint32_t s_rect[NT];  // rx | ry << 12 | w << 24
  __shared__ float s_rcpw[NT];
  __shared__ int s_xy[6][NT];
  __shared__ float s_zr[4][NT];    // z2, z0mz2, z1mz2, onebybary
  __shared__ uint32_t s_tri[NT];
  __shared__ float s_u8[256];
  __shared__ uint32_t sh[NT / 64 + 1];

  const int t = threadIdx.x;
  const int bin = blockIdx.x;
  const int bx = bin % fa.nBx, by = bin / fa.nBx;
  const int xlo = bx * T, ylo = by * T;
  const int xhi = min(xlo + T, fa.W), yhi = min(ylo + T, fa.H);
  const int S = fa.S;
  for (int q = t; q < T * T; q += NT) vis[q] = kVisSentinel;
  for (int q = t; q < 256; q += NT) s_u8[q] = (float)q / 255.0f;  // fromABGR, math.h:323-330
  const uint32_t beg = bin_start[bin], end = bin_start[bin + 1];
  __syncthreads();

  for (uint32_t b0 = beg; b0 < end; b0 += NT) {
    const uint32_t j = b0 + t;
    uint32_t nc = 0;
    if (j < end) {
      const uint32_t tri = bin_tris[j];
      const int4 a = geo[3 * (size_t)tri + 0], b = geo[3 * (size_t)tri + 1], c = geo[3 * (size_t)tri + 2];
      const int p[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
      // pixel range of the rounded bbox (rasterMacros.h:58-63) inside the bin
      const int lox = (imin3(p[0], p[2], p[4]) & (int)0xffffff00) >> S;
      const int loy = (imin3(p[1], p[3], p[5]) & (int)0xffffff00) >> S;
      const int hix = ((imax3(p[0], p[2], p[4]) + 0xff) & (int)0xffffff00) >> S;
      const int hiy = ((imax3(p[1], p[3], p[5]) + 0xff) & (int)0xffffff00) >> S;
      const int x0 = max(lox, xlo), x1 = min(hix, xhi);
      const int y0 = max(loy, ylo), y1 = min(hiy, yhi);
      const int w = max(0, x1 - x0), h = max(0, y1 - y0);
      nc = (uint32_t)(w * h);
      s_rect[t] = (uint32_t)(x0 & 0xfff) | ((uint32_t)(y0 & 0xfff) << 12) | ((uint32_t)w << 24);
      s_rcpw[t] = w ? 1.0f / (float)w : 0.0f;
#pragma unroll
      for (int k = 0; k < 6; ++k) s_xy[k][t] = p[k];
      s_zr[0][t] = __int_as_float(b.z);
      s_zr[1][t] = __int_as_float(b.w);
      s_zr[2][t] = __int_as_float(c.x);
      s_zr[3][t] = __int_as_float(c.y);
      s_tri[t] = tri;
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<NT>(nc, sh, tot);
    coff[t] = ex;
    if (t == 0) coff[NT] = tot;
    __syncthreads();
    for (uint32_t cidx = t; cidx < tot; cidx += NT) {
      // owner: largest o with coff[o] <= cidx (coff is non-decreasing)
      int lo = 0, hi = NT - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (coff[mid] <= cidx) lo = mid;
        else hi = mid - 1;
      }
      const int o = lo;
      const uint32_t k = cidx - coff[o];
      const uint32_t rect = s_rect[o];
      const int w = (int)(rect >> 24);
      const int q = (int)(((float)k + 0.5f) * s_rcpw[o]);
      const int r = (int)k - q * w;
      const int px = (int)(rect & 0xfff) + r, py = (int)((rect >> 12) & 0xfff) + q;
      Edges e;
      const int p[6] = {s_xy[0][o], s_xy[1][o], s_xy[2][o], s_xy[3][o], s_xy[4][o], s_xy[5][o]};
      e.set(p);
      long long e0, e1, e2;
      pixel_edges(e, px, py, S, e0, e1, e2);
      if ((e0 | e1 | e2) >= 0) {
        const float rcp = s_zr[3][o];
        const float alpha = (float)e1 * rcp;  // raster.pikostage:269-272
        const float beta = (float)e2 * rcp;
        const float z = interp(s_zr[1][o], s_zr[2][o], s_zr[0][o], alpha, beta);
        atomicMin((long long *)&vis[(py - ylo) * T + (px - xlo)], (long long)vis_key(z, s_tri[o]));
      }
    }
    __syncthreads();
  }

  // resolve: depth = winning z, colour = Gouraud of the winner (raster.pikostage:289-316)
  for (int q = t; q < T * T; q += NT) {
    const int px = xlo + (q % T), py = ylo + (q / T);
    if (px >= xhi || py >= yhi) continue;
    const unsigned long long key = vis[q];
    uint32_t col = kClearColor;
    float d = 1.0f;
    if (key != kVisSentinel) {
      const uint32_t tri = 0xFFFFFFFEu - (uint32_t)(key & 0xFFFFFFFFull);
      d = __int_as_float((int)(key >> 32));
      const int4 a = geo[3 * (size_t)tri + 0], b = geo[3 * (size_t)tri + 1], c = geo[3 * (size_t)tri + 2];
      const int p[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
      Edges e;
      e.set(p);
      long long e0, e1, e2;
      pixel_edges(e, px, py, S, e0, e1, e2);
      const float rcp = __int_as_float(c.y);
      const float alpha = (float)e1 * rcp;
      const float beta = (float)e2 * rcp;
      const uint4 cc = colr[tri];
      float rgb[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        const float v0 = s_u8[(cc.x >> (8 * ch)) & 255u];
        const float v1 = s_u8[(cc.y >> (8 * ch)) & 255u];
        const float v2 = s_u8[(cc.z >> (8 * ch)) & 255u];
        rgb[ch] = interp(v0 - v2, v1 - v2, v2, alpha, beta);
      }
      col = to_abgr(rgb[0], rgb[1], rgb[2]);
    }
    const size_t pid = (size_t)py * fa.W + px;
    color[pid] = col;
    depth[pid] = d;
  }
}

inline int ilog2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace

// ---- pipe object --------------------------------------------------------

struct piko_raster_pipe {
  int W = 0, H = 0, S = 0, T = 0, lgT = 0, nBx = 0, nBy = 0, nbins = 0, kind = 0, count = 0;
  FrameArgs fa{};
  hipStream_t stream = nullptr;
  const float *d_pos = nullptr;
  const void *d_attr = nullptr;
  bool owns_input = false;
  int4 *d_geo = nullptr;
  uint4 *d_col = nullptr;
  uint32_t *d_cnt = nullptr;   // [n]
  uint32_t *d_off = nullptr;   // [n+1]
  uint32_t *d_blk = nullptr;   // scan block sums
  long long blk_cap = 0;
  uint32_t *d_keys[2] = {nullptr, nullptr};
  uint32_t *d_vals[2] = {nullptr, nullptr};
  long long cap = 0;
  uint32_t *d_hist = nullptr;  // [256 * nblk_cap + 1]
  int nblk_cap = 0;
  uint32_t *d_start = nullptr; // [nbins+1]
  uint32_t *d_flags = nullptr; // [0] overflow
  uint32_t *d_keep = nullptr;  // per vertex-stage block: emitted triangles
  uint32_t *d_color = nullptr;
  float *d_depth = nullptr;
  int passes = 0;
  int final_buf = 0;
  long long device_bytes = 0;
};

namespace {

#define HIPOK(x)                                  \
  do {                                            \
    if ((x) != hipSuccess) return PIKO_ERR_HIP;   \
  } while (0)

template <typename T>
int dalloc(piko_raster_pipe *p, T **ptr, long long elems) {
  const size_t bytes = (size_t)std::max<long long>(elems, 1) * sizeof(T);
  if (hipMalloc((void **)ptr, bytes) != hipSuccess) {
    *ptr = nullptr;
    return PIKO_ERR_OOM;
  }
  p->device_bytes += (long long)bytes;
  return PIKO_OK;
}

void free_pairs(piko_raster_pipe *p) {
  for (int k = 0; k < 2; ++k) {
    if (p->d_keys[k]) (void)hipFree(p->d_keys[k]);
    if (p->d_vals[k]) (void)hipFree(p->d_vals[k]);
    p->d_keys[k] = p->d_vals[k] = nullptr;
  }
  if (p->d_hist) (void)hipFree(p->d_hist);
  p->d_hist = nullptr;
  if (p->d_blk) (void)hipFree(p->d_blk);
  p->d_blk = nullptr;
  p->device_bytes -= (long long)(4 * p->cap * sizeof(uint32_t)) +
                     (long long)((256ll * p->nblk_cap + 1) * sizeof(uint32_t)) +
                     (long long)(std::max<long long>(p->blk_cap, 1) * sizeof(uint32_t));
}

int alloc_pairs(piko_raster_pipe *p, long long cap) {
  p->cap = cap;
  p->nblk_cap = (int)cdiv(cap, kTile);
  int rc;
  for (int k = 0; k < 2; ++k) {
    if ((rc = dalloc(p, &p->d_keys[k], cap)) != PIKO_OK) return rc;
    if ((rc = dalloc(p, &p->d_vals[k], cap)) != PIKO_OK) return rc;
  }
  if ((rc = dalloc(p, &p->d_hist, 256ll * p->nblk_cap + 1)) != PIKO_OK) return rc;
  const long long scan_n = std::max<long long>(p->count, 256ll * p->nblk_cap);
  p->blk_cap = cdiv(scan_n, kTile) + 1;
  if ((rc = dalloc(p, &p->d_blk, p->blk_cap)) != PIKO_OK) return rc;
  return PIKO_OK;
}

void scan(piko_raster_pipe *p, const uint32_t *in, long long n, uint32_t *out) {
  const int nb = (int)std::max<long long>(cdiv(n, kTile), 1);
  k_scan_reduce<<<nb, kThreads, 0, p->stream>>>(in, n, p->d_blk);
  k_scan_top<<<1, 1024, 0, p->stream>>>(p->d_blk, nb, out, n);
  k_scan_down<<<nb, kThreads, 0, p->stream>>>(in, n, p->d_blk, out);
}

// Kernel names in launch order, for profiling.
const char *kNames[] = {"k_vs_count", "k_scan_reduce", "k_scan_top", "k_scan_down", "k_emit_pairs",
                        "k_radix_hist", "k_scan_reduce", "k_scan_top", "k_scan_down", "k_radix_scatter",
                        "k_radix_hist", "k_scan_reduce", "k_scan_top", "k_scan_down", "k_radix_scatter",
                        "k_radix_hist", "k_scan_reduce", "k_scan_top", "k_scan_down", "k_radix_scatter",
                        "k_bin_bounds", "k_raster"};

// Enqueues one frame.  ev (optional) receives an event after each kernel.
int enqueue_frame(piko_raster_pipe *p, hipEvent_t *ev, int *nev) {
  int e = 0;
  auto mark = [&]() {
    if (ev) (void)hipEventRecord(ev[e], p->stream);
    ++e;
  };
  const int n = p->count;
  const int gv = (int)std::max<long long>(cdiv(n, kThreads), 1);
  if (p->kind == PIKO_INPUT_PRETRANSFORMED)
    k_vs_count<PIKO_INPUT_PRETRANSFORMED><<<gv, kThreads, 0, p->stream>>>(
        p->d_pos, p->d_attr, n, p->fa, p->d_geo, p->d_col, p->d_cnt, p->d_keep);
  else
    k_vs_count<PIKO_INPUT_WORLD><<<gv, kThreads, 0, p->stream>>>(p->d_pos, p->d_attr, n, p->fa, p->d_geo,
                                                                 p->d_col, p->d_cnt, p->d_keep);
  mark();
  {
    const int nb = (int)std::max<long long>(cdiv(n, kTile), 1);
    k_scan_reduce<<<nb, kThreads, 0, p->stream>>>(p->d_cnt, n, p->d_blk);
    mark();
    k_scan_top<<<1, 1024, 0, p->stream>>>(p->d_blk, nb, p->d_off, n);
    mark();
    k_scan_down<<<nb, kThreads, 0, p->stream>>>(p->d_cnt, n, p->d_blk, p->d_off);
    mark();
  }
  k_emit_pairs<<<gv, kThreads, 0, p->stream>>>(p->d_geo, p->d_off, n, p->fa, p->cap, p->d_keys[0],
                                               p->d_vals[0], p->d_flags);
  mark();
  const uint32_t *total = p->d_off + n;
  int cur = 0;
  const long long hn = 256ll * p->nblk_cap;
  for (int pass = 0; pass < p->passes; ++pass) {
    k_radix_hist<<<p->nblk_cap, kThreads, 0, p->stream>>>(p->d_keys[cur], total, p->cap, p->d_hist,
                                                          p->nblk_cap, 8 * pass);
    mark();
    const int nb = (int)std::max<long long>(cdiv(hn, kTile), 1);
    k_scan_reduce<<<nb, kThreads, 0, p->stream>>>(p->d_hist, hn, p->d_blk);
    mark();
    k_scan_top<<<1, 1024, 0, p->stream>>>(p->d_blk, nb, p->d_hist, hn);
    mark();
    k_scan_down<<<nb, kThreads, 0, p->stream>>>(p->d_hist, hn, p->d_blk, p->d_hist);
    mark();
    k_radix_scatter<<<p->nblk_cap, kThreads, 0, p->stream>>>(p->d_keys[cur], p->d_vals[cur], p->d_keys[cur ^ 1],
                                                             p->d_vals[cur ^ 1], total, p->cap, p->d_hist,
                                                             p->nblk_cap, 8 * pass);
    mark();
    cur ^= 1;
  }
  p->final_buf = cur;
  {
    const long long nthreads = std::max<long long>(p->cap, p->nbins + 1);
    k_bin_bounds<<<(int)cdiv(nthreads, kThreads), kThreads, 0, p->stream>>>(p->d_keys[cur], total, p->cap,
                                                                           p->d_start, p->nbins);
    mark();
  }
  switch (p->T) {
    case 8:
      k_raster<8, 64><<<p->nbins, 64, 0, p->stream>>>(p->d_geo, p->d_col, p->d_start, p->d_vals[cur], p->fa,
                                                      p->d_color, p->d_depth);
      break;
    case 16:
      k_raster<16, 256><<<p->nbins, 256, 0, p->stream>>>(p->d_geo, p->d_col, p->d_start, p->d_vals[cur], p->fa,
                                                         p->d_color, p->d_depth);
      break;
    case 32:
      k_raster<32, 256><<<p->nbins, 256, 0, p->stream>>>(p->d_geo, p->d_col, p->d_start, p->d_vals[cur], p->fa,
                                                         p->d_color, p->d_depth);
      break;
    default:
      k_raster<64, 256><<<p->nbins, 256, 0, p->stream>>>(p->d_geo, p->d_col, p->d_start, p->d_vals[cur], p->fa,
                                                         p->d_color, p->d_depth);
      break;
  }
  mark();
  if (nev) *nev = e;
  return hipGetLastError() == hipSuccess ? PIKO_OK : PIKO_ERR_HIP;
}

// Reads the overflow flag and the pair total; grows the queues when they
// overflowed.  Returns 1 if the frame must be re-run.
int check_capacity(piko_raster_pipe *p, int *rerun) {
  uint32_t flags[1], total;
  HIPOK(hipMemcpyAsync(flags, p->d_flags, sizeof(flags), hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipMemcpyAsync(&total, p->d_off + p->count, sizeof(total), hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  *rerun = 0;
  if (flags[0] || (long long)total > p->cap) {
    free_pairs(p);
    int rc = alloc_pairs(p, (long long)total + (long long)total / 4 + 4096);
    if (rc != PIKO_OK) return rc;
    HIPOK(hipMemsetAsync(p->d_flags, 0, sizeof(uint32_t), p->stream));
    *rerun = 1;
  }
  return PIKO_OK;
}

int validate_opts(const piko_constant_state *cs, int count, const piko_raster_options *opt) {
  if (!cs || !opt || count < 0) return PIKO_ERR_INVALID;
  if (cs->screenSizeX <= 0 || cs->screenSizeY <= 0 || cs->screenSizeX > 4095 || cs->screenSizeY > 4095)
    return PIKO_ERR_INVALID;
  const int T = opt->bin_size;
  if (T != 8 && T != 16 && T != 32 && T != 64) return PIKO_ERR_INVALID;
  if (opt->subpixel_bits < 1 || opt->subpixel_bits > 8) return PIKO_ERR_INVALID;
  if (opt->input_kind != PIKO_INPUT_PRETRANSFORMED && opt->input_kind != PIKO_INPUT_WORLD)
    return PIKO_ERR_INVALID;
  return PIKO_OK;
}

void set_frame_args(piko_raster_pipe *p, const piko_constant_state *cs) {
  FrameArgs &fa = p->fa;
  fa.W = p->W;
  fa.H = p->H;
  fa.S = p->S;
  fa.lgT = p->lgT;
  fa.nBx = p->nBx;
  fa.nBy = p->nBy;
  fa.halfW = cs->halfW;
  fa.halfH = cs->halfH;
  std::memcpy(fa.vp, cs->viewProjMatrix, sizeof(fa.vp));
  std::memcpy(fa.light, cs->light, sizeof(fa.light));
  std::memcpy(fa.matcol, cs->matcol, sizeof(fa.matcol));
  std::memcpy(fa.ambient, cs->ambient, sizeof(fa.ambient));
}

int allocate_common(piko_raster_pipe **out, const piko_constant_state *cs, int count,
                    const piko_raster_options *opt, piko_raster_pipe *p) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PIKO_ERR_NO_DEVICE;
  p->W = cs->screenSizeX;
  p->H = cs->screenSizeY;
  p->S = opt->subpixel_bits;
  p->T = opt->bin_size;
  p->lgT = ilog2i(p->T);
  p->nBx = (p->W + p->T - 1) / p->T;
  p->nBy = (p->H + p->T - 1) / p->T;
  p->nbins = p->nBx * p->nBy;
  p->kind = opt->input_kind;
  p->count = count;
  int bits = 0;
  while ((1ll << bits) < (long long)p->nbins) ++bits;
  p->passes = (bits + 7) / 8;
  set_frame_args(p, cs);
  HIPOK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  int rc;
  if ((rc = dalloc(p, &p->d_geo, 3ll * count)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_col, (long long)count)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_cnt, (long long)count)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_off, (long long)count + 1)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_start, (long long)p->nbins + 1)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_flags, 4)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_keep, std::max<long long>(cdiv(count, kThreads), 1))) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_color, (long long)p->W * p->H)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_depth, (long long)p->W * p->H)) != PIKO_OK) return rc;
  const long long cap = opt->pair_capacity > 0 ? opt->pair_capacity
                                               : std::max<long long>(2ll * count + 1024, 1 << 16);
  if ((rc = alloc_pairs(p, cap)) != PIKO_OK) return rc;
  HIPOK(hipMemsetAsync(p->d_flags, 0, 4 * sizeof(uint32_t), p->stream));
  HIPOK(hipMemsetAsync(p->d_start, 0, ((size_t)p->nbins + 1) * sizeof(uint32_t), p->stream));
  HIPOK(hipMemsetAsync(p->d_off, 0, ((size_t)count + 1) * sizeof(uint32_t), p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  *out = p;
  return PIKO_OK;
}

}  // namespace

extern "C" {

const char *piko_status_string(int s) {
  switch (s) {
    case PIKO_OK: return "ok";
    case PIKO_ERR_INVALID: return "invalid argument";
    case PIKO_ERR_HIP: return "HIP runtime error";
    case PIKO_ERR_NO_DEVICE: return "no HIP device";
    case PIKO_ERR_CAPACITY: return "queue capacity exceeded";
    case PIKO_ERR_OOM: return "device out of memory";
    case PIKO_ERR_BUSY: return "admission refused";
    default: return "unknown status";
  }
}

int piko_raster_destroy(piko_raster_pipe *p) {
  if (!p) return PIKO_ERR_INVALID;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  free_pairs(p);
  void *bufs[] = {p->d_geo, p->d_col, p->d_cnt, p->d_off, p->d_start, p->d_flags, p->d_keep, p->d_color, p->d_depth};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (p->owns_input) {
    if (p->d_pos) (void)hipFree((void *)p->d_pos);
    if (p->d_attr) (void)hipFree((void *)p->d_attr);
  }
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return PIKO_OK;
}

int piko_raster_allocate(piko_raster_pipe **out, const piko_constant_state *cs, const float *pos,
                         const void *attr, int count, const piko_raster_options *opt) {
  if (!out) return PIKO_ERR_INVALID;
  *out = nullptr;
  int rc = validate_opts(cs, count, opt);
  if (rc != PIKO_OK) return rc;
  if (count > 0 && (!pos || !attr)) return PIKO_ERR_INVALID;
  piko_raster_pipe *p = new (std::nothrow) piko_raster_pipe();
  if (!p) return PIKO_ERR_OOM;
  rc = allocate_common(out, cs, count, opt, p);
  if (rc == PIKO_OK) {
    const size_t pb = (size_t)count * 9 * sizeof(float);
    const size_t ab = (size_t)count * (opt->input_kind == PIKO_INPUT_WORLD ? 9 * sizeof(float) : 3 * sizeof(uint32_t));
    float *dp = nullptr;
    void *da = nullptr;
    p->owns_input = true;
    if (hipMalloc((void **)&dp, std::max<size_t>(pb, 4)) != hipSuccess) rc = PIKO_ERR_OOM;
    p->d_pos = dp;
    if (rc == PIKO_OK && hipMalloc(&da, std::max<size_t>(ab, 4)) != hipSuccess) rc = PIKO_ERR_OOM;
    p->d_attr = da;
    if (rc == PIKO_OK) {
      p->device_bytes += (long long)(pb + ab);
      if (count > 0 && (hipMemcpy(dp, pos, pb, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(da, attr, ab, hipMemcpyHostToDevice) != hipSuccess))
        rc = PIKO_ERR_HIP;
    }
  }
  if (rc != PIKO_OK) {
    *out = nullptr;
    piko_raster_destroy(p);
  }
  return rc;
}

int piko_raster_allocate_device(piko_raster_pipe **out, const piko_constant_state *cs, const float *d_pos,
                                const void *d_attr, int count, const piko_raster_options *opt) {
  if (!out) return PIKO_ERR_INVALID;
  *out = nullptr;
  int rc = validate_opts(cs, count, opt);
  if (rc != PIKO_OK) return rc;
  if (count > 0 && (!d_pos || !d_attr)) return PIKO_ERR_INVALID;
  piko_raster_pipe *p = new (std::nothrow) piko_raster_pipe();
  if (!p) return PIKO_ERR_OOM;
  p->d_pos = d_pos;
  p->d_attr = d_attr;
  rc = allocate_common(out, cs, count, opt, p);
  if (rc != PIKO_OK) {
    *out = nullptr;
    piko_raster_destroy(p);
  }
  return rc;
}

int piko_raster_prepare(piko_raster_pipe *p, const piko_constant_state *cs) {
  if (!p || !cs) return PIKO_ERR_INVALID;
  if (cs->screenSizeX != p->W || cs->screenSizeY != p->H) return PIKO_ERR_INVALID;
  set_frame_args(p, cs);
  return PIKO_OK;
}

int piko_raster_run_frames(piko_raster_pipe *p, int nframes) {
  if (!p || nframes < 0) return PIKO_ERR_INVALID;
  for (int attempt = 0; attempt < 3; ++attempt) {
    for (int f = 0; f < nframes; ++f) {
      int rc = enqueue_frame(p, nullptr, nullptr);
      if (rc != PIKO_OK) return rc;
    }
    int rerun = 0;
    int rc = check_capacity(p, &rerun);
    if (rc != PIKO_OK) return rc;
    if (!rerun) return PIKO_OK;
  }
  return PIKO_ERR_CAPACITY;
}

int piko_raster_run_single(piko_raster_pipe *p) { return piko_raster_run_frames(p, 1); }

int piko_raster_profile_frames(piko_raster_pipe *p, int nframes, float *kernel_ms, const char **names, int cap,
                               int *nk, float *frame_ms) {
  if (!p || nframes <= 0 || !kernel_ms || !nk) return PIKO_ERR_INVALID;
  int rc = piko_raster_run_frames(p, 1);  // settles queue capacity
  if (rc != PIKO_OK) return rc;
  constexpr int kMaxEv = 64;
  hipEvent_t ev0;
  hipEvent_t ev[kMaxEv];
  HIPOK(hipEventCreate(&ev0));
  for (int k = 0; k < kMaxEv; ++k) HIPOK(hipEventCreate(&ev[k]));
  double acc[kMaxEv] = {0};
  double facc = 0;
  int ne = 0;
  for (int f = 0; f < nframes; ++f) {
    HIPOK(hipEventRecord(ev0, p->stream));
    rc = enqueue_frame(p, ev, &ne);
    if (rc != PIKO_OK) return rc;
    HIPOK(hipEventSynchronize(ev[ne - 1]));
    float ms;
    HIPOK(hipEventElapsedTime(&ms, ev0, ev[0]));
    acc[0] += ms;
    for (int k = 1; k < ne; ++k) {
      HIPOK(hipEventElapsedTime(&ms, ev[k - 1], ev[k]));
      acc[k] += ms;
    }
    HIPOK(hipEventElapsedTime(&ms, ev0, ev[ne - 1]));
    facc += ms;
  }
  // map event slots to kernel names (the radix section repeats per pass)
  int n = std::min(ne, cap);
  for (int k = 0; k < n; ++k) {
    kernel_ms[k] = (float)(acc[k] / nframes);
    if (names) {
      if (k < 5) names[k] = kNames[k];
      else if (k < 5 + 5 * p->passes) names[k] = kNames[5 + ((k - 5) % 5)];
      else names[k] = k == ne - 1 ? "k_raster" : "k_bin_bounds";
    }
  }
  *nk = n;
  if (frame_ms) *frame_ms = (float)(facc / nframes);
  (void)hipEventDestroy(ev0);
  for (int k = 0; k < kMaxEv; ++k) (void)hipEventDestroy(ev[k]);
  int rerun = 0;
  rc = check_capacity(p, &rerun);
  return rc != PIKO_OK ? rc : (rerun ? PIKO_ERR_CAPACITY : PIKO_OK);
}

int piko_raster_get_data(piko_raster_pipe *p, uint32_t *color) {
  if (!p || !color) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(color, p->d_color, (size_t)p->W * p->H * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

int piko_raster_get_depth(piko_raster_pipe *p, float *depth) {
  if (!p || !depth) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(depth, p->d_depth, (size_t)p->W * p->H * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

long long piko_raster_get_bins(piko_raster_pipe *p, int64_t *bin_start, int32_t *bin_tris, long long cap) {
  if (!p || !bin_start) return PIKO_ERR_INVALID;
  uint32_t total = 0;
  HIPOK(hipMemcpyAsync(&total, p->d_off + p->count, 4, hipMemcpyDeviceToHost, p->stream));
  uint32_t *st = (uint32_t *)std::malloc(((size_t)p->nbins + 1) * 4);
  if (!st) return PIKO_ERR_OOM;
  if (hipMemcpyAsync(st, p->d_start, ((size_t)p->nbins + 1) * 4, hipMemcpyDeviceToHost, p->stream) !=
          hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess) {
    std::free(st);
    return PIKO_ERR_HIP;
  }
  for (int b = 0; b <= p->nbins; ++b) bin_start[b] = st[b];
  std::free(st);
  const long long n = std::min<long long>(std::min<long long>(total, p->cap), cap);
  if (bin_tris && n > 0) {
    HIPOK(hipMemcpyAsync(bin_tris, p->d_vals[p->final_buf], (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
    HIPOK(hipStreamSynchronize(p->stream));
  }
  return (long long)total;
}

int piko_raster_get_stats(piko_raster_pipe *p, piko_raster_stats *st) {
  if (!p || !st) return PIKO_ERR_INVALID;
  uint32_t total = 0;
  const int nkb = (int)std::max<long long>(cdiv(p->count, kThreads), 1);
  uint32_t *kb = (uint32_t *)std::malloc((size_t)nkb * 4);
  if (!kb) return PIKO_ERR_OOM;
  if (hipMemcpyAsync(&total, p->d_off + p->count, 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipMemcpyAsync(kb, p->d_keep, (size_t)nkb * 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess) {
    std::free(kb);
    return PIKO_ERR_HIP;
  }
  long long emitted = 0;
  for (int k = 0; k < nkb; ++k) emitted += kb[k];
  std::free(kb);
  st->triangles = p->count;
  st->emitted = p->count ? emitted : 0;
  st->pairs = total;
  st->pair_capacity = p->cap;
  st->bins_x = p->nBx;
  st->bins_y = p->nBy;
  st->device_bytes = p->device_bytes;
  return PIKO_OK;
}

}  // extern "C"
