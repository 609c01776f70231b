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
#include <mutex>
#include <new>
#include <vector>

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

// ---- frame state ----------------------------------------------------------

struct FrameState {
  uint32_t total;        // (triangle, bin) pairs of this frame (binned path)
  uint32_t overflow;     // sticky: bit 0 pairs exceeded the queues, bit 1 the large-triangle list,
                         // bit 2 the large triangles' 8x8 tiles exceeded 2^32 (cannot grow),
                         // bit 3 a tile queue of the queue path
  uint32_t large_peak;   // most large triangles of one frame since the host last looked
  uint32_t tiles_peak;   // most 8x8 tiles of large triangles of one frame since then
  unsigned long long large_acc;  // direct path, this frame: large triangles listed (low
                                 // half) and their 8x8 tiles (high half)
};

// One 1024-thread block scans n values in place (exclusive) and returns the
// total.  Element i lives at addr(i).  Each thread owns a contiguous run
// (sum, one block-wide scan of the sums, rescan), so the block synchronises
// once however large n is.
template <typename Addr>
__device__ uint32_t block_scan_inplace(uint32_t *__restrict__ v, int n, Addr addr, uint32_t *sh) {
  const int per = (n + 1023) / 1024;
  const int b = threadIdx.x * per, e = min(n, b + per);
  uint32_t s = 0;
  for (int i = b; i < e; ++i) s += v[addr(i)];
  uint32_t tot;
  uint32_t run = block_excl_scan<1024>(s, sh, tot);
  for (int i = b; i < e; ++i) {
    const size_t a = addr(i);
    const uint32_t x = v[a];
    v[a] = run;
    run += x;
  }
  return tot;
}

// Setup record of an emitted triangle, 48 bytes (3 x int4), written by the
// vertex stage and gathered by the raster and resolve kernels:
//   [0] x0 y0 x1 y1 (fixed point)
//   [1] x2 y2, 1/barydenom (raster.pikostage:174), z2
//   [2] z0 - z2, z1 - z2
// Colour needs the first 32 bytes only.
// colour_only: write the first 32 bytes (a triangle whose visibility is
// decided without gathering its depths again).
__device__ __forceinline__ void put_setup(int4 *__restrict__ geo, size_t i, const int p[6], float rcp,
                                          const float z[3], bool colour_only = false) {
  geo[3 * i + 0] = make_int4(p[0], p[1], p[2], p[3]);
  geo[3 * i + 1] = make_int4(p[4], p[5], __float_as_int(rcp), __float_as_int(z[2]));
  if (!colour_only) geo[3 * i + 2] = make_int4(__float_as_int(z[0] - z[2]), __float_as_int(z[1] - z[2]), 0, 0);
}

// ---- vertex stage + bin count ---------------------------------------------

constexpr int kVsThreads = 512;
constexpr int kVsPer = 4;                      // triangles per thread
constexpr int kVsTile = kVsThreads * kVsPer;   // triangles per block

// Job of virtual triangle i: the last job whose vbase <= i (jobs sorted by
// vbase).  `hint` is a job index known to be <= the answer.
__device__ __forceinline__ int job_of_tri(const FrameArgs *__restrict__ jobs, int njobs, int hint, int i) {
  int j = hint;
  while (j + 1 < njobs && jobs[j + 1].vbase <= (uint32_t)i) ++j;
  return j;
}
__device__ __forceinline__ int job_search(const FrameArgs *__restrict__ jobs, int njobs, uint32_t key,
                                          bool by_bin) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const uint32_t v = by_bin ? jobs[mid].bin_base : jobs[mid].vbase;
    if (v <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// VertexShaderStage::process (vertexShader.pikostage:36-109) and the number of
// bins RasterStage::assignBin gives each emitted triangle (raster.pikostage:
// 40-117).  Triangle i of tile b is lane (i % 512) of round (i / 512) % 4.
// MULTI: triangles belong to the jobs of a batch (per-job camera, screen,
// bins); otherwise `fa` describes the single job.
template <int KIND, bool MULTI>
__global__ void __launch_bounds__(kVsThreads) k_vs_count(const float *__restrict__ pos, const void *__restrict__ attr,
                                                          int n, FrameArgs fa, const FrameArgs *__restrict__ jobs,
                                                          int njobs, int4 *__restrict__ geo,
                                                          uint4 *__restrict__ colr, uint32_t *__restrict__ cnt,
                                                          uint32_t *__restrict__ blk_pairs,
                                                          uint32_t *__restrict__ blk_keep) {
  __shared__ uint32_t sh[kVsThreads / 64 + 1];
  __shared__ int s_job0;
  const int t = threadIdx.x;
  if (MULTI) {
    if (t == 0) s_job0 = job_search(jobs, njobs, (uint32_t)(blockIdx.x * kVsTile), false);
    __syncthreads();
  }
  uint32_t pairs = 0, keepc = 0;
#pragma unroll
  for (int k = 0; k < kVsPer; ++k) {
    const int i = blockIdx.x * kVsTile + k * kVsThreads + t;
    if (i >= n) continue;
    const FrameArgs &ja = MULTI ? jobs[job_of_tri(jobs, njobs, s_job0, i)] : fa;
    const int bsb = ja.S + ja.lgT;
    const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(i - (int)ja.vbase) : (size_t)i;
    int p[6];
    float z[3];
    uint32_t c[3];
    const float *v = pos + 9 * src;
    if (KIND == PIKO_INPUT_PRETRANSFORMED) {
      const uint32_t *cc = (const uint32_t *)attr + 3 * src;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        p[2 * q] = fixpt(v[3 * q], ja.S);
        p[2 * q + 1] = fixpt(v[3 * q + 1], ja.S);
        z[q] = v[3 * q + 2];
        c[q] = cc[q];
      }
    } else {
      const float *nn = (const float *)attr + 9 * src;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        transform_vertex(v + 3 * q, ja, p[2 * q], p[2 * q + 1], z[q]);
        c[q] = light_vertex(nn[3 * q], nn[3 * q + 1], nn[3 * q + 2], ja);
      }
    }
    uint32_t m = 0;
    if (vs_keep(p, ja.W, ja.H, ja.S)) {
      ++keepc;
      Edges e;
      e.set(p);
      const float rcp = 1.0f / (float)e.denom();  // raster.pikostage:174
      put_setup(geo, (size_t)i, p, rcp, z);
      colr[i] = make_uint4(c[0], c[1], c[2], 0u);
      int bx1, by1, bx2, by2;
      bin_rect(p, bsb, ja.nBx, ja.nBy, bx1, by1, bx2, by2);
      for (int by = by1; by <= by2; ++by)
        for (int bx = bx1; bx <= bx2; ++bx) m += tr_accept(e, bx, by, bsb) ? 1u : 0u;
    }
    cnt[i] = m;
    pairs += m;
  }
  uint32_t tp, tk;
  block_excl_scan<kVsThreads>(pairs, sh, tp);
  block_excl_scan<kVsThreads>(keepc, sh, tk);
  if (t == 0) {
    blk_pairs[blockIdx.x] = tp;
    blk_keep[blockIdx.x] = tk;
  }
}

// Exclusive scan of the tiles' pair counts; total and overflow flag.
__global__ void __launch_bounds__(1024) k_scan_tiles(uint32_t *__restrict__ blk_pairs, int ntiles,
                                                     FrameState *__restrict__ fs, long long cap) {
  __shared__ uint32_t sh[1024 / 64 + 1];
  const uint32_t tot = block_scan_inplace(blk_pairs, ntiles, [](int i) { return (size_t)i; }, sh);
  if (threadIdx.x == 0) {
    fs->total = tot;
    if ((long long)tot > cap) fs->overflow |= 1u;
  }
}

// (bin, triangle) pairs at their offsets: submission order across triangles,
// walk order (rows, then columns) within one.  Bins are numbered per batch
// (job bin_base + local bin).
template <bool MULTI>
__global__ void __launch_bounds__(kVsThreads) k_emit_pairs(const int4 *__restrict__ geo,
                                                            const uint32_t *__restrict__ cnt, int n, FrameArgs fa,
                                                            const FrameArgs *__restrict__ jobs, int njobs,
                                                            const uint32_t *__restrict__ blk_off, long long cap,
                                                            uint32_t *__restrict__ keys,
                                                            uint32_t *__restrict__ vals) {
  __shared__ uint32_t sh[kVsThreads / 64 + 1];
  __shared__ int s_job0;
  const int t = threadIdx.x;
  if (MULTI) {
    if (t == 0) s_job0 = job_search(jobs, njobs, (uint32_t)(blockIdx.x * kVsTile), false);
    __syncthreads();
  }
  uint32_t c[kVsPer], off[kVsPer];
#pragma unroll
  for (int k = 0; k < kVsPer; ++k) {
    const int i = blockIdx.x * kVsTile + k * kVsThreads + t;
    c[k] = i < n ? cnt[i] : 0u;
  }
  uint32_t run = blk_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kVsPer; ++k) {
    uint32_t tot;
    off[k] = run + block_excl_scan<kVsThreads>(c[k], sh, tot);
    run += tot;
  }
#pragma unroll
  for (int k = 0; k < kVsPer; ++k) {
    if (!c[k]) continue;
    const int i = blockIdx.x * kVsTile + k * kVsThreads + t;
    const FrameArgs &ja = MULTI ? jobs[job_of_tri(jobs, njobs, s_job0, i)] : fa;
    const int bsb = ja.S + ja.lgT;
    const uint32_t bbase = MULTI ? ja.bin_base : 0u;
    const int4 a = geo[3 * (size_t)i + 0], b = geo[3 * (size_t)i + 1];
    const int p[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
    Edges e;
    e.set(p);
    int bx1, by1, bx2, by2;
    bin_rect(p, bsb, ja.nBx, ja.nBy, bx1, by1, bx2, by2);
    uint32_t o = off[k];
    for (int by = by1; by <= by2; ++by)
      for (int bx = bx1; bx <= bx2; ++bx)
        if (tr_accept(e, bx, by, bsb)) {
          if ((long long)o < cap) {
            keys[o] = bbase + (uint32_t)(by * ja.nBx + bx);
            vals[o] = (uint32_t)i;
          }
          ++o;
        }
  }
}

// ---- stable LSD radix sort of the pairs by bin ----------------------------
//
// Per 8-bit digit: k_sort_count writes each tile's digit histogram into a
// digit-major matrix, k_sort_rows scans each digit row (one block per digit),
// k_sort_scatter ranks its tile stably (wave ballots + per-wave digit counts
// in LDS), stages it in LDS in digit order and writes it out coalesced.  The
// bins' queues end up in submission order; no global atomics touch a bin.

constexpr int kSortItems = 16;
constexpr int kSortTile = kThreads * kSortItems;  // keys per block

__global__ void __launch_bounds__(kThreads) k_sort_count(const uint32_t *__restrict__ keys,
                                                          const FrameState *__restrict__ fs, long long cap,
                                                          uint32_t *__restrict__ hist, int nblk, int shift) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const long long P = min((long long)fs->total, cap);
  const long long base = (long long)blockIdx.x * kSortTile;
  if (base < P) {
#pragma unroll 4
    for (int r = 0; r < kSortItems; ++r) {
      const long long idx = base + r * kThreads + t;
      if (idx < P) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
  }
  __syncthreads();
  hist[(size_t)t * nblk + blockIdx.x] = h[t];
}

// Exclusive scan of each digit's row of the matrix (tiles at or beyond the
// pair total are empty) and the row totals; block d handles digit d.
__global__ void __launch_bounds__(kThreads) k_sort_rows(uint32_t *__restrict__ hist, const FrameState *__restrict__ fs,
                                                         long long cap, int nblk, uint32_t *__restrict__ rowtot) {
  __shared__ uint32_t sh[kThreads / 64 + 1];
  const long long P = min((long long)fs->total, cap);
  const int used = (int)((P + kSortTile - 1) / kSortTile);
  uint32_t *row = hist + (size_t)blockIdx.x * nblk;
  uint32_t carry = 0;
  for (int base = 0; base < used; base += kThreads) {
    const int i = base + threadIdx.x;
    const uint32_t x = i < used ? row[i] : 0u;
    uint32_t tot;
    const uint32_t e = block_excl_scan<kThreads>(x, sh, tot);
    if (i < used) row[i] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0) rowtot[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kThreads) k_sort_scatter(const uint32_t *__restrict__ kin,
                                                            const uint32_t *__restrict__ vin,
                                                            uint32_t *__restrict__ kout,
                                                            uint32_t *__restrict__ vout,
                                                            const FrameState *__restrict__ fs, long long cap,
                                                            const uint32_t *__restrict__ hist,
                                                            const uint32_t *__restrict__ rowtot, int nblk, int shift) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[kThreads / 64][256];
  __shared__ uint32_t s_gb[256];
  __shared__ uint32_t s_loff[256];
  __shared__ uint32_t s_k[kSortTile], s_v[kSortTile];
  __shared__ uint32_t sh[kThreads / 64 + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long P = min((long long)fs->total, cap);
  const long long base = (long long)blockIdx.x * kSortTile;
  if (base >= P) return;
  run[t] = 0;
  {
    uint32_t tot;  // digit base = totals of the smaller digits + this tile's row prefix
    s_gb[t] = block_excl_scan<kThreads>(rowtot[t], sh, tot) + hist[(size_t)t * nblk + blockIdx.x];
  }
  const unsigned long long lt = (1ull << lane) - 1ull;
  uint32_t mk[kSortItems], mv[kSortItems], mr[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const long long idx = base + r * kThreads + t;
    const bool valid = idx < P;
    mk[r] = valid ? kin[idx] : 0u;
    mv[r] = valid ? vin[idx] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
#pragma unroll
    for (int q = 0; q < kThreads / 64; ++q) wcnt[q][t] = 0;
    __syncthreads();
    const bool valid = base + r * kThreads + t < P;
    const uint32_t d = valid ? ((mk[r] >> shift) & 255u) : 256u;
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
      for (int q = 0; q < w; ++q) o += wcnt[q][d];
      mr[r] = o;
    } else {
      mr[r] = 0xFFFFFFFFu;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int q = 0; q < kThreads / 64; ++q) add += wcnt[q][t];
    run[t] += add;
  }
  __syncthreads();
  uint32_t tot;
  s_loff[t] = block_excl_scan<kThreads>(run[t], sh, tot);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (mr[r] == 0xFFFFFFFFu) continue;
    const uint32_t lp = s_loff[(mk[r] >> shift) & 255u] + mr[r];
    s_k[lp] = mk[r];
    s_v[lp] = mv[r];
  }
  __syncthreads();
  const int ntile = (int)min((long long)kSortTile, P - base);
  for (int i = t; i < ntile; i += kThreads) {
    const uint32_t k = s_k[i];
    const uint32_t d = (k >> shift) & 255u;
    const uint32_t g = s_gb[d] + ((uint32_t)i - s_loff[d]);
    kout[g] = k;
    vout[g] = s_v[i];
  }
}

// Queue bounds from the sorted keys: bstart[b] / bend[b] for every non-empty
// bin (empty bins keep 0/0; the raster kernel clears each bin's pair after use).
__global__ void __launch_bounds__(kThreads) k_bounds(const uint32_t *__restrict__ keys,
                                                      const FrameState *__restrict__ fs, long long cap,
                                                      uint32_t *__restrict__ bstart, uint32_t *__restrict__ bend) {
  const long long P = min((long long)fs->total, cap);
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < P; i += (long long)gridDim.x * kThreads) {
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) bstart[k] = (uint32_t)i;
    if (i == P - 1 || keys[i + 1] != k) bend[k] = (uint32_t)(i + 1);
  }
}

// ---- tile raster ----------------------------------------------------------
//
// One wave per bin (bins laid out so that the blocks an XCD receives cover a
// contiguous band of bins).  The bin's queue is processed in batches of 64
// triangles staged in LDS.  Every (triangle, pixel) candidate -- the pixels
// of the triangle's rounded bbox inside the bin -- is spread evenly over the
// lanes (exclusive scan of per-triangle candidate counts + binary search);
// covered candidates min-reduce a visibility key (z bits, 0xFFFFFFFE - queue
// position) into LDS.  Queue position orders like submission order because
// the queues are stably sorted.  The winner of each pixel is then shaded from
// the staged batch (single-batch bins, the common case) or from global memory.

constexpr int kRT = 64;  // raster lanes per bin

// fromABGR's byte/255.0f (math.h:323-330) for every byte, folded at compile
// time: constant evaluation rounds the division exactly like IEEE runtime
// division, so the table holds the same floats the reference computes.
struct U8Table {
  float v[256];
  constexpr U8Table() : v() {
    for (int i = 0; i < 256; ++i) v[i] = (float)i / 255.0f;
  }
};
__constant__ U8Table kU8Table = U8Table();

struct RasterBatch {
  uint32_t coff[kRT + 1];
  uint32_t rect[kRT];  // x | y << 12 | w << 24
  int xy[6][kRT];
  float zr[4][kRT];    // z2, z0mz2, z1mz2, onebybary
  uint32_t col[3][kRT];
  uint32_t small[kRT];
};

// Edge functions at the centre of (px, py); int32 when the triangle's rounded
// bbox is below 2^(30-2S) pixels (then |e| < 2^31 everywhere inside it, so the
// values equal the int64 ones), else int64.
__device__ __forceinline__ bool covered_at(const int *xy, bool small, int px, int py, int S, float rcp, float &alpha,
                                           float &beta) {
  const int half = 1 << (S - 1);
  const int cx = (px << S) + half, cy = (py << S) + half;
  if (small) {
    const int e0 = (cy - xy[1]) * (xy[2] - xy[0]) - (cx - xy[0]) * (xy[3] - xy[1]);
    const int e1 = (cy - xy[3]) * (xy[4] - xy[2]) - (cx - xy[2]) * (xy[5] - xy[3]);
    const int e2 = (cy - xy[5]) * (xy[0] - xy[4]) - (cx - xy[4]) * (xy[1] - xy[5]);
    if ((e0 | e1 | e2) < 0) return false;
    alpha = (float)e1 * rcp;
    beta = (float)e2 * rcp;
    return true;
  }
  Edges e;
  e.set(xy);
  long long e0, e1, e2;
  pixel_edges(e, px, py, S, e0, e1, e2);
  if ((e0 | e1 | e2) < 0) return false;
  alpha = (float)e1 * rcp;  // raster.pikostage:269-270
  beta = (float)e2 * rcp;
  return true;
}

__device__ __forceinline__ bool small_bbox(const int *p, int S) {
  const int lox = (imin3(p[0], p[2], p[4]) & (int)0xffffff00) >> S;
  const int loy = (imin3(p[1], p[3], p[5]) & (int)0xffffff00) >> S;
  const int hix = ((imax3(p[0], p[2], p[4]) + 0xff) & (int)0xffffff00) >> S;
  const int hiy = ((imax3(p[1], p[3], p[5]) + 0xff) & (int)0xffffff00) >> S;
  return (long long)(hix - lox) * (long long)(hiy - loy) < (1ll << (30 - 2 * S));
}

__device__ __forceinline__ uint32_t gouraud(const float *u8, uint32_t c0, uint32_t c1, uint32_t c2, float alpha,
                                            float beta) {
  float rgb[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {  // raster.pikostage:183-185, 254-255, 291-293
    const float v0 = u8[(c0 >> (8 * ch)) & 255u];
    const float v1 = u8[(c1 >> (8 * ch)) & 255u];
    const float v2 = u8[(c2 >> (8 * ch)) & 255u];
    rgb[ch] = interp(v0 - v2, v1 - v2, v2, alpha, beta);
  }
  return to_abgr(rgb[0], rgb[1], rgb[2]);
}

// Setup record of one triangle as the raster kernel consumes it.
struct TriRec {
  int4 a, b, c;  // the setup record (put_setup)
  uint4 col;     // c0 c1 c2 -
};

__device__ __forceinline__ TriRec load_tri(const int4 *__restrict__ geo, const uint4 *__restrict__ colr,
                                           uint32_t tri) {
  TriRec r;
  r.a = geo[3 * (size_t)tri + 0];
  r.b = geo[3 * (size_t)tri + 1];
  r.c = geo[3 * (size_t)tri + 2];
  r.col = colr[tri];
  return r;
}

// Stages lane t's triangle (if `valid`) into the batch arrays and returns its
// candidate count: pixels of its rounded bbox (rasterMacros.h:58-63) inside
// [xlo,xhi) x [ylo,yhi).
__device__ __forceinline__ uint32_t stage_tri(RasterBatch &bt, int t, bool valid, const TriRec &r, int S, int xlo,
                                              int xhi, int ylo, int yhi) {
  if (!valid) return 0;
  const int p[6] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y};
  const int lox = (imin3(p[0], p[2], p[4]) & (int)0xffffff00) >> S;
  const int loy = (imin3(p[1], p[3], p[5]) & (int)0xffffff00) >> S;
  const int hix = ((imax3(p[0], p[2], p[4]) + 0xff) & (int)0xffffff00) >> S;
  const int hiy = ((imax3(p[1], p[3], p[5]) + 0xff) & (int)0xffffff00) >> S;
  const int x0 = max(lox, xlo), x1 = min(hix, xhi);
  const int y0 = max(loy, ylo), y1 = min(hiy, yhi);
  const int w = max(0, x1 - x0), h = max(0, y1 - y0);
  bt.rect[t] = (uint32_t)(x0 & 0xfff) | ((uint32_t)(y0 & 0xfff) << 12) | ((uint32_t)w << 24);
#pragma unroll
  for (int k = 0; k < 6; ++k) bt.xy[k][t] = p[k];
  bt.zr[0][t] = __int_as_float(r.b.w);
  bt.zr[1][t] = __int_as_float(r.c.x);
  bt.zr[2][t] = __int_as_float(r.c.y);
  bt.zr[3][t] = __int_as_float(r.b.z);
  bt.col[0][t] = r.col.x;
  bt.col[1][t] = r.col.y;
  bt.col[2][t] = r.col.z;
  bt.small[t] = (long long)(hix - lox) * (long long)(hiy - loy) < (1ll << (30 - 2 * S));
  return (uint32_t)(w * h);
}

// Candidates of the staged batch: covered ones min-reduce their key into vis.
template <int T>
__device__ __forceinline__ void raster_batch(RasterBatch &bt, unsigned long long *vis, uint32_t nc, int t, int S,
                                             int xlo, int ylo, uint32_t pos0) {
  const uint32_t inc = wave_incl_scan(nc);
  bt.coff[t] = inc - nc;
  const uint32_t tot = __shfl(inc, 63, 64);
  __syncthreads();
  for (uint32_t cidx = t; cidx < tot; cidx += kRT) {
    int lo = 0, hi = kRT - 1;  // owner: largest o with coff[o] <= cidx
#pragma unroll
    for (int step = 0; step < 6; ++step) {
      const int mid = (lo + hi + 1) >> 1;
      if (bt.coff[mid] <= cidx) lo = mid;
      else hi = mid - 1;
    }
    const int o = lo;
    const uint32_t k = cidx - bt.coff[o];
    const uint32_t rect = bt.rect[o];
    const int w = (int)(rect >> 24);
    const int q = (int)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)w));
    const int r = (int)k - q * w;
    const int px = (int)(rect & 0xfff) + r, py = (int)((rect >> 12) & 0xfff) + q;
    const int xy[6] = {bt.xy[0][o], bt.xy[1][o], bt.xy[2][o], bt.xy[3][o], bt.xy[4][o], bt.xy[5][o]};
    float alpha, beta;
    if (covered_at(xy, bt.small[o] != 0, px, py, S, bt.zr[3][o], alpha, beta)) {
      const float z = interp(bt.zr[1][o], bt.zr[2][o], bt.zr[0][o], alpha, beta);  // raster.pikostage:272
      atomicMin((long long *)&vis[(py - ylo) * T + (px - xlo)], (long long)vis_key(z, pos0 + (uint32_t)o));
    }
  }
  __syncthreads();
}

// Software pipeline per wave: while bin j is rasterised, the setup records of
// bin j+1, the queue entries of bin j+2 and the queue range of bin j+3 are
// already in flight.
template <int T, bool MULTI>
__global__ void __launch_bounds__(kRT) k_raster(const int4 *__restrict__ geo, const uint4 *__restrict__ colr,
                                                uint32_t *__restrict__ bstart, uint32_t *__restrict__ bend,
                                                const uint32_t *__restrict__ bin_tris, FrameArgs fa,
                                                const FrameArgs *__restrict__ jobs, int njobs, int bin_lo,
                                                int bin_hi, uint32_t *__restrict__ color, float *__restrict__ depth,
                                                int persistent) {
  __shared__ unsigned long long vis[T * T];
  __shared__ RasterBatch bt;
  __shared__ float s_u8[256];

  const int t = threadIdx.x;
  const int nb = bin_hi - bin_lo;
  // Bin order over [bin_lo, bin_hi).  One bin per block: block i -> XCD i % 8
  // gets a contiguous band of bins.  Persistent: wave w takes bins w,
  // w + grid, w + 2 grid, ... (interleaved across the screen, which evens out
  // empty and dense regions).
  int jn = 0;
  auto next_bin = [&]() -> int {
    const int j = jn++;
    if (!persistent) {
      if (j > 0) return -1;
      const int per = (nb + 7) / 8;
      const int b = (blockIdx.x % 8) * per + blockIdx.x / 8;
      return b < nb ? bin_lo + b : -1;
    }
    const int b = blockIdx.x + j * gridDim.x;
    return b < nb ? bin_lo + b : -1;
  };
  for (int q = t; q < 256; q += kRT) s_u8[q] = kU8Table.v[q];

  // pipeline prologue
  int bin0 = next_bin();
  int bin1 = bin0 >= 0 ? next_bin() : -1;
  int bin2 = bin1 >= 0 ? next_bin() : -1;
  uint32_t b0 = 0, e0 = 0, b1 = 0, e1 = 0, b2 = 0, e2 = 0;
  if (bin0 >= 0) { b0 = bstart[bin0]; e0 = bend[bin0]; }
  if (bin1 >= 0) { b1 = bstart[bin1]; e1 = bend[bin1]; }
  if (bin2 >= 0) { b2 = bstart[bin2]; e2 = bend[bin2]; }
  uint32_t tri0 = b0 + t < e0 ? bin_tris[b0 + t] : 0u;
  uint32_t tri1 = b1 + t < e1 ? bin_tris[b1 + t] : 0u;
  TriRec g0;
  if (b0 + t < e0) g0 = load_tri(geo, colr, tri0);

  while (bin0 >= 0) {
    // issue the next stages' loads
    const int bin3 = bin2 >= 0 ? next_bin() : -1;
    uint32_t b3 = 0, e3 = 0;
    if (bin3 >= 0) { b3 = bstart[bin3]; e3 = bend[bin3]; }
    const uint32_t tri2 = b2 + t < e2 ? bin_tris[b2 + t] : 0u;
    TriRec g1;
    if (b1 + t < e1) g1 = load_tri(geo, colr, tri1);

    // rasterise bin0
    const FrameArgs &ja = MULTI ? jobs[job_search(jobs, njobs, (uint32_t)bin0, true)] : fa;
    const int S = ja.S;
    const int lbin = MULTI ? bin0 - (int)ja.bin_base : bin0;
    const size_t fbb = MULTI ? (size_t)ja.fb_base : 0;
    const int bx = lbin % ja.nBx, by = lbin / ja.nBx;
    const int xlo = bx * T, ylo = by * T;
    const int xhi = min(xlo + T, ja.W), yhi = min(ylo + T, ja.H);
    if (t == 0) {  // leave the queue range cleared for the next frame
      bstart[bin0] = 0;
      bend[bin0] = 0;
    }
    for (int q = t; q < T * T; q += kRT) vis[q] = kVisSentinel;
    const uint32_t nq = e0 - b0;
    const uint32_t nc = stage_tri(bt, t, t < (int)nq, g0, S, xlo, xhi, ylo, yhi);
    raster_batch<T>(bt, vis, nc, t, S, xlo, ylo, 0u);
    for (uint32_t off = kRT; off < nq; off += kRT) {  // queues longer than one batch
      const bool v = off + t < nq;
      TriRec r;
      if (v) r = load_tri(geo, colr, bin_tris[b0 + off + t]);
      const uint32_t ncx = stage_tri(bt, t, v, r, S, xlo, xhi, ylo, yhi);
      raster_batch<T>(bt, vis, ncx, t, S, xlo, ylo, off);
    }
    // resolve: depth = winning z, colour = Gouraud of the winner, clear elsewhere
    for (int q = t; q < T * T; q += kRT) {
      const int px = xlo + (q % T), py = ylo + (q / T);
      if (px >= xhi || py >= yhi) continue;
      const unsigned long long key = vis[q];
      const size_t pid = fbb + (size_t)py * ja.W + px;
      uint32_t col = kClearColor;
      if (key != kVisSentinel) {
        const uint32_t pos = 0xFFFFFFFEu - (uint32_t)(key & 0xFFFFFFFFull);
        int xy[6];
        float rcp;
        uint32_t c0, c1, c2;
        if (nq <= kRT) {
#pragma unroll
          for (int k = 0; k < 6; ++k) xy[k] = bt.xy[k][pos];
          rcp = bt.zr[3][pos];
          c0 = bt.col[0][pos];
          c1 = bt.col[1][pos];
          c2 = bt.col[2][pos];
        } else {
          const TriRec r = load_tri(geo, colr, bin_tris[b0 + pos]);
          xy[0] = r.a.x; xy[1] = r.a.y; xy[2] = r.a.z; xy[3] = r.a.w; xy[4] = r.b.x; xy[5] = r.b.y;
          rcp = __int_as_float(r.b.z);
          c0 = r.col.x;
          c1 = r.col.y;
          c2 = r.col.z;
        }
        float alpha = 0.0f, beta = 0.0f;
        covered_at(xy, small_bbox(xy, S), px, py, S, rcp, alpha, beta);
        col = gouraud(s_u8, c0, c1, c2, alpha, beta);
      }
      depth[pid] = __int_as_float((int)(key >> 32));
      color[pid] = col;
    }
    __syncthreads();
    // rotate the pipeline
    bin0 = bin1; b0 = b1; e0 = e1; g0 = g1;
    bin1 = bin2; b1 = b2; e1 = e2; tri1 = tri2;
    bin2 = bin3; b2 = b3; e2 = e3;
  }
}

// Vertex positions of input triangle `src` -> fixed-point xy and z.
template <int KIND>
__device__ __forceinline__ void tri_positions(const float *__restrict__ pos, size_t src, const FrameArgs &ja,
                                              int p[6], float z[3]) {
  const float *v = pos + 9 * src;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (KIND == PIKO_INPUT_PRETRANSFORMED) {
      p[2 * q] = fixpt(v[3 * q], ja.S);
      p[2 * q + 1] = fixpt(v[3 * q + 1], ja.S);
      z[q] = v[3 * q + 2];
    } else {
      transform_vertex(v + 3 * q, ja, p[2 * q], p[2 * q + 1], z[q]);
    }
  }
}

// Packed vertex colours of input triangle `src` (lit normals for WORLD input).
template <int KIND>
__device__ __forceinline__ void tri_colours(const void *__restrict__ attr, size_t src, const FrameArgs &ja,
                                            uint32_t c[3]) {
  if (KIND == PIKO_INPUT_PRETRANSFORMED) {
    const uint32_t *cc = (const uint32_t *)attr + 3 * src;
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = cc[q];
  } else {
    const float *nn = (const float *)attr + 9 * src;
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = light_vertex(nn[3 * q], nn[3 * q + 1], nn[3 * q + 2], ja);
  }
}

// ---- direct path --------------------------------------------------------
//
// For triangles whose clipped rounded bbox (rasterMacros.h:48-49, 58-63) has
// at most kDirectMaxPx pixels and whose edge functions fit the int32 path
// (raster.pikostage:223-241) -- nearly all of them in dense meshes -- the
// vertex stage rasterises the triangle itself.  The bbox columns are cut into
// spans of kSpan pixels; the spans of a wave's 64 triangles are spread evenly
// over its lanes (neighbouring lanes on neighbouring columns, so one atomic
// instruction touches few cache lines), each lane steps the three edge
// functions down its span, and every covered pixel takes a 64-bit atomicMin
// of (z bits, 0xFFFFFFFE - triangle index) in a per-pixel visibility buffer.  Other triangles are
// listed and rasterised in 8x8 tiles by every wave of k_large_raster.  A
// per-pixel resolve shades the winner (the same arithmetic k_raster uses)
// and resets the buffer.  With the submission index in the key the result
// is the sequential one whatever the order of the atomics, so no bin queues
// are needed; the binned path remains for the queues themselves.

constexpr int kDirThreads = 256;     // one triangle per lane
constexpr int kDirectMaxPx = 16384;  // larger bboxes (never int32-path ones) take the tile path
constexpr int kSpan = 8;             // pixels per work unit (one bbox column segment)
constexpr int kTileSpan = 8;         // the same in k_tile_raster
constexpr int kLargeTile = 8;

// One staged triangle (64 bytes, read with 16-byte LDS loads).  Edge function
// k at the centre of pixel (x0 + i, y0 + j) is e[k] + i*dx[k] + j*dy[k], exact
// in 32-bit arithmetic for triangles of the int32 path.
// One wave's staged triangles, as structure of arrays (lanes reading
// neighbouring triangles hit neighbouring LDS banks).  Edge function k at the
// centre of pixel (x0 + i, y0 + j) is e[k] + i*dx[k] + j*dy[k], exact in
// 32-bit arithmetic for triangles of the int32 path.
template <bool MULTI>
struct DirStage {
  uint32_t pre[64];  // exclusive prefix of the triangles' span counts
  uint4 r0[64];      // x0 | y0 << 16, w | h << 16, submission index, 1/barydenom
  int4 re[64];       // e[0..2] at (x0, y0), z2 (float bits)
  int4 rx[64];       // dx[0..2], z0 - z2
  int4 ry[64];       // dy[0..2], z1 - z2
};
template <>
struct DirStage<true> {
  uint32_t pre[64];
  uint4 r0[64];
  int4 re[64];
  int4 rx[64];
  int4 ry[64];
  unsigned long long fb[64];  // framebuffer base of the triangle's job
  uint32_t fbw[64];           // framebuffer width of the triangle's job
};

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rounded pixel bbox (rasterMacros.h:48-49, 58-63) clipped to the screen.
__device__ __forceinline__ void clip_rect(const int *p, int S, int W, int H, int &x0, int &y0, int &x1, int &y1) {
  x0 = max((imin3(p[0], p[2], p[4]) & (int)0xffffff00) >> S, 0);
  y0 = max((imin3(p[1], p[3], p[5]) & (int)0xffffff00) >> S, 0);
  x1 = min(((imax3(p[0], p[2], p[4]) + 0xff) & (int)0xffffff00) >> S, W);
  y1 = min(((imax3(p[1], p[3], p[5]) + 0xff) & (int)0xffffff00) >> S, H);
}

// The jobs of one block's items (triangles or pixels, consecutive in the
// batch numbering): the job of the block's first item and the next ones,
// copied to LDS once, so each item finds its job with a short forward scan
// and reads the job's fields from LDS.  Jobs past the staged ones (a block
// spanning more than kStageJobs tiny jobs) are read from global memory.
constexpr int kStageJobs = 4;
struct JobStage {
  FrameArgs a[kStageJobs];
  int first, count;
};

// by_pixel: items are pixels (fb_base order), else triangles (vbase order).
__device__ __forceinline__ void stage_jobs(JobStage &js, const FrameArgs *__restrict__ jobs, int njobs,
                                           unsigned long long first_item, bool by_pixel) {
  if (threadIdx.x == 0) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      const unsigned long long v = by_pixel ? jobs[mid].fb_base : (unsigned long long)jobs[mid].vbase;
      if (v <= first_item) lo = mid;
      else hi = mid - 1;
    }
    js.first = lo;
    js.count = min(kStageJobs, njobs - lo);
  }
  __syncthreads();
  const int nw = (int)(sizeof(FrameArgs) / 4);
  for (int q = threadIdx.x; q < js.count * nw; q += blockDim.x)
    ((uint32_t *)js.a)[q] = ((const uint32_t *)(jobs + js.first))[q];
  __syncthreads();
}

// The job of `item` (at or after the block's first item).
__device__ __forceinline__ const FrameArgs &staged_job(const JobStage &js, const FrameArgs *__restrict__ jobs,
                                                       int njobs, unsigned long long item, bool by_pixel) {
  int j = js.first;
  for (;;) {
    if (j + 1 >= njobs) break;
    const int k = j + 1 - js.first;
    const FrameArgs &nx = k < js.count ? js.a[k] : jobs[j + 1];
    if ((by_pixel ? nx.fb_base : (unsigned long long)nx.vbase) > item) break;
    ++j;
  }
  return j - js.first < js.count ? js.a[j - js.first] : jobs[j];
}

__device__ __forceinline__ int job_by_pixel(const FrameArgs *__restrict__ jobs, int njobs, unsigned long long g) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].fb_base <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int KIND, bool MULTI>
__global__ void __launch_bounds__(kDirThreads) k_vs_direct(const float *__restrict__ pos, int n, int S, FrameArgs fa,
                                                            const FrameArgs *__restrict__ jobs, int njobs,
                                                            int4 *__restrict__ geo,
                                                            unsigned long long *__restrict__ vis,
                                                            uint32_t *__restrict__ large,
                                                            uint32_t *__restrict__ tile_off, FrameState *__restrict__ fs,
                                                            uint32_t large_cap, int inline_large) {
  // every wave works on its own 64 triangles: no block barrier
  __shared__ DirStage<MULTI> st[kDirThreads / 64];
  __shared__ JobStage js;
  const int t = threadIdx.x, lane = t & 63;
  DirStage<MULTI> &s = st[t >> 6];
  const int i = blockIdx.x * kDirThreads + t;
  if (MULTI) stage_jobs(js, jobs, njobs, (unsigned long long)blockIdx.x * kDirThreads, false);
  uint32_t units = 0, tiles = 0;
  bool big = false;  // listed large triangle; drawn inline from its setup record
  if (i < n) {
    // vertex positions only: the winner's colours are computed in k_resolve
    const FrameArgs &ja = MULTI ? staged_job(js, jobs, njobs, (unsigned long long)i, false) : fa;
    const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(i - (int)ja.vbase) : (size_t)i;
    int p[6];
    float z[3];
    tri_positions<KIND>(pos, src, ja, p, z);
    if (vs_keep(p, ja.W, ja.H, ja.S)) {
      Edges e;
      e.set(p);
      const float rcp = 1.0f / (float)e.denom();  // raster.pikostage:174
      int x0, y0, x1, y1;
      clip_rect(p, ja.S, ja.W, ja.H, x0, y0, x1, y1);
      const int w = max(0, x1 - x0), h = max(0, y1 - y0);
      const uint32_t npx = (uint32_t)w * (uint32_t)h;
      const bool listed = npx > (uint32_t)kDirectMaxPx || !small_bbox(p, ja.S);
      // the large-triangle raster gathers the depths again; spans and the
      // resolve need only the colour part
      put_setup(geo, (size_t)i, p, rcp, z, !listed);
      if (listed) {
        big = npx > 0;
        tiles = (uint32_t)((w + kLargeTile - 1) / kLargeTile) * (uint32_t)((h + kLargeTile - 1) / kLargeTile);
      } else if (npx > 0) {
        units = (uint32_t)w * ((uint32_t)(h + kSpan - 1) / kSpan);
        // edge functions at the centre of pixel (x0, y0) and their steps;
        // unsigned arithmetic, equal to the int32 evaluation for int32-path
        // triangles (no intermediate leaves the int32 range)
        const uint32_t half = 1u << (S - 1);
        const uint32_t cx = ((uint32_t)x0 << S) + half, cy = ((uint32_t)y0 << S) + half;
        uint32_t ee[3], dx[3], dy[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const uint32_t xa = (uint32_t)p[2 * k], ya = (uint32_t)p[2 * k + 1];
          const uint32_t xb = (uint32_t)p[(2 * k + 2) % 6], yb = (uint32_t)p[(2 * k + 3) % 6];
          ee[k] = (cy - ya) * (xb - xa) - (cx - xa) * (yb - ya);
          dx[k] = (0u - (yb - ya)) << S;
          dy[k] = (xb - xa) << S;
        }
        s.r0[lane] = make_uint4((uint32_t)x0 | ((uint32_t)y0 << 16), (uint32_t)w | ((uint32_t)h << 16), (uint32_t)i,
                                __float_as_uint(rcp));
        s.re[lane] = make_int4((int)ee[0], (int)ee[1], (int)ee[2], __float_as_int(z[2]));
        s.rx[lane] = make_int4((int)dx[0], (int)dx[1], (int)dx[2], __float_as_int(z[0] - z[2]));
        s.ry[lane] = make_int4((int)dy[0], (int)dy[1], (int)dy[2], __float_as_int(z[1] - z[2]));
        if constexpr (MULTI) {
          s.fb[lane] = ja.fb_base;
          s.fbw[lane] = (uint32_t)ja.W;
        }
      }
    }
  }
  // large triangles: one append per wave.  A single 64-bit atomic reserves
  // list slots (low half) and 8x8 tile numbers (high half) together, so the
  // first tile of the listed triangles increases with the slot.  With
  // inline_large (the last frame had none) the counts are still added, for
  // the host's next launch decision, and the wave draws them itself.
  const unsigned long long bm = __ballot(big);
  if (bm) {
    const int first = __ffsll((long long)bm) - 1;
    const uint32_t tl = big ? tiles : 0u;
    const uint32_t tinc = wave_incl_scan(tl);
    const uint32_t tsum = __shfl(tinc, 63, 64);
    uint32_t bt = 0, btile = 0;
    if (lane == first) {
      const unsigned long long old =
          atomicAdd(&fs->large_acc, ((unsigned long long)tsum << 32) | (unsigned long long)__popcll(bm));
      bt = (uint32_t)old;
      btile = (uint32_t)(old >> 32);
      if (!inline_large && (unsigned long long)btile + tsum > 0xFFFFFFFFull) atomicOr(&fs->overflow, 4u);
    }
    if (inline_large) {
      // the wave's own setup records of its listed triangles (full 48-byte
      // records, written above): their stores before these loads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      for (unsigned long long m = bm; m; m &= m - 1) {
        const int L = __ffsll((long long)m) - 1;
        const uint32_t tri = (uint32_t)(blockIdx.x * kDirThreads + (t & ~63) + L);
        const FrameArgs &jl = MULTI ? staged_job(js, jobs, njobs, (unsigned long long)tri, false) : fa;
        const int4 ga = geo[3 * (size_t)tri], gb = geo[3 * (size_t)tri + 1], gc = geo[3 * (size_t)tri + 2];
        const int xy[6] = {ga.x, ga.y, ga.z, ga.w, gb.x, gb.y};
        const float rcp = __int_as_float(gb.z), z2 = __int_as_float(gb.w);
        const float z0mz2 = __int_as_float(gc.x), z1mz2 = __int_as_float(gc.y);
        const int W = jl.W;
        const unsigned long long fbb = MULTI ? jl.fb_base : 0ull;
        int x0, y0, x1, y1;
        clip_rect(xy, jl.S, jl.W, jl.H, x0, y0, x1, y1);
        const int w = x1 - x0, h = y1 - y0;
        const bool sm = small_bbox(xy, S);
        const int tw = (w + kLargeTile - 1) / kLargeTile, th = (h + kLargeTile - 1) / kLargeTile;
        for (int g = 0; g < tw * th; ++g) {
          const int ty = g / tw, tx = g - ty * tw;
          const int px = x0 + tx * kLargeTile + (lane & (kLargeTile - 1));
          const int py = y0 + ty * kLargeTile + lane / kLargeTile;
          float alpha, beta;
          if (px < x0 + w && py < y0 + h && covered_at(xy, sm, px, py, S, rcp, alpha, beta)) {
            const float z = interp(z0mz2, z1mz2, z2, alpha, beta);
            atomicMin((long long *)&vis[fbb + (size_t)py * W + px], (long long)vis_key(z, tri));
          }
        }
      }
    } else {
      bt = __shfl(bt, first, 64);
      btile = __shfl(btile, first, 64);
      if (big) {
        const uint32_t at = bt + (uint32_t)__popcll(bm & ((1ull << lane) - 1ull));
        if (at < large_cap) {
          large[at] = (uint32_t)i;
          tile_off[at] = btile + tinc - tl;
        } else {
          atomicOr(&fs->overflow, 2u);
        }
      }
    }
  }
  const uint32_t inc = wave_incl_scan(units);
  s.pre[lane] = inc - units;
  const uint32_t tot = __shfl(inc, 63, 64);
  wave_fence();
  for (uint32_t cidx = lane; cidx < tot; cidx += 64) {
    int o = 0;  // owner: largest o with pre[o] <= cidx (o + step stays <= 63)
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
      if (s.pre[o + step] <= cidx) o += step;
    const uint4 r0 = s.r0[o];
    const int4 re = s.re[o];
    const int4 rx = s.rx[o];
    const int4 ry = s.ry[o];
    const uint32_t k = cidx - s.pre[o];
    const uint32_t w = r0.y & 0xffffu, h = r0.y >> 16;
    // band and column of unit k: (k + 0.5) / w is at least 0.5 / w from an
    // integer, and the float product is within 3 * 2^-24 * (k + 0.5) / w of
    // it (k + 0.5 is exact, rcpf and the product round once each), which is
    // below 0.5 / w for every k < 2^21: the quotient truncates exactly
    const uint32_t q = (uint32_t)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)w));
    const uint32_t sx = k - q * w, sy = q * kSpan;
    const uint32_t px = (r0.x & 0xffffu) + sx, py = (r0.x >> 16) + sy;
    const uint32_t nv = min((uint32_t)kSpan, h - sy);
    uint32_t e0 = (uint32_t)re.x + sx * (uint32_t)rx.x + sy * (uint32_t)ry.x;
    uint32_t e1 = (uint32_t)re.y + sx * (uint32_t)rx.y + sy * (uint32_t)ry.y;
    uint32_t e2 = (uint32_t)re.z + sx * (uint32_t)rx.z + sy * (uint32_t)ry.z;
    const float rcp = __uint_as_float(r0.w);
    const float z2 = __int_as_float(re.w), z0mz2 = __int_as_float(rx.w), z1mz2 = __int_as_float(ry.w);
    const uint32_t low = 0xFFFFFFFEu - r0.z;
    unsigned long long *col;
    uint32_t pitch;
    if constexpr (MULTI) {
      pitch = s.fbw[o];
      col = vis + s.fb[o] + (size_t)py * pitch + px;
    } else {
      pitch = (uint32_t)fa.W;
      col = vis + (size_t)py * pitch + px;
    }
#pragma unroll
    for (uint32_t j = 0; j < (uint32_t)kSpan; ++j) {
      if (j < nv && (int)(e0 | e1 | e2) >= 0) {
        const float alpha = (float)(int)e1 * rcp, beta = (float)(int)e2 * rcp;  // raster.pikostage:269-270
        const float z = interp(z0mz2, z1mz2, z2, alpha, beta);                    // raster.pikostage:272
        atomicMin((long long *)(col + (size_t)j * pitch),
                  (long long)(((unsigned long long)__float_as_uint(z) << 32) | (unsigned long long)low));
      }
      e0 += (uint32_t)ry.x;
      e1 += (uint32_t)ry.y;
      e2 += (uint32_t)ry.z;
    }
  }
}

// One wave per 8x8 tile of a large triangle.  Each wave of the grid takes a
// contiguous run of the frame's tiles and walks it in order, so it looks its
// triangle up (and, in a batch, the triangle's job) once per run and again
// only where the run enters the next triangle's tiles.  Tiles handed out one
// at a time over the grid repeated both searches for every tile.
template <bool MULTI>
__global__ void __launch_bounds__(256) k_large_raster(const int4 *__restrict__ geo,
                                                      const uint32_t *__restrict__ large,
                                                      const uint32_t *__restrict__ tile_off,
                                                      const FrameState *__restrict__ fs, uint32_t large_cap,
                                                      FrameArgs fa, const FrameArgs *__restrict__ jobs, int njobs,
                                                      unsigned long long *__restrict__ vis) {
  const unsigned long long acc = fs->large_acc;
  const uint32_t nl32 = (uint32_t)acc;
  if (nl32 > large_cap || (fs->overflow & 4u)) return;  // the list overflowed: the frame is re-run
  const int nl = (int)nl32;
  const uint32_t total = (uint32_t)(acc >> 32);
  const int lane = threadIdx.x & 63;
  const unsigned long long nw = (unsigned long long)gridDim.x * (blockDim.x / 64);
  const unsigned long long w = (unsigned long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const unsigned long long per = ((unsigned long long)total + nw - 1) / nw;
  const unsigned long long gbeg = w * per, gend = min(gbeg + per, (unsigned long long)total);
  if (gbeg >= gend) return;
  int k = 0, hi = nl - 1;  // largest k with tile_off[k] <= gbeg
  while (k < hi) {
    const int mid = (k + hi + 1) >> 1;
    if (tile_off[mid] <= gbeg) k = mid;
    else hi = mid - 1;
  }
  uint32_t kbeg = 0, kend = 0, tri = 0, tw = 0;
  int p[6], x0 = 0, y0 = 0, x1 = 0, y1 = 0, S = 0, W = 0;
  bool sm = false;
  float rcp = 0.0f, z0mz2 = 0.0f, z1mz2 = 0.0f, z2 = 0.0f;
  unsigned long long fbb = 0;
  for (unsigned long long g = gbeg; g < gend; ++g) {
    if (g == gbeg || g >= kend) {
      // the triangle whose tiles hold g: the ranges [tile_off[k], tile_off[k + 1])
      // are consecutive and end at total, so this stops at k < nl with g inside
      while (k + 1 < nl && tile_off[k + 1] <= g) ++k;
      kbeg = tile_off[k];
      kend = k + 1 < nl ? tile_off[k + 1] : total;
      tri = large[k];
      const FrameArgs *ja = &fa;
      if (MULTI) ja = &jobs[job_search(jobs, njobs, tri, false)];
      S = ja->S;
      W = ja->W;
      fbb = MULTI ? ja->fb_base : 0ull;
      const int4 a = geo[3 * (size_t)tri], b = geo[3 * (size_t)tri + 1], c = geo[3 * (size_t)tri + 2];
      p[0] = a.x; p[1] = a.y; p[2] = a.z; p[3] = a.w; p[4] = b.x; p[5] = b.y;
      clip_rect(p, S, W, ja->H, x0, y0, x1, y1);
      tw = (uint32_t)(x1 - x0 + kLargeTile - 1) / kLargeTile;
      sm = small_bbox(p, S);
      rcp = __int_as_float(b.z);
      z2 = __int_as_float(b.w);
      z0mz2 = __int_as_float(c.x);
      z1mz2 = __int_as_float(c.y);
    }
    const uint32_t local = (uint32_t)(g - kbeg);
    const int ty = (int)(local / tw), tx = (int)(local - (uint32_t)ty * tw);
    const int px = x0 + tx * kLargeTile + (lane & (kLargeTile - 1)), py = y0 + ty * kLargeTile + lane / kLargeTile;
    if (px >= x1 || py >= y1) continue;
    float alpha, beta;
    if (covered_at(p, sm, px, py, S, rcp, alpha, beta)) {
      const float z = interp(z0mz2, z1mz2, z2, alpha, beta);
      atomicMin((long long *)&vis[fbb + (size_t)py * W + px], (long long)vis_key(z, tri));
    }
  }
}

// Per pixel: depth = the winner's z, colour = Gouraud of the winner (the clear
// colour where nothing is drawn); the visibility word is reset for the next
// frame.
template <int KIND, bool MULTI>
__global__ void __launch_bounds__(256) k_resolve(const int4 *__restrict__ geo, const void *__restrict__ attr,
                                                 unsigned long long *__restrict__ vis, FrameArgs fa,
                                                 const FrameArgs *__restrict__ jobs, int njobs, long long npix,
                                                 uint32_t *__restrict__ color, float *__restrict__ depth,
                                                 FrameState *__restrict__ fs) {
  __shared__ float s_u8[256];
  __shared__ JobStage js;
  const int t = threadIdx.x;
  // block i runs on XCD i % 8: each XCD resolves a contiguous run of pixel
  // blocks (whole bands of rows), so the records of a triangle that covers
  // pixels in neighbouring rows are gathered into one XCD's L2
  const unsigned nblk = (unsigned)((npix + 255) / 256);
  const unsigned per = (nblk + 7) / 8;
  const unsigned b = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (b >= nblk) return;
  s_u8[t] = kU8Table.v[t];
  if (MULTI) stage_jobs(js, jobs, njobs, (unsigned long long)b * blockDim.x, true);
  __syncthreads();
  if (b == 0 && t == 0) {  // the large list of this frame is consumed
    fs->large_peak = max(fs->large_peak, (uint32_t)fs->large_acc);
    fs->tiles_peak = max(fs->tiles_peak, (uint32_t)(fs->large_acc >> 32));
    fs->large_acc = 0;
  }
  const long long g = (long long)b * blockDim.x + t;
  if (g >= npix) return;
  const FrameArgs &ja = MULTI ? staged_job(js, jobs, njobs, (unsigned long long)g, true) : fa;
  const unsigned long long key = vis[g];
  uint32_t col = kClearColor;
  if (key != kVisSentinel) {
    vis[g] = kVisSentinel;
    const long long local = MULTI ? g - (long long)ja.fb_base : g;
    const int py = (int)(local / ja.W), px = (int)(local - (long long)py * ja.W);
    const uint32_t tri = 0xFFFFFFFEu - (uint32_t)(key & 0xFFFFFFFFull);
    const int4 a = geo[3 * (size_t)tri], b = geo[3 * (size_t)tri + 1];
    const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(tri - ja.vbase) : (size_t)tri;
    uint32_t cc[3];
    tri_colours<KIND>(attr, src, ja, cc);
    const int xy[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
    float alpha = 0.0f, beta = 0.0f;
    covered_at(xy, small_bbox(xy, ja.S), px, py, ja.S, __int_as_float(b.z), alpha, beta);
    col = gouraud(s_u8, cc[0], cc[1], cc[2], alpha, beta);
  }
  depth[g] = __int_as_float((int)(key >> 32));
  color[g] = col;
}

// ---- tiled path ---------------------------------------------------------
//
// Bins are screen tiles of T x T pixels; every fragment is resolved in the
// tile's LDS, and no global atomic is used.
//
//   k_tile_bin     one block per chunk of up to kChunk consecutive triangles of
//                  one job.  Vertex stage on the positions only (vertexShader.
//                  pikostage:36-109), then RasterStage::assignBin's trivial-
//                  reject test per tile (raster.pikostage:40-117) for triangles
//                  whose clipped bbox spans at most 2x2 tiles: each accepted
//                  (tile, triangle) pair takes a slot from an LDS counter of
//                  its tile.  The counters are scanned, the scan is written as
//                  the chunk's column of the offset table, and the triangle
//                  indices are scattered into the chunk's own pair region,
//                  grouped by tile.  Larger triangles (or ones whose edge
//                  functions need 64 bits) are listed at the region's end.
//   k_tile_raster  one block per tile.  It reads where its pairs lie in each
//                  chunk of its job from the table, recomputes each listed
//                  triangle's setup from its positions, rasterises it into
//                  the tile's LDS visibility words (z bits, 0xFFFFFFFE -
//                  triangle), does the same for the chunks' large triangles
//                  clipped to the tile, then shades each pixel's winner from
//                  the winner's vertices and writes colour and depth once.
//
// A chunk region holds 4 slots per triangle and a listed triangle takes at
// most 4, so no queue can overflow.  With the submission index in the key the
// result is the sequential one whatever the order of pairs within a tile.

constexpr int kChunk = 4096;                      // triangles per chunk
constexpr int kBinThreads = 1024;
constexpr int kChunkPer = kChunk / kBinThreads;   // triangles per thread
constexpr uint32_t kChunkCap = 4u * kChunk;       // pair slots per chunk region
constexpr int kMaxTilesJob = 8192;                // tiles of one job (LDS counters)
constexpr int kTRThreads = 256;                   // k_tile_raster block
constexpr int kTRWinPer = 2;                      // chunks each k_tile_raster thread reads per window
constexpr int kTRWin = kTRThreads * kTRWinPer;    // chunks read per window
constexpr int kExactPx = 16;                      // k_tile_bin: exact tile mask up to this many bbox pixels

struct ChunkDesc {
  uint32_t job;      // index in the job table
  uint32_t tri0;     // first triangle (batch numbering)
  uint32_t count;    // triangles in the chunk
  uint32_t tab_off;  // first entry of the chunk's column of the offset table
};

template <int KIND, bool MULTI>
__global__ void __launch_bounds__(kBinThreads) k_tile_bin(const float *__restrict__ pos, FrameArgs fa,
                                                           const FrameArgs *__restrict__ jobs,
                                                           const ChunkDesc *__restrict__ chunks,
                                                           uint32_t *__restrict__ tab, uint32_t *__restrict__ pairs,
                                                           uint32_t *__restrict__ chunk_keep,
                                                           int4 *__restrict__ geo) {
  __shared__ uint32_t cnt[kMaxTilesJob + 2];
  __shared__ uint32_t sh[kBinThreads / 64 + 1];
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  const ChunkDesc cd = chunks[c];
  const FrameArgs &ja = MULTI ? jobs[cd.job] : fa;
  const int ntiles = ja.nBx * ja.nBy;
  for (int q = t; q < ntiles + 2; q += kBinThreads) cnt[q] = 0;
  __syncthreads();
  const int S = ja.S, lgT = ja.lgT, bsb = S + lgT;
  // all of the thread's vertex positions first, so their loads overlap
  float vp[kChunkPer][9];
#pragma unroll
  for (int k = 0; k < kChunkPer; ++k) {
    const uint32_t il = (uint32_t)(k * kBinThreads + t);
    const uint32_t v = cd.tri0 + min(il, cd.count - 1u);
    const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(v - ja.vbase) : (size_t)v;
#pragma unroll
    for (int q = 0; q < 9; ++q) vp[k][q] = il < cd.count ? pos[9 * src + q] : 0.0f;
  }
  // per triangle: ra = kind (bit 0 listed per tile, bit 1 large) | tiles it
  // is listed in << 2 | tx0 << 8 | ty0 << 20; rb/rc = slots (16 bits each)
  uint32_t ra[kChunkPer], rb[kChunkPer], rc[kChunkPer];
  uint32_t keepc = 0;
#pragma unroll
  for (int k = 0; k < kChunkPer; ++k) {
    ra[k] = 0;
    rb[k] = 0;
    rc[k] = 0;
    const uint32_t il = (uint32_t)(k * kBinThreads + t);
    if (il >= cd.count) continue;
    int p[6];
    float z[3];
    tri_positions<KIND>(vp[k], 0, ja, p, z);
    if (!vs_keep(p, ja.W, ja.H, S)) continue;
    ++keepc;
    int x0, y0, x1, y1;
    clip_rect(p, S, ja.W, ja.H, x0, y0, x1, y1);
    if (x1 <= x0 || y1 <= y0) continue;
    const int tx0 = x0 >> lgT, ty0 = y0 >> lgT, tx1 = (x1 - 1) >> lgT, ty1 = (y1 - 1) >> lgT;
    if (tx1 - tx0 <= 1 && ty1 - ty0 <= 1 && small_bbox(p, S)) {
      uint32_t acc = 0;  // tiles that may hold a covered pixel
      if ((x1 - x0) * (y1 - y0) <= kExactPx) {
        // few pixels: the tiles that hold a covered pixel centre, exactly
        // (raster.pikostage:223-241); a triangle covering none is dropped
        const uint32_t half = 1u << (S - 1);
        for (int py = y0; py < y1; ++py)
          for (int px = x0; px < x1; ++px) {
            const uint32_t cx = ((uint32_t)px << S) + half, cy = ((uint32_t)py << S) + half;
            bool in = true;
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
              const uint32_t xa = (uint32_t)p[2 * kk], ya = (uint32_t)p[2 * kk + 1];
              const uint32_t xb = (uint32_t)p[(2 * kk + 2) % 6], yb = (uint32_t)p[(2 * kk + 3) % 6];
              in &= (int)((cy - ya) * (xb - xa) - (cx - xa) * (yb - ya)) >= 0;
            }
            if (in) acc |= 1u << (((px >> lgT) - tx0) | (((py >> lgT) - ty0) << 1));
          }
      } else {
        // RasterStage::assignBin's trivial-reject test (raster.pikostage:77-114)
        Edges e;
        e.set(p);
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int bx = tx0 + (d & 1), by = ty0 + (d >> 1);
          if (bx <= tx1 && by <= ty1 && tr_accept(e, bx, by, bsb)) acc |= 1u << d;
        }
      }
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        if (!((acc >> d) & 1u)) continue;
        const uint32_t s = atomicAdd(&cnt[(ty0 + (d >> 1)) * ja.nBx + tx0 + (d & 1)], 1u);
        if (d < 2) rb[k] |= s << (16 * d);
        else rc[k] |= s << (16 * (d - 2));
      }
      ra[k] = 1u | (acc << 2) | ((uint32_t)tx0 << 8) | ((uint32_t)ty0 << 20);
      if (!acc) continue;
    } else {
      rb[k] = atomicAdd(&cnt[ntiles + 1], 1u);
      ra[k] = 2u;
    }
    // the setup record k_tile_raster reads for each listed triangle
    Edges e;
    e.set(p);
    put_setup(geo, (size_t)(cd.tri0 + il), p, 1.0f / (float)e.denom(), z);  // raster.pikostage:174
  }
  __syncthreads();
  {  // exclusive scan of cnt[0..ntiles]; cnt[ntiles] becomes the chunk's pair total
    const int m = ntiles + 1;
    const int per = (m + kBinThreads - 1) / kBinThreads;
    const int b = t * per, e = min(m, b + per);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    uint32_t tot;
    uint32_t run = block_excl_scan<kBinThreads>(s, sh, tot);
    for (int i = b; i < e; ++i) {
      const uint32_t x = cnt[i];
      cnt[i] = run;
      run += x;
    }
  }
  __syncthreads();
  for (int q = t; q < ntiles + 2; q += kBinThreads) tab[cd.tab_off + q] = cnt[q];
  uint32_t *reg = pairs + (size_t)c * kChunkCap;
#pragma unroll
  for (int k = 0; k < kChunkPer; ++k) {
    if (!ra[k]) continue;
    const uint32_t v = cd.tri0 + (uint32_t)(k * kBinThreads + t);
    if (ra[k] & 1u) {
      const int tx0 = (int)((ra[k] >> 8) & 0xfffu), ty0 = (int)(ra[k] >> 20);
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        if (!((ra[k] >> (2 + d)) & 1u)) continue;
        const int tile = (ty0 + (d >> 1)) * ja.nBx + tx0 + (d & 1);
        const uint32_t s = ((d < 2 ? rb[k] : rc[k]) >> (16 * (d & 1))) & 0xffffu;
        reg[cnt[tile] + s] = v;
      }
    } else {
      reg[kChunkCap - 1u - rb[k]] = v;
    }
  }
  uint32_t kt;
  block_excl_scan<kBinThreads>(keepc, sh, kt);
  if (t == 0) chunk_keep[c] = kt;
}

// One wave's staged triangles in k_tile_raster (structure of arrays so that
// lanes reading neighbouring triangles hit neighbouring banks).
struct TileStage {
  uint32_t pre[64];  // exclusive prefix of the triangles' span counts
  int4 xy01[64];     // x0 y0 x1 y1 (fixed point)
  int4 xy2z[64];     // x2 y2, z2 and z0 - z2 (float bits)
  uint4 misc[64];    // z1 - z2 and 1/barydenom (float bits), rect x | y << 16,
                     // rect w | h << 16 | (int32 edges) << 31
  uint32_t tri[64];
};

// Rasterises staged triangle o's span k (a column segment of kTileSpan pixels of
// its bbox clipped to the tile) into the tile's visibility words.
template <int T>
__device__ __forceinline__ void tile_span(const TileStage &s, int o, uint32_t k, int S, int xlo, int ylo,
                                          unsigned long long *vis) {
  const int4 a = s.xy01[o];
  const int4 b = s.xy2z[o];
  const uint4 m = s.misc[o];
  const uint32_t w = m.w & 0xffffu, h = (m.w >> 16) & 0x7fffu;
  // band and column of span k (the quotient truncates exactly: see k_vs_direct)
  const uint32_t q = (uint32_t)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)w));
  const uint32_t sx = k - q * w, sy = q * kTileSpan;
  const int px = (int)(m.z & 0xffffu) + (int)sx, py0 = (int)(m.z >> 16) + (int)sy;
  const uint32_t nv = min((uint32_t)kTileSpan, h - sy);
  const float rcp = __uint_as_float(m.y);
  const float z2 = __int_as_float(b.z), z0mz2 = __int_as_float(b.w), z1mz2 = __uint_as_float(m.x);
  const unsigned long long low = (unsigned long long)(0xFFFFFFFEu - s.tri[o]);
  unsigned long long *col = vis + (py0 - ylo) * T + (px - xlo);
  const int xy[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
  if (m.w >> 31) {
    // int32-path triangle: the edge functions at the span's pixel centres in
    // wrapping 32-bit arithmetic equal the exact values (all fit in int32)
    const uint32_t half = 1u << (S - 1);
    const uint32_t cx = ((uint32_t)px << S) + half, cy = ((uint32_t)py0 << S) + half;
    uint32_t e[3], dy[3];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      const uint32_t xa = (uint32_t)xy[2 * kk], ya = (uint32_t)xy[2 * kk + 1];
      const uint32_t xb = (uint32_t)xy[(2 * kk + 2) % 6], yb = (uint32_t)xy[(2 * kk + 3) % 6];
      e[kk] = (cy - ya) * (xb - xa) - (cx - xa) * (yb - ya);
      dy[kk] = (xb - xa) << S;
    }
#pragma unroll
    for (uint32_t j = 0; j < (uint32_t)kTileSpan; ++j) {
      if (j < nv && (int)(e[0] | e[1] | e[2]) >= 0) {
        const float alpha = (float)(int)e[1] * rcp, beta = (float)(int)e[2] * rcp;  // raster.pikostage:269-270
        const float z = interp(z0mz2, z1mz2, z2, alpha, beta);                      // raster.pikostage:272
        atomicMin((long long *)(col + j * T), (long long)(((unsigned long long)__float_as_uint(z) << 32) | low));
      }
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) e[kk] += dy[kk];
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < (uint32_t)kTileSpan; ++j) {
      float alpha, beta;
      if (j < nv && covered_at(xy, false, px, py0 + (int)j, S, rcp, alpha, beta)) {
        const float z = interp(z0mz2, z1mz2, z2, alpha, beta);
        atomicMin((long long *)(col + j * T), (long long)(((unsigned long long)__float_as_uint(z) << 32) | low));
      }
    }
  }
}

template <int T, int KIND, bool MULTI>
__global__ void __launch_bounds__(kTRThreads) k_tile_raster(const int4 *__restrict__ geo,
                                                             const void *__restrict__ attr, FrameArgs fa,
                                                             const FrameArgs *__restrict__ jobs, int njobs,
                                                             const uint32_t *__restrict__ tab,
                                                             const uint32_t *__restrict__ pairs, int tile_lo,
                                                             int tile_hi, uint32_t *__restrict__ color,
                                                             float *__restrict__ depth) {
  __shared__ unsigned long long vis[T * T];
  __shared__ uint32_t w_pre[2][kTRWin];   // window prefix: pairs, large triangles
  __shared__ uint32_t w_base[2][kTRWin];  // where the chunk's run starts in `pairs`
  __shared__ TileStage st[kTRThreads / 64];
  __shared__ uint32_t sh[kTRThreads / 64 + 1];
  __shared__ float s_u8[256];
  const int t = threadIdx.x, lane = t & 63;
  TileStage &s = st[t >> 6];
  // block i runs on XCD i % 8, so consecutive tiles go to different XCDs:
  // a dense band of the screen (a horizon) is spread over the whole GPU.
  // Giving each XCD a contiguous range of tiles (neighbouring tiles' table
  // entries in one L2) was slower: 87 vs 79 us on the 2M-triangle height
  // field at T=32, 4.7 vs 4.5 ms on the batch.
  const int nt = tile_hi - tile_lo;
  const int tb = (int)blockIdx.x;
  if (tb >= nt) return;
  const int tile = tile_lo + tb;
  const FrameArgs &ja = MULTI ? jobs[job_search(jobs, njobs, (uint32_t)tile, true)] : fa;
  const int lt = MULTI ? tile - (int)ja.bin_base : tile;
  const int ntiles = ja.nBx * ja.nBy;
  const int tx = lt % ja.nBx, ty = lt / ja.nBx;
  const int xlo = tx * T, ylo = ty * T, xhi = min(xlo + T, ja.W), yhi = min(ylo + T, ja.H);
  const int S = ja.S;
  for (int q = t; q < 256; q += kTRThreads) s_u8[q] = kU8Table.v[q];
  for (int q = t; q < T * T; q += kTRThreads) vis[q] = kVisSentinel;
  const uint32_t nch = ja.nchunks, col_h = (uint32_t)ntiles + 2u;
  for (uint32_t w0 = 0; w0 < nch; w0 += kTRWin) {
    // thread t reads chunks w0 + t*kTRWinPer ... (consecutive, so the scan
    // keeps chunk order)
    uint32_t cnt[2][kTRWinPer], base[2][kTRWinPer], sum[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < kTRWinPer; ++k) {
      const uint32_t c = w0 + (uint32_t)(t * kTRWinPer + k);
      cnt[0][k] = cnt[1][k] = base[0][k] = base[1][k] = 0;
      if (c < nch) {
        const uint32_t *colm = tab + ja.tab_base + (size_t)c * col_h;
        const uint32_t o = colm[lt], o1 = colm[lt + 1], nl = colm[ntiles + 1];
        const uint32_t reg = (ja.chunk_base + c) * kChunkCap;
        cnt[0][k] = o1 - o;
        base[0][k] = reg + o;
        cnt[1][k] = nl;
        base[1][k] = reg + kChunkCap - nl;
      }
      sum[0] += cnt[0][k];
      sum[1] += cnt[1][k];
    }
    uint32_t tot[2];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      uint32_t run = block_excl_scan<kTRThreads>(sum[l], sh, tot[l]);
#pragma unroll
      for (int k = 0; k < kTRWinPer; ++k) {
        w_pre[l][t * kTRWinPer + k] = run;
        w_base[l][t * kTRWinPer + k] = base[l][k];
        run += cnt[l][k];
      }
    }
    __syncthreads();
    for (int l = 0; l < 2; ++l) {  // the tile's pairs, then the large triangles
      // entry j of the window's list -> triangle index (~0u past the end)
      auto entry = [&](uint32_t j) -> uint32_t {
        if (j >= tot[l]) return ~0u;
        int lo = 0;  // chunk of entry j: largest k with w_pre[k] <= j
#pragma unroll
        for (int step = kTRWin / 2; step > 0; step >>= 1)
          if (w_pre[l][lo + step] <= j) lo += step;
        return pairs[w_base[l][lo] + (j - w_pre[l][lo])];
      };
      auto setup = [&](uint32_t v, int4 *out) {
        if (v != ~0u) {
#pragma unroll
          for (int q = 0; q < 3; ++q) out[q] = geo[3 * (size_t)v + q];
        } else {
#pragma unroll
          for (int q = 0; q < 3; ++q) out[q] = make_int4(0, 0, 0, 0);
        }
      };
      // software pipeline: the setup records of the next batch and the
      // entries of the one after are in flight while a batch is rasterised
      uint32_t v_cur = entry((uint32_t)t);
      int4 gc[3];
      setup(v_cur, gc);
      uint32_t v_nxt = entry((uint32_t)(kTRThreads + t));
      for (uint32_t b0 = 0; b0 < tot[l]; b0 += kTRThreads) {
        int4 gn[3];
        setup(v_nxt, gn);
        const uint32_t v_nn = entry(b0 + 2u * kTRThreads + (uint32_t)t);
        uint32_t units = 0;
        if (v_cur != ~0u) {
          const int p[6] = {gc[0].x, gc[0].y, gc[0].z, gc[0].w, gc[1].x, gc[1].y};
          int x0, y0, x1, y1;
          clip_rect(p, S, ja.W, ja.H, x0, y0, x1, y1);
          x0 = max(x0, xlo);
          y0 = max(y0, ylo);
          x1 = min(x1, xhi);
          y1 = min(y1, yhi);
          const int w = max(0, x1 - x0), h = max(0, y1 - y0);
          if (w > 0 && h > 0) {
            // staged for the wave's balanced spans (tile_span)
            units = (uint32_t)w * (uint32_t)((h + kTileSpan - 1) / kTileSpan);
            s.xy01[lane] = make_int4(p[0], p[1], p[2], p[3]);
            s.xy2z[lane] = make_int4(p[4], p[5], gc[1].w, gc[2].x);
            s.misc[lane] = make_uint4((uint32_t)gc[2].y, (uint32_t)gc[1].z, (uint32_t)x0 | ((uint32_t)y0 << 16),
                                      (uint32_t)w | ((uint32_t)h << 16) | (small_bbox(p, S) ? 0x80000000u : 0u));
            s.tri[lane] = v_cur;
          }
        }
        v_cur = v_nxt;
#pragma unroll
        for (int q = 0; q < 3; ++q) gc[q] = gn[q];
        v_nxt = v_nn;
        const uint32_t inc = wave_incl_scan(units);
        s.pre[lane] = inc - units;
        const uint32_t wtot = __shfl(inc, 63, 64);
        wave_fence();
        for (uint32_t cidx = (uint32_t)lane; cidx < wtot; cidx += 64) {
          int o = 0;  // owner: largest o with pre[o] <= cidx
#pragma unroll
          for (int step = 32; step > 0; step >>= 1)
            if (s.pre[o + step] <= cidx) o += step;
          tile_span<T>(s, o, cidx - s.pre[o], S, xlo, ylo, vis);
        }
        wave_fence();
      }
    }
    __syncthreads();
  }
  __syncthreads();
  // resolve: depth = the winner's z, colour = Gouraud of the winner
  const size_t fbb = MULTI ? (size_t)ja.fb_base : 0;
  for (int q = t; q < T * T; q += kTRThreads) {
    const int px = xlo + (q % T), py = ylo + (q / T);
    if (px >= xhi || py >= yhi) continue;
    const unsigned long long key = vis[q];
    uint32_t col = kClearColor;
    if (key != kVisSentinel) {
      const uint32_t v = 0xFFFFFFFEu - (uint32_t)(key & 0xFFFFFFFFull);
      const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(v - ja.vbase) : (size_t)v;
      const int4 a = geo[3 * (size_t)v], b = geo[3 * (size_t)v + 1];
      const int p[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
      uint32_t c[3];
      tri_colours<KIND>(attr, src, ja, c);
      float alpha = 0.0f, beta = 0.0f;
      covered_at(p, small_bbox(p, S), px, py, S, __int_as_float(b.z), alpha, beta);
      col = gouraud(s_u8, c[0], c[1], c[2], alpha, beta);
    }
    const size_t pid = fbb + (size_t)py * ja.W + px;
    depth[pid] = __int_as_float((int)(key >> 32));
    color[pid] = col;
  }
}

#include "raster_queue.h"

inline int ilog2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace

// ---- pipe object --------------------------------------------------------
//
// One object serves both forms.  A plain pipe is one job described by
// kernel arguments (RasterPipe of rasterPipe.h:16-45).  A batch pipe holds a
// device job table: every job is a screen with its own camera, resolution and
// bin size drawing a range of the resident input triangles; all jobs of a
// batch run through one kernel sequence (one raster launch per bin size).

struct RasterClass {
  int T, bin_lo, bin_hi, waves;  // waves: persistent raster grid for this class
};

struct piko_raster_pipe {
  bool multi = false;
  // A plain queue pipe's second set of frame buffers and stream, made on the
  // first run of two or more frames: run_frames keeps two frames in flight,
  // alternating between the pipe and its twin (see piko_raster_run_frames).
  piko_raster_pipe *twin = nullptr;
  bool twin_tried = false;
  piko_constant_state cs_cur{};  // the plain pipe's current uniforms and options (for the twin)
  piko_raster_options opt0{};
  int kind = 0, S = 0;
  int W = 0, H = 0, T = 0, lgT = 0, nBx = 0, nBy = 0;  // plain pipe view
  FrameArgs fa{};                                         // plain pipe job
  // current frame
  int ntris = 0, nbins = 0, passes = 0;
  long long npix = 0;
  std::vector<FrameArgs> jobs;     // device order: grouped by bin size, then submission order
  std::vector<int> job_slot;       // submitted job index -> position in `jobs`
  std::vector<RasterClass> classes;
  bool in_flight = false;
  // capacities
  long long cap_tris = 0, cap_pix = 0;
  int cap_bins = 0, cap_jobs = 0;
  long long cap = 0;  // (triangle, bin) pairs
  int nblk_cap = 0;
  long long input_tris = 0;
  hipStream_t stream = nullptr;
  const float *d_pos = nullptr;
  const void *d_attr = nullptr;
  bool owns_input = false;
  FrameArgs *d_jobs = nullptr;
  FrameArgs *h_jobs = nullptr;  // pinned staging for the job table
  int4 *d_geo = nullptr;
  uint4 *d_col = nullptr;
  uint32_t *d_cnt = nullptr;         // per triangle: bins assigned
  uint32_t *d_keep = nullptr;        // per vertex-stage tile: emitted triangles
  uint32_t *d_tile_pairs = nullptr;  // per vertex-stage tile: pairs, then their offsets
  FrameState *d_fs = nullptr;
  uint32_t *d_bstart = nullptr, *d_bend = nullptr;  // [cap_bins]
  uint32_t *d_keys[2] = {nullptr, nullptr};
  uint32_t *d_vals[2] = {nullptr, nullptr};
  uint32_t *d_hist = nullptr;    // [256][nblk_cap] digit matrix
  uint32_t *d_rowtot = nullptr;  // [256] digit totals
  uint32_t *d_color = nullptr;
  float *d_depth = nullptr;
  unsigned long long *d_vis = nullptr;  // [cap_pix] direct-path visibility words
  uint32_t *d_large = nullptr;          // [large_cap] large triangles of the frame
  uint32_t *d_tile_off = nullptr;       // [large_cap] their first 8x8 tile
  uint32_t large_cap = 0;
  // tiled path: chunk table, offset table, per-chunk pair regions
  ChunkDesc *d_chunks = nullptr;
  ChunkDesc *h_chunks = nullptr;        // pinned staging for the chunk table
  uint32_t *d_tab = nullptr;            // [cap_tab] chunk-major tile offsets
  uint32_t *d_tpairs = nullptr;         // [cap_chunks * kChunkCap] triangle indices
  uint32_t *d_chunk_keep = nullptr;     // [cap_chunks] emitted triangles per chunk
  long long cap_chunks = 0, cap_tab = 0, chunk_bytes = 0;
  int nchunks = 0;
  FrameArgs fa_t{};                     // plain pipe job in tile numbering
  RasterClass tclass{};                 // plain pipe: its one tile class
  int mode = PIKO_RASTER_TILED;         // frame path
  int last_mode = PIKO_RASTER_TILED;    // path of the last enqueued frame
  bool bins_current = false;            // the bin queues hold the last frame's pairs
  int final_buf = 0;
  int raster_waves = 0;       // persistent raster grid (default)
  int cus = 256;              // compute units of the device
  int raster_persistent = 1;  // 1: persistent pipelined waves; 0: one block per bin
  // queue path: bin items (chunks), per-tile queue ranges and counters, part
  // items, the pair buffer and the split tiles' scratch slots
  QChunk *d_qchunks = nullptr;
  QChunk *h_qchunks = nullptr;          // pinned staging for the chunk table
  long long cap_qchunks = 0;
  int q_nchunks = 0;
  uint32_t *d_qbase = nullptr, *d_qcap = nullptr, *d_qfill = nullptr, *d_qneed = nullptr;
  unsigned long long *d_qcost = nullptr;  // the counting pass's raster cost estimate per tile
  long long cap_qtiles = 0;
  int q_ntiles = 0;
  std::vector<uint32_t> q_cap;          // host copy of the tiles' queue capacities
  std::vector<unsigned long long> q_cost;  // the tiles' cost estimates (the counting pass's; 0: none)
  bool q_caps_known = false;            // q_cap holds counts of a frame of the current jobs
  QItem *d_qitems = nullptr;
  long long cap_qitems = 0;
  int q_nitems = 0;
  // Frames in flight (a plain pipe and its twin in run_frames): one part per
  // tile, costliest first, and a persistent grid of kQFramesWpc waves per CU,
  // which leaves CU slots for the other frame's launches.
  bool q_frames = false;                // launch with the frames-in-flight items and grid
  QItem *d_qitems_f = nullptr;
  long long cap_qitems_f = 0;
  int q_nitems_f = 0;
  uint32_t *d_qsplits = nullptr;        // first part's item of each split tile
  long long cap_qsplits = 0;
  int q_nsplits = 0;
  uint32_t *d_qpairs = nullptr;
  long long cap_qpairs = 0;
  unsigned long long *d_qscr = nullptr;
  long long cap_qscr = 0;               // scratch words
  QState *d_qs = nullptr;
  int q_QT = 32;                        // LDS tile edge of the launch (largest job tile)
  int q_part_pairs = 1024;
  int q_wpc = 16;                       // persistent waves per CU
  uint32_t q_walk = 1;                  // PIKO_COVER_WALK (1) or SPANS (0) for jobs naming none
  FrameArgs fa_q{};                     // plain pipe job in queue tile numbering
  long long q_bytes = 0;
  long long large_tiles_hint = -1;  // 8x8 tiles of large triangles last seen (-1: unknown)
  bool last_inline = false;         // the last direct frame drew large triangles in the vertex stage
  long long device_bytes = 0;
  long long pair_bytes = 0, tri_bytes = 0, pix_bytes = 0, bin_bytes = 0;
};

namespace {

#define HIPOK(x)                                  \
  do {                                            \
    if ((x) != hipSuccess) return PIKO_ERR_HIP;   \
  } while (0)

template <typename T>
int dalloc(piko_raster_pipe *p, T **ptr, long long elems, long long *acct = nullptr) {
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
void dfree(T *&ptr) {
  if (ptr) (void)hipFree((void *)ptr);
  ptr = nullptr;
}

void free_pairs(piko_raster_pipe *p) {
  for (int k = 0; k < 2; ++k) {
    dfree(p->d_keys[k]);
    dfree(p->d_vals[k]);
  }
  dfree(p->d_hist);
  p->device_bytes -= p->pair_bytes;
  p->pair_bytes = 0;
}

int alloc_pairs(piko_raster_pipe *p, long long cap) {
  p->cap = cap;
  p->nblk_cap = (int)cdiv(cap, kSortTile);
  int rc;
  for (int k = 0; k < 2; ++k) {
    if ((rc = dalloc(p, &p->d_keys[k], cap, &p->pair_bytes)) != PIKO_OK) return rc;
    if ((rc = dalloc(p, &p->d_vals[k], cap, &p->pair_bytes)) != PIKO_OK) return rc;
  }
  if ((rc = dalloc(p, &p->d_hist, 256ll * p->nblk_cap, &p->pair_bytes)) != PIKO_OK) return rc;
  return PIKO_OK;
}

void free_tris(piko_raster_pipe *p) {
  dfree(p->d_geo);
  dfree(p->d_col);
  dfree(p->d_cnt);
  dfree(p->d_keep);
  dfree(p->d_tile_pairs);
  dfree(p->d_large);
  dfree(p->d_tile_off);
  p->device_bytes -= p->tri_bytes;
  p->tri_bytes = 0;
}

int alloc_tris(piko_raster_pipe *p, long long ntris) {
  p->cap_tris = std::max<long long>(ntris, 1);
  const long long tiles = cdiv(p->cap_tris, kVsTile);
  int rc;
  if ((rc = dalloc(p, &p->d_geo, 3 * p->cap_tris, &p->tri_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_col, p->cap_tris, &p->tri_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_cnt, p->cap_tris, &p->tri_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_keep, tiles, &p->tri_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_tile_pairs, tiles, &p->tri_bytes)) != PIKO_OK) return rc;
  p->large_cap = (uint32_t)std::min<long long>(std::max<long long>({(long long)p->large_cap, p->cap_tris / 32, 4096}), 1 << 30);
  if ((rc = dalloc(p, &p->d_large, p->large_cap, &p->tri_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_tile_off, p->large_cap, &p->tri_bytes)) != PIKO_OK) return rc;
  return PIKO_OK;
}

__global__ void k_fill_vis(unsigned long long *p, long long n) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = kVisSentinel;
}

void free_pix(piko_raster_pipe *p) {
  dfree(p->d_color);
  dfree(p->d_depth);
  dfree(p->d_vis);
  p->device_bytes -= p->pix_bytes;
  p->pix_bytes = 0;
}

int alloc_pix(piko_raster_pipe *p, long long npix) {
  p->cap_pix = std::max<long long>(npix, 1);
  int rc;
  if ((rc = dalloc(p, &p->d_color, p->cap_pix, &p->pix_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_depth, p->cap_pix, &p->pix_bytes)) != PIKO_OK) return rc;
  if (p->mode != PIKO_RASTER_DIRECT) return PIKO_OK;  // per-pixel visibility words: direct path only
  if ((rc = dalloc(p, &p->d_vis, p->cap_pix, &p->pix_bytes)) != PIKO_OK) return rc;
  k_fill_vis<<<(unsigned)std::min<long long>(cdiv(p->cap_pix, 256), 8192), 256, 0, p->stream>>>(p->d_vis, p->cap_pix);
  return hipGetLastError() == hipSuccess ? PIKO_OK : PIKO_ERR_HIP;
}

void free_chunks(piko_raster_pipe *p) {
  dfree(p->d_chunks);
  dfree(p->d_tab);
  dfree(p->d_tpairs);
  dfree(p->d_chunk_keep);
  if (p->h_chunks) (void)hipHostFree(p->h_chunks);
  p->h_chunks = nullptr;
  p->device_bytes -= p->chunk_bytes;
  p->chunk_bytes = 0;
  p->cap_chunks = p->cap_tab = 0;
}

int alloc_chunks(piko_raster_pipe *p, long long nch, long long ntab) {
  p->cap_chunks = std::max<long long>(nch, 1);
  p->cap_tab = std::max<long long>(ntab, 1);
  int rc;
  if ((rc = dalloc(p, &p->d_chunks, p->cap_chunks, &p->chunk_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_tab, p->cap_tab, &p->chunk_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_tpairs, p->cap_chunks * (long long)kChunkCap, &p->chunk_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_chunk_keep, p->cap_chunks, &p->chunk_bytes)) != PIKO_OK) return rc;
  if (hipHostMalloc((void **)&p->h_chunks, (size_t)p->cap_chunks * sizeof(ChunkDesc)) != hipSuccess) {
    (void)hipGetLastError();
    p->h_chunks = nullptr;
    return PIKO_ERR_OOM;
  }
  return PIKO_OK;
}

// Tile edge of the tiled path for a job asked to bin at T: at least 16, and
// large enough that one job has at most kMaxTilesJob tiles (64 always is for
// screens up to 4095 x 4095).
int tiled_T(int T, int W, int H) {
  int tt = std::max(T, 16);
  while (tt < 64 && cdiv(W, tt) * cdiv(H, tt) > kMaxTilesJob) tt *= 2;
  return tt;
}

// Chunks of the jobs (device order; triangle counts from the vbase steps and
// the batch total): sets each job's chunk_base / nchunks / tab_base, grows
// the tables when needed and enqueues the chunk table's upload.
int setup_chunks(piko_raster_pipe *p, FrameArgs *jobs, int njobs, long long total_tris) {
  long long nch = 0, ntab = 0;
  for (int j = 0; j < njobs; ++j) {
    const long long end = j + 1 < njobs ? (long long)jobs[j + 1].vbase : total_tris;
    const long long cnt = end - (long long)jobs[j].vbase;
    const long long nc = cdiv(cnt, kChunk);
    jobs[j].chunk_base = (uint32_t)nch;
    jobs[j].nchunks = (uint32_t)nc;
    jobs[j].tab_base = (uint32_t)ntab;
    nch += nc;
    ntab += nc * ((long long)jobs[j].nBx * jobs[j].nBy + 2);
  }
  if (nch > 0xFFFFFFFFll / kChunkCap || ntab > 0xFFFFFFFFll) return PIKO_ERR_CAPACITY;
  HIPOK(hipStreamSynchronize(p->stream));  // the staging buffer may still be in use
  if (nch > p->cap_chunks || ntab > p->cap_tab || !p->h_chunks) {
    free_chunks(p);
    int rc = alloc_chunks(p, nch, ntab);
    if (rc != PIKO_OK) return rc;
  }
  long long k = 0;
  for (int j = 0; j < njobs; ++j) {
    const uint32_t ntl = (uint32_t)(jobs[j].nBx * jobs[j].nBy) + 2u;
    const long long end = j + 1 < njobs ? (long long)jobs[j + 1].vbase : total_tris;
    for (uint32_t c = 0; c < jobs[j].nchunks; ++c, ++k) {
      ChunkDesc &d = p->h_chunks[k];
      d.job = (uint32_t)j;
      d.tri0 = jobs[j].vbase + c * (uint32_t)kChunk;
      d.count = (uint32_t)std::min<long long>(kChunk, end - (long long)d.tri0);
      d.tab_off = jobs[j].tab_base + c * ntl;
    }
  }
  p->nchunks = (int)nch;
  if (nch > 0)
    HIPOK(hipMemcpyAsync(p->d_chunks, p->h_chunks, (size_t)nch * sizeof(ChunkDesc), hipMemcpyHostToDevice,
                         p->stream));
  return PIKO_OK;
}

// ---- queue path (host) ---------------------------------------------------

// Tile edge of the queue path for a job asked to bin at T: at least 16, and
// large enough that one job has at most kQMaxTiles tiles (64 always is for
// screens up to 4095 x 4095).
int queue_T(int T, int W, int H) {
  int tt = std::max(T, 16);
  while (tt < 64 && cdiv(W, tt) * cdiv(H, tt) > kQMaxTiles) tt *= 2;
  return tt;
}

template <typename T>
void q_free(piko_raster_pipe *p, T *&ptr, long long &cap) {
  if (ptr) {
    const long long b = std::max<long long>(cap, 1) * (long long)sizeof(T);
    p->device_bytes -= b;
    p->q_bytes -= b;
  }
  dfree(ptr);
  cap = 0;
}

// Grows a queue-path buffer to at least `need` elements (contents dropped).
template <typename T>
int q_grow(piko_raster_pipe *p, T *&ptr, long long &cap, long long need) {
  if (ptr && need <= cap) return PIKO_OK;
  q_free(p, ptr, cap);
  const long long n = std::max<long long>(need + need / 4, 1);
  int rc = dalloc(p, &ptr, n, &p->q_bytes);
  if (rc != PIKO_OK) return rc;
  cap = n;
  return PIKO_OK;
}

void free_queue(piko_raster_pipe *p) {
  q_free(p, p->d_qchunks, p->cap_qchunks);
  if (p->h_qchunks) (void)hipHostFree(p->h_qchunks);
  p->h_qchunks = nullptr;
  long long tc = p->cap_qtiles;
  q_free(p, p->d_qbase, tc);
  tc = p->cap_qtiles;
  q_free(p, p->d_qcap, tc);
  tc = p->cap_qtiles;
  q_free(p, p->d_qfill, tc);
  tc = p->cap_qtiles;
  q_free(p, p->d_qneed, tc);
  tc = p->cap_qtiles;
  q_free(p, p->d_qcost, tc);
  p->cap_qtiles = 0;
  q_free(p, p->d_qitems, p->cap_qitems);
  q_free(p, p->d_qitems_f, p->cap_qitems_f);
  q_free(p, p->d_qsplits, p->cap_qsplits);
  q_free(p, p->d_qpairs, p->cap_qpairs);
  q_free(p, p->d_qscr, p->cap_qscr);
  long long one = 1;
  q_free(p, p->d_qs, one);
}

// Part items and queue ranges from the tiles' capacities: ranges in tile
// order, split tiles with their scratch slots, a tile cut into parts of at
// most part_pairs entries (equal slices of its range).  Items go costliest
// first, by the counting pass's cost estimate per part (q_cost) where there
// is one, else by entries: a near tile of a few hundred large triangles costs
// as much as a horizon tile of thousands of small ones, and started last it
// set the length of the launch (k_qraster 79.8 -> 76.8 us on the 2M height
// field).  Cutting the costliest tiles into more parts balanced the
// workgroups better but added as much per-item time as it saved.
int q_build(piko_raster_pipe *p) {
  const int n = p->q_ntiles;
  std::vector<uint32_t> base((size_t)n);
  unsigned long long total = 0, total_cost = 0;
  for (int t = 0; t < n; ++t) {
    base[t] = (uint32_t)total;
    total += p->q_cap[t];
    total_cost += t < (int)p->q_cost.size() ? p->q_cost[t] : 0ull;
  }
  if (total >= (unsigned long long)kQNoSlot) return PIKO_ERR_CAPACITY;
  const long long P = p->q_part_pairs;
  std::vector<uint32_t> np((size_t)n);
  std::vector<unsigned long long> key((size_t)n);  // per-part cost (entries where no cost is known)
  for (int t = 0; t < n; ++t) {
    const long long cap = p->q_cap[t];
    const unsigned long long c = t < (int)p->q_cost.size() ? p->q_cost[t] : 0ull;
    np[t] = (uint32_t)std::max<long long>(1, cdiv(cap, P));
    key[t] = total_cost > 0 ? c / np[t] : (unsigned long long)cap;
  }
  std::vector<int> order((size_t)n);
  for (int t = 0; t < n; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] > key[b]; });
  std::vector<QItem> items;
  std::vector<uint32_t> splits;
  items.reserve((size_t)n + (size_t)(total / (unsigned long long)P) + 1);
  long long slots = 0;
  for (int t : order) {
    const uint32_t nparts = np[t];
    const uint32_t slot = nparts > 1 ? (uint32_t)slots : 0u;
    if (nparts > 1) {
      slots += nparts;
      splits.push_back((uint32_t)items.size());
    }
    const unsigned long long cap = p->q_cap[t];
    for (uint32_t k = 0; k < nparts; ++k)
      items.push_back(QItem{(uint32_t)t, k, nparts, slot, base[t], p->q_cap[t], (uint32_t)(cap * k / nparts),
                            (uint32_t)(cap * (k + 1) / nparts)});
  }
  if (items.size() > 0x7FFFFFFFull || slots > 0xFFFFFFFFll) return PIKO_ERR_CAPACITY;
  // the frames-in-flight items (plain pipes): whole tiles, costliest first
  std::vector<QItem> items_f;
  if (!p->multi) {
    std::vector<int> ord((size_t)n);
    for (int t = 0; t < n; ++t) ord[t] = t;
    auto cost = [&](int t) {
      return total_cost > 0 ? (t < (int)p->q_cost.size() ? p->q_cost[t] : 0ull) : (unsigned long long)p->q_cap[t];
    };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cost(a) > cost(b); });
    items_f.reserve((size_t)n);
    for (int t : ord) items_f.push_back(QItem{(uint32_t)t, 0u, 1u, 0u, base[t], p->q_cap[t], 0u, p->q_cap[t]});
  }
  int rc;
  HIPOK(hipStreamSynchronize(p->stream));  // buffers may still be in use
  if ((rc = q_grow(p, p->d_qpairs, p->cap_qpairs, (long long)total)) != PIKO_OK) return rc;
  if ((rc = q_grow(p, p->d_qitems, p->cap_qitems, (long long)items.size())) != PIKO_OK) return rc;
  if ((rc = q_grow(p, p->d_qitems_f, p->cap_qitems_f, (long long)items_f.size())) != PIKO_OK) return rc;
  if ((rc = q_grow(p, p->d_qscr, p->cap_qscr, slots * p->q_QT * p->q_QT)) != PIKO_OK) return rc;
  if ((rc = q_grow(p, p->d_qsplits, p->cap_qsplits, (long long)splits.size())) != PIKO_OK) return rc;
  if (n > 0) {
    HIPOK(hipMemcpyAsync(p->d_qbase, base.data(), (size_t)n * 4, hipMemcpyHostToDevice, p->stream));
    HIPOK(hipMemcpyAsync(p->d_qcap, p->q_cap.data(), (size_t)n * 4, hipMemcpyHostToDevice, p->stream));
  }
  if (!items.empty())
    HIPOK(hipMemcpyAsync(p->d_qitems, items.data(), items.size() * sizeof(QItem), hipMemcpyHostToDevice,
                         p->stream));
  if (!splits.empty())
    HIPOK(hipMemcpyAsync(p->d_qsplits, splits.data(), splits.size() * 4, hipMemcpyHostToDevice, p->stream));
  if (!items_f.empty())
    HIPOK(hipMemcpyAsync(p->d_qitems_f, items_f.data(), items_f.size() * sizeof(QItem), hipMemcpyHostToDevice,
                         p->stream));
  p->q_nitems = (int)items.size();
  p->q_nsplits = (int)splits.size();
  p->q_nitems_f = (int)items_f.size();
  HIPOK(hipStreamSynchronize(p->stream));  // the host vectors go out of scope
  return PIKO_OK;
}

// Chunks and tiles of the jobs (device order, tile numbering from bin_base);
// capacities are kept when the jobs are the ones already set up.
int q_setup(piko_raster_pipe *p, const FrameArgs *jobs, int njobs, long long total_tris) {
  long long nch = 0, ntiles = 0;
  int QT = 32;
  for (int j = 0; j < njobs; ++j) {
    const long long end = j + 1 < njobs ? (long long)jobs[j + 1].vbase : total_tris;
    nch += cdiv(end - (long long)jobs[j].vbase, kQChunk);
    ntiles = std::max<long long>(ntiles, (long long)jobs[j].bin_base + jobs[j].nbins);
    if ((1 << jobs[j].lgT) > QT) QT = 1 << jobs[j].lgT;
  }
  if (nch > 0x7FFFFFFFll || ntiles > 0x7FFFFFFFll) return PIKO_ERR_CAPACITY;
  HIPOK(hipStreamSynchronize(p->stream));  // staging and counters may still be in use
  int rc;
  if (nch > p->cap_qchunks || !p->h_qchunks) {
    q_free(p, p->d_qchunks, p->cap_qchunks);
    if (p->h_qchunks) (void)hipHostFree(p->h_qchunks);
    p->h_qchunks = nullptr;
    long long cap = 0;
    if ((rc = q_grow(p, p->d_qchunks, cap, nch)) != PIKO_OK) return rc;
    p->cap_qchunks = cap;
    if (hipHostMalloc((void **)&p->h_qchunks, (size_t)cap * sizeof(QChunk)) != hipSuccess) {
      (void)hipGetLastError();
      p->h_qchunks = nullptr;
      return PIKO_ERR_OOM;
    }
  }
  long long k = 0;
  for (int j = 0; j < njobs; ++j) {
    const long long end = j + 1 < njobs ? (long long)jobs[j + 1].vbase : total_tris;
    for (long long t0 = jobs[j].vbase; t0 < end; t0 += kQChunk, ++k)
      p->h_qchunks[k] = QChunk{(uint32_t)j, (uint32_t)t0, (uint32_t)std::min<long long>(kQChunk, end - t0), 0u};
  }
  p->q_nchunks = (int)nch;
  if (nch > 0)
    HIPOK(hipMemcpyAsync(p->d_qchunks, p->h_qchunks, (size_t)nch * sizeof(QChunk), hipMemcpyHostToDevice,
                         p->stream));
  const bool same_tiles = p->q_caps_known && ntiles == p->q_ntiles && QT == p->q_QT;
  if (!p->d_qs) {
    long long one = 0;
    if ((rc = q_grow(p, p->d_qs, one, 1)) != PIKO_OK) return rc;
    HIPOK(hipMemsetAsync(p->d_qs, 0, sizeof(QState), p->stream));
  }
  if (ntiles > p->cap_qtiles) {
    long long tc = p->cap_qtiles;
    q_free(p, p->d_qbase, tc);
    tc = p->cap_qtiles;
    q_free(p, p->d_qcap, tc);
    tc = p->cap_qtiles;
    q_free(p, p->d_qfill, tc);
    tc = p->cap_qtiles;
    q_free(p, p->d_qneed, tc);
    tc = p->cap_qtiles;
    q_free(p, p->d_qcost, tc);
    const long long cap = std::max<long long>(ntiles, 1);
    for (uint32_t **a : {&p->d_qbase, &p->d_qcap, &p->d_qfill, &p->d_qneed})
      if ((rc = dalloc(p, a, cap, &p->q_bytes)) != PIKO_OK) return rc;
    if ((rc = dalloc(p, &p->d_qcost, cap, &p->q_bytes)) != PIKO_OK) return rc;
    p->cap_qtiles = cap;
    HIPOK(hipMemsetAsync(p->d_qfill, 0, (size_t)cap * 4, p->stream));
  }
  p->q_QT = QT;
  if (!same_tiles) {
    p->q_ntiles = (int)ntiles;
    p->q_cap.assign((size_t)ntiles, 0u);
    p->q_cost.assign((size_t)ntiles, 0ull);
    p->q_caps_known = false;
  }
  return q_build(p);
}

void free_bins(piko_raster_pipe *p) {
  dfree(p->d_bstart);
  dfree(p->d_bend);
  p->device_bytes -= p->bin_bytes;
  p->bin_bytes = 0;
}

int alloc_bins(piko_raster_pipe *p, int nbins) {
  p->cap_bins = std::max(nbins, 1);
  int rc;
  if ((rc = dalloc(p, &p->d_bstart, (long long)p->cap_bins, &p->bin_bytes)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_bend, (long long)p->cap_bins, &p->bin_bytes)) != PIKO_OK) return rc;
  // queue ranges start cleared; the raster kernel clears each one after use
  HIPOK(hipMemsetAsync(p->d_bstart, 0, (size_t)p->cap_bins * 4, p->stream));
  HIPOK(hipMemsetAsync(p->d_bend, 0, (size_t)p->cap_bins * 4, p->stream));
  return PIKO_OK;
}

int radix_passes(int nbins) {
  int bits = 0;
  while ((1ll << bits) < (long long)nbins) ++bits;
  return (bits + 7) / 8;
}

// Kernel names in launch order (the sort section repeats once per pass, the
// raster once per bin-size class).
const char *kernel_name(const piko_raster_pipe *p, int k) {
  static const char *direct[] = {"k_vs_direct", "k_large_raster", "k_resolve"};
  if (p->last_mode == PIKO_RASTER_DIRECT) {
    if (p->last_inline) return k == 0 ? direct[0] : (k == 1 ? direct[2] : "");
    return k < 3 ? direct[k] : "";
  }
  if (p->last_mode == PIKO_RASTER_TILED) return k == 0 ? "k_tile_bin" : "k_tile_raster";
  if (p->last_mode == PIKO_RASTER_QUEUE) return k == 0 ? "k_qbin" : (k == 1 ? "k_qraster" : "k_qmerge");
  static const char *head[] = {"k_vs_count", "k_scan_tiles", "k_emit_pairs"};
  static const char *pass[] = {"k_sort_count", "k_sort_rows", "k_sort_scatter"};
  if (k < 3) return head[k];
  if (k < 3 + 3 * p->passes) return pass[(k - 3) % 3];
  if (k == 3 + 3 * p->passes) return "k_bounds";
  return "k_raster";
}

template <bool MULTI>
void launch_raster(piko_raster_pipe *p, int T, int bin_lo, int bin_hi, int waves_override,
                   const uint32_t *bin_tris) {
  const int nb = bin_hi - bin_lo;
  if (nb <= 0) return;
  const int waves = waves_override > 0 ? waves_override : p->raster_waves;
  const int rgrid = p->raster_persistent ? std::max(1, std::min(waves, nb)) : 8 * ((nb + 7) / 8);
#define PIKO_RASTER_LAUNCH(TT)                                                                                   \
  k_raster<TT, MULTI><<<rgrid, kRT, 0, p->stream>>>(p->d_geo, p->d_col, p->d_bstart, p->d_bend, bin_tris, p->fa, \
                                                    p->d_jobs, (int)p->jobs.size(), bin_lo, bin_hi, p->d_color,  \
                                                    p->d_depth, p->raster_persistent)
  switch (T) {
    case 8: PIKO_RASTER_LAUNCH(8); break;
    case 16: PIKO_RASTER_LAUNCH(16); break;
    case 32: PIKO_RASTER_LAUNCH(32); break;
    default: PIKO_RASTER_LAUNCH(64); break;
  }
#undef PIKO_RASTER_LAUNCH
}

// Records an event after each kernel when profiling.
struct Marks {
  hipEvent_t *ev;
  hipStream_t stream;
  int n = 0;
  void operator()() {
    if (ev) (void)hipEventRecord(ev[n], stream);
    ++n;
  }
};

// Vertex stage, (triangle, bin) pairs, sort by bin and bin bounds: the raster
// stage's queues for the current job set.
void enqueue_binning(piko_raster_pipe *p, Marks &mark) {
  const int n = p->ntris;
  const int tiles = (int)std::max<long long>(cdiv(n, kVsTile), 1);
  const int nj = (int)p->jobs.size();
#define PIKO_VS_LAUNCH(KIND, MULTI)                                                                              \
  k_vs_count<KIND, MULTI><<<tiles, kVsThreads, 0, p->stream>>>(p->d_pos, p->d_attr, n, p->fa, p->d_jobs, nj,     \
                                                               p->d_geo, p->d_col, p->d_cnt, p->d_tile_pairs,    \
                                                               p->d_keep)
  if (p->kind == PIKO_INPUT_PRETRANSFORMED) {
    if (p->multi) PIKO_VS_LAUNCH(PIKO_INPUT_PRETRANSFORMED, true);
    else PIKO_VS_LAUNCH(PIKO_INPUT_PRETRANSFORMED, false);
  } else {
    if (p->multi) PIKO_VS_LAUNCH(PIKO_INPUT_WORLD, true);
    else PIKO_VS_LAUNCH(PIKO_INPUT_WORLD, false);
  }
#undef PIKO_VS_LAUNCH
  mark();
  k_scan_tiles<<<1, 1024, 0, p->stream>>>(p->d_tile_pairs, tiles, p->d_fs, p->cap);
  mark();
  if (p->multi)
    k_emit_pairs<true><<<tiles, kVsThreads, 0, p->stream>>>(p->d_geo, p->d_cnt, n, p->fa, p->d_jobs, nj,
                                                           p->d_tile_pairs, p->cap, p->d_keys[0], p->d_vals[0]);
  else
    k_emit_pairs<false><<<tiles, kVsThreads, 0, p->stream>>>(p->d_geo, p->d_cnt, n, p->fa, p->d_jobs, nj,
                                                            p->d_tile_pairs, p->cap, p->d_keys[0], p->d_vals[0]);
  mark();
  int cur = 0;
  for (int pass = 0; pass < p->passes; ++pass) {
    k_sort_count<<<p->nblk_cap, kThreads, 0, p->stream>>>(p->d_keys[cur], p->d_fs, p->cap, p->d_hist,
                                                          p->nblk_cap, 8 * pass);
    mark();
    k_sort_rows<<<256, kThreads, 0, p->stream>>>(p->d_hist, p->d_fs, p->cap, p->nblk_cap, p->d_rowtot);
    mark();
    k_sort_scatter<<<p->nblk_cap, kThreads, 0, p->stream>>>(p->d_keys[cur], p->d_vals[cur], p->d_keys[cur ^ 1],
                                                            p->d_vals[cur ^ 1], p->d_fs, p->cap, p->d_hist,
                                                            p->d_rowtot, p->nblk_cap, 8 * pass);
    mark();
    cur ^= 1;
  }
  p->final_buf = cur;
  const int gstride = (int)std::min<long long>(std::max<long long>(cdiv(p->cap, kThreads * 8), 1), 1024);
  k_bounds<<<gstride, kThreads, 0, p->stream>>>(p->d_keys[cur], p->d_fs, p->cap, p->d_bstart, p->d_bend);
  mark();
}

// Sort-free frame: vertex stage + small-triangle raster, large-triangle tiles,
// resolve.
void enqueue_direct(piko_raster_pipe *p, Marks &mark) {
  const int n = p->ntris;
  const int nj = (int)p->jobs.size();
  const int blocks = (int)std::max<long long>(cdiv(n, kDirThreads), 1);
  // the last frame looked at had no large triangle: draw any in the vertex
  // stage and skip the large-triangle launch
  const int inl = p->large_tiles_hint == 0 ? 1 : 0;
  p->last_inline = inl != 0;
#define PIKO_VSD_LAUNCH(KIND, MULTI)                                                                           \
  k_vs_direct<KIND, MULTI><<<blocks, kDirThreads, 0, p->stream>>>(p->d_pos, n, p->S, p->fa, p->d_jobs, nj,     \
                                                                  p->d_geo, p->d_vis, p->d_large, p->d_tile_off, \
                                                                  p->d_fs, p->large_cap, inl)
  if (p->kind == PIKO_INPUT_PRETRANSFORMED) {
    if (p->multi) PIKO_VSD_LAUNCH(PIKO_INPUT_PRETRANSFORMED, true);
    else PIKO_VSD_LAUNCH(PIKO_INPUT_PRETRANSFORMED, false);
  } else {
    if (p->multi) PIKO_VSD_LAUNCH(PIKO_INPUT_WORLD, true);
    else PIKO_VSD_LAUNCH(PIKO_INPUT_WORLD, false);
  }
#undef PIKO_VSD_LAUNCH
  mark();
  if (!inl) {
  // grid-stride over the frame's large-triangle tiles: sized from the last
  // frame that was looked at (4 tiles per block in flight), the full grid
  // when unknown
  const int full = std::max(p->cus, 1) * 2;
  const int lgrid = p->large_tiles_hint < 0
                        ? full
                        : (int)std::min<long long>(std::max<long long>(cdiv(p->large_tiles_hint, 4), 1), full);
  if (p->multi)
    k_large_raster<true><<<lgrid, 256, 0, p->stream>>>(p->d_geo, p->d_large, p->d_tile_off, p->d_fs, p->large_cap,
                                                       p->fa, p->d_jobs, nj, p->d_vis);
  else
    k_large_raster<false><<<lgrid, 256, 0, p->stream>>>(p->d_geo, p->d_large, p->d_tile_off, p->d_fs, p->large_cap,
                                                        p->fa, p->d_jobs, nj, p->d_vis);
  mark();
  }
  const long long rblocks = 8 * cdiv(std::max<long long>(cdiv(p->npix, 256), 1), 8);  // see k_resolve
#define PIKO_RES_LAUNCH(KIND, MULTI)                                                                          \
  k_resolve<KIND, MULTI><<<(unsigned)rblocks, 256, 0, p->stream>>>(p->d_geo, p->d_attr, p->d_vis, p->fa, p->d_jobs, \
                                                                   nj, p->npix, p->d_color, p->d_depth, p->d_fs)
  if (p->kind == PIKO_INPUT_PRETRANSFORMED) {
    if (p->multi) PIKO_RES_LAUNCH(PIKO_INPUT_PRETRANSFORMED, true);
    else PIKO_RES_LAUNCH(PIKO_INPUT_PRETRANSFORMED, false);
  } else {
    if (p->multi) PIKO_RES_LAUNCH(PIKO_INPUT_WORLD, true);
    else PIKO_RES_LAUNCH(PIKO_INPUT_WORLD, false);
  }
#undef PIKO_RES_LAUNCH
  mark();
}

template <bool MULTI>
void launch_tile_raster(piko_raster_pipe *p, const RasterClass &c, const FrameArgs &fa) {
  const int nt = c.bin_hi - c.bin_lo;
  if (nt <= 0) return;
  const unsigned grid = (unsigned)nt;
  const int nj = (int)p->jobs.size();
#define PIKO_TR_LAUNCH(TT, KIND)                                                                                 \
  k_tile_raster<TT, KIND, MULTI><<<grid, kTRThreads, 0, p->stream>>>(p->d_geo, p->d_attr, fa, p->d_jobs, nj,    \
                                                                     p->d_tab, p->d_tpairs, c.bin_lo, c.bin_hi, \
                                                                     p->d_color, p->d_depth)
#define PIKO_TR_KIND(TT)                                                       \
  if (p->kind == PIKO_INPUT_PRETRANSFORMED) PIKO_TR_LAUNCH(TT, PIKO_INPUT_PRETRANSFORMED); \
  else PIKO_TR_LAUNCH(TT, PIKO_INPUT_WORLD)
  switch (c.T) {
    case 16: PIKO_TR_KIND(16); break;
    case 32: PIKO_TR_KIND(32); break;
    default: PIKO_TR_KIND(64); break;
  }
#undef PIKO_TR_KIND
#undef PIKO_TR_LAUNCH
}

// Tiled frame: chunk binning, then one raster launch per tile-size class.
void enqueue_tiled(piko_raster_pipe *p, Marks &mark) {
  const FrameArgs &fa = p->multi ? p->fa : p->fa_t;
  if (p->nchunks > 0) {
#define PIKO_TB_LAUNCH(KIND, MULTI)                                                                          \
  k_tile_bin<KIND, MULTI><<<p->nchunks, kBinThreads, 0, p->stream>>>(p->d_pos, fa, p->d_jobs, p->d_chunks, \
                                                                     p->d_tab, p->d_tpairs, p->d_chunk_keep,   \
                                                                     p->d_geo)
    if (p->kind == PIKO_INPUT_PRETRANSFORMED) {
      if (p->multi) PIKO_TB_LAUNCH(PIKO_INPUT_PRETRANSFORMED, true);
      else PIKO_TB_LAUNCH(PIKO_INPUT_PRETRANSFORMED, false);
    } else {
      if (p->multi) PIKO_TB_LAUNCH(PIKO_INPUT_WORLD, true);
      else PIKO_TB_LAUNCH(PIKO_INPUT_WORLD, false);
    }
#undef PIKO_TB_LAUNCH
  }
  mark();
  if (p->multi) {
    for (const RasterClass &c : p->classes) {
      launch_tile_raster<true>(p, c, fa);
      if (c.bin_hi > c.bin_lo) mark();
    }
  } else {
    launch_tile_raster<false>(p, p->tclass, fa);
    mark();
  }
}

// Persistent raster waves per CU with two frames in flight: three 256-thread
// workgroups of one frame's raster leave two slots per CU to the other
// frame's launches (2M height field with whole-tile items: 12 waves 81 us
// per frame, 16 85 us, 20 87 us, 8 93 us).
constexpr int kQFramesWpc = 12;

// The queue frame's launches: the bin launch, then (raster) the persistent
// raster launch over the part items and the merge of the split tiles.

// Workgroups of one k_qraster instance that a CU holds at once (registers
// and LDS), asked of the runtime once per instance.
template <int QT, int KIND, bool MULTI>
int q_resident() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_qraster<QT, KIND, MULTI>, kQThreads, 0) != hipSuccess)
      (void)hipGetLastError();
    return b > 0 ? b : 1 << 20;
  }();
  return n;
}

void launch_queue(piko_raster_pipe *p, bool raster, Marks *mark) {
  const bool fr = p->q_frames && !p->multi;
  const int nitems = fr ? p->q_nitems_f : p->q_nitems, nsplits = fr ? 0 : p->q_nsplits;
  QArgs qa{p->d_pos,    p->d_attr,  p->d_geo,    p->d_qpairs,   p->d_qbase, p->d_qcap,
           p->d_qfill,  p->d_qneed, fr ? p->d_qitems_f : p->d_qitems, p->d_qsplits,  p->d_qchunks, p->d_qscr,
           p->d_jobs,   p->d_qs,    p->d_fs,     p->d_color,    p->d_depth,
           raster ? nullptr : p->d_qcost,
           (int)p->jobs.size(), p->q_nchunks, nitems, nsplits};
  const FrameArgs &fa = p->fa_q;
  const unsigned bgrid = (unsigned)std::max(p->q_nchunks, 1);
#define PIKO_QB_LAUNCH(KIND)                                                                          \
  do {                                                                                                \
    if (p->multi && raster) k_qbin<KIND, true, false><<<bgrid, kQThreads, 0, p->stream>>>(qa, fa);   \
    else if (p->multi) k_qbin<KIND, true, true><<<bgrid, kQThreads, 0, p->stream>>>(qa, fa);         \
    else if (raster) k_qbin<KIND, false, false><<<bgrid, kQThreads, 0, p->stream>>>(qa, fa);         \
    else k_qbin<KIND, false, true><<<bgrid, kQThreads, 0, p->stream>>>(qa, fa);                      \
  } while (0)
  if (p->kind == PIKO_INPUT_PRETRANSFORMED) PIKO_QB_LAUNCH(PIKO_INPUT_PRETRANSFORMED);
  else PIKO_QB_LAUNCH(PIKO_INPUT_WORLD);
#undef PIKO_QB_LAUNCH
  if (mark) (*mark)();
  if (!raster) return;
  const long long per_cu = std::max(1, (fr ? kQFramesWpc : p->q_wpc) / (kQThreads / 64));
  // the persistent grid never exceeds what the CUs hold at once
  auto rgrid = [&](int resident) {
    return (int)std::max<long long>(
        1, std::min<long long>(nitems, (long long)p->cus * std::min<long long>(per_cu, resident)));
  };
#define PIKO_QR_LAUNCH(QT, KIND)                                                                    \
  do {                                                                                              \
    if (p->multi) k_qraster<QT, KIND, true><<<rgrid(q_resident<QT, KIND, true>()), kQThreads, 0, p->stream>>>(qa, fa); \
    else k_qraster<QT, KIND, false><<<rgrid(q_resident<QT, KIND, false>()), kQThreads, 0, p->stream>>>(qa, fa); \
    if (mark) (*mark)();                                                                            \
    if (nsplits > 0) {                                                                              \
      if (p->multi) k_qmerge<QT, KIND, true><<<nsplits, kQMergeThreads, 0, p->stream>>>(qa, fa);     \
      else k_qmerge<QT, KIND, false><<<nsplits, kQMergeThreads, 0, p->stream>>>(qa, fa);             \
      if (mark) (*mark)();                                                                          \
    }                                                                                               \
  } while (0)
  if (p->q_QT > 32) {
    if (p->kind == PIKO_INPUT_PRETRANSFORMED) PIKO_QR_LAUNCH(64, PIKO_INPUT_PRETRANSFORMED);
    else PIKO_QR_LAUNCH(64, PIKO_INPUT_WORLD);
  } else {
    if (p->kind == PIKO_INPUT_PRETRANSFORMED) PIKO_QR_LAUNCH(32, PIKO_INPUT_PRETRANSFORMED);
    else PIKO_QR_LAUNCH(32, PIKO_INPUT_WORLD);
  }
#undef PIKO_QR_LAUNCH
}

// First frame of a job set on the queue path: the bin items alone count each
// tile's entries (every run is refused, the ranges being empty), and the
// ranges are sized from those counts with some room for camera changes.
int q_count(piko_raster_pipe *p) {
  const int n = p->q_ntiles;
  if (n > 0) HIPOK(hipMemsetAsync(p->d_qcost, 0, (size_t)n * sizeof(unsigned long long), p->stream));
  launch_queue(p, false, nullptr);
  HIPOK(hipGetLastError());
  std::vector<uint32_t> cnt((size_t)std::max(n, 1), 0u);
  p->q_cost.assign((size_t)std::max(n, 1), 0ull);
  if (n > 0) {
    HIPOK(hipMemcpyAsync(cnt.data(), p->d_qfill, (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
    HIPOK(hipMemcpyAsync(p->q_cost.data(), p->d_qcost, (size_t)n * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, p->stream));
  }
  HIPOK(hipStreamSynchronize(p->stream));
  if (n > 0) HIPOK(hipMemsetAsync(p->d_qfill, 0, (size_t)n * 4, p->stream));
  HIPOK(hipMemsetAsync(&p->d_fs->overflow, 0, 4, p->stream));
  for (int t = 0; t < n; ++t) p->q_cap[t] = cnt[t] ? cnt[t] + cnt[t] / 8 + 32 : 0u;
  p->q_caps_known = true;
  return q_build(p);
}

// Queue frame: one persistent launch.
int enqueue_queue(piko_raster_pipe *p, Marks &mark) {
  if (!p->q_caps_known) {
    int rc = q_count(p);
    if (rc != PIKO_OK) return rc;
  }
  launch_queue(p, true, &mark);
  return PIKO_OK;
}

// Grows the tile queues that overflowed in the last queue frame, from the
// counts its parts recorded.
int q_regrow(piko_raster_pipe *p) {
  const int n = p->q_ntiles;
  std::vector<uint32_t> need((size_t)std::max(n, 1), 0u);
  if (n > 0) HIPOK(hipMemcpyAsync(need.data(), p->d_qneed, (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  for (int t = 0; t < n; ++t)
    if (need[t] > p->q_cap[t]) p->q_cap[t] = need[t] + need[t] / 8 + 32;
  return q_build(p);
}

// Enqueues one frame of the current job set.  ev (optional) receives an event
// after each kernel.
int enqueue_frame(piko_raster_pipe *p, hipEvent_t *ev, int *nev) {
  Marks mark{ev, p->stream};
  p->last_mode = p->mode;
  if (p->mode == PIKO_RASTER_QUEUE) {
    const int rc = enqueue_queue(p, mark);
    if (rc != PIKO_OK) return rc;
    p->bins_current = false;
  } else if (p->mode == PIKO_RASTER_TILED) {
    enqueue_tiled(p, mark);
    p->bins_current = false;
  } else if (p->mode == PIKO_RASTER_DIRECT) {
    enqueue_direct(p, mark);
    p->bins_current = false;
  } else {
    enqueue_binning(p, mark);
    for (const RasterClass &c : p->classes) {
      if (p->multi) launch_raster<true>(p, c.T, c.bin_lo, c.bin_hi, c.waves, p->d_vals[p->final_buf]);
      else launch_raster<false>(p, c.T, c.bin_lo, c.bin_hi, c.waves, p->d_vals[p->final_buf]);
      if (c.bin_hi > c.bin_lo) mark();
    }
    p->bins_current = true;
  }
  if (nev) *nev = mark.n;
  return hipGetLastError() == hipSuccess ? PIKO_OK : PIKO_ERR_HIP;
}

int read_frame_state(piko_raster_pipe *p, FrameState *fs) {
  HIPOK(hipMemcpyAsync(fs, p->d_fs, sizeof(FrameState), hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

// Reads the overflow flags; grows the pair queues or the large-triangle list
// when they overflowed.  *rerun = 1 if the frames must be re-run.
int check_capacity(piko_raster_pipe *p, int *rerun) {
  FrameState fs;
  int rc = read_frame_state(p, &fs);
  if (rc != PIKO_OK) return rc;
  *rerun = 0;
  if (fs.overflow & 4u) return PIKO_ERR_CAPACITY;
  const bool queue_short = (fs.overflow & 8u) != 0;
  if (queue_short && (rc = q_regrow(p)) != PIKO_OK) return rc;
  const bool pairs_short = (fs.overflow & 1u) || (p->bins_current && (long long)fs.total > p->cap);
  const bool large_short = (fs.overflow & 2u) != 0;
  if (pairs_short) {
    free_pairs(p);
    rc = alloc_pairs(p, (long long)fs.total + (long long)fs.total / 4 + 4096);
    if (rc != PIKO_OK) return rc;
  }
  if (large_short) {
    dfree(p->d_large);
    dfree(p->d_tile_off);
    p->tri_bytes -= 8ll * p->large_cap;
    p->device_bytes -= 8ll * p->large_cap;
    p->large_cap = (uint32_t)std::min<long long>((long long)fs.large_peak + fs.large_peak / 4 + 4096, 1 << 30);
    if ((rc = dalloc(p, &p->d_large, p->large_cap, &p->tri_bytes)) != PIKO_OK) return rc;
    if ((rc = dalloc(p, &p->d_tile_off, p->large_cap, &p->tri_bytes)) != PIKO_OK) return rc;
  }
  if (p->last_mode == PIKO_RASTER_DIRECT) p->large_tiles_hint = fs.tiles_peak;
  if (fs.overflow || fs.large_peak || fs.tiles_peak) {
    const uint32_t zero[3] = {0, 0, 0};
    HIPOK(hipMemcpyAsync(&p->d_fs->overflow, &zero[0], 4, hipMemcpyHostToDevice, p->stream));
    HIPOK(hipMemcpyAsync(&p->d_fs->large_peak, &zero[1], 8, hipMemcpyHostToDevice, p->stream));
    HIPOK(hipStreamSynchronize(p->stream));
  }
  *rerun = pairs_short || large_short || queue_short;
  return PIKO_OK;
}

// Builds the bin queues of the current frame when the last frame did not
// (a direct frame), growing them until they fit.
int ensure_bins(piko_raster_pipe *p) {
  if (p->bins_current) return PIKO_OK;
  for (int attempt = 0; attempt < 3; ++attempt) {
    Marks mark{nullptr, p->stream};
    enqueue_binning(p, mark);
    HIPOK(hipGetLastError());
    int rerun = 0;
    int rc = check_capacity(p, &rerun);
    if (rc != PIKO_OK) return rc;
    if (!rerun) {
      p->bins_current = true;
      return PIKO_OK;
    }
  }
  return PIKO_ERR_CAPACITY;
}

bool valid_bin(int T) { return T == 8 || T == 16 || T == 32 || T == 64; }

bool valid_screen(int W, int H) { return W > 0 && H > 0 && W <= 4095 && H <= 4095; }

int validate_opts(const piko_constant_state *cs, int count, const piko_raster_options *opt) {
  if (!cs || !opt || count < 0) return PIKO_ERR_INVALID;
  if (!valid_screen(cs->screenSizeX, cs->screenSizeY)) return PIKO_ERR_INVALID;
  if (!valid_bin(opt->bin_size)) return PIKO_ERR_INVALID;
  if (opt->subpixel_bits < 1 || opt->subpixel_bits > 8) return PIKO_ERR_INVALID;
  if (opt->input_kind != PIKO_INPUT_PRETRANSFORMED && opt->input_kind != PIKO_INPUT_WORLD)
    return PIKO_ERR_INVALID;
  if (opt->raster_mode < PIKO_RASTER_AUTO || opt->raster_mode > PIKO_RASTER_QUEUE) return PIKO_ERR_INVALID;
  if (opt->part_pairs < 0 || opt->part_pairs > (1 << 20) || opt->waves_per_cu < 0) return PIKO_ERR_INVALID;
  if (opt->queue_cover < PIKO_COVER_AUTO || opt->queue_cover > PIKO_COVER_SPANS) return PIKO_ERR_INVALID;
  return PIKO_OK;
}

FrameArgs make_args(const piko_constant_state *cs, int S, int T) {
  FrameArgs fa{};
  fa.W = cs->screenSizeX;
  fa.H = cs->screenSizeY;
  fa.S = S;
  fa.lgT = ilog2i(T);
  fa.nBx = (fa.W + T - 1) / T;
  fa.nBy = (fa.H + T - 1) / T;
  fa.halfW = cs->halfW;
  fa.halfH = cs->halfH;
  std::memcpy(fa.vp, cs->viewProjMatrix, sizeof(fa.vp));
  std::memcpy(fa.light, cs->light, sizeof(fa.light));
  std::memcpy(fa.matcol, cs->matcol, sizeof(fa.matcol));
  std::memcpy(fa.ambient, cs->ambient, sizeof(fa.ambient));
  fa.nbins = (uint32_t)(fa.nBx * fa.nBy);
  return fa;
}

int create_common(piko_raster_pipe *p, int kind, int S, long long tris, long long pix, int bins,
                  long long pair_capacity, int mode) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PIKO_ERR_NO_DEVICE;
  p->kind = kind;
  p->S = S;
  if (mode == PIKO_RASTER_AUTO) {
    // plain frames: direct (fewest passes over one scene); batches: tiled
    // (no counting pass when a new job set is submitted).  QUEUE is faster
    // for repeated frames of one job set (4.3 against 4.5 ms on the 256-scene
    // batch) but its first frame after a submit counts first (8.7 against
    // 4.7 ms)
    const char *m = std::getenv("PIKO_RASTER_MODE");
    mode = p->multi ? PIKO_RASTER_TILED : PIKO_RASTER_DIRECT;
    if (m && std::strcmp(m, "binned") == 0) mode = PIKO_RASTER_BINNED;
    if (m && std::strcmp(m, "direct") == 0) mode = PIKO_RASTER_DIRECT;
    if (m && std::strcmp(m, "tiled") == 0) mode = PIKO_RASTER_TILED;
    if (m && std::strcmp(m, "queue") == 0) mode = PIKO_RASTER_QUEUE;
  }
  p->mode = p->last_mode = mode;
  {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    p->cus = cus;
    p->raster_waves = cus * 20;  // ~20 resident waves per CU
    const char *m = std::getenv("PIKO_RASTER_PERSISTENT");
    if (m) p->raster_persistent = m[0] == '1';
  }
  HIPOK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  int rc;
  if ((rc = dalloc(p, &p->d_rowtot, 256)) != PIKO_OK) return rc;
  if ((rc = dalloc(p, &p->d_fs, 1)) != PIKO_OK) return rc;
  if ((rc = alloc_tris(p, tris)) != PIKO_OK) return rc;
  if ((rc = alloc_pix(p, pix)) != PIKO_OK) return rc;
  if ((rc = alloc_bins(p, bins)) != PIKO_OK) return rc;
  // (triangle, bin) queues: sized for the frame in the binned path; the
  // other paths build them only when asked for, growing them then
  const long long cap = pair_capacity > 0 ? pair_capacity
                                          : (mode == PIKO_RASTER_BINNED ? std::max<long long>(2ll * tris + 1024, 1 << 16)
                                                                        : (1 << 16));
  if ((rc = alloc_pairs(p, cap)) != PIKO_OK) return rc;
  HIPOK(hipMemsetAsync(p->d_fs, 0, sizeof(FrameState), p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

int upload_input(piko_raster_pipe *p, const float *pos, const void *attr, long long tris, int kind) {
  const size_t pb = (size_t)tris * 9 * sizeof(float);
  const size_t ab = (size_t)tris * (kind == PIKO_INPUT_WORLD ? 9 * sizeof(float) : 3 * sizeof(uint32_t));
  float *dp = nullptr;
  void *da = nullptr;
  p->owns_input = true;
  if (hipMalloc((void **)&dp, std::max<size_t>(pb, 4)) != hipSuccess) return PIKO_ERR_OOM;
  p->d_pos = dp;
  if (hipMalloc(&da, std::max<size_t>(ab, 4)) != hipSuccess) return PIKO_ERR_OOM;
  p->d_attr = da;
  p->device_bytes += (long long)(pb + ab);
  if (tris > 0 && (hipMemcpy(dp, pos, pb, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(da, attr, ab, hipMemcpyHostToDevice) != hipSuccess))
    return PIKO_ERR_HIP;
  return PIKO_OK;
}

int plain_pipe(piko_raster_pipe **out, const piko_constant_state *cs, const float *pos, const void *attr, int count,
               const piko_raster_options *opt, bool on_device) {
  if (!out) return PIKO_ERR_INVALID;
  *out = nullptr;
  int rc = validate_opts(cs, count, opt);
  if (rc != PIKO_OK) return rc;
  if (count > 0 && (!pos || !attr)) return PIKO_ERR_INVALID;
  piko_raster_pipe *p = new (std::nothrow) piko_raster_pipe();
  if (!p) return PIKO_ERR_OOM;
  p->W = cs->screenSizeX;
  p->H = cs->screenSizeY;
  p->T = opt->bin_size;
  p->fa = make_args(cs, opt->subpixel_bits, p->T);
  p->lgT = p->fa.lgT;
  p->nBx = p->fa.nBx;
  p->nBy = p->fa.nBy;
  p->ntris = count;
  p->input_tris = count;
  p->nbins = (int)p->fa.nbins;
  p->npix = (long long)p->W * p->H;
  p->passes = radix_passes(p->nbins);
  p->jobs.assign(1, p->fa);
  rc = create_common(p, opt->input_kind, opt->subpixel_bits, count, p->npix, p->nbins, opt->pair_capacity,
                     opt->raster_mode);
  p->classes.assign(1, RasterClass{p->T, 0, p->nbins, opt->waves_per_cu > 0 ? opt->waves_per_cu * p->cus : 0});
  if (rc == PIKO_OK && p->mode == PIKO_RASTER_QUEUE) {
    p->fa_q = make_args(cs, opt->subpixel_bits, queue_T(p->T, p->W, p->H));
    p->q_part_pairs = opt->part_pairs > 0 ? opt->part_pairs : 2048;
    p->q_wpc = opt->waves_per_cu > 0 ? opt->waves_per_cu : 20;
    p->q_walk = opt->queue_cover == PIKO_COVER_SPANS ? 0u : 1u;
    p->fa_q.qwalk = p->q_walk;
    rc = q_setup(p, &p->fa_q, 1, count);
  }
  if (rc == PIKO_OK && p->mode == PIKO_RASTER_TILED) {
    const int tt = tiled_T(p->T, p->W, p->H);
    p->fa_t = make_args(cs, opt->subpixel_bits, tt);
    p->tclass = RasterClass{tt, 0, (int)p->fa_t.nbins, 0};
    rc = setup_chunks(p, &p->fa_t, 1, count);
  }
  if (rc == PIKO_OK) {
    if (on_device) {
      p->d_pos = pos;
      p->d_attr = attr;
    } else {
      rc = upload_input(p, pos, attr, count, opt->input_kind);
    }
  }
  if (rc != PIKO_OK) {
    piko_raster_destroy(p);
    return rc;
  }
  p->cs_cur = *cs;
  p->opt0 = *opt;
  *out = p;
  return PIKO_OK;
}

// The twin of a plain queue pipe: the same options and uniforms over the
// same input (not copied), with its own frame buffers and stream.  Made once;
// where it cannot be (device memory), frames run one after another.
piko_raster_pipe *frame_twin(piko_raster_pipe *p) {
  if (p->twin) return p->twin;
  if (p->twin_tried || p->multi || p->mode != PIKO_RASTER_QUEUE || p->input_tris <= 0) return nullptr;
  p->twin_tried = true;
  piko_raster_pipe *t = nullptr;
  if (plain_pipe(&t, &p->cs_cur, p->d_pos, p->d_attr, p->input_tris, &p->opt0, true) != PIKO_OK) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (t->mode != p->mode) {
    piko_raster_destroy(t);
    return nullptr;
  }
  p->twin = t;
  return t;
}

// ---- allocation-free frame ---------------------------------------------------
//
// piko_raster_frame_noalloc renders one direct-path frame with every per-frame
// buffer in `scratch`, a file-scope device array of fixed size: the frame
// state, a 48-byte setup record per triangle, a visibility word per pixel and
// the large-triangle list.  Each device holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device.
constexpr long long kNaMaxTris = 1ll << 20;          // triangles per call
constexpr long long kNaMaxPix = 1ll << 21;           // pixels per call (1920 x 1080 fits)
constexpr uint32_t kNaMaxLarge = 1u << 16;           // large triangles per frame
constexpr size_t kNaGeoOff = 256;                    // after the frame state
constexpr size_t kNaVisOff = kNaGeoOff + (size_t)kNaMaxTris * 48;
constexpr size_t kNaLargeOff = kNaVisOff + (size_t)kNaMaxPix * 8;
constexpr size_t kNaTileOff = kNaLargeOff + (size_t)kNaMaxLarge * 4;
constexpr size_t kNaBytes = kNaTileOff + (size_t)kNaMaxLarge * 4;
static_assert(sizeof(FrameState) <= kNaGeoOff, "frame state fits its slot");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// The frame state zeroed and the frame's visibility words set to empty: the
// array keeps whatever the last call on the device left there.
__global__ void k_na_clear(FrameState *fs, unsigned long long *vis, long long npix) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) *fs = FrameState{};
  for (long long k = g; k < npix; k += (long long)gridDim.x * blockDim.x) vis[k] = kVisSentinel;
}

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
unsigned char *g_na_ptr[kNaMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
int na_scratch(unsigned char **out) {
  int dev = 0;
  HIPOK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) return PIKO_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *ptr = nullptr;
    HIPOK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return PIKO_ERR_HIP;
    g_na_ptr[dev] = (unsigned char *)ptr;
  }
  *out = g_na_ptr[dev];
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
    case PIKO_ERR_SCRATCH: return "larger than the fixed scratch of the allocation-free frame";
    default: return "unknown status";
  }
}

int piko_raster_destroy(piko_raster_pipe *p) {
  if (!p) return PIKO_ERR_INVALID;
  if (p->twin) piko_raster_destroy(p->twin);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  free_pairs(p);
  free_tris(p);
  free_pix(p);
  free_bins(p);
  free_chunks(p);
  free_queue(p);
  dfree(p->d_rowtot);
  dfree(p->d_fs);
  dfree(p->d_jobs);
  if (p->h_jobs) (void)hipHostFree(p->h_jobs);
  if (p->owns_input) {
    dfree(p->d_pos);
    dfree(p->d_attr);
  }
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return PIKO_OK;
}

int piko_raster_allocate(piko_raster_pipe **out, const piko_constant_state *cs, const float *pos,
                         const void *attr, int count, const piko_raster_options *opt) {
  return plain_pipe(out, cs, pos, attr, count, opt, false);
}

int piko_raster_allocate_device(piko_raster_pipe **out, const piko_constant_state *cs, const float *d_pos,
                                const void *d_attr, int count, const piko_raster_options *opt) {
  return plain_pipe(out, cs, d_pos, d_attr, count, opt, true);
}

int piko_raster_prepare(piko_raster_pipe *p, const piko_constant_state *cs) {
  if (!p || !cs || p->multi) return PIKO_ERR_INVALID;
  if (cs->screenSizeX != p->W || cs->screenSizeY != p->H) return PIKO_ERR_INVALID;
  p->fa = make_args(cs, p->S, p->T);
  p->jobs.assign(1, p->fa);
  if (p->mode == PIKO_RASTER_QUEUE) {
    p->fa_q = make_args(cs, p->S, 1 << p->fa_q.lgT);
    p->fa_q.qwalk = p->q_walk;
  }
  if (p->mode == PIKO_RASTER_TILED) {
    FrameArgs ft = make_args(cs, p->S, p->tclass.T);
    ft.chunk_base = p->fa_t.chunk_base;
    ft.nchunks = p->fa_t.nchunks;
    ft.tab_base = p->fa_t.tab_base;
    p->fa_t = ft;
  }
  p->cs_cur = *cs;
  if (p->twin) return piko_raster_prepare(p->twin, cs);
  return PIKO_OK;
}

// Frames of the current state.  On the queue path a run of two or more
// frames alternates between the pipe and its twin, each on its own stream and
// frame buffers, so two frames are in flight: one frame's bin launch runs
// beside the other's raster.  These frames take whole tiles as items and a
// persistent grid of kQFramesWpc waves per CU, which leaves CU slots to the
// other frame (2M-triangle height field: 117 us per frame one at a time, 93
// in flight with the single-frame items and grid, 80 with these).  The last
// frame's image ends in the pipe's buffers.
int piko_raster_run_frames(piko_raster_pipe *p, int nframes) {
  if (!p || nframes < 0) return PIKO_ERR_INVALID;
  piko_raster_pipe *q = nframes >= 2 ? frame_twin(p) : nullptr;
  for (int attempt = 0; attempt < 3; ++attempt) {
    p->q_frames = q != nullptr;
    if (q) q->q_frames = true;
    for (int f = 0; f < nframes; ++f) {
      int rc = enqueue_frame(q && (f & 1) ? q : p, nullptr, nullptr);
      if (rc != PIKO_OK) {
        p->q_frames = false;
        return rc;
      }
    }
    p->q_frames = false;  // single frames keep the split parts and the wider grid
    int rerun = 0, rerun_q = 0;
    int rc = check_capacity(p, &rerun);  // waits for the pipe's frames
    if (rc == PIKO_OK && q) rc = check_capacity(q, &rerun_q);  // and the twin's
    if (rc != PIKO_OK) return rc;
    if (rerun || rerun_q) continue;
    if (q && !(nframes & 1)) {  // the last frame ran on the twin
      HIPOK(hipMemcpyAsync(p->d_color, q->d_color, (size_t)p->npix * 4, hipMemcpyDeviceToDevice, p->stream));
      HIPOK(hipMemcpyAsync(p->d_depth, q->d_depth, (size_t)p->npix * 4, hipMemcpyDeviceToDevice, p->stream));
      HIPOK(hipStreamSynchronize(p->stream));
    }
    return PIKO_OK;
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
  const int n = std::min(ne, cap);
  for (int k = 0; k < n; ++k) {
    kernel_ms[k] = (float)(acc[k] / nframes);
    if (names) names[k] = kernel_name(p, k);
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
  if (!p || !color || p->multi) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(color, p->d_color, (size_t)p->npix * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

int piko_raster_frame_noalloc(const piko_constant_state *cs, const float *d_pos, const void *d_attr, int count,
                              int input_kind, int subpixel_bits, uint32_t *d_color, float *d_depth) {
  if (!cs || count < 0 || (count > 0 && (!d_pos || !d_attr)) || !d_color || !d_depth) return PIKO_ERR_INVALID;
  if (!valid_screen(cs->screenSizeX, cs->screenSizeY)) return PIKO_ERR_INVALID;
  if (subpixel_bits < 1 || subpixel_bits > 8) return PIKO_ERR_INVALID;
  if (input_kind != PIKO_INPUT_PRETRANSFORMED && input_kind != PIKO_INPUT_WORLD) return PIKO_ERR_INVALID;
  const long long npix = (long long)cs->screenSizeX * cs->screenSizeY;
  if ((long long)count > kNaMaxTris || npix > kNaMaxPix) return PIKO_ERR_SCRATCH;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PIKO_ERR_NO_DEVICE;
  unsigned char *base = nullptr;
  int rc = na_scratch(&base);
  if (rc != PIKO_OK) return rc;
  FrameState *fs = (FrameState *)base;
  int4 *geo = (int4 *)(base + kNaGeoOff);
  unsigned long long *vis = (unsigned long long *)(base + kNaVisOff);
  uint32_t *large = (uint32_t *)(base + kNaLargeOff);
  uint32_t *tile_off = (uint32_t *)(base + kNaTileOff);
  const FrameArgs fa = make_args(cs, subpixel_bits, 16);
  int dev = 0, cus = 256;
  HIPOK(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  hipStream_t st = nullptr;  // the null stream; the call returns when the frame is done
  k_na_clear<<<(unsigned)std::min<long long>(std::max<long long>(cdiv(npix, 256), 1), 2048), 256, 0, st>>>(fs, vis,
                                                                                                          npix);
  const int blocks = (int)std::max<long long>(cdiv(count, kDirThreads), 1);
  const int S = subpixel_bits;
  if (input_kind == PIKO_INPUT_PRETRANSFORMED) {
    k_vs_direct<PIKO_INPUT_PRETRANSFORMED, false><<<blocks, kDirThreads, 0, st>>>(
        d_pos, count, S, fa, nullptr, 0, geo, vis, large, tile_off, fs, kNaMaxLarge, 0);
  } else {
    k_vs_direct<PIKO_INPUT_WORLD, false><<<blocks, kDirThreads, 0, st>>>(d_pos, count, S, fa, nullptr, 0, geo, vis,
                                                                        large, tile_off, fs, kNaMaxLarge, 0);
  }
  k_large_raster<false><<<cus * 2, 256, 0, st>>>(geo, large, tile_off, fs, kNaMaxLarge, fa, nullptr, 0, vis);
  const unsigned rblocks = (unsigned)(8 * cdiv(std::max<long long>(cdiv(npix, 256), 1), 8));
  if (input_kind == PIKO_INPUT_PRETRANSFORMED)
    k_resolve<PIKO_INPUT_PRETRANSFORMED, false><<<rblocks, 256, 0, st>>>(geo, d_attr, vis, fa, nullptr, 0, npix,
                                                                         d_color, d_depth, fs);
  else
    k_resolve<PIKO_INPUT_WORLD, false><<<rblocks, 256, 0, st>>>(geo, d_attr, vis, fa, nullptr, 0, npix, d_color,
                                                                d_depth, fs);
  HIPOK(hipGetLastError());
  FrameState h{};
  HIPOK(hipMemcpy(&h, fs, sizeof(FrameState), hipMemcpyDeviceToHost));  // waits for the frame
  // the large-triangle list cannot grow here: a frame that overflowed it is reported
  if (h.overflow & 6u) return PIKO_ERR_SCRATCH;
  return PIKO_OK;
}

int piko_raster_get_depth(piko_raster_pipe *p, float *depth) {
  if (!p || !depth || p->multi) return PIKO_ERR_INVALID;
  HIPOK(hipMemcpyAsync(depth, p->d_depth, (size_t)p->npix * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

long long piko_raster_get_bins(piko_raster_pipe *p, int64_t *bin_start, int32_t *bin_tris, long long cap) {
  if (!p || !bin_start) return PIKO_ERR_INVALID;
  int rc = ensure_bins(p);
  if (rc != PIKO_OK) return rc;
  FrameState fs;
  rc = read_frame_state(p, &fs);
  if (rc != PIKO_OK) return rc;
  const long long P = std::min<long long>(fs.total, p->cap);
  uint32_t *keys = (uint32_t *)std::malloc((size_t)std::max<long long>(P, 1) * 4);
  if (!keys) return PIKO_ERR_OOM;
  if (P > 0 && (hipMemcpyAsync(keys, p->d_keys[p->final_buf], (size_t)P * 4, hipMemcpyDeviceToHost, p->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(p->stream) != hipSuccess)) {
    std::free(keys);
    return PIKO_ERR_HIP;
  }
  for (int b = 0; b <= p->nbins; ++b) bin_start[b] = 0;
  for (long long i = 0; i < P; ++i) bin_start[keys[i] + 1]++;
  for (int b = 0; b < p->nbins; ++b) bin_start[b + 1] += bin_start[b];
  std::free(keys);
  const long long n = std::min<long long>(P, cap);
  if (bin_tris && n > 0) {
    HIPOK(hipMemcpyAsync(bin_tris, p->d_vals[p->final_buf], (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
    HIPOK(hipStreamSynchronize(p->stream));
  }
  return (long long)fs.total;
}

int piko_raster_get_stats(piko_raster_pipe *p, piko_raster_stats *st) {
  if (!p || !st) return PIKO_ERR_INVALID;
  int rc = ensure_bins(p);
  if (rc != PIKO_OK) return rc;
  FrameState fs;
  rc = read_frame_state(p, &fs);
  if (rc != PIKO_OK) return rc;
  const int nkb = (int)std::max<long long>(cdiv(p->ntris, kVsTile), 1);
  uint32_t *kb = (uint32_t *)std::malloc((size_t)nkb * 4);
  if (!kb) return PIKO_ERR_OOM;
  if (hipMemcpyAsync(kb, p->d_keep, (size_t)nkb * 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess) {
    std::free(kb);
    return PIKO_ERR_HIP;
  }
  long long emitted = 0;
  for (int k = 0; k < nkb; ++k) emitted += kb[k];
  std::free(kb);
  st->triangles = p->ntris;
  st->emitted = p->ntris ? emitted : 0;
  st->pairs = fs.total;
  st->pair_capacity = p->cap;
  st->bins_x = p->multi ? 0 : p->nBx;
  st->bins_y = p->multi ? 0 : p->nBy;
  st->device_bytes = p->device_bytes;
  return PIKO_OK;
}

// ---- batches --------------------------------------------------------------

int piko_batch_create(piko_raster_pipe **out, const float *pos, const void *attr, long long input_tris,
                      int on_device, const piko_batch_options *opt) {
  if (!out || !opt || input_tris < 0) return PIKO_ERR_INVALID;
  *out = nullptr;
  if (opt->subpixel_bits < 1 || opt->subpixel_bits > 8) return PIKO_ERR_INVALID;
  if (opt->input_kind != PIKO_INPUT_PRETRANSFORMED && opt->input_kind != PIKO_INPUT_WORLD) return PIKO_ERR_INVALID;
  if (opt->max_jobs <= 0 || opt->max_tris < 0 || opt->max_pixels <= 0) return PIKO_ERR_INVALID;
  if (opt->raster_mode < PIKO_RASTER_AUTO || opt->raster_mode > PIKO_RASTER_QUEUE) return PIKO_ERR_INVALID;
  if (opt->part_pairs < 0 || opt->part_pairs > (1 << 20) || opt->waves_per_cu < 0) return PIKO_ERR_INVALID;
  if (opt->queue_cover < PIKO_COVER_AUTO || opt->queue_cover > PIKO_COVER_SPANS) return PIKO_ERR_INVALID;
  if (input_tris > 0 && (!pos || !attr)) return PIKO_ERR_INVALID;
  piko_raster_pipe *p = new (std::nothrow) piko_raster_pipe();
  if (!p) return PIKO_ERR_OOM;
  p->multi = true;
  p->input_tris = input_tris;
  p->cap_jobs = opt->max_jobs;
  // a batch's queue path: whole tiles as parts and a wider persistent grid
  // (the 256-scene batch: 4.05 ms per batch frame with 65536-entry parts,
  // 4.07 with 8192, 4.13 with 4096, 4.22 with 2048; a single 1080p frame
  // tunes to 2048 / 20)
  p->q_part_pairs = opt->part_pairs > 0 ? opt->part_pairs : 65536;
  p->q_wpc = opt->waves_per_cu > 0 ? opt->waves_per_cu : 24;
  p->q_walk = opt->queue_cover == PIKO_COVER_WALK ? 1u : 0u;  // AUTO: spans (the batched soups)
  const int bins = (int)std::min<long long>(opt->max_pixels / 64 + 1024, 1ll << 30);
  int rc = create_common(p, opt->input_kind, opt->subpixel_bits, opt->max_tris, opt->max_pixels, bins,
                         opt->pair_capacity, opt->raster_mode);
  if (rc == PIKO_OK) rc = dalloc(p, &p->d_jobs, (long long)p->cap_jobs);
  if (rc == PIKO_OK && hipHostMalloc((void **)&p->h_jobs, (size_t)p->cap_jobs * sizeof(FrameArgs)) != hipSuccess) {
    (void)hipGetLastError();  // reported as PIKO_ERR_OOM
    rc = PIKO_ERR_OOM;
  }
  if (rc == PIKO_OK) {
    if (on_device) {
      p->d_pos = pos;
      p->d_attr = attr;
    } else {
      rc = upload_input(p, pos, attr, input_tris, opt->input_kind);
    }
  }
  if (rc != PIKO_OK) {
    piko_raster_destroy(p);
    return rc;
  }
  *out = p;
  return PIKO_OK;
}

int piko_batch_wait(piko_raster_pipe *p) {
  if (!p || !p->multi) return PIKO_ERR_INVALID;
  if (!p->in_flight) return PIKO_OK;
  for (int attempt = 0; attempt < 3; ++attempt) {
    int rerun = 0;
    int rc = check_capacity(p, &rerun);
    if (rc != PIKO_OK) return rc;
    if (!rerun) {
      p->in_flight = false;
      return PIKO_OK;
    }
    rc = enqueue_frame(p, nullptr, nullptr);
    if (rc != PIKO_OK) return rc;
  }
  return PIKO_ERR_CAPACITY;
}

int piko_batch_query(piko_raster_pipe *p) {
  if (!p || !p->multi) return PIKO_ERR_INVALID;
  if (!p->in_flight) return 1;
  return hipStreamQuery(p->stream) == hipSuccess ? 1 : 0;
}

// Bin (tile) edge a job is rasterised with: its own in the binned and direct
// paths, the tiled path's tile edge otherwise.
static int job_T(const piko_raster_pipe *p, const piko_raster_job &jb) {
  if (p->mode == PIKO_RASTER_QUEUE) return queue_T(jb.bin_size, jb.cs.screenSizeX, jb.cs.screenSizeY);
  return p->mode == PIKO_RASTER_TILED ? tiled_T(jb.bin_size, jb.cs.screenSizeX, jb.cs.screenSizeY) : jb.bin_size;
}

int piko_batch_submit(piko_raster_pipe *p, const piko_raster_job *jobs, int njobs) {
  if (!p || !p->multi || !jobs || njobs <= 0 || njobs > p->cap_jobs) return PIKO_ERR_INVALID;
  int rc = piko_batch_wait(p);  // one batch in flight per pipe
  if (rc != PIKO_OK) return rc;
  long long tris = 0, pix = 0, bins = 0;
  for (int j = 0; j < njobs; ++j) {
    const piko_raster_job &jb = jobs[j];
    if (jb.count < 0 || jb.src_base < 0 || jb.src_base + jb.count > p->input_tris) return PIKO_ERR_INVALID;
    if (!valid_bin(jb.bin_size) || !valid_screen(jb.cs.screenSizeX, jb.cs.screenSizeY)) return PIKO_ERR_INVALID;
    if (jb.queue_cover < PIKO_COVER_AUTO || jb.queue_cover > PIKO_COVER_SPANS) return PIKO_ERR_INVALID;
    const int T = job_T(p, jb);
    tris += jb.count;
    pix += (long long)jb.cs.screenSizeX * jb.cs.screenSizeY;
    bins += (long long)((jb.cs.screenSizeX + T - 1) / T) * ((jb.cs.screenSizeY + T - 1) / T);
  }
  if (tris > 0x7FFFFFFFll || bins > 0x7FFFFFFFll) return PIKO_ERR_CAPACITY;
  // grow per-triangle, per-pixel and per-bin storage when this batch needs it
  HIPOK(hipStreamSynchronize(p->stream));
  if (tris > p->cap_tris) {
    free_tris(p);
    if ((rc = alloc_tris(p, tris)) != PIKO_OK) return rc;
  }
  if (pix > p->cap_pix) {
    free_pix(p);
    if ((rc = alloc_pix(p, pix)) != PIKO_OK) return rc;
  }
  if (bins > p->cap_bins) {
    free_bins(p);
    if ((rc = alloc_bins(p, (int)bins)) != PIKO_OK) return rc;
  }
  // Jobs are stored grouped by bin size (stable in submission order), so that
  // each raster launch covers one contiguous bin range and both the triangle
  // and the bin numbering increase with the stored index (the kernels find a
  // triangle's or a bin's job by binary search).
  p->jobs.clear();
  p->job_slot.assign(njobs, -1);
  p->classes.clear();
  uint32_t vbase = 0, bb = 0;
  unsigned long long fbb = 0;
  long long want_pairs = 0;
  for (int T : {8, 16, 32, 64}) {
    const uint32_t lo = bb;
    int wpc = 0;
    for (int j = 0; j < njobs; ++j) {
      if (job_T(p, jobs[j]) != T) continue;
      wpc = std::max(wpc, jobs[j].waves_per_cu);
      const double qd = jobs[j].queue_depth > 0.0f ? (double)jobs[j].queue_depth : 2.0;
      want_pairs += (long long)(qd * jobs[j].count) + 64;
      FrameArgs fa = make_args(&jobs[j].cs, p->S, T);
      fa.vbase = vbase;
      fa.src_base = (uint32_t)jobs[j].src_base;
      fa.bin_base = bb;
      fa.fb_base = fbb;
      fa.qwalk = jobs[j].queue_cover == PIKO_COVER_AUTO ? p->q_walk : (jobs[j].queue_cover == PIKO_COVER_WALK ? 1u : 0u);
      vbase += (uint32_t)jobs[j].count;
      bb += fa.nbins;
      fbb += (unsigned long long)fa.W * fa.H;
      p->job_slot[j] = (int)p->jobs.size();
      p->jobs.push_back(fa);
    }
    if (bb > lo) p->classes.push_back(RasterClass{T, (int)lo, (int)bb, wpc > 0 ? wpc * p->cus : 0});
  }
  if (want_pairs > p->cap && p->mode == PIKO_RASTER_BINNED) {  // preset the queues from the jobs' expected depth
    free_pairs(p);
    if ((rc = alloc_pairs(p, want_pairs)) != PIKO_OK) return rc;
  }
  p->ntris = (int)tris;
  p->nbins = (int)bb;
  p->npix = (long long)fbb;
  p->passes = radix_passes(p->nbins);
  if (p->mode == PIKO_RASTER_TILED && (rc = setup_chunks(p, p->jobs.data(), njobs, tris)) != PIKO_OK) return rc;
  if (p->mode == PIKO_RASTER_QUEUE && (rc = q_setup(p, p->jobs.data(), njobs, tris)) != PIKO_OK) return rc;
  std::memcpy(p->h_jobs, p->jobs.data(), (size_t)njobs * sizeof(FrameArgs));
  HIPOK(hipMemcpyAsync(p->d_jobs, p->h_jobs, (size_t)njobs * sizeof(FrameArgs), hipMemcpyHostToDevice, p->stream));
  rc = enqueue_frame(p, nullptr, nullptr);
  if (rc != PIKO_OK) return rc;
  p->in_flight = true;
  return PIKO_OK;
}

int piko_batch_get_color(piko_raster_pipe *p, int job, uint32_t *color) {
  if (!p || !p->multi || !color || job < 0 || job >= (int)p->job_slot.size()) return PIKO_ERR_INVALID;
  int rc = piko_batch_wait(p);
  if (rc != PIKO_OK) return rc;
  const FrameArgs &fa = p->jobs[p->job_slot[job]];
  HIPOK(hipMemcpyAsync(color, p->d_color + fa.fb_base, (size_t)fa.W * fa.H * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

int piko_batch_get_depth(piko_raster_pipe *p, int job, float *depth) {
  if (!p || !p->multi || !depth || job < 0 || job >= (int)p->job_slot.size()) return PIKO_ERR_INVALID;
  int rc = piko_batch_wait(p);
  if (rc != PIKO_OK) return rc;
  const FrameArgs &fa = p->jobs[p->job_slot[job]];
  HIPOK(hipMemcpyAsync(depth, p->d_depth + fa.fb_base, (size_t)fa.W * fa.H * 4, hipMemcpyDeviceToHost, p->stream));
  HIPOK(hipStreamSynchronize(p->stream));
  return PIKO_OK;
}

}  // extern "C"
