// raster_queue.h -- the queue path of the raster pipeline.  Included by
// piko_raster.hip inside its kernel namespace, after the helpers it uses
// (vertex stage, culls, trivial-reject test, spans, Gouraud resolve).
//
// A frame is three launches:
//   k_qbin     one workgroup per chunk of kQChunk consecutive input triangles
//              of one job: VertexShaderStage::process and its culls
//              (vertexShader.pikostage:36-109), then RasterStage::assignBin's
//              trivial-reject test per screen tile of the triangle's bbox
//              (raster.pikostage:40-117).  The chunk counts its (tile,
//              triangle) pairs in an LDS histogram over the job's tiles,
//              reserves one run per touched tile with a single global add on
//              the tile's fill counter, and scatters the triangle indices into
//              the runs through LDS cursors.  No global atomic per pair.
//   k_qraster  the raster stage as one persistent launch: its workgroups pull
//              part items from a device queue (one global counter), densest
//              tile first.  A tile whose queue holds more than part_pairs
//              entries is cut into parts of at most that many, so a dense tile
//              (a horizon band) does not set the length of the frame.  A part
//              rasterises its entries (RasterStage::process, raster.pikostage:
//              124-321) into 64-bit LDS visibility words (z bits, 0xFFFFFFFE -
//              triangle).  A tile of one part resolves from LDS: the winner of
//              each pixel is shaded and colour and depth are written once
//              (stage.h:57-61).  The parts of a split tile leave their words
//              in a scratch slot each.
//   k_qmerge   one workgroup per split tile: the minimum over its parts' slots,
//              then the same resolve.  Not launched when no tile is split.
// No workgroup waits for another inside a launch: on this GPU each XCD has
// its own L2, and an agent-scope acquire or release inside a kernel writes
// back or invalidates the whole L2, so every hand-over of data goes through a
// launch boundary.
//
// A tile's queue lives at a fixed range [base, base + cap) of the pair buffer,
// sized by the host from the counts of an earlier frame.  A run that does not
// fit is dropped and flagged; the fill counter keeps counting, the host reads
// the counts, grows the ranges and runs the frame again.  The frame returns
// its counters to zero itself: the part of a one-part tile, or the merge of a
// split tile, clears the tile's fill counter after recording it; the bin
// launch clears the queue head before the raster launch takes items.

constexpr int kQThreads = 256;                      // workgroup of the queue kernel
constexpr int kQPer = 4;                            // triangles per thread in a bin item
constexpr int kQChunk = kQThreads * kQPer;          // triangles per bin item
constexpr int kQMaxTiles = 4096;                    // tiles of one job (LDS histogram)
constexpr uint32_t kQNoSlot = 0x80000000u;          // a run that did not fit its tile's range

// Queue head of the raster launch.
struct QState {
  uint32_t next;        // next part item to hand out
  uint32_t pad[3];
};

// One part item: tile (batch tile numbering), part index, parts of the tile,
// first scratch slot of the tile's parts (split tiles only).  The split tiles
// are listed for k_qmerge as their first part's item.
struct QItem {
  uint32_t tile, part, nparts, slot;
  uint32_t base, cap, pad0, pad1;  // the tile's queue range (as tile_base / tile_cap)
};

struct QChunk {
  uint32_t job, tri0, count, pad;
};

struct QArgs {
  const float *pos;
  const void *attr;
  int4 *geo;
  uint32_t *pairs;
  const uint32_t *tile_base;
  const uint32_t *tile_cap;
  uint32_t *tile_fill;
  uint32_t *tile_need;
  const QItem *items;
  const uint32_t *splits;  // item index of the first part of each split tile
  const QChunk *chunks;
  unsigned long long *scratch;
  const FrameArgs *jobs;
  QState *qs;
  FrameState *fs;
  uint32_t *color;
  float *depth;
  int njobs, nchunks, nitems, part_pairs, nsplits;
};

// Wave-level span rasteriser of the queue path: span k (a column segment of
// kTileSpan pixels of staged triangle o's bbox clipped to the tile) into the
// tile's visibility words, row stride T.
__device__ __forceinline__ void q_span(const TileStage &s, int o, uint32_t k, int S, int xlo, int ylo, int T,
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
    // int32-path triangle: edge functions at the span's pixel centres in
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

// Inclusive prefix sum over the 64 lanes of a wave: DPP row shifts inside
// each row of 16 lanes, then row broadcasts of lanes 15 and 31 into the
// rows above (a lane without a source adds 0).  Every lane must be active.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15 -> rows 1, 3
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31 -> rows 2, 3
  return x;
}

__device__ __forceinline__ uint32_t wave_last(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Inclusive max-scan over the 64 lanes, with the shifts of wave_scan_dpp (a
// lane without a source takes -1, below every value scanned here).  Every
// lane must be active.
__device__ __forceinline__ int wave_max_scan_dpp(int x) {
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x118, 0xf, 0xf, false));  // row_shr:8
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x142, 0xa, 0xf, false));  // row_bcast:15 -> rows 1, 3
  x = max(x, __builtin_amdgcn_update_dpp(-1, x, 0x143, 0xc, 0xf, false));  // row_bcast:31 -> rows 2, 3
  return x;
}

// The wave's items are cut into consecutive runs, one per lane: lane i owns
// [start_i, start_i + len_i) (starts non-decreasing with i, empty runs
// allowed).  Returns the lane owning item w0 + lane, for every item of the
// window [w0, w0 + 64) below the total (beyond it, the last run's lane).
// Each run that reaches into the window marks its first item there in
// `mark` (64 words of the wave's own), and a max-scan carries each mark up
// to the next: two LDS round trips, where a binary search over the starts
// costs six dependent ones.  Every lane must be active.
__device__ __forceinline__ int wave_owner(uint32_t start, uint32_t len, uint32_t w0, uint32_t *mark) {
  const int lane = (int)(threadIdx.x & 63);
  mark[lane] = ~0u;
  wave_fence();
  if (len && start < w0 + 64u && start + len > w0) mark[(start > w0 ? start : w0) - w0] = (uint32_t)lane;
  wave_fence();
  return wave_max_scan_dpp((int)mark[lane]);
}

// Rows j in [0, h) of one bbox column where the edge function E + j*D is
// >= 0 (exact int32 values: an int32-path triangle inside its bbox), as
// [lo, hi] (empty: lo > hi).  The float quotient locates the boundary to
// within one row (its relative error is a few 2^-24 and the boundary is
// only looked for inside [-2, h + 2]); one exact test on each side fixes it.
__device__ __forceinline__ void q_edge_rows(int E, int D, int h, int &lo, int &hi) {
  if (D > 0) {  // increasing: j >= ceil(-E / D)
    float q = (float)(-E) * __builtin_amdgcn_rcpf((float)D);
    q = fminf(fmaxf(q, -1.0f), (float)(h + 1));
    int j = min(max((int)ceilf(q), 0), h);
    if (j < h && (int)((uint32_t)E + (uint32_t)j * (uint32_t)D) < 0) ++j;
    if (j > 0 && (int)((uint32_t)E + (uint32_t)(j - 1) * (uint32_t)D) >= 0) --j;
    lo = max(lo, j);
  } else if (D < 0) {  // decreasing: j <= floor(E / -D)
    float q = (float)E * __builtin_amdgcn_rcpf((float)(-D));
    q = fminf(fmaxf(q, -2.0f), (float)h);
    int j = min(max((int)floorf(q), -1), h - 1);
    if (j >= 0 && (int)((uint32_t)E + (uint32_t)j * (uint32_t)D) < 0) --j;
    if (j < h - 1 && (int)((uint32_t)E + (uint32_t)(j + 1) * (uint32_t)D) >= 0) ++j;
    hi = min(hi, j);
  } else if (E < 0) {
    lo = h;
  }
}

// One wave's bbox columns of int32-path triangles: per column, the first
// covered row's pixel and edge values and the exclusive prefix of the
// columns' covered-pixel counts.
struct QColumns {
  uint32_t pre[64];   // exclusive prefix of the columns' covered-pixel counts
  uint32_t info[64];  // staged triangle | px << 6 | first covered row << 18
  int e1[64], e2[64];
  uint32_t mark[64];  // wave_owner's marks
};

// Tiles of a kept triangle's clipped bbox that pass the trivial-reject test
// (the single tile of a bbox inside one tile needs no test).
__device__ __forceinline__ bool q_tile_ok(const Edges &e, int tx, int ty, bool one, int bsb) {
  return one || tr_accept(e, tx, ty, bsb);
}

// Resolve of one tile from its visibility words: depth = the winner's z,
// colour = Gouraud of the winner (the clear colour where nothing is drawn).
template <int KIND, bool MULTI>
__device__ __forceinline__ void q_resolve(const QArgs &qa, const FrameArgs &ja, const unsigned long long *vis,
                                          const float *s_u8, int xlo, int ylo, int xhi, int yhi) {
  const int T = 1 << ja.lgT, S = ja.S;
  const size_t fbb = MULTI ? (size_t)ja.fb_base : 0;
  constexpr int kBatch = 4;  // pixels per thread whose gathers are in flight together
  for (int q0 = 0; q0 < T * T; q0 += kBatch * kQThreads) {
    unsigned long long key[kBatch];
    int4 ga[kBatch], gb[kBatch];
    uint32_t c[kBatch][3];
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const int q = q0 + k * kQThreads + (int)threadIdx.x;
      key[k] = q < T * T ? vis[q] : kVisSentinel;
      ga[k] = gb[k] = make_int4(0, 0, 0, 0);
      c[k][0] = c[k][1] = c[k][2] = 0;
      if (key[k] != kVisSentinel) {
        const uint32_t v = 0xFFFFFFFEu - (uint32_t)(key[k] & 0xFFFFFFFFull);
        const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(v - ja.vbase) : (size_t)v;
        ga[k] = qa.geo[3 * (size_t)v];
        gb[k] = qa.geo[3 * (size_t)v + 1];
        tri_colours<KIND>(qa.attr, src, ja, c[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const int q = q0 + k * kQThreads + (int)threadIdx.x;
      if (q >= T * T) continue;
      const int px = xlo + (q & (T - 1)), py = ylo + (q >> ja.lgT);
      if (px >= xhi || py >= yhi) continue;
      uint32_t col = kClearColor;
      if (key[k] != kVisSentinel) {
        const int p[6] = {ga[k].x, ga[k].y, ga[k].z, ga[k].w, gb[k].x, gb[k].y};
        float alpha = 0.0f, beta = 0.0f;
        covered_at(p, small_bbox(p, S), px, py, S, __int_as_float(gb[k].z), alpha, beta);
        col = gouraud(s_u8, c[k][0], c[k][1], c[k][2], alpha, beta);
      }
      const size_t pid = fbb + (size_t)py * ja.W + px;
      qa.depth[pid] = __int_as_float((int)(key[k] >> 32));
      qa.color[pid] = col;
    }
  }
}

// Tile geometry of a batch tile: its job, and its pixel rectangle.
struct QTile {
  int lt, xlo, ylo, xhi, yhi;
};
__device__ __forceinline__ QTile q_tile(const FrameArgs &ja, int gt) {
  QTile r;
  const int T = 1 << ja.lgT;
  r.lt = gt - (int)ja.bin_base;
  const int tx = r.lt % ja.nBx, ty = r.lt / ja.nBx;
  r.xlo = tx * T;
  r.ylo = ty * T;
  r.xhi = min(r.xlo + T, ja.W);
  r.yhi = min(r.ylo + T, ja.H);
  return r;
}

// ---- k_qbin ----------------------------------------------------------------
template <int KIND, bool MULTI>
__global__ void __launch_bounds__(kQThreads) k_qbin(QArgs qa, FrameArgs fa) {
  __shared__ uint32_t cnt[kQMaxTiles];
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t == 0) qa.qs->next = 0;  // the raster launch's queue head
  if ((int)blockIdx.x >= qa.nchunks) return;       // a frame without triangles runs one block
  const QChunk cd = qa.chunks[blockIdx.x];
  const FrameArgs &ja = MULTI ? qa.jobs[cd.job] : fa;
  const int ntiles = ja.nBx * ja.nBy;
  const int S = ja.S, lgT = ja.lgT, bsb = S + lgT;
  for (int q = t; q < ntiles; q += kQThreads) cnt[q] = 0;
  // all of the thread's vertex positions first, so their loads overlap
  float vp[kQPer][9];
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    const uint32_t il = (uint32_t)(k * kQThreads + t);
    const uint32_t v = cd.tri0 + min(il, cd.count - 1u);
    const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(v - ja.vbase) : (size_t)v;
#pragma unroll
    for (int q = 0; q < 9; ++q) vp[k][q] = il < cd.count ? qa.pos[9 * src + q] : 0.0f;
  }
  __syncthreads();  // histogram cleared
  // per triangle: its tile rectangle (tx0 | ty0 << 16, tx1 | ty1 << 16), or
  // rt0 = ~0u when the triangle is in no tile's queue
  uint32_t rt0[kQPer], rt1[kQPer];
  int pk[kQPer][6];
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    rt0[k] = ~0u;
    rt1[k] = 0;
    const uint32_t il = (uint32_t)(k * kQThreads + t);
    float z[3];
    tri_positions<KIND>(vp[k], 0, ja, pk[k], z);
    if (il >= cd.count || !vs_keep(pk[k], ja.W, ja.H, S)) continue;
    int x0, y0, x1, y1;
    clip_rect(pk[k], S, ja.W, ja.H, x0, y0, x1, y1);
    if (x1 <= x0 || y1 <= y0) continue;
    const int tx0 = x0 >> lgT, ty0 = y0 >> lgT, tx1 = (x1 - 1) >> lgT, ty1 = (y1 - 1) >> lgT;
    Edges e;
    e.set(pk[k]);
    const bool one = tx0 == tx1 && ty0 == ty1;
    bool any = false;
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tx = tx0; tx <= tx1; ++tx)
        if (q_tile_ok(e, tx, ty, one, bsb)) {
          atomicAdd(&cnt[ty * ja.nBx + tx], 1u);
          any = true;
        }
    if (!any) continue;
    rt0[k] = (uint32_t)tx0 | ((uint32_t)ty0 << 16);
    rt1[k] = (uint32_t)tx1 | ((uint32_t)ty1 << 16);
    put_setup(qa.geo, (size_t)(cd.tri0 + il), pk[k], 1.0f / (float)e.denom(), z);  // raster.pikostage:174
  }
  __syncthreads();
  // one run per touched tile: a global add on the tile's fill counter
  for (int q = t; q < ntiles; q += kQThreads) {
    const uint32_t n = cnt[q];
    if (!n) continue;
    const uint32_t gt = (MULTI ? ja.bin_base : 0u) + (uint32_t)q;
    const uint32_t o = atomicAdd(&qa.tile_fill[gt], n);
    if ((unsigned long long)o + n <= qa.tile_cap[gt]) {
      cnt[q] = qa.tile_base[gt] + o;
    } else {
      cnt[q] = kQNoSlot;
      atomicOr(&qa.fs->overflow, 8u);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    if (rt0[k] == ~0u) continue;
    const uint32_t v = cd.tri0 + (uint32_t)(k * kQThreads + t);
    const int tx0 = (int)(rt0[k] & 0xffffu), ty0 = (int)(rt0[k] >> 16);
    const int tx1 = (int)(rt1[k] & 0xffffu), ty1 = (int)(rt1[k] >> 16);
    Edges e;
    e.set(pk[k]);
    const bool one = tx0 == tx1 && ty0 == ty1;
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tx = tx0; tx <= tx1; ++tx)
        if (q_tile_ok(e, tx, ty, one, bsb)) {
          const uint32_t s = atomicAdd(&cnt[ty * ja.nBx + tx], 1u);
          if (s < kQNoSlot) qa.pairs[s] = v;
        }
  }
}

// ---- k_qraster -------------------------------------------------------------
template <int QT>
struct QRasterShared {
  unsigned long long vis[QT * QT];
  TileStage st[kQThreads / 64];
  QColumns cols[kQThreads / 64];
  float u8[256];
  uint32_t item;
};

template <int QT, int KIND, bool MULTI>
__global__ void __launch_bounds__(kQThreads) k_qraster(QArgs qa, FrameArgs fa) {
  __shared__ QRasterShared<QT> sm;
  unsigned long long *const vis = sm.vis;
  const int t = threadIdx.x, lane = t & 63;
  for (int q = t; q < 256; q += kQThreads) sm.u8[q] = kU8Table.v[q];
  // The first item of a workgroup is its own index; later ones come from the
  // queue head, one global add each, taken only when the workgroup is free
  // (taking the next one early, while working on this one, was slower: a
  // busy workgroup held an item other workgroups could have started).
  bool first = true;
  for (;;) {
    __syncthreads();  // the previous item is done with the shared state
    if (t == 0) sm.item = first ? blockIdx.x : gridDim.x + atomicAdd(&qa.qs->next, 1u);
    first = false;
    __syncthreads();
    const uint32_t item = sm.item;
    if (item >= (uint32_t)qa.nitems) break;
    const QItem it = qa.items[item];
    const int gt = (int)it.tile;
    const FrameArgs &ja = MULTI ? qa.jobs[job_search(qa.jobs, qa.njobs, (uint32_t)gt, true)] : fa;
    const QTile tl = q_tile(ja, gt);
    const int T = 1 << ja.lgT, S = ja.S;
    for (int q = t; q < T * T; q += kQThreads) vis[q] = kVisSentinel;
    // the part's slice of the tile's range: its entries load beside the fill
    // count (a slot past the count holds a stale index and is dropped below)
    const uint32_t pbeg = min(it.cap, it.part * (uint32_t)qa.part_pairs);
    const uint32_t pend = min(it.cap, pbeg + (uint32_t)qa.part_pairs);
    const uint32_t *list = qa.pairs + it.base;
    const uint32_t fill = qa.tile_fill[gt];
    // Every load of the pipeline below is issued unconditionally (an index
    // out of range reads a valid word, the pair buffer's first or the setup
    // buffer's first record, and the value is then dropped).  A load under
    // a branch that a whole wave can skip leaves the compiler unable to count
    // the loads in flight, and it then waits for all of them, prefetches
    // included, at the first use of any.
    auto raw = [&](uint32_t j) -> uint32_t {
      const uint32_t x = *(j < pend ? list + j : qa.pairs);
      return j < pend ? x : ~0u;
    };
    uint32_t r_cur = raw(pbeg + (uint32_t)t), r_nxt = raw(pbeg + (uint32_t)(kQThreads + t));
    // a tile whose queue overflowed has runs that were refused and left
    // unwritten slots behind them: it draws nothing (the frame is re-run)
    const uint32_t nq = fill <= it.cap ? fill : 0u;
    const uint32_t beg = min(nq, pbeg);
    const uint32_t end = min(nq, pend);
    __syncthreads();  // visibility words initialised
    TileStage &s = sm.st[t >> 6];
    {
      auto entry = [&](uint32_t j) -> uint32_t {
        const uint32_t x = *(j < end ? list + j : qa.pairs);
        return j < end ? x : ~0u;
      };
      // the record of an absent triangle (~0u) is loaded from record 0 and
      // never read
      auto setup = [&](uint32_t v, int4 *out) {
        const int4 *g = qa.geo + 3 * (size_t)(v != ~0u ? v : 0u);
#pragma unroll
        for (int q = 0; q < 3; ++q) out[q] = g[q];
      };
      // software pipeline: the setup records of the next batch and the
      // entries of the one after are in flight while a batch is rasterised
      uint32_t v_cur = beg + (uint32_t)t < end ? r_cur : ~0u;
      int4 gc[3];
      setup(v_cur, gc);
      uint32_t v_nxt = beg + (uint32_t)(kQThreads + t) < end ? r_nxt : ~0u;
      for (uint32_t b0 = beg; b0 < end; b0 += kQThreads) {
        int4 gn[3];
        setup(v_nxt, gn);
        const uint32_t v_nn = entry(b0 + 2u * kQThreads + (uint32_t)t);
        // each staged triangle goes by bbox columns or by 8-row spans
        uint32_t ucol = 0, uspan = 0;
        if (v_cur != ~0u) {
          const int p[6] = {gc[0].x, gc[0].y, gc[0].z, gc[0].w, gc[1].x, gc[1].y};
          int x0, y0, x1, y1;
          clip_rect(p, S, ja.W, ja.H, x0, y0, x1, y1);
          x0 = max(x0, tl.xlo);
          y0 = max(y0, tl.ylo);
          x1 = min(x1, tl.xhi);
          y1 = min(y1, tl.yhi);
          const int w = max(0, x1 - x0), h = max(0, y1 - y0);
          if (w > 0 && h > 0) {
            const bool small = small_bbox(p, S);
            const uint32_t spans = (uint32_t)w * (uint32_t)((h + kTileSpan - 1) / kTileSpan);
            // int32-path triangles take the cheaper of the two: 8-row spans
            // test every bbox pixel, columns find each column's covered rows
            // and then touch only covered pixels.  Measured per unit, a column
            // costs about as much as a span and a covered pixel of the column
            // path about 0.45 of one (cost weights 2 : 1 : 0.45 mis-sent the
            // small soup triangles of the 256-scene batch to columns: 4.96 ms
            // against 4.31 ms per batch).  Tall slivers, whose bbox is mostly
            // empty, go by columns; compact triangles by spans.  The area in
            // pixels is |2 x signed area| / 2^(2S+1) of the fixed-point edges.
            bool by_col = false;
            if (small) {
              const float area = fabsf(__int_as_float(gc[1].z)) > 0.0f
                                     ? 0.5f / (fabsf(__int_as_float(gc[1].z)) * (float)(1 << (2 * S)))
                                     : 0.0f;
              const float area_clip = fminf(area, (float)(w * h));
              by_col = (float)spans > (float)w + 0.45f * area_clip;
            }
            if (by_col) ucol = (uint32_t)w;
            else uspan = spans;
            s.xy01[lane] = make_int4(p[0], p[1], p[2], p[3]);
            s.xy2z[lane] = make_int4(p[4], p[5], gc[1].w, gc[2].x);
            s.misc[lane] = make_uint4((uint32_t)gc[2].y, (uint32_t)gc[1].z, (uint32_t)x0 | ((uint32_t)y0 << 16),
                                      (uint32_t)w | ((uint32_t)h << 16) | (small ? 0x80000000u : 0u));
            s.tri[lane] = v_cur;
          }
        }
        v_cur = v_nxt;
#pragma unroll
        for (int q = 0; q < 3; ++q) gc[q] = gn[q];
        v_nxt = v_nn;
        // 8-row spans: every bbox pixel tested (int64-path triangles always)
        const uint32_t sinc = wave_scan_dpp(uspan);
        const uint32_t stot = wave_last(sinc);
        QColumns &cl = sm.cols[t >> 6];
        if (stot) {
          s.pre[lane] = sinc - uspan;
          for (uint32_t c0 = 0; c0 < stot; c0 += 64) {
            const uint32_t cidx = c0 + (uint32_t)lane;
            const int o = wave_owner(sinc - uspan, uspan, c0, cl.mark);  // fences s.pre too
            if (cidx < stot) q_span(s, o, cidx - s.pre[o], S, tl.xlo, tl.ylo, T, vis);
          }
          wave_fence();
        }
        // columns (int32-path triangles): each lane takes a bbox column and
        // finds its covered rows exactly, then the wave's covered pixels are
        // spread evenly over the lanes
        const uint32_t cinc = wave_scan_dpp(ucol);
        const uint32_t ctot = wave_last(cinc);
        const uint32_t cpre = cinc - ucol;
        s.pre[lane] = cpre;
        const uint32_t half = 1u << (S - 1);
        for (uint32_t c0 = 0; c0 < ctot; c0 += 64) {
          const uint32_t cidx = c0 + (uint32_t)lane;
          const int o = wave_owner(cpre, ucol, c0, cl.mark);  // fences s.pre too
          uint32_t cnt = 0;
          if (cidx < ctot) {
            const int4 a = s.xy01[o];
            const int4 b = s.xy2z[o];
            const uint4 m = s.misc[o];
            const int px = (int)(m.z & 0xffffu) + (int)(cidx - s.pre[o]);
            const int py0 = (int)(m.z >> 16), h = (int)((m.w >> 16) & 0x7fffu);
            const uint32_t cx = ((uint32_t)px << S) + half, cy = ((uint32_t)py0 << S) + half;
            const int xy[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
            int E[3], D[3];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {  // raster.pikostage:176-180, 223-241
              const uint32_t xa = (uint32_t)xy[2 * kk], ya = (uint32_t)xy[2 * kk + 1];
              const uint32_t xb = (uint32_t)xy[(2 * kk + 2) % 6], yb = (uint32_t)xy[(2 * kk + 3) % 6];
              E[kk] = (int)((cy - ya) * (xb - xa) - (cx - xa) * (yb - ya));
              D[kk] = (int)((xb - xa) << S);
            }
            int lo = 0, hi = h - 1;
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) q_edge_rows(E[kk], D[kk], h, lo, hi);
            if (lo <= hi) {
              cnt = (uint32_t)(hi - lo + 1);
              cl.info[lane] = (uint32_t)o | ((uint32_t)px << 6) | ((uint32_t)(py0 + lo) << 18);
              cl.e1[lane] = (int)((uint32_t)E[1] + (uint32_t)lo * (uint32_t)D[1]);
              cl.e2[lane] = (int)((uint32_t)E[2] + (uint32_t)lo * (uint32_t)D[2]);
            }
          }
          const uint32_t finc = wave_scan_dpp(cnt);
          const uint32_t ftot = wave_last(finc);
          const uint32_t fpre = finc - cnt;
          cl.pre[lane] = fpre;
          for (uint32_t f0 = 0; f0 < ftot; f0 += 64) {
            const uint32_t f = f0 + (uint32_t)lane;
            const int c = wave_owner(fpre, cnt, f0, cl.mark);  // fences cl.pre, info, e1, e2 too
            if (f < ftot) {
              const uint32_t r = f - cl.pre[c];
              const uint32_t info = cl.info[c];
              const int oc = (int)(info & 63u);
              const int px = (int)((info >> 6) & 0xfffu), py = (int)(info >> 18) + (int)r;
              const int4 a = s.xy01[oc];
              const int4 b = s.xy2z[oc];
              const uint4 m = s.misc[oc];
              const int e1 = (int)((uint32_t)cl.e1[c] + r * (((uint32_t)b.x - (uint32_t)a.z) << S));
              const int e2 = (int)((uint32_t)cl.e2[c] + r * (((uint32_t)a.x - (uint32_t)b.x) << S));
              const float rcp = __uint_as_float(m.y);
              const float alpha = (float)e1 * rcp, beta = (float)e2 * rcp;  // raster.pikostage:269-270
              const float z = interp(__int_as_float(b.w), __uint_as_float(m.x), __int_as_float(b.z), alpha,
                                     beta);  // raster.pikostage:272
              atomicMin((long long *)(vis + (py - tl.ylo) * T + (px - tl.xlo)),
                        (long long)(((unsigned long long)__float_as_uint(z) << 32) |
                                    (unsigned long long)(0xFFFFFFFEu - s.tri[oc])));
            }
          }
          wave_fence();  // the columns are read before the next window writes them
        }
        wave_fence();  // the staged triangles are read before the next batch stages
      }
    }
    __syncthreads();  // every fragment of the part is in the visibility words
    if (it.nparts > 1) {
      // split tile: this part's words to its scratch slot, merged by k_qmerge
      unsigned long long *mine = qa.scratch + (size_t)(it.slot + it.part) * (size_t)(QT * QT);
      for (int q = t; q < T * T; q += kQThreads) mine[q] = vis[q];
      continue;
    }
    if (t == 0) {
      qa.tile_need[gt] = fill;
      qa.tile_fill[gt] = 0;  // the tile's only reader is done with it
    }
    q_resolve<KIND, MULTI>(qa, ja, vis, sm.u8, tl.xlo, tl.ylo, tl.xhi, tl.yhi);
  }
}

// ---- k_qmerge --------------------------------------------------------------
template <int QT, int KIND, bool MULTI>
__global__ void __launch_bounds__(kQThreads) k_qmerge(QArgs qa, FrameArgs fa) {
  __shared__ unsigned long long vis[QT * QT];
  __shared__ float s_u8[256];
  const int t = threadIdx.x;
  for (int q = t; q < 256; q += kQThreads) s_u8[q] = kU8Table.v[q];
  const QItem it = qa.items[qa.splits[blockIdx.x]];
  const int gt = (int)it.tile;
  const FrameArgs &ja = MULTI ? qa.jobs[job_search(qa.jobs, qa.njobs, (uint32_t)gt, true)] : fa;
  const QTile tl = q_tile(ja, gt);
  const int T = 1 << ja.lgT;
  // the minimum over the parts' words: a thread's kPix pixels of two parts at
  // a time, all their loads issued before any is used (a loop over one part
  // word at a time waited for each load in turn; four parts at a time were
  // no faster)
  constexpr int kPix = 4;
  const unsigned long long *slots = qa.scratch + (size_t)it.slot * (size_t)(QT * QT);
  for (int q0 = 0; q0 < T * T; q0 += kPix * kQThreads) {
    long long m[kPix], a[kPix], b[kPix];
#pragma unroll
    for (int k = 0; k < kPix; ++k) m[k] = (long long)kVisSentinel;
    for (uint32_t j = 0; j < it.nparts; j += 2) {
      const bool two = j + 1 < it.nparts;
#pragma unroll
      for (int k = 0; k < kPix; ++k) {
        const int q = min(q0 + k * kQThreads + t, T * T - 1);
        a[k] = (long long)slots[(size_t)j * (size_t)(QT * QT) + q];
        b[k] = (long long)slots[(size_t)(two ? j + 1 : j) * (size_t)(QT * QT) + q];
      }
#pragma unroll
      for (int k = 0; k < kPix; ++k) m[k] = min(m[k], min(a[k], b[k]));
    }
#pragma unroll
    for (int k = 0; k < kPix; ++k) {
      const int q = q0 + k * kQThreads + t;
      if (q < T * T) vis[q] = (unsigned long long)m[k];
    }
  }
  if (t == 0) {
    qa.tile_need[gt] = qa.tile_fill[gt];
    qa.tile_fill[gt] = 0;
  }
  __syncthreads();
  q_resolve<KIND, MULTI>(qa, ja, vis, s_u8, tl.xlo, tl.ylo, tl.xhi, tl.yhi);
}
