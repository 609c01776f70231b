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

// Queue heads of the raster launch: the part items are dealt round-robin to
// kQQueues queues, one per XCD (workgroup b runs on XCD b % 8), each head in
// a line of its own.  With one head for the whole launch, the workgroups'
// global adds serialised at one address: a launch drawing nothing (the items'
// loads, barriers and hand-out alone) took 41 us on the 2M height field, 19
// us with a head per XCD.
constexpr int kQQueues = 8;
constexpr int kQHeadStride = 64;  // words between heads (256 bytes)
struct QState {
  uint32_t next[kQQueues * kQHeadStride];  // next position of each queue to hand out
};

// One part item: tile (batch tile numbering), part index, parts of the tile,
// first scratch slot of the tile's parts (split tiles only), the tile's queue
// range (as tile_base / tile_cap) and the part's slice [lo, hi) of it.  The
// split tiles are listed for k_qmerge as their first part's item.
struct QItem {
  uint32_t tile, part, nparts, slot;
  uint32_t base, cap, lo, hi;
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
  unsigned long long *tile_cost;  // the counting pass's output: each tile's estimated raster cost
  int njobs, nchunks, nitems, nsplits;
};

// Staged triangles of one wave of the queue raster (structure of arrays, so
// lanes reading neighbouring triangles hit neighbouring LDS banks).
//   fa, fb   what a fragment needs: the row steps of edge functions 1 and 2,
//            1/barydenom, z2; z0 - z2, z1 - z2 and the key's low word
//            0xFFFFFFFE - triangle (raster.pikostage:173-174, 269-272)
//   xy01, xr the fixed-point vertices and the bbox clipped to the tile
//            (x2, y2, x | y << 16, w | h << 16 | int32 edges << 31), read by
//            the span path only, which runs while the ring is empty (so the
//            two share their memory)
//   pre      the span path's nonempty runs in lane order: first span << 6 |
//            lane
//   rec      the column ring: per nonempty bbox column whose fragments are not
//            all drawn, its first fragment's number in the wave's running
//            count, slot | column << 6 | first covered row << 12 (tile-local),
//            and edge functions 1 and 2 at that pixel
//   mark     the marks of a window of 64 spans or fragments, tagged with
//            the window's sequence number
constexpr uint32_t kQRing = 128;
struct QStage {
  uint4 fa[64];
  uint4 fb[64];
  union {
    struct {
      int4 xy01[64];
      int4 xr[64];
    } sp;
    uint4 rec[kQRing];
  };
  uint32_t pre[64];
  uint32_t mark[64];
};

// Wave-level span rasteriser of the queue path: span k (a column segment of
// kTileSpan pixels of staged triangle o's bbox clipped to the tile) into the
// tile's visibility words, row stride T.  Every bbox pixel of the span is
// tested.  Used for the triangles the column walk does not take (int64 edge
// functions, or an edge wider than the 24-bit products allow).
__device__ __forceinline__ void q_span(const QStage &s, int o, uint32_t k, int S, int xlo, int ylo, int T,
                                       unsigned long long *vis) {
  const int4 a = s.sp.xy01[o];
  const int4 b = s.sp.xr[o];
  const uint4 fa = s.fa[o];
  const uint4 fb = s.fb[o];
  const uint32_t w = (uint32_t)b.w & 0xffffu, h = ((uint32_t)b.w >> 16) & 0x7fffu;
  // band and column of span k (the quotient truncates exactly: see k_vs_direct)
  const uint32_t q = (uint32_t)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)w));
  const uint32_t sx = k - q * w, sy = q * kTileSpan;
  const int px = (int)((uint32_t)b.z & 0xffffu) + (int)sx, py0 = (int)((uint32_t)b.z >> 16) + (int)sy;
  const uint32_t nv = min((uint32_t)kTileSpan, h - sy);
  const float rcp = __uint_as_float(fa.z);
  const float z2 = __uint_as_float(fa.w), z0mz2 = __uint_as_float(fb.x), z1mz2 = __uint_as_float(fb.y);
  const unsigned long long low = (unsigned long long)fb.z;
  unsigned long long *col = vis + (py0 - ylo) * T + (px - xlo);
  const int xy[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
  if ((uint32_t)b.w >> 31) {
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

// floor(n / d) and its remainder n - d * floor(n / d) in [0, d), exact, for
// 0 < d < 2^30 and |n| < 2^31.  The float quotient (relative error at most
// 2^-22: the conversion, the reciprocal and the product) errs by at most
// |n / d| * 2^-22 + 1, so the remainder of the estimate lies within
// |n| * 2^-22 + 2d < 2^31 of zero and wrapping 32-bit arithmetic gives it
// exactly; each loop step moves it by d.  With d >= 2^8 (8 subpixel bits) the
// estimate is off by at most 2.
__device__ __forceinline__ int q_floordiv(int n, int d, float rd, int &rem) {
  int q = (int)floorf((float)n * rd);
  int r = (int)((uint32_t)n - (uint32_t)q * (uint32_t)d);
  while (r < 0) {
    --q;
    r += d;
  }
  while (r >= d) {
    ++q;
    r -= d;
  }
  rem = r;
  return q;
}

// One edge of a column walk.  Along a bbox column the edge function is
// E(c) + j * D (row j, column c, E(c + 1) = E(c) + A).  With F = floor(E(c) /
// |D|) the rows where it is >= 0 are j >= -F (D > 0) or j <= F (D < 0).  F and
// its remainder R step from column to column by A = Q * |D| + Rr: F += Q, R +=
// Rr, and a carry when R reaches |D| (Bresenham's stepping, all integer).
struct QEdgeWalk {
  int F, R, Q, Rr, Dabs;
  __device__ __forceinline__ void init(int E, int A, int D) {
    Dabs = D < 0 ? -D : D;
    const float rd = __builtin_amdgcn_rcpf((float)Dabs);
    F = q_floordiv(E, Dabs, rd, R);
    Q = q_floordiv(A, Dabs, rd, Rr);
  }
  __device__ __forceinline__ void step() {
    R += Rr;
    F += Q;
    const bool c = R >= Dabs;
    R = c ? R - Dabs : R;
    F += c ? 1 : 0;
  }
};

// Tiles of a kept triangle's clipped bbox that pass the trivial-reject test
// (the single tile of a bbox inside one tile needs no test).
__device__ __forceinline__ bool q_tile_ok(const Edges &e, int tx, int ty, bool one, int bsb) {
  return one || tr_accept(e, tx, ty, bsb);
}

// Resolve of one tile from its visibility words by a workgroup of NT threads:
// depth = the winner's z, colour = Gouraud of the winner (the clear colour
// where nothing is drawn).
template <int KIND, bool MULTI, int NT = kQThreads>
__device__ __forceinline__ void q_resolve(const QArgs &qa, const FrameArgs &ja, const unsigned long long *vis,
                                          const float *s_u8, int xlo, int ylo, int xhi, int yhi) {
  const int T = 1 << ja.lgT, S = ja.S;
  const size_t fbb = MULTI ? (size_t)ja.fb_base : 0;
  constexpr int kBatch = NT >= 1024 ? 1 : 4;  // pixels per thread whose gathers are in flight together
  for (int q0 = 0; q0 < T * T; q0 += kBatch * NT) {
    unsigned long long key[kBatch];
    int4 ga[kBatch], gb[kBatch];
    uint32_t c[kBatch][3];
#pragma unroll
    for (int k = 0; k < kBatch; ++k) {
      const int q = q0 + k * NT + (int)threadIdx.x;
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
      const int q = q0 + k * NT + (int)threadIdx.x;
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
// A vertex position as stored in the input (three floats, 4-byte aligned).
struct __attribute__((packed, aligned(4))) QVec3 {
  float x, y, z;
};

// COUNT: the counting pass (the ranges are empty, so every run is refused),
// which also sums each tile's estimated raster cost into qa.tile_cost.
template <int KIND, bool MULTI, bool COUNT>
__global__ void __launch_bounds__(kQThreads) k_qbin(QArgs qa, FrameArgs fa) {
  __shared__ uint32_t cnt[kQMaxTiles];
  __shared__ uint32_t cst[COUNT ? kQMaxTiles : 1];  // the chunk's raster cost estimate per tile
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t < kQQueues) qa.qs->next[t * kQHeadStride] = 0;  // the raster launch's queue heads
  if ((int)blockIdx.x >= qa.nchunks) return;       // a frame without triangles runs one block
  const QChunk cd = qa.chunks[blockIdx.x];
  const FrameArgs &ja = MULTI ? qa.jobs[cd.job] : fa;
  const int ntiles = ja.nBx * ja.nBy;
  const int S = ja.S, lgT = ja.lgT, bsb = S + lgT;
  constexpr bool costs = COUNT;
  for (int q = t; q < ntiles; q += kQThreads) {
    cnt[q] = 0;
    if (costs) cst[q] = 0;
  }
  // all of the thread's vertex positions first, so their loads overlap: one
  // 12-byte load per vertex where nine 4-byte loads touched the same cache
  // lines nine times over (a triangle's floats are only 4-byte aligned)
  float vp[kQPer][9];
#pragma unroll
  for (int k = 0; k < kQPer; ++k) {
    const uint32_t il = (uint32_t)(k * kQThreads + t);
    const uint32_t v = cd.tri0 + min(il, cd.count - 1u);
    const size_t src = MULTI ? (size_t)ja.src_base + (size_t)(v - ja.vbase) : (size_t)v;
    const QVec3 *p3 = reinterpret_cast<const QVec3 *>(qa.pos + 9 * src);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const QVec3 a = p3[q];
      vp[k][3 * q] = il < cd.count ? a.x : 0.0f;
      vp[k][3 * q + 1] = il < cd.count ? a.y : 0.0f;
      vp[k][3 * q + 2] = il < cd.count ? a.z : 0.0f;
    }
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
          if (costs) {
            // a part's time per entry: a fixed share plus its bbox columns
            // and pixels in the tile (fragments ~ a fraction of the pixels),
            // in units of about a quarter fragment
            const int wx = min(x1, (tx + 1) << lgT) - max(x0, tx << lgT);
            const int hy = min(y1, (ty + 1) << lgT) - max(y0, ty << lgT);
            atomicAdd(&cst[ty * ja.nBx + tx], (uint32_t)(32 + 2 * wx + wx * hy));
          }
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
    if (costs) atomicAdd(&qa.tile_cost[gt], (unsigned long long)cst[q]);
    const uint32_t o = atomicAdd(&qa.tile_fill[gt], n);
    if ((unsigned long long)o + n <= qa.tile_cap[gt]) {
      cnt[q] = qa.tile_base[gt] + o;
    } else {
      cnt[q] = kQNoSlot;
      atomicOr(&qa.fs->overflow, 8u);
    }
  }
  if (costs) return;  // the counting pass: every run was refused
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
  QStage st[kQThreads / 64];
  float u8[256];
  uint32_t item;
};

template <int QT, int KIND, bool MULTI>
__global__ void __launch_bounds__(kQThreads) k_qraster(QArgs qa, FrameArgs fa) {
  __shared__ QRasterShared<QT> sm;
  unsigned long long *const vis = sm.vis;
  const int t = threadIdx.x, lane = t & 63;
  for (int q = t; q < 256; q += kQThreads) sm.u8[q] = kU8Table.v[q];
  // Windows of spans and of column-walk fragments tag their marks with a
  // sequence number, never ~0u.
  sm.st[t >> 6].mark[lane] = ~0u;
  uint32_t wseq = 0;
  // The first item of a workgroup is its own index; later ones come from its
  // XCD's queue (item i is position i / nqs of queue i % nqs), one global add
  // each, taken only when the workgroup is free.  Taking the next one earlier
  // (at the start of an item, or at its last batch) was slower: a busy
  // workgroup held an item other workgroups could have started, and the
  // global add's return held up every load the wave issued after it.  Every
  // queue has a workgroup (nqs is at most the grid), and its first positions
  // are the static items of its workgroups.
  const uint32_t nqs = min((uint32_t)kQQueues, gridDim.x);
  const uint32_t myq = blockIdx.x % nqs;
  const uint32_t statics = (gridDim.x - myq + nqs - 1) / nqs;
  bool first = true;
  for (;;) {
    __syncthreads();  // the previous item is done with the shared state
    if (t == 0)
      sm.item = first ? blockIdx.x : (statics + atomicAdd(&qa.qs->next[myq * kQHeadStride], 1u)) * nqs + myq;
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
    const uint32_t pbeg = min(it.cap, it.lo);
    const uint32_t pend = min(it.cap, it.hi);
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
    QStage &s = sm.st[t >> 6];
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
      // The wave's column ring: records (head, tail) hold fragments not yet
      // drawn, `head` the record holding fragment fcons - 1 (at a batch's
      // start, the record before its first); fragments [0, ftot) appended in
      // record order, [0, fcons) drawn.  Fragments are drawn 64 at a time, so
      // every window but a batch's last is full.
      uint32_t head = ~0u, tail = 0, ftot = 0, fcons = 0;
      // Draws fragments [fcons, fcons + 64) (those below ftot).  The records
      // starting inside the window are head + 1, head + 2, ... (each holds at
      // least one fragment), at most 64 of them: each tags the mark of its
      // first fragment with the window's sequence number, a ballot of the
      // tagged marks gives the window's record starts, and a fragment's record
      // is head plus the starts at or below it.  Live records number at most
      // 64 before an append of at most 64, so they fit the ring.
      auto window = [&]() {
        const uint32_t idx = head + 1u + (uint32_t)lane;
        if (idx < tail) {
          const uint32_t d = s.rec[idx & (kQRing - 1)].x - fcons;
          if (d < 64u) s.mark[d] = wseq;
        }
        wave_fence();
        const unsigned long long starts = __builtin_amdgcn_ballot_w64(s.mark[lane] == wseq);
        const uint32_t o = head + (uint32_t)((starts >> lane) & 1ull) +
                           __builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u));
        const uint32_t f = fcons + (uint32_t)lane;
        if (f < ftot) {
          const uint4 rc = s.rec[o & (kQRing - 1)];
          const uint32_t r = f - rc.x;
          const uint32_t slot = rc.y & 63u;
          const uint4 a = s.fa[slot];
          const uint4 b = s.fb[slot];
          // edge functions 1 and 2 down the column (exact: |row step| < 2^23)
          const int e1 = (int)rc.z + __mul24((int)r, (int)a.x);
          const int e2 = (int)rc.w + __mul24((int)r, (int)a.y);
          const float rcp = __uint_as_float(a.z);
          const float alpha = (float)e1 * rcp, beta = (float)e2 * rcp;  // raster.pikostage:269-270
          const float z = interp(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(a.w), alpha,
                                 beta);  // raster.pikostage:272
          const uint32_t lx = (rc.y >> 6) & 63u, ly = ((rc.y >> 12) & 63u) + r;
          atomicMin((long long *)(vis + (ly << ja.lgT) + lx),
                    (long long)(((unsigned long long)__float_as_uint(z) << 32) | (unsigned long long)b.z));
        }
        head += (uint32_t)__builtin_popcountll(starts);
        ++wseq;
      };
      // software pipeline: the setup records of the next batch and the
      // entries of the one after are in flight while a batch is rasterised
      uint32_t v_cur = beg + (uint32_t)t < end ? r_cur : ~0u;
      int4 gc[3];
      setup(v_cur, gc);
      uint32_t v_nxt = beg + (uint32_t)(kQThreads + t) < end ? r_nxt : ~0u;
      const uint32_t half = 1u << (S - 1);
      for (uint32_t b0 = beg; b0 < end; b0 += kQThreads) {
        int4 gn[3];
        setup(v_nxt, gn);
        const uint32_t v_nn = entry(b0 + 2u * kQThreads + (uint32_t)t);
        // Each staged triangle is drawn by a column walk (int32 edge
        // functions and |row steps| of edges 1 and 2 below 2^23, so 24-bit
        // products are exact) or by 8-row spans (the others).
        uint32_t uspan = 0;
        int walk = 0;  // bbox columns this lane's triangle walks
        QEdgeWalk wa = {0, 0, 0, 0, 1}, wb = {0, 0, 0, 0, 1}, wc = {0, 0, 0, 0, 1};
        bool bl = false, bh = false;  // the middle edge bounds the rows from below / above
        int E1 = 0, E2 = 0, A1 = 0, A2 = 0, D1 = 0, D2 = 0, hm1 = -1, lxc = 0, ly0 = 0;
        const int p[6] = {gc[0].x, gc[0].y, gc[0].z, gc[0].w, gc[1].x, gc[1].y};
        int x0 = 0, y0 = 0, w = 0, h = 0;  // the bbox clipped to the tile (w = 0: nothing staged)
        bool small = false, colw = false;
        uint32_t spans = 0;
        if (v_cur != ~0u) {
          int x1, y1;
          clip_rect(p, S, ja.W, ja.H, x0, y0, x1, y1);
          x0 = max(x0, tl.xlo);
          y0 = max(y0, tl.ylo);
          x1 = min(x1, tl.xhi);
          y1 = min(y1, tl.yhi);
          w = max(0, x1 - x0);
          h = max(0, y1 - y0);
          if (w > 0 && h > 0) {
            small = small_bbox(p, S);
            const int lim = 1 << (23 - S);
            spans = (uint32_t)w * (uint32_t)((h + kTileSpan - 1) / kTileSpan);
            // Which int32-path triangles are walked (ja.qwalk, PIKO_COVER_*):
            // all of them, or (spans) only tall slivers.  Spans test every
            // bbox pixel; the walk touches only covered pixels but sets up
            // each triangle's edge walks and hands its pixels out 64 at a
            // time.  With spans, a triangle is walked when its spans outnumber
            // its columns plus 0.45 per covered pixel, the measured exchange
            // rate.  The area in pixels is |2 x signed area| / 2^(2S+1) of the
            // fixed-point edges.
            colw = small && abs(p[4] - p[2]) < lim && abs(p[0] - p[4]) < lim;
            if (colw && !ja.qwalk) {
              const float rd = fabsf(__int_as_float(gc[1].z));
              const float area = rd > 0.0f ? 0.5f / (rd * (float)(1 << (2 * S))) : 0.0f;
              colw = (float)spans > (float)w + 0.45f * fminf(area, (float)(w * h));
            }
            s.fa[lane] = make_uint4((uint32_t)(p[4] - p[2]) << S, (uint32_t)(p[0] - p[4]) << S, (uint32_t)gc[1].z,
                                    (uint32_t)gc[1].w);
            s.fb[lane] = make_uint4((uint32_t)gc[2].x, (uint32_t)gc[2].y, 0xFFFFFFFEu - v_cur, 0u);
          } else {
            w = 0;
          }
        }
        if (w > 0) {
          if (colw) {
            // edge k at the centre of bbox pixel (c, j): E[k] + c*A[k] + j*D[k]
            // (raster.pikostage:176-180, 223-241)
            const uint32_t cx0 = ((uint32_t)x0 << S) + half, cy0 = ((uint32_t)y0 << S) + half;
            int E[3], A[3], D[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const uint32_t xa = (uint32_t)p[2 * k], ya = (uint32_t)p[2 * k + 1];
              const uint32_t xb = (uint32_t)p[(2 * k + 2) % 6], yb = (uint32_t)p[(2 * k + 3) % 6];
              E[k] = (int)((cy0 - ya) * (xb - xa) - (cx0 - xa) * (yb - ya));
              A[k] = (int)((ya - yb) << S);
              D[k] = (int)((xb - xa) << S);
            }
            // a vertical edge (D = 0) limits the columns: E + c*A >= 0
            int c0 = 0, c1 = w - 1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              if (D[k] != 0) continue;
              if (A[k] == 0) {
                if (E[k] < 0) c1 = -1;
                continue;
              }
              const int aa = A[k] < 0 ? -A[k] : A[k];
              int rr;
              const int fq = q_floordiv(E[k], aa, __builtin_amdgcn_rcpf((float)aa), rr);
              if (A[k] > 0) c0 = max(c0, -fq);  // c >= ceil(-E / A)
              else c1 = min(c1, fq);            // c <= floor(E / -A)
            }
            if (c0 <= c1) {
#pragma unroll
              for (int k = 0; k < 3; ++k) E[k] = (int)((uint32_t)E[k] + (uint32_t)c0 * (uint32_t)A[k]);
              // a front-facing triangle has an edge with D > 0 (a lower row
              // bound) and one with D < 0 (an upper one); the third is either
              // or vertical
              const int iA = D[0] > 0 ? 0 : (D[1] > 0 ? 1 : 2);
              const int iC = D[0] < 0 ? 0 : (D[1] < 0 ? 1 : 2);
              const int iB = 3 - iA - iC;
              auto pick = [](const int *v, int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); };
              wa.init(pick(E, iA), pick(A, iA), pick(D, iA));
              wc.init(pick(E, iC), pick(A, iC), pick(D, iC));
              const int DB = pick(D, iB);
              bl = DB > 0;
              bh = DB < 0;
              if (DB != 0) wb.init(pick(E, iB), pick(A, iB), DB);
              E1 = E[1];
              E2 = E[2];
              A1 = A[1];
              A2 = A[2];
              D1 = D[1];
              D2 = D[2];
              hm1 = h - 1;
              lxc = x0 + c0 - tl.xlo;
              ly0 = y0 - tl.ylo;
              walk = c1 - c0 + 1;
            }
          } else {
            uspan = spans;
            s.sp.xy01[lane] = make_int4(p[0], p[1], p[2], p[3]);
            s.sp.xr[lane] = make_int4(p[4], p[5], (int)((uint32_t)x0 | ((uint32_t)y0 << 16)),
                                   (int)((uint32_t)w | ((uint32_t)h << 16) | (small ? 0x80000000u : 0u)));
          }
        }
        v_cur = v_nxt;
#pragma unroll
        for (int q = 0; q < 3; ++q) gc[q] = gn[q];
        v_nxt = v_nn;
        wave_fence();  // staged
        // 8-row spans: every bbox pixel tested
        const uint32_t sinc = wave_scan_dpp(uspan);
        const uint32_t stot = wave_last(sinc);
        if (stot) {
          // Lane i's spans are [start, start + uspan).  The nonempty runs are
          // listed in lane order; those starting inside a window of 64 spans
          // follow the run holding the window's first span minus one, and
          // are found as the column walk finds its records (tagged marks and
          // a ballot).
          const uint32_t start = sinc - uspan;
          const unsigned long long nz = __builtin_amdgcn_ballot_w64(uspan != 0u);
          if (uspan) s.pre[__builtin_amdgcn_mbcnt_hi((uint32_t)(nz >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nz, 0u))] =
              start << 6 | (uint32_t)lane;
          uint32_t base = ~0u;  // the run holding span c0 - 1
          for (uint32_t c0 = 0; c0 < stot; c0 += 64) {
            if (uspan && start - c0 < 64u) s.mark[start - c0] = wseq;
            wave_fence();
            const unsigned long long st = __builtin_amdgcn_ballot_w64(s.mark[lane] == wseq);
            const uint32_t rank = base + (uint32_t)((st >> lane) & 1ull) +
                                  __builtin_amdgcn_mbcnt_hi((uint32_t)(st >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)st, 0u));
            const uint32_t run = s.pre[rank & 63u];
            base += (uint32_t)__builtin_popcountll(st);
            ++wseq;
            const uint32_t cidx = c0 + (uint32_t)lane;
            if (cidx < stot) q_span(s, (int)(run & 63u), cidx - (run >> 6), S, tl.xlo, tl.ylo, T, vis);
          }
          wave_fence();
        }
        // column walk: each lane steps its triangle's bbox columns, finding
        // each column's covered rows exactly; the nonempty columns go to the
        // ring, whose fragments are spread evenly over the lanes
        const int wmax = (int)wave_last((uint32_t)wave_max_scan_dpp(walk));
        for (int it = 0; it < wmax; ++it) {
          uint32_t cnt = 0, info = 0;
          int e1lo = 0, e2lo = 0;
          if (it < walk) {
            int lo = max(-wa.F, 0), hi = min(wc.F, hm1);
            lo = bl ? max(lo, -wb.F) : lo;
            hi = bh ? min(hi, wb.F) : hi;
            if (lo <= hi) {
              cnt = (uint32_t)(hi - lo + 1);
              info = (uint32_t)lane | ((uint32_t)(lxc + it) << 6) | ((uint32_t)(ly0 + lo) << 12);
              e1lo = (int)((uint32_t)E1 + (uint32_t)__mul24(lo, D1));
              e2lo = (int)((uint32_t)E2 + (uint32_t)__mul24(lo, D2));
            }
            E1 = (int)((uint32_t)E1 + (uint32_t)A1);
            E2 = (int)((uint32_t)E2 + (uint32_t)A2);
            wa.step();
            wb.step();
            wc.step();
          }
          const unsigned long long bal = __builtin_amdgcn_ballot_w64(cnt != 0u);
          const uint32_t rank =
              __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
          const uint32_t inc = wave_scan_dpp(cnt);
          if (cnt) s.rec[(tail + rank) & (kQRing - 1)] = make_uint4(ftot + inc - cnt, info, (uint32_t)e1lo, (uint32_t)e2lo);
          tail += (uint32_t)__builtin_popcountll(bal);
          ftot += wave_last(inc);
          wave_fence();
          while (ftot - fcons >= 64u) {
            window();
            fcons += 64u;
          }
        }
        if (ftot != fcons) window();  // the batch's last fragments: its slots are restaged next
        fcons = ftot;  // head is now the batch's last record
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
// One workgroup of kQMergeThreads per split tile: a thread per pixel of a
// 32-pixel tile, its parts' words loaded four at a time (256 threads with
// four pixels each waited for their loads over more rounds: 10.7 us on the
// 2M height field).
constexpr int kQMergeThreads = 1024;
template <int QT, int KIND, bool MULTI>
__global__ void __launch_bounds__(kQMergeThreads) k_qmerge(QArgs qa, FrameArgs fa) {
  __shared__ unsigned long long vis[QT * QT];
  __shared__ float s_u8[256];
  const int t = threadIdx.x;
  if (t < 256) s_u8[t] = kU8Table.v[t];
  const QItem it = qa.items[qa.splits[blockIdx.x]];
  const int gt = (int)it.tile;
  const FrameArgs &ja = MULTI ? qa.jobs[job_search(qa.jobs, qa.njobs, (uint32_t)gt, true)] : fa;
  const QTile tl = q_tile(ja, gt);
  const int T = 1 << ja.lgT;
  const unsigned long long *slots = qa.scratch + (size_t)it.slot * (size_t)(QT * QT);
  for (int q = t; q < T * T; q += kQMergeThreads) {
    long long m = (long long)kVisSentinel;
    uint32_t j = 0;
    for (; j + 4 <= it.nparts; j += 4) {
      long long w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = (long long)slots[(size_t)(j + k) * (size_t)(QT * QT) + q];
      m = min(m, min(min(w[0], w[1]), min(w[2], w[3])));
    }
    for (; j < it.nparts; ++j) m = min(m, (long long)slots[(size_t)j * (size_t)(QT * QT) + q]);
    vis[q] = (unsigned long long)m;
  }
  if (t == 0) {
    qa.tile_need[gt] = qa.tile_fill[gt];
    qa.tile_fill[gt] = 0;
  }
  __syncthreads();
  q_resolve<KIND, MULTI, kQMergeThreads>(qa, ja, vis, s_u8, tl.xlo, tl.ylo, tl.xhi, tl.yhi);
}
