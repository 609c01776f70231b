/*
 * piko/stage_exec.h -- executes stage bodies compiled from a pipe's own
 * source on gfx950: the device side of Stage (bins, assignToBin, emit), the
 * schedule kernels, and the host executor of the two stage graphs pikoc_hip
 * compiles.  The reference compiles stage bodies into kernels plus a
 * schedule of launches (pikoc/src/pikoc.cpp:69-318,
 * pikoc/src/Backend/PTXBackend.cpp:356-647); pikoc_hip writes a
 * __pikoStages.hip that includes the pipe's stage headers, then this file,
 * and builds it into a shared library the generated host methods call.
 *
 * Schedule (the serial order of the oracle and of the stage bodies run on
 * the host one primitive at a time):
 *  - streaming stage (every stage but the last): one thread per input
 *    primitive runs assignBin (its bins only spread work in the reference;
 *    here they are range-checked) then process.  Each emit is appended
 *    with (primitive, emit number); the emits are then put back in
 *    primitive order, so the next stage sees its input in the order a serial
 *    run emits it.  Reyes-shaped pipes run the looping stage level by level:
 *    port 1 feeds the next level, port 0 collects leaves across levels.
 *  - binned stage (the last, writing the PikoScreen): assignBin per
 *    primitive, (bin, primitive) entries in emission order, a stable radix
 *    sort by bin, then one thread per bin runs process over the bin's queue
 *    in order and its emits land on the screen at once.
 * A binned stage whose effects stay inside its bin (the depth words and the
 * screen pixels of the bin) therefore produces the serial result bit for
 * bit: a bin's queue is processed in serial order and no other thread
 * touches its pixels.  Bodies whose effects leave their bin race, as they
 * do in the reference's parallel schedule.
 *
 * The file is compiled by hipcc as part of __pikoStages.hip only: the stage
 * headers are included between `#pragma clang force_cuda_host_device
 * begin/end`, which makes every function they declare callable from device
 * code, and the Stage members declared there are defined below.
 */
#ifndef PIKO_STAGE_EXEC_H
#define PIKO_STAGE_EXEC_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "piko_hip.h"

namespace pikox {

enum : uint32_t {
  kErrBin = 1u,     // assignToBin outside the stage's bin grid
  kErrPort = 2u,    // emit to an unconnected output port
  kErrScreen = 4u,  // a pixel or sample outside the screen
  kErrType = 8u,    // the last stage emits a type the screen does not take
  kErrPolicy = 16u  // assignToBin with a policy (not supported here)
};

enum : int { kAssignCheck = 0, kAssign = 1, kStream = 2, kScreen = 3 };

constexpr uint32_t kClearColor = 0xff663313u;  // PikoScreen clear colour (ABGR8)

// An append-only record stream: each record is [primitive, emit number,
// payload words].  `fill` counts every append, also those past `cap`, which
// are dropped (the host then grows the stream and runs the frame again).
struct Sink {
  uint32_t *rec;
  uint32_t *fill;
  uint32_t cap;
};

// What one launch hands its threads.
struct Launch {
  ConstantState cs;
  float *zbuf;
  int nbx, nby;    // the stage's bin grid
  int W, H, spp;   // the screen
  int ports;       // bit p: output port p is connected
  Sink out0, out1, asg;
  uint32_t *cnt0, *cnt1, *acnt;  // per primitive: emits on port 0, 1; bins assigned
  uint32_t *color;               // PikoScreen words (Pixel emits)
  float *samples;                // PikoScreen samples, W*H*spp RGB (sample emits)
  uint32_t *err;
};

// One thread's state while a stage body runs.
struct Ctx {
  const Launch *l;
  int mode, bin;
  uint32_t src, seq0, seq1, nasg;
};

__device__ inline Ctx *ctx(void *p) { return static_cast<Ctx *>(p); }

// Appends `words` payload words for the calling lane; the active lanes of a
// wave take their slots with one atomic (the stage body may call this from
// divergent code).
__device__ inline void append(const Sink &s, uint32_t src, uint32_t seq, const uint32_t *payload, int words) {
  const uint64_t active = __ballot(1);
  const int lane = (int)__lane_id();
  const int leader = __builtin_ctzll(active);
  const uint32_t below = (uint32_t)__popcll(active & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(s.fill, (uint32_t)__popcll(active));
  base = __shfl(base, leader, 64);
  const uint32_t slot = base + below;
  if (slot >= s.cap) return;
  uint32_t *r = s.rec + (size_t)slot * (size_t)(2 + words);
  r[0] = src;
  r[1] = seq;
  for (int k = 0; k < words; ++k) r[2 + k] = payload[k];
}

template <class O, class = void>
struct IsSample : std::false_type {};
template <class O>
struct IsSample<O, std::void_t<decltype(std::declval<O>().rgb[2]), decltype(std::declval<O>().index),
                               decltype(std::declval<O>().pos.x)>> : std::true_type {};

// PikoScreen: a Pixel sets the pixel's colour word; a sample (pos, index,
// rgb[3]) sets one of the pixel's spp RGB samples, averaged at the end.
template <class O>
__device__ void screen(Ctx *c, const O &p) {
  const Launch &l = *c->l;
  if constexpr (std::is_same<O, Pixel>::value) {
    if ((unsigned)p.pos.x >= (unsigned)l.W || (unsigned)p.pos.y >= (unsigned)l.H) {
      atomicOr(l.err, kErrScreen);
      return;
    }
    l.color[(size_t)p.pos.y * (size_t)l.W + (size_t)p.pos.x] = p.color;
  } else if constexpr (IsSample<O>::value) {
    if ((unsigned)p.pos.x >= (unsigned)l.W || (unsigned)p.pos.y >= (unsigned)l.H || (unsigned)p.index >= (unsigned)l.spp) {
      atomicOr(l.err, kErrScreen);
      return;
    }
    const size_t slot = ((size_t)p.pos.y * (size_t)l.W + (size_t)p.pos.x) * (size_t)l.spp + (size_t)p.index;
    for (int k = 0; k < 3; ++k) l.samples[3 * slot + k] = p.rgb[k];
  } else {
    atomicOr(l.err, kErrType);
  }
}

template <class O>
__device__ void emit(Ctx *c, const O &p, int port) {
  static_assert(sizeof(O) % 4 == 0, "emitted primitives are whole 32-bit words");
  if (c->mode == kScreen) {
    screen(c, p);
    return;
  }
  const Launch &l = *c->l;
  if (c->mode != kStream || (port != 0 && port != 1) || !((l.ports >> port) & 1)) {
    atomicOr(l.err, kErrPort);
    return;
  }
  uint32_t w[sizeof(O) / 4];
  __builtin_memcpy(w, &p, sizeof(O));
  if (port == 0) append(l.out0, c->src, c->seq0++, w, (int)(sizeof(O) / 4));
  else append(l.out1, c->src, c->seq1++, w, (int)(sizeof(O) / 4));
}

__device__ inline void assign(Ctx *c, int bin) {
  const Launch &l = *c->l;
  if (bin < 0 || bin >= l.nbx * l.nby) {
    atomicOr(l.err, kErrBin);
    return;
  }
  if (c->mode != kAssign) return;  // a streaming stage's bins only spread its work
  const uint32_t b = (uint32_t)bin;
  append(l.asg, c->src, c->nasg++, &b, 1);
}

// The stage object a thread runs: the pipe's stage class with its uniforms,
// mutable state and this thread's context bound.
template <class S>
struct Bound : S {
  __device__ void bind(const ConstantState &cs, MutableState *ms, Ctx *c) {
    this->constState = cs;
    this->mutableState = ms;
    this->pikoRun_ = c;
  }
};

// ---- stage kernels -----------------------------------------------------------
constexpr int kThreads = 256;

template <class S>
__global__ void __launch_bounds__(kThreads) k_stage_stream(const typename S::input_type *in, uint32_t n, Launch l) {
  const uint32_t i = blockIdx.x * (uint32_t)kThreads + threadIdx.x;
  if (i >= n) return;
  Ctx c{&l, kAssignCheck, 0, i, 0u, 0u, 0u};
  MutableState ms;
  ms.zBuffer = l.zbuf;
  Bound<S> st;
  st.bind(l.cs, &ms, &c);
  const typename S::input_type p = in[i];
  st.assignBin(p);
  c.mode = kStream;
  st.process(p);
  l.cnt0[i] = c.seq0;
  l.cnt1[i] = c.seq1;
}

template <class S>
__global__ void __launch_bounds__(kThreads) k_stage_assign(const typename S::input_type *in, uint32_t n, Launch l) {
  const uint32_t i = blockIdx.x * (uint32_t)kThreads + threadIdx.x;
  if (i >= n) return;
  Ctx c{&l, kAssign, 0, i, 0u, 0u, 0u};
  MutableState ms;
  ms.zBuffer = l.zbuf;
  Bound<S> st;
  st.bind(l.cs, &ms, &c);
  st.assignBin(in[i]);
  l.acnt[i] = c.nasg;
}

// One thread per bin: the bin's queue in order, emits to the screen.
template <class S>
__global__ void __launch_bounds__(kThreads) k_stage_bins(const typename S::input_type *in, const uint32_t *src,
                                                         const uint32_t *start, const uint32_t *end, uint32_t nbins,
                                                         Launch l) {
  const uint32_t b = blockIdx.x * (uint32_t)kThreads + threadIdx.x;
  if (b >= nbins) return;
  Ctx c{&l, kScreen, (int)b, 0u, 0u, 0u, 0u};
  MutableState ms;
  ms.zBuffer = l.zbuf;
  Bound<S> st;
  st.bind(l.cs, &ms, &c);
  for (uint32_t e = start[b]; e < end[b]; ++e) {
    c.src = src[e];
    st.process(in[c.src]);
  }
}

// ---- order, scan, sort ------------------------------------------------------------
// Record r of a stream goes to base[primitive] + emit number: primitive order,
// and within a primitive the order of its emits.
__global__ void __launch_bounds__(kThreads) k_stage_order(const uint32_t *rec, uint32_t nrec, int words,
                                                          const uint32_t *base, uint32_t *out) {
  for (uint32_t r = blockIdx.x * (uint32_t)kThreads + threadIdx.x; r < nrec; r += gridDim.x * (uint32_t)kThreads) {
    const uint32_t *x = rec + (size_t)r * (size_t)(2 + words);
    const size_t d = (size_t)base[x[0]] + x[1];
    for (int k = 0; k < words; ++k) out[d * (size_t)words + k] = x[2 + k];
  }
}

// The assignment stream in order: key = bin, value = primitive.
__global__ void __launch_bounds__(kThreads) k_stage_order_bins(const uint32_t *rec, uint32_t nrec, const uint32_t *base,
                                                               uint32_t *keys, uint32_t *vals) {
  for (uint32_t r = blockIdx.x * (uint32_t)kThreads + threadIdx.x; r < nrec; r += gridDim.x * (uint32_t)kThreads) {
    const uint32_t *x = rec + (size_t)r * 3u;
    const uint32_t d = base[x[0]] + x[1];
    keys[d] = x[2];
    vals[d] = x[0];
  }
}

constexpr uint32_t kScanItems = 8, kScanBlock = kThreads * kScanItems;

// Exclusive scan of each 2,048-value block; the block's total to sums[].
__global__ void __launch_bounds__(kThreads) k_stage_scan(const uint32_t *in, uint32_t *out, uint32_t n,
                                                         uint32_t *sums) {
  __shared__ uint32_t wsum[kThreads / 64];
  const uint32_t first = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  uint32_t v[kScanItems], t = 0;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    v[k] = first + k < n ? in[first + k] : 0u;
    t += v[k];
  }
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  uint32_t x = t;  // inclusive scan of the thread totals across the wave
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, (unsigned)d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int k = 0; k < w; ++k) off += wsum[k];
  uint32_t run = off + x - t;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    if (first + k < n) out[first + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = off + x;
}

__global__ void __launch_bounds__(kThreads) k_stage_scan_add(uint32_t *out, uint32_t n, const uint32_t *offs) {
  const uint32_t first = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  const uint32_t o = offs[blockIdx.x];
  for (uint32_t k = 0; k < kScanItems; ++k)
    if (first + k < n) out[first + k] += o;
}

constexpr uint32_t kSortBlock = 4u * kThreads;  // keys per block: four rounds of 256

// Per block, how many of its keys carry each 8-bit digit; hist[digit][block].
__global__ void __launch_bounds__(kThreads) k_stage_sort_hist(const uint32_t *keys, uint32_t m, int shift,
                                                              uint32_t *hist, uint32_t nblk) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t i = blockIdx.x * kSortBlock + r * kThreads + threadIdx.x;
    if (i < m) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter by digit: a key's place is its block's start for the digit
// (the exclusive scan of hist), plus the keys of that digit earlier in the
// block -- earlier rounds, earlier waves of this round, lower lanes.
__global__ void __launch_bounds__(kThreads) k_stage_sort_scatter(const uint32_t *keys, const uint32_t *vals, uint32_t m,
                                                                 int shift, const uint32_t *offs, uint32_t nblk,
                                                                 uint32_t *keys_out, uint32_t *vals_out) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[kThreads / 64][256];
  const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
  run[t] = offs[(size_t)t * nblk + blockIdx.x];
  for (uint32_t r = 0; r < 4; ++r) {
    for (int q = 0; q < kThreads / 64; ++q) wc[q][t] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * kSortBlock + r * kThreads + (uint32_t)t;
    const bool valid = i < m;
    const uint32_t key = valid ? keys[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t same = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      same &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      keys_out[pos] = key;
      vals_out[pos] = vals[i];
    }
    __syncthreads();
    uint32_t add = 0;
    for (int q = 0; q < kThreads / 64; ++q) add += wc[q][t];
    run[t] += add;
    __syncthreads();
  }
}

// Queue bounds of each bin from the sorted keys (bins without entries keep
// the 0..0 they were cleared to).
__global__ void __launch_bounds__(kThreads) k_stage_bin_bounds(const uint32_t *keys, uint32_t m, uint32_t *start,
                                                               uint32_t *end) {
  for (uint32_t i = blockIdx.x * (uint32_t)kThreads + threadIdx.x; i < m; i += gridDim.x * (uint32_t)kThreads) {
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = i;
    if (i + 1 == m || keys[i + 1] != k) end[k] = i + 1;
  }
}

__global__ void __launch_bounds__(kThreads) k_stage_fill(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) p[i] = v;
}

// Screen samples: every sample starts at the clear colour.
__global__ void __launch_bounds__(kThreads) k_stage_clear_samples(float *rgb, size_t nsamples, float r, float g,
                                                                  float b) {
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < nsamples; i += (size_t)gridDim.x * kThreads) {
    rgb[3 * i] = r;
    rgb[3 * i + 1] = g;
    rgb[3 * i + 2] = b;
  }
}

// PikoScreen of sample emits: the box filter over a pixel's samples, in
// sample order, alpha 1.
__global__ void __launch_bounds__(kThreads) k_stage_resolve(const float *rgb, int spp, size_t npix, float *rgba) {
  for (size_t px = blockIdx.x * (size_t)kThreads + threadIdx.x; px < npix; px += (size_t)gridDim.x * kThreads) {
    const float inv = 1.0f / (float)spp;
    for (int k = 0; k < 3; ++k) {
      float a = rgb[3 * (px * (size_t)spp) + k];
      for (int s = 1; s < spp; ++s) a = a + rgb[3 * (px * (size_t)spp + (size_t)s) + k];
      rgba[4 * px + k] = a * inv;
    }
    rgba[4 * px + 3] = 1.0f;
  }
}

// ---- host executor -----------------------------------------------------------------
inline unsigned grid_for(size_t n, size_t per = kThreads, size_t cap = 16384) {
  return (unsigned)std::max<size_t>(1, std::min<size_t>((n + per - 1) / per, cap));
}

// A grow-only device buffer.
struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    const size_t want = std::max<size_t>(b + b / 4, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

#define PIKOX_HIP(call)                                 \
  do {                                                  \
    const hipError_t e_ = (call);                       \
    if (e_ != hipSuccess) {                             \
      (void)hipGetLastError();                          \
      return e_ == hipErrorOutOfMemory ? PIKO_ERR_OOM : PIKO_ERR_HIP; \
    }                                                   \
  } while (0)

#define PIKOX_RC(call)              \
  do {                              \
    const int rc_ = (call);         \
    if (rc_ != PIKO_OK) return rc_; \
  } while (0)

// A record stream with its counter, and the per-primitive counts of the
// stage that feeds it.
struct Stream {
  Buf rec, fill;
  uint32_t cap = 0;
  int words = 0;
  Sink sink() const { return Sink{rec.as<uint32_t>(), fill.as<uint32_t>(), cap}; }
};

class Exec {
 public:
  virtual ~Exec() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (Buf *b : owned_) b->release();
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  int prepare(const ConstantState *cs) {
    if (!cs || cs->screenSizeX != cs_.screenSizeX || cs->screenSizeY != cs_.screenSizeY) return PIKO_ERR_INVALID;
    cs_ = *cs;
    return PIKO_OK;
  }
  // One frame; a stream that overflowed is grown and the frame run again.
  int run() {
    for (int attempt = 0; attempt < 8; ++attempt) {
      bool again = false;
      PIKOX_RC(frame(again));
      if (!again) return check_errors();
    }
    return PIKO_ERR_CAPACITY;
  }
  int profile(float *ms, const char **names, int cap, int *nk, float *frame_ms) {
    if (!ms || !nk || cap <= 0) return PIKO_ERR_INVALID;
    PIKOX_RC(run());  // settles the stream capacities
    profiling_ = true;
    launches_.clear();
    if (!ev0_) PIKOX_HIP(hipEventCreate(&ev0_));
    PIKOX_HIP(hipEventRecord(ev0_, stream_));
    const int rc = run();
    profiling_ = false;
    PIKOX_RC(rc);
    PIKOX_HIP(hipStreamSynchronize(stream_));
    // per kernel (in order of first launch): the summed time of its launches
    std::vector<const char *> order;
    std::vector<float> sum;
    hipEvent_t prev = ev0_;
    for (size_t k = 0; k < launches_.size(); ++k) {
      float t = 0.0f;
      PIKOX_HIP(hipEventElapsedTime(&t, prev, events_[k]));
      prev = events_[k];
      size_t j = 0;
      while (j < order.size() && strcmp(order[j], launches_[k]) != 0) ++j;
      if (j == order.size()) {
        order.push_back(launches_[k]);
        sum.push_back(0.0f);
      }
      sum[j] += t;
    }
    const int n = std::min<int>((int)order.size(), cap);
    for (int k = 0; k < n; ++k) {
      ms[k] = sum[k];
      if (names) names[k] = order[k];
    }
    if (frame_ms && !launches_.empty()) {
      float t = 0.0f;
      PIKOX_HIP(hipEventElapsedTime(&t, ev0_, events_[launches_.size() - 1]));
      *frame_ms = t;
    }
    *nk = n;
    return PIKO_OK;
  }
  int W() const { return cs_.screenSizeX; }
  int H() const { return cs_.screenSizeY; }

 protected:
  virtual int frame(bool &again) = 0;

  int init(const ConstantState *cs) {
    if (!cs || cs->screenSizeX <= 0 || cs->screenSizeY <= 0 || cs->screenSizeX > 16384 || cs->screenSizeY > 16384)
      return PIKO_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      return PIKO_ERR_NO_DEVICE;
    }
    cs_ = *cs;
    PIKOX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    PIKOX_HIP(err_.need(4));
    owned_.push_back(&err_);
    return PIKO_OK;
  }

  // After a launch: its name, and an event when a frame is being profiled.
  int mark(const char *name) {
    PIKOX_HIP(hipGetLastError());
    if (!profiling_) return PIKO_OK;
    if (launches_.size() == events_.size()) {
      hipEvent_t e;
      PIKOX_HIP(hipEventCreate(&e));
      events_.push_back(e);
    }
    PIKOX_HIP(hipEventRecord(events_[launches_.size()], stream_));
    launches_.push_back(name);
    return PIKO_OK;
  }

  int fill(Buf &b, size_t words, uint32_t v) {
    if (!words) return PIKO_OK;
    k_stage_fill<<<grid_for(words), kThreads, 0, stream_>>>(b.as<uint32_t>(), words, v);
    return mark("k_stage_fill");
  }

  int read_u32(const Buf &b, uint32_t *out, size_t n = 1) {
    PIKOX_HIP(hipMemcpyAsync(out, b.p, n * 4, hipMemcpyDeviceToHost, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    return PIKO_OK;
  }

  // Exclusive scan of n counts into out (levels of 2,048-value blocks).
  int scan(const uint32_t *in, uint32_t *out, uint32_t n, int level = 0) {
    if (!n) return PIKO_OK;
    const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (2 * level + 1 >= kScanLevels) return PIKO_ERR_CAPACITY;
    Buf &sums = scan_tmp_[2 * level], &offs = scan_tmp_[2 * level + 1];
    PIKOX_HIP(sums.need((size_t)nb * 4));
    PIKOX_HIP(offs.need((size_t)nb * 4));
    k_stage_scan<<<nb, kThreads, 0, stream_>>>(in, out, n, sums.as<uint32_t>());
    PIKOX_RC(mark("k_stage_scan"));
    if (nb > 1) {
      PIKOX_RC(scan(sums.as<uint32_t>(), offs.as<uint32_t>(), nb, level + 1));
      k_stage_scan_add<<<nb, kThreads, 0, stream_>>>(out, n, offs.as<uint32_t>());
      PIKOX_RC(mark("k_stage_scan_add"));
    }
    return PIKO_OK;
  }

  // Appended records of `s` in primitive order into out (s.words each); the
  // feeding stage wrote its per-primitive counts to `cnt`.
  int order(Stream &s, uint32_t nrec, const Buf &cnt, uint32_t nsrc, Buf &out) {
    PIKOX_HIP(base_.need((size_t)std::max<uint32_t>(nsrc, 1) * 4));
    PIKOX_HIP(out.need((size_t)std::max<uint32_t>(nrec, 1) * (size_t)s.words * 4));
    if (!nrec) return PIKO_OK;
    PIKOX_RC(scan(cnt.as<uint32_t>(), base_.as<uint32_t>(), nsrc));
    k_stage_order<<<grid_for(nrec), kThreads, 0, stream_>>>(s.rec.as<uint32_t>(), nrec, s.words, base_.as<uint32_t>(),
                                                            out.as<uint32_t>());
    return mark("k_stage_order");
  }

  // Stream capacity for a launch over n primitives; counter cleared.
  int arm(Stream &s, size_t want) {
    if (s.cap < want) s.cap = (uint32_t)std::min<size_t>(want, 0x7fffffffu / (size_t)(2 + s.words));
    PIKOX_HIP(s.rec.need((size_t)std::max<uint32_t>(s.cap, 1) * (size_t)(2 + s.words) * 4));
    PIKOX_HIP(s.fill.need(4));
    return fill(s.fill, 1, 0u);
  }

  // After a launch that appended to s: the number of records, or `again`
  // with s grown to hold them all.
  int settle(Stream &s, uint32_t &n, bool &again) {
    PIKOX_RC(read_u32(s.fill, &n));
    if (n > s.cap) {
      if ((size_t)n * (size_t)(2 + s.words) > 0x7fffffffull) return PIKO_ERR_CAPACITY;
      s.cap = n + n / 4;
      again = true;
    }
    return PIKO_OK;
  }

  // Binned last stage S over `items` (n of them): assignment stream in
  // order, stable sort by bin, one thread per bin onto the screen.
  template <class S>
  int binned(const Buf &items, uint32_t n, Launch l, bool &again) {
    const uint32_t nbins = (uint32_t)(l.nbx * l.nby);
    PIKOX_HIP(acnt_.need((size_t)std::max<uint32_t>(n, 1) * 4));
    PIKOX_RC(arm(asg_, (size_t)n * 4 + 1024));
    l.asg = asg_.sink();
    l.acnt = acnt_.as<uint32_t>();
    if (n) {
      k_stage_assign<S><<<grid_for(n, kThreads, 1u << 30), kThreads, 0, stream_>>>(
          items.as<typename S::input_type>(), n, l);
      PIKOX_RC(mark("k_stage_assign"));
    }
    uint32_t m = 0;
    PIKOX_RC(settle(asg_, m, again));
    if (again) return PIKO_OK;
    PIKOX_HIP(keys_[0].need((size_t)std::max<uint32_t>(m, 1) * 4));
    PIKOX_HIP(keys_[1].need((size_t)std::max<uint32_t>(m, 1) * 4));
    PIKOX_HIP(vals_[0].need((size_t)std::max<uint32_t>(m, 1) * 4));
    PIKOX_HIP(vals_[1].need((size_t)std::max<uint32_t>(m, 1) * 4));
    PIKOX_HIP(start_.need((size_t)nbins * 4));
    PIKOX_HIP(end_.need((size_t)nbins * 4));
    PIKOX_RC(fill(start_, nbins, 0u));
    PIKOX_RC(fill(end_, nbins, 0u));
    if (m) {
      PIKOX_HIP(base_.need((size_t)std::max<uint32_t>(n, 1) * 4));
      PIKOX_RC(scan(acnt_.as<uint32_t>(), base_.as<uint32_t>(), n));
      k_stage_order_bins<<<grid_for(m), kThreads, 0, stream_>>>(asg_.rec.as<uint32_t>(), m, base_.as<uint32_t>(),
                                                                keys_[0].as<uint32_t>(), vals_[0].as<uint32_t>());
      PIKOX_RC(mark("k_stage_order_bins"));
      int bits = 0;
      while (bits < 32 && (nbins - 1) >> bits) ++bits;
      const uint32_t nblk = (m + kSortBlock - 1) / kSortBlock;
      PIKOX_HIP(hist_.need((size_t)nblk * 256 * 4));
      PIKOX_HIP(hoff_.need((size_t)nblk * 256 * 4));
      int cur = 0;
      for (int shift = 0; shift < bits; shift += 8) {
        k_stage_sort_hist<<<nblk, kThreads, 0, stream_>>>(keys_[cur].as<uint32_t>(), m, shift, hist_.as<uint32_t>(),
                                                          nblk);
        PIKOX_RC(mark("k_stage_sort_hist"));
        PIKOX_RC(scan(hist_.as<uint32_t>(), hoff_.as<uint32_t>(), nblk * 256));
        k_stage_sort_scatter<<<nblk, kThreads, 0, stream_>>>(keys_[cur].as<uint32_t>(), vals_[cur].as<uint32_t>(), m,
                                                             shift, hoff_.as<uint32_t>(), nblk,
                                                             keys_[cur ^ 1].as<uint32_t>(),
                                                             vals_[cur ^ 1].as<uint32_t>());
        PIKOX_RC(mark("k_stage_sort_scatter"));
        cur ^= 1;
      }
      k_stage_bin_bounds<<<grid_for(m), kThreads, 0, stream_>>>(keys_[cur].as<uint32_t>(), m, start_.as<uint32_t>(),
                                                                end_.as<uint32_t>());
      PIKOX_RC(mark("k_stage_bin_bounds"));
      k_stage_bins<S><<<grid_for(nbins, kThreads, 1u << 30), kThreads, 0, stream_>>>(
          items.as<typename S::input_type>(), vals_[cur].as<uint32_t>(), start_.as<uint32_t>(), end_.as<uint32_t>(),
          nbins, l);
      PIKOX_RC(mark("k_stage_bins"));
    }
    return PIKO_OK;
  }

  Launch launch_for(int binSize) const {
    Launch l;
    memset(&l, 0, sizeof(l));
    l.cs = cs_;
    l.nbx = binSize > 0 ? (cs_.screenSizeX + binSize - 1) / binSize : 1;
    l.nby = binSize > 0 ? (cs_.screenSizeY + binSize - 1) / binSize : 1;
    l.W = cs_.screenSizeX;
    l.H = cs_.screenSizeY;
    l.err = err_.as<uint32_t>();
    return l;
  }

  int check_errors() {
    uint32_t e = 0;
    PIKOX_RC(read_u32(err_, &e));
    if (!e) return PIKO_OK;
    fprintf(stderr, "piko stage code:%s%s%s%s%s\n", (e & kErrBin) ? " assignToBin outside the bin grid;" : "",
            (e & kErrPort) ? " emit to an unconnected port;" : "",
            (e & kErrScreen) ? " emit outside the screen;" : "",
            (e & kErrType) ? " the last stage emits a type the PikoScreen does not take;" : "",
            (e & kErrPolicy) ? " assignToBin with a policy (not supported);" : "");
    return PIKO_ERR_INVALID;
  }

  template <class T>
  int upload(Buf &b, const T *host, size_t n) {
    PIKOX_HIP(b.need(std::max<size_t>(n, 1) * sizeof(T)));
    if (n) PIKOX_HIP(hipMemcpyAsync(b.p, host, n * sizeof(T), hipMemcpyHostToDevice, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    return PIKO_OK;
  }

  void own(std::initializer_list<Buf *> bs) { owned_.insert(owned_.end(), bs.begin(), bs.end()); }
  void own(Stream &s) { own({&s.rec, &s.fill}); }

  ConstantState cs_{};
  hipStream_t stream_ = nullptr;
  Buf err_, base_, acnt_, keys_[2], vals_[2], hist_, hoff_, start_, end_;
  Stream asg_;
  static constexpr int kScanLevels = 16;
  Buf scan_tmp_[kScanLevels];

 private:
  std::vector<Buf *> owned_;
  std::vector<hipEvent_t> events_;
  std::vector<const char *> launches_;
  hipEvent_t ev0_ = nullptr;
  bool profiling_ = false;

 public:
  void own_common() {
    asg_.words = 1;
    own({&base_, &acnt_, &keys_[0], &keys_[1], &vals_[0], &vals_[1], &hist_, &hoff_, &start_, &end_});
    own(asg_);
    for (Buf &b : scan_tmp_) owned_.push_back(&b);
  }
};

// VertexShade-shaped pipe: streaming stage A -> binned stage B -> PikoScreen
// (Pixel emits; depth words W*H).
template <class A, class B>
class RasterExec : public Exec {
 public:
  using In = typename A::input_type;
  using Mid = typename A::output_type;
  static_assert(std::is_same<Mid, typename B::input_type>::value, "A's output is B's input");

  int create(const ConstantState *cs, const In *input, int count) {
    if (count < 0 || (count > 0 && !input)) return PIKO_ERR_INVALID;
    PIKOX_RC(init(cs));
    own_common();
    own({&in_, &cnt0_, &cnt1_, &mid_, &color_, &zbuf_});
    own(out_);
    out_.words = (int)(sizeof(Mid) / 4);
    n_ = (uint32_t)count;
    return upload(in_, input, (size_t)count);
  }
  int get_color(uint32_t *out) {
    PIKOX_HIP(hipMemcpyAsync(out, color_.p, (size_t)W() * H() * 4, hipMemcpyDeviceToHost, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    return PIKO_OK;
  }
  int get_depth(float *out) {
    PIKOX_HIP(hipMemcpyAsync(out, zbuf_.p, (size_t)W() * H() * 4, hipMemcpyDeviceToHost, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    return PIKO_OK;
  }

 protected:
  int frame(bool &again) override {
    const size_t npix = (size_t)W() * H();
    PIKOX_HIP(color_.need(npix * 4));
    PIKOX_HIP(zbuf_.need(npix * 4));
    PIKOX_RC(fill(err_, 1, 0u));
    PIKOX_RC(fill(color_, npix, kClearColor));
    PIKOX_RC(fill(zbuf_, npix, 0x3f800000u));  // depth 1.0
    PIKOX_HIP(cnt0_.need((size_t)std::max<uint32_t>(n_, 1) * 4));
    PIKOX_HIP(cnt1_.need((size_t)std::max<uint32_t>(n_, 1) * 4));
    PIKOX_RC(arm(out_, (size_t)n_ + 1024));
    Launch la = launch_for(A::binSizeX);
    la.zbuf = zbuf_.as<float>();
    la.ports = 1;
    la.out0 = out_.sink();
    la.cnt0 = cnt0_.as<uint32_t>();
    la.cnt1 = cnt1_.as<uint32_t>();
    if (n_) {
      k_stage_stream<A><<<grid_for(n_, kThreads, 1u << 30), kThreads, 0, stream_>>>(in_.as<In>(), n_, la);
      PIKOX_RC(mark("k_stage_stream"));
    }
    uint32_t nmid = 0;
    PIKOX_RC(settle(out_, nmid, again));
    if (again) return PIKO_OK;
    PIKOX_RC(order(out_, nmid, cnt0_, n_, mid_));
    Launch lb = launch_for(B::binSizeX);
    lb.zbuf = zbuf_.as<float>();
    lb.color = color_.as<uint32_t>();
    return binned<B>(mid_, nmid, lb, again);
  }

 private:
  uint32_t n_ = 0;
  Buf in_, cnt0_, cnt1_, mid_, color_, zbuf_;
  Stream out_;
};

// Reyes-shaped pipe: stage Sp loops on itself through port 1 (level by
// level) and sends leaves out of port 0 -> streaming stage Di -> binned
// stage Sh -> PikoScreen of `spp` samples per pixel (depth words W*H*spp).
template <class Sp, class Di, class Sh>
class ReyesExec : public Exec {
 public:
  using In = typename Sp::input_type;
  using Leaf = typename Sp::output_type;
  using Mp = typename Di::output_type;
  static_assert(std::is_same<In, Leaf>::value, "the looping stage emits its own input type");
  static_assert(std::is_same<Leaf, typename Di::input_type>::value, "leaves feed the second stage");
  static_assert(std::is_same<Mp, typename Sh::input_type>::value, "the second stage feeds the last");

  int create(const ConstantState *cs, const In *input, int count, int spp, int max_levels) {
    if (count < 0 || (count > 0 && !input) || spp < 1 || spp > 64 || max_levels < 1) return PIKO_ERR_INVALID;
    PIKOX_RC(init(cs));
    own_common();
    own({&in_, &lvl_[0], &lvl_[1], &leaves_, &tmp_, &cnt0_, &cnt1_, &mp_, &samples_, &zbuf_, &rgba_});
    own(next_);
    own(leaf_);
    own(mpout_);
    next_.words = leaf_.words = (int)(sizeof(Leaf) / 4);
    mpout_.words = (int)(sizeof(Mp) / 4);
    spp_ = spp;
    max_levels_ = max_levels;
    n_ = (uint32_t)count;
    return upload(in_, input, (size_t)count);
  }
  // FP32 RGBA per pixel
  int get_color(float *out) {
    PIKOX_HIP(hipMemcpyAsync(out, rgba_.p, (size_t)W() * H() * 16, hipMemcpyDeviceToHost, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    return PIKO_OK;
  }
  // per pixel, the smallest depth word of its samples
  int get_depth(float *out) {
    const size_t npix = (size_t)W() * H();
    std::vector<int32_t> z(npix * (size_t)spp_);
    PIKOX_HIP(hipMemcpyAsync(z.data(), zbuf_.p, z.size() * 4, hipMemcpyDeviceToHost, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    for (size_t p = 0; p < npix; ++p) {
      int32_t m = z[p * (size_t)spp_];
      for (int s = 1; s < spp_; ++s) m = std::min(m, z[p * (size_t)spp_ + (size_t)s]);
      memcpy(out + p, &m, 4);
    }
    return PIKO_OK;
  }
  int spp() const { return spp_; }

 protected:
  int frame(bool &again) override {
    const size_t npix = (size_t)W() * H(), nsamp = npix * (size_t)spp_;
    PIKOX_HIP(samples_.need(nsamp * 12));
    PIKOX_HIP(zbuf_.need(nsamp * 4));
    PIKOX_HIP(rgba_.need(npix * 16));
    PIKOX_RC(fill(err_, 1, 0u));
    PIKOX_RC(fill(zbuf_, nsamp, 0x3f800000u));
    const float cr = (float)(kClearColor & 255u) / 255.0f, cg = (float)((kClearColor >> 8) & 255u) / 255.0f,
                cb = (float)((kClearColor >> 16) & 255u) / 255.0f;
    k_stage_clear_samples<<<grid_for(nsamp), kThreads, 0, stream_>>>(samples_.as<float>(), nsamp, cr, cg, cb);
    PIKOX_RC(mark("k_stage_clear_samples"));
    // the looping stage, level by level
    const Buf *level = &in_;
    uint32_t nlev = n_, nleaves = 0;
    int li = 0;
    for (int depth = 0; nlev > 0; ++depth) {
      if (depth >= max_levels_) {
        fprintf(stderr, "piko stage code: the looping stage still emits after %d levels\n", max_levels_);
        return PIKO_ERR_CAPACITY;
      }
      PIKOX_HIP(cnt0_.need((size_t)nlev * 4));
      PIKOX_HIP(cnt1_.need((size_t)nlev * 4));
      PIKOX_RC(arm(next_, 2 * (size_t)nlev + 1024));
      PIKOX_RC(arm(leaf_, (size_t)nlev + 1024));
      Launch l = launch_for(Sp::binSizeX);
      l.zbuf = zbuf_.as<float>();
      l.ports = 3;
      l.out0 = leaf_.sink();
      l.out1 = next_.sink();
      l.cnt0 = cnt0_.as<uint32_t>();
      l.cnt1 = cnt1_.as<uint32_t>();
      k_stage_stream<Sp><<<grid_for(nlev, kThreads, 1u << 30), kThreads, 0, stream_>>>(level->as<In>(), nlev, l);
      PIKOX_RC(mark("k_stage_stream"));
      uint32_t nnext = 0, nleaf = 0;
      PIKOX_RC(settle(next_, nnext, again));
      PIKOX_RC(settle(leaf_, nleaf, again));
      if (again) return PIKO_OK;
      // this level's leaves after those of the levels before
      PIKOX_HIP(tmp_.need((size_t)std::max<uint32_t>(nleaf, 1) * sizeof(Leaf)));
      PIKOX_RC(order(leaf_, nleaf, cnt0_, nlev, tmp_));
      PIKOX_RC(grow_keep(leaves_, (size_t)nleaves * sizeof(Leaf), (size_t)(nleaves + nleaf) * sizeof(Leaf)));
      if (nleaf)
        PIKOX_HIP(hipMemcpyAsync(leaves_.as<char>() + (size_t)nleaves * sizeof(Leaf), tmp_.p, (size_t)nleaf * sizeof(Leaf),
                                 hipMemcpyDeviceToDevice, stream_));
      nleaves += nleaf;
      Buf &dst = lvl_[li];
      PIKOX_RC(order(next_, nnext, cnt1_, nlev, dst));
      level = &dst;
      li ^= 1;
      nlev = nnext;
    }
    // the second stage over the leaves
    PIKOX_HIP(cnt0_.need((size_t)std::max<uint32_t>(nleaves, 1) * 4));
    PIKOX_HIP(cnt1_.need((size_t)std::max<uint32_t>(nleaves, 1) * 4));
    PIKOX_RC(arm(mpout_, (size_t)nleaves * 16 + 1024));
    Launch ld = launch_for(Di::binSizeX);
    ld.zbuf = zbuf_.as<float>();
    ld.ports = 1;
    ld.out0 = mpout_.sink();
    ld.cnt0 = cnt0_.as<uint32_t>();
    ld.cnt1 = cnt1_.as<uint32_t>();
    if (nleaves) {
      k_stage_stream<Di><<<grid_for(nleaves, kThreads, 1u << 30), kThreads, 0, stream_>>>(leaves_.as<Leaf>(), nleaves,
                                                                                           ld);
      PIKOX_RC(mark("k_stage_stream"));
    }
    uint32_t nmp = 0;
    PIKOX_RC(settle(mpout_, nmp, again));
    if (again) return PIKO_OK;
    PIKOX_RC(order(mpout_, nmp, cnt0_, nleaves, mp_));
    Launch ls = launch_for(Sh::binSizeX);
    ls.zbuf = zbuf_.as<float>();
    ls.spp = spp_;
    ls.samples = samples_.as<float>();
    PIKOX_RC(binned<Sh>(mp_, nmp, ls, again));
    if (again) return PIKO_OK;
    k_stage_resolve<<<grid_for(npix), kThreads, 0, stream_>>>(samples_.as<float>(), spp_, npix, rgba_.as<float>());
    return mark("k_stage_resolve");
  }

 private:
  // grows b to `bytes`, keeping its first `keep` bytes
  int grow_keep(Buf &b, size_t keep, size_t bytes) {
    if (bytes <= b.bytes) return PIKO_OK;
    Buf nb;
    PIKOX_HIP(nb.need(bytes));
    if (keep) PIKOX_HIP(hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, stream_));
    PIKOX_HIP(hipStreamSynchronize(stream_));
    b.release();
    b = nb;
    return PIKO_OK;
  }

  uint32_t n_ = 0;
  int spp_ = 1, max_levels_ = 64;
  Buf in_, lvl_[2], leaves_, tmp_, cnt0_, cnt1_, mp_, samples_, zbuf_, rgba_;
  Stream next_, leaf_, mpout_;
};

}  // namespace pikox

// ---- Stage members a body calls, on the device -------------------------------------
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ void Stage<BX, BY, TT, I, O>::assignToBin(const I &, int bin) {
  pikox::assign(pikox::ctx(pikoRun_), bin);
}
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ void Stage<BX, BY, TT, I, O>::assignToBin(const I &, AssignPolicy) {
  atomicOr(pikox::ctx(pikoRun_)->l->err, pikox::kErrPolicy);
}
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ int Stage<BX, BY, TT, I, O>::getBinID() const {
  return pikox::ctx(pikoRun_)->bin;
}
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ int Stage<BX, BY, TT, I, O>::getNumBins() const {
  return pikox::ctx(pikoRun_)->l->nbx * pikox::ctx(pikoRun_)->l->nby;
}
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ int Stage<BX, BY, TT, I, O>::getNumBinsX() const {
  return pikox::ctx(pikoRun_)->l->nbx;
}
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ int Stage<BX, BY, TT, I, O>::getNumBinsY() const {
  return pikox::ctx(pikoRun_)->l->nby;
}
template <int BX, int BY, int TT, class I, class O>
__host__ __device__ void Stage<BX, BY, TT, I, O>::emit(const O &p, int port) {
  pikox::emit(pikox::ctx(pikoRun_), p, port);
}

namespace piko {
// inline: emitted only where device code calls it
__host__ __device__ inline int atomicMinInt(int *address, int value) { return atomicMin(address, value); }
}  // namespace piko

#endif
