// piko_serve.hip -- serving loop over resident scenes.
//
// Scene geometry is uploaded once and stays resident in HBM; K batch pipes
// (each with its own HIP stream) render from it.  Requests arrive over time
// (a trace of arrival times); whenever a pipe is free, it takes the requests
// that have arrived, in arrival order, as long as the batch fits the pipe's
// memory budget (framebuffer + per-triangle + queue footprint), and launches
// them as one batch.  While one batch runs, the next forms and launches on
// another pipe, so batches overlap on the device.  The reference has no
// serving path; this sits above the pipe methods (rasterPipelineFixPt/
// main.cpp:172-193 drives one frame at a time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <new>
#include <thread>
#include <vector>

#include "piko_hip.h"

struct piko_server {
  int kind = 0, S = 8;
  long long tris = 0;
  std::vector<long long> scene_base;
  std::vector<int> scene_count;
  std::vector<piko_raster_pipe *> pipes;
  long long budget = 0;  // bytes per batch
  int max_jobs = 0;
  float *d_pos = nullptr;
  void *d_attr = nullptr;
  bool owns = false;
};

namespace {

// Device bytes one request adds to a batch: colour + depth, setup record
// (64 B) + count/queue entries per triangle (2 x 8 B pairs, double-buffered
// radix keys/values).
long long footprint(const piko_server *s, const piko_request &r) {
  // colour + depth + the direct path's visibility word per pixel; setup
  // record, bin count and queue entries per triangle
  const long long px = (long long)r.cs.screenSizeX * r.cs.screenSizeY;
  const long long n = s->scene_count[r.scene];
  return px * 16 + n * (64 + 4 + 4 * 16);
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int piko_server_create(piko_server **out, const float *pos, const void *attr, long long tris, int on_device,
                       const long long *scene_base, const int *scene_count, int nscenes,
                       const piko_server_options *opt) {
  if (!out || !opt || nscenes <= 0 || !scene_base || !scene_count || tris < 0) return PIKO_ERR_INVALID;
  *out = nullptr;
  if (opt->pipes <= 0 || opt->pipes > 16 || opt->max_jobs <= 0 || opt->batch_budget_bytes <= 0)
    return PIKO_ERR_INVALID;
  for (int k = 0; k < nscenes; ++k)
    if (scene_base[k] < 0 || scene_count[k] < 0 || scene_base[k] + scene_count[k] > tris) return PIKO_ERR_INVALID;
  piko_server *s = new (std::nothrow) piko_server();
  if (!s) return PIKO_ERR_OOM;
  s->kind = opt->input_kind;
  s->S = opt->subpixel_bits;
  s->tris = tris;
  s->scene_base.assign(scene_base, scene_base + nscenes);
  s->scene_count.assign(scene_count, scene_count + nscenes);
  s->budget = opt->batch_budget_bytes;
  s->max_jobs = opt->max_jobs;
  int rc = PIKO_OK;
  if (on_device) {
    s->d_pos = (float *)pos;
    s->d_attr = (void *)attr;
  } else {
    const size_t pb = (size_t)tris * 9 * sizeof(float);
    const size_t ab = (size_t)tris * (s->kind == PIKO_INPUT_WORLD ? 9 * sizeof(float) : 3 * sizeof(uint32_t));
    s->owns = true;
    if (hipMalloc((void **)&s->d_pos, std::max<size_t>(pb, 4)) != hipSuccess ||
        hipMalloc(&s->d_attr, std::max<size_t>(ab, 4)) != hipSuccess)
      rc = PIKO_ERR_OOM;
    else if (tris > 0 && (hipMemcpy(s->d_pos, pos, pb, hipMemcpyHostToDevice) != hipSuccess ||
                          hipMemcpy(s->d_attr, attr, ab, hipMemcpyHostToDevice) != hipSuccess))
      rc = PIKO_ERR_HIP;
  }
  long long max_scene = 1;
  for (int c : s->scene_count) max_scene = std::max<long long>(max_scene, c);
  piko_batch_options bo{};
  bo.input_kind = s->kind;
  bo.subpixel_bits = s->S;
  bo.max_jobs = s->max_jobs;
  // each pipe is sized up front for the largest batch admission lets through
  // (max_jobs requests of the largest scene at 1920x1080, within the batch
  // budget), so serving does not grow a pipe's buffers: a free synchronises
  // the whole device, stalling every pipe's stream.  Larger batches still
  // grow it.
  const long long per_tri = 64 + 4 + 4 * 16, per_px = 16;  // as footprint()
  bo.max_tris = std::max<long long>(max_scene, std::min<long long>((long long)s->max_jobs * max_scene,
                                                                     s->budget / per_tri));
  bo.max_pixels = std::max<long long>(1920ll * 1080, std::min<long long>((long long)s->max_jobs * 1920 * 1080,
                                                                          s->budget / per_px));
  bo.pair_capacity = 0;
  // a request is one screen of one resident scene: the direct path, which
  // measured the lower latency on such frames; PIKO_RASTER_MODE overrides
  bo.raster_mode = std::getenv("PIKO_RASTER_MODE") ? PIKO_RASTER_AUTO : PIKO_RASTER_DIRECT;
  for (int k = 0; rc == PIKO_OK && k < opt->pipes; ++k) {
    piko_raster_pipe *p = nullptr;
    rc = piko_batch_create(&p, s->d_pos, s->d_attr, tris, 1, &bo);
    if (rc == PIKO_ERR_OOM && (bo.max_tris > max_scene || bo.max_pixels > 1920ll * 1080)) {
      (void)hipGetLastError();
      // no room to size every pipe up front: this pipe and the rest start at
      // one request and grow with their batches
      bo.max_tris = max_scene;
      bo.max_pixels = 1920ll * 1080;
      rc = piko_batch_create(&p, s->d_pos, s->d_attr, tris, 1, &bo);
    }
    if (rc == PIKO_OK) s->pipes.push_back(p);
  }
  if (rc != PIKO_OK) {
    piko_server_destroy(s);
    return rc;
  }
  *out = s;
  return PIKO_OK;
}

int piko_server_destroy(piko_server *s) {
  if (!s) return PIKO_ERR_INVALID;
  for (piko_raster_pipe *p : s->pipes) piko_raster_destroy(p);
  if (s->owns) {
    if (s->d_pos) (void)hipFree(s->d_pos);
    if (s->d_attr) (void)hipFree(s->d_attr);
  }
  delete s;
  return PIKO_OK;
}

int piko_server_replay(piko_server *s, const piko_request *reqs, int n, double *done_s, uint32_t **keep_color,
                       piko_server_stats *st) {
  if (!s || (n > 0 && (!reqs || !done_s))) return PIKO_ERR_INVALID;
  for (int i = 0; i < n; ++i) {
    const piko_request &r = reqs[i];
    if (r.scene < 0 || r.scene >= (int)s->scene_count.size()) return PIKO_ERR_INVALID;
    if (i > 0 && r.arrival_s < reqs[i - 1].arrival_s) return PIKO_ERR_INVALID;  // trace in arrival order
    if (footprint(s, r) > s->budget) return PIKO_ERR_BUSY;  // never admissible
  }
  const int K = (int)s->pipes.size();
  std::vector<std::vector<int>> inflight(K);
  std::deque<int> pending;
  int next = 0, done = 0, batches = 0;
  long long batched_jobs = 0;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<piko_raster_job> jobs;
  while (done < n) {
    double t = now_s(t0);
    while (next < n && reqs[next].arrival_s <= t) pending.push_back(next++);
    bool progressed = false;
    // retire finished batches
    for (int k = 0; k < K; ++k) {
      if (inflight[k].empty()) continue;
      const int q = piko_batch_query(s->pipes[k]);
      if (q < 0) return q;
      if (q == 0) continue;
      int rc = piko_batch_wait(s->pipes[k]);
      if (rc != PIKO_OK) return rc;
      const double tf = now_s(t0);
      for (size_t j = 0; j < inflight[k].size(); ++j) {
        const int r = inflight[k][j];
        if (keep_color && keep_color[r] && (rc = piko_batch_get_color(s->pipes[k], (int)j, keep_color[r])) != PIKO_OK)
          return rc;
        done_s[r] = tf;
      }
      done += (int)inflight[k].size();
      inflight[k].clear();
      progressed = true;
    }
    // launch: each free pipe takes the arrived requests that fit its budget
    for (int k = 0; k < K && !pending.empty(); ++k) {
      if (!inflight[k].empty()) continue;
      long long used = 0;
      jobs.clear();
      while (!pending.empty() && (int)jobs.size() < s->max_jobs) {
        const piko_request &r = reqs[pending.front()];
        const long long f = footprint(s, r);
        if (used + f > s->budget) break;
        used += f;
        piko_raster_job jb{};
        jb.count = s->scene_count[r.scene];
        jb.src_base = s->scene_base[r.scene];
        jb.bin_size = r.bin_size;
        jb.cs = r.cs;
        jobs.push_back(jb);
        inflight[k].push_back(pending.front());
        pending.pop_front();
      }
      if (jobs.empty()) continue;
      const int rc = piko_batch_submit(s->pipes[k], jobs.data(), (int)jobs.size());
      if (rc != PIKO_OK) return rc;
      ++batches;
      batched_jobs += (long long)jobs.size();
      progressed = true;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  if (st) {
    st->requests = n;
    st->batches = batches;
    st->wall_s = now_s(t0);
    st->mean_batch = batches ? (double)batched_jobs / batches : 0.0;
  }
  return PIKO_OK;
}

}  // extern "C"
