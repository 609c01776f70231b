// piko_serve.hip -- serving loop over resident scenes.
//
// Scene geometry is uploaded once and stays resident in HBM; K batch pipes
// (each with its own HIP stream) render from it.  Clients submit render
// requests (piko_server_submit) from any thread while batches run and collect
// finished frames with piko_server_poll.  One worker thread owns the pipes:
// whenever a pipe is free it takes the requests that have arrived, in arrival
// order, as long as the batch fits the per-batch device-memory budget
// (framebuffer + per-triangle + queue footprint; the budget asked for, at
// most each pipe's share of the device memory free at creation), and
// launches them as one
// batch; while one batch runs the next forms and launches on another pipe, so
// batches overlap on the device.  A finished batch's frames are copied to the
// buffers their requests named and reported as completions.
// piko_server_replay drives the same loop from a trace of arrival times (the
// benchmark's harness).  The reference has no serving path; this sits above
// the pipe methods (rasterPipelineFixPt/main.cpp:172-193 drives one frame at
// a time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "piko_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
  long long id;
  piko_request req;
  uint32_t *color;  // host framebuffer of the request, or null
};

}  // namespace

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
  // requests and completions, guarded by mu
  std::mutex mu;
  std::condition_variable cv_work;  // the worker: new requests or stop
  std::condition_variable cv_done;  // pollers: new completions
  std::deque<Pending> incoming;
  std::deque<piko_completion> completed;
  long long next_id = 0;
  long long outstanding = 0;  // submitted, not yet completed
  long long batches = 0, batched_jobs = 0;
  bool stop = false;
  std::thread worker;
  bool worker_started = false;
  Clock::time_point t0 = Clock::now();
};

namespace {

// Device bytes one request adds to a batch: colour + depth + the direct
// path's visibility word per pixel; setup record, bin count and queue entries
// per triangle.
long long footprint(const piko_server *s, const piko_request &r) {
  const long long px = (long long)r.cs.screenSizeX * r.cs.screenSizeY;
  const long long n = s->scene_count[r.scene];
  return px * 16 + n * (64 + 4 + 4 * 16);
}

double now_s(const piko_server *s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

void complete(piko_server *s, const std::vector<Pending> &batch, int status) {
  const double t = now_s(s);
  std::lock_guard<std::mutex> lk(s->mu);
  for (const Pending &p : batch) s->completed.push_back(piko_completion{p.id, t, status});
  s->outstanding -= (long long)batch.size();
  s->cv_done.notify_all();
}

// The worker: retire finished batches, then give every free pipe the
// requests that have arrived and fit its budget.  It is the only thread
// that touches the pipes while it runs.
void work(piko_server *s) {
  const int K = (int)s->pipes.size();
  std::vector<std::vector<Pending>> inflight(K);
  std::deque<Pending> pending;
  std::vector<piko_raster_job> jobs;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      while (!s->incoming.empty()) {
        pending.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
      bool busy = !pending.empty();
      for (int k = 0; k < K; ++k) busy = busy || !inflight[k].empty();
      if (!busy) {
        if (s->stop) return;
        s->cv_work.wait(lk, [&] { return s->stop || !s->incoming.empty(); });
        continue;
      }
    }
    bool progressed = false;
    for (int k = 0; k < K; ++k) {
      if (inflight[k].empty()) continue;
      const int q = piko_batch_query(s->pipes[k]);
      if (q == 0) continue;
      int rc = q < 0 ? q : piko_batch_wait(s->pipes[k]);
      for (size_t j = 0; rc == PIKO_OK && j < inflight[k].size(); ++j)
        if (inflight[k][j].color) rc = piko_batch_get_color(s->pipes[k], (int)j, inflight[k][j].color);
      complete(s, inflight[k], rc);
      inflight[k].clear();
      progressed = true;
    }
    for (int k = 0; k < K && !pending.empty(); ++k) {
      if (!inflight[k].empty()) continue;
      long long used = 0;
      jobs.clear();
      while (!pending.empty() && (int)jobs.size() < s->max_jobs) {
        const piko_request &r = pending.front().req;
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
      if (rc != PIKO_OK) {
        complete(s, inflight[k], rc);
        inflight[k].clear();
      } else {
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->batches;
        s->batched_jobs += (long long)jobs.size();
      }
      progressed = true;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

int start_worker(piko_server *s) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->worker_started) return PIKO_OK;
  // the worker issues HIP calls: the device of the creating thread
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return PIKO_ERR_HIP;
  try {
    s->worker = std::thread([s, dev] {
      (void)hipSetDevice(dev);
      work(s);
    });
  } catch (...) {
    return PIKO_ERR_OOM;
  }
  s->worker_started = true;
  return PIKO_OK;
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
  // admission never promises more than the device holds: a pipe's batch
  // budget is at most its share of the memory free once the geometry is
  // resident (nine tenths of it, leaving room for the runtime)
  size_t free_b = 0, total_b = 0;
  if (rc == PIKO_OK && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    const long long share = (long long)(free_b / 10 * 9) / opt->pipes;
    if (share > 0) s->budget = std::min(s->budget, share);
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
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();  // finishes the requests already submitted
  for (piko_raster_pipe *p : s->pipes) piko_raster_destroy(p);
  if (s->owns) {
    if (s->d_pos) (void)hipFree(s->d_pos);
    if (s->d_attr) (void)hipFree(s->d_attr);
  }
  delete s;
  return PIKO_OK;
}

int piko_server_submit(piko_server *s, const piko_request *req, uint32_t *color, long long *id) {
  if (!s || !req || !id) return PIKO_ERR_INVALID;
  if (req->scene < 0 || req->scene >= (int)s->scene_count.size()) return PIKO_ERR_INVALID;
  if (req->cs.screenSizeX <= 0 || req->cs.screenSizeY <= 0) return PIKO_ERR_INVALID;
  if (footprint(s, *req) > s->budget) return PIKO_ERR_BUSY;  // never admissible
  int rc = start_worker(s);
  if (rc != PIKO_OK) return rc;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return PIKO_ERR_INVALID;
  *id = s->next_id++;
  s->incoming.push_back(Pending{*id, *req, color});
  ++s->outstanding;
  s->cv_work.notify_one();
  return PIKO_OK;
}

int piko_server_poll(piko_server *s, piko_completion *out, int cap, int wait_ms) {
  if (!s || cap < 0 || (cap > 0 && !out)) return PIKO_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && wait_ms > 0 && s->outstanding > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms),
                        [&] { return !s->completed.empty() || s->outstanding == 0; });
  int n = 0;
  while (n < cap && !s->completed.empty()) {
    out[n++] = s->completed.front();
    s->completed.pop_front();
  }
  return n;
}

int piko_server_replay(piko_server *s, const piko_request *reqs, int n, double *done_s, uint32_t **keep_color,
                       piko_server_stats *st) {
  if (!s || n < 0 || (n > 0 && (!reqs || !done_s))) return PIKO_ERR_INVALID;
  for (int i = 0; i < n; ++i) {
    const piko_request &r = reqs[i];
    if (r.scene < 0 || r.scene >= (int)s->scene_count.size()) return PIKO_ERR_INVALID;
    if (i > 0 && r.arrival_s < reqs[i - 1].arrival_s) return PIKO_ERR_INVALID;  // trace in arrival order
    if (footprint(s, r) > s->budget) return PIKO_ERR_BUSY;  // never admissible
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->outstanding > 0 || !s->completed.empty()) return PIKO_ERR_INVALID;  // the server is in use
  }
  long long b0, j0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    b0 = s->batches;
    j0 = s->batched_jobs;
  }
  std::unordered_map<long long, int> index;
  index.reserve((size_t)n * 2 + 1);
  const double start = now_s(s);
  int next = 0, done = 0, rc = PIKO_OK;
  std::vector<piko_completion> got(64);
  while (done < n) {
    // submit every request whose arrival time has passed
    const double t = now_s(s) - start;
    while (next < n && reqs[next].arrival_s <= t) {
      long long id = -1;
      const int r = piko_server_submit(s, &reqs[next], keep_color ? keep_color[next] : nullptr, &id);
      if (r != PIKO_OK) return r;
      index[id] = next++;
    }
    // wait for completions, at most until the next arrival
    int wait_ms = 1;
    if (next < n) wait_ms = (int)std::max(0.0, (reqs[next].arrival_s - (now_s(s) - start)) * 1e3);
    wait_ms = std::min(wait_ms, 5);
    const int k = piko_server_poll(s, got.data(), (int)got.size(), wait_ms);
    if (k < 0) return k;
    for (int i = 0; i < k; ++i) {
      const auto it = index.find(got[i].id);
      if (it == index.end()) continue;
      done_s[it->second] = got[i].done_s - start;
      if (got[i].status != PIKO_OK && rc == PIKO_OK) rc = got[i].status;
      ++done;
    }
    if (next < n && wait_ms == 0 && k == 0) std::this_thread::yield();
  }
  if (st) {
    std::lock_guard<std::mutex> lk(s->mu);
    st->requests = n;
    st->batches = s->batches - b0;
    st->wall_s = now_s(s) - start;
    st->mean_batch = st->batches ? (double)(s->batched_jobs - j0) / st->batches : 0.0;
  }
  return rc;
}

}  // extern "C"
