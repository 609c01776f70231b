/*
 * piko_hip.h -- C-ABI of the MI355X (gfx950) Piko backend, libpiko_hip.so.
 *
 * All plain C (pointers and sizes, no C++ or torch types):
 *
 * 1. The raster pipeline (piko_raster_*): the pipe methods pikoc generates
 *    for RasterPipe (rasterPipelineFixPt/rasterPipe.h:16-45) --
 *    allocate/prepare/run_single/destroy plus PikoScreen::getData -- backed
 *    by hand-written HIP kernels instead of generated PTX; its batched form
 *    (piko_batch_*: many screens in one kernel sequence) and a serving loop
 *    over resident scenes (piko_server_*).
 *
 * 2. The Reyes pipeline (piko_reyes_*): the same methods for ReyesPipe
 *    (reyesPipe/reyesPipe.h:12-40).
 *
 * 3. The runtime shim (piko_rt_*): a 1:1 functional cover of the CUDA driver
 *    calls the generated host code of the PTX backend makes
 *    (pikoc/src/Backend/PTXBackend.cpp), over hipModule*, for code emitted
 *    against a HIP target.
 *
 * Errors: every call returns 0 (PIKO_OK) on success and a negative
 * piko_status otherwise; nothing aborts.  The reference prints and exits
 * through CUDACHECK (api/include/internal/cudaMacros.h:10-20); its C++
 * wrapper keeps that behaviour with PIKO_CHECK.
 * Threading: one pipe may be driven by one host thread at a time; distinct
 * pipes are independent (each owns its device memory and HIP stream).
 */
#ifndef PIKO_HIP_H
#define PIKO_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PIKO_OK = 0,
  PIKO_ERR_INVALID = -1,     /* bad argument */
  PIKO_ERR_HIP = -2,         /* a HIP runtime call failed */
  PIKO_ERR_NO_DEVICE = -3,   /* no GPU (the reference silently returns,
                                PTXBackend.cpp:685-690) */
  PIKO_ERR_CAPACITY = -4,    /* a queue would overflow: reported, never
                                wrapped (Bin::insert wraps, datatypes.h:88-93) */
  PIKO_ERR_OOM = -5,         /* device allocation failed */
  PIKO_ERR_BUSY = -6,        /* admission refused (serving budget) */
  PIKO_ERR_SCRATCH = -7      /* larger than the fixed scratch of
                                piko_raster_frame_noalloc */
} piko_status;

const char *piko_status_string(int status);

/* ---- raster pipeline --------------------------------------------------- */

/* Per-frame uniforms: ConstantState (api/include/piko/builtinTypes.h:20-33)
 * reduced to the fields the raster path reads, plus the Gouraud lighting
 * constants the reference hard-codes (main.cpp:317-328). */
typedef struct {
  int screenSizeX, screenSizeY;
  float halfW, halfH;          /* 0.5f*(float)W, 0.5f*(float)H (main.cpp:104-105) */
  float viewProjMatrix[16];    /* column-major, matmultfloat4 (math.h:365-373) */
  float light[3];              /* unit direction towards the light */
  float matcol[3];
  float ambient[3];
} piko_constant_state;

enum { PIKO_INPUT_PRETRANSFORMED = 0, PIKO_INPUT_WORLD = 1 };

/* How a frame is rasterised.  All give the reference's image bit for bit.
 *   BINNED: vertex stage -> (triangle, bin) queues sorted by bin -> one raster
 *           wave per bin (the reference's schedule, rasterPipe.h:16-45).
 *   DIRECT: the vertex stage rasterises small triangles itself into a per-pixel
 *           (depth, submission index) visibility word; large ones are drawn
 *           in 8x8 tiles; a resolve shades the winner.
 *   TILED:  bins are screen tiles (the bin size, at least 16, larger where a
 *           screen would have more than 8192 tiles).  The vertex stage
 *           counts, scans and scatters each chunk of 4096 triangles into
 *           per-tile lists in LDS; one workgroup per tile then resolves every
 *           fragment in LDS and writes each pixel once.  No global atomics.
 *   QUEUE:  one persistent launch per frame pulls its work from a device
 *           queue: bin items (vertex stage and trivial-reject test of a
 *           chunk of 1024 triangles, an LDS histogram over the screen tiles,
 *           one global add per touched tile reserving the chunk's run in
 *           that tile's queue), then part items (a tile's queue cut into
 *           parts of at most part_pairs entries, densest tile first; each
 *           part resolves its fragments in LDS, the parts of a split tile
 *           merge through scratch slots).  Tiles are the bin size, at least
 *           16, larger where a screen would have more than 4096 tiles.  Tile
 *           queues are sized from the counts of the first frame and grow
 *           (the frame re-runs) when a later frame needs more.
 *   DIRECT, TILED and QUEUE keep no sorted queues: the bin queues are built
 *   on demand when asked for (piko_raster_get_bins).
 *   AUTO:   DIRECT for a plain pipe, TILED for a batch (QUEUE is faster
 *           for repeated frames of one job set, but the first frame after a
 *           submit runs a counting pass first).  The environment
 *           variable PIKO_RASTER_MODE=binned|direct|tiled|queue overrides AUTO. */
enum {
  PIKO_RASTER_AUTO = 0,
  PIKO_RASTER_BINNED = 1,
  PIKO_RASTER_DIRECT = 2,
  PIKO_RASTER_TILED = 3,
  PIKO_RASTER_QUEUE = 4
};

/* QUEUE: how a tile part covers a triangle whose edge functions fit int32.
 *   WALK:  each lane steps its triangle's bbox columns (exact integer edge
 *          walks, no division per column) and the wave's covered pixels are
 *          spread evenly over its lanes, 64 to a window; fastest where
 *          triangles are thin slivers (the height field, whose bboxes are
 *          about a sixth covered).
 *   SPANS: 8-row column spans test every bbox pixel, except tall slivers,
 *          which are walked; fastest where triangles are compact (the
 *          batched soups, whose bboxes are about half covered).
 *   AUTO:  WALK for a plain pipe, SPANS for a batch.  Both give the same
 *          image. */
enum { PIKO_COVER_AUTO = 0, PIKO_COVER_WALK = 1, PIKO_COVER_SPANS = 2 };

typedef struct {
  int bin_size;           /* raster bin edge T in pixels: 8, 16, 32 or 64 */
  int subpixel_bits;      /* S in [1, 8] (reference: 4) */
  int input_kind;         /* PIKO_INPUT_* */
  long long pair_capacity;/* (triangle, bin) queue capacity; 0 = automatic,
                             grows on demand between frames */
  int waves_per_cu;       /* persistent raster waves per CU; 0 = default (20;
                             QUEUE: 20, five 256-thread workgroups, at most
                             the workgroups a CU holds at once) */
  int raster_mode;        /* PIKO_RASTER_*; 0 = AUTO */
  int part_pairs;         /* QUEUE: queue entries per part item; 0 = 2048 */
  int queue_cover;        /* QUEUE: PIKO_COVER_*; 0 = AUTO (WALK) */
} piko_raster_options;

typedef struct piko_raster_pipe piko_raster_pipe;

/* RasterPipe::allocate (PTXBackend.cpp:649-789 emits it): device memory for
 * the input and every stage, input upload (PikoArray::copyData,
 * datatypes.h:122-135, without its silent 15M truncation).
 *   pos  : [count*9] float  -- screen xyz per vertex (PRETRANSFORMED) or
 *                              world xyz per vertex (WORLD)
 *   attr : [count*3] uint32 -- packed ABGR8 per vertex (PRETRANSFORMED), or
 *          [count*9] float  -- unit normal per vertex (WORLD) */
int piko_raster_allocate(piko_raster_pipe **pipe, const piko_constant_state *cs,
                         const float *pos, const void *attr, int count,
                         const piko_raster_options *opt);

/* Same, with inputs already resident in device memory (not copied; must
 * outlive the pipe). */
int piko_raster_allocate_device(piko_raster_pipe **pipe, const piko_constant_state *cs,
                                const float *d_pos, const void *d_attr, int count,
                                const piko_raster_options *opt);

/* RasterPipe::prepare (PTXBackend.cpp:791-809): per-frame constant state
 * (camera).  The depth reset of resetDepthBuffer (main.cpp:402-409) and the
 * colour clear are folded into the raster kernel. */
int piko_raster_prepare(piko_raster_pipe *pipe, const piko_constant_state *cs);

/* RasterPipe::run_single (PTXBackend.cpp:811-829): one frame, synchronous. */
int piko_raster_run_single(piko_raster_pipe *pipe);

/* nframes frames back to back, one synchronisation at the end (throughput
 * measurement).  On the queue path a run of two or more frames keeps two
 * frames in flight: the first such run makes a twin of the pipe (its own
 * stream and frame buffers over the same input, about the pipe's device
 * memory again) and frames alternate between the two; these frames take
 * whole tiles as part items and a persistent grid of 12 waves per CU (the
 * options' part_pairs and waves_per_cu apply to single frames); the last
 * frame's image ends in the pipe's buffers.  Elsewhere frames run on the
 * pipe's stream one after another. */
int piko_raster_run_frames(piko_raster_pipe *pipe, int nframes);

/* Like run_frames, timing each kernel of the frame with HIP events on the
 * pipe's stream.  kernel_ms[k] receives the mean duration of kernel k over
 * the frames; names[k] its name (static strings); *nk the number of kernels
 * (at most cap).  frame_ms receives the mean event-to-event frame time. */
int piko_raster_profile_frames(piko_raster_pipe *pipe, int nframes, float *kernel_ms,
                               const char **names, int cap, int *nk, float *frame_ms);

/* PikoScreen::getData (stage.h:108-117): colour buffer, W*H ABGR8 words,
 * row 0 at the bottom (builtinTypes.h:54-55). */
int piko_raster_get_data(piko_raster_pipe *pipe, uint32_t *color);
/* Depth buffer, W*H floats (MutableState::zBuffer, builtinTypes.h:35-40). */
int piko_raster_get_depth(piko_raster_pipe *pipe, float *depth);

/* One frame that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same frame as piko_raster_allocate_device + run_single + get_data /
 * get_depth on the DIRECT path, with the inputs (d_pos, d_attr as for
 * piko_raster_allocate_device) and the outputs (d_color: W*H ABGR8 words,
 * d_depth: W*H floats, row 0 at the bottom) in caller-owned device memory.
 * Every per-frame buffer lives in one file-scope device array of about
 * 65 MB that each GPU holds until the process ends: at most 1,048,576
 * triangles, 2,097,152 pixels and 65,536 large triangles (bbox over 16,384
 * pixels) per call; larger calls return PIKO_ERR_SCRATCH.  Calls that
 * overlap on one GPU share that array and race: issue them one at a time.
 * Runs on the null stream of the current device and returns when the frame
 * is done. */
int piko_raster_frame_noalloc(const piko_constant_state *cs, const float *d_pos, const void *d_attr, int count,
                              int input_kind, int subpixel_bits, uint32_t *d_color, float *d_depth);

/* Raster-stage bin queues of the last frame (Stage::getData, stage.h:
 * 331-333): CSR, bin_start[nbins+1], bin_tris[pairs] in submission order.
 * Returns the pair count (>=0) or a piko_status; writes at most cap entries.
 * After a DIRECT frame the queues are built here, on the pipe's stream. */
long long piko_raster_get_bins(piko_raster_pipe *pipe, int64_t *bin_start, int32_t *bin_tris,
                               long long cap);

typedef struct {
  long long triangles;       /* input count */
  long long emitted;         /* triangles the vertex stage emitted */
  long long pairs;           /* (triangle, bin) queue entries (built on demand
                                after a DIRECT frame) */
  long long pair_capacity;
  int bins_x, bins_y;
  long long device_bytes;    /* device memory held by the pipe */
} piko_raster_stats;

int piko_raster_get_stats(piko_raster_pipe *pipe, piko_raster_stats *st);

/* RasterPipe::destroy (PTXBackend.cpp:831-861). */
int piko_raster_destroy(piko_raster_pipe *pipe);

/* ---- batches (many screens in one launch sequence) --------------------- */

/* One job of a batch: a screen (size, camera, lighting in `cs`) rendering
 * `count` triangles starting at triangle `src_base` of the batch input, with
 * raster bin size `bin_size`.  A scene of the batched configuration, or a
 * request on a resident scene when serving. */
typedef struct {
  int count;
  long long src_base;
  int bin_size;
  int waves_per_cu;        /* raster waves per CU for this job's bin-size class; 0 = default */
  float queue_depth;       /* expected (triangle, bin) queue entries per triangle; 0 = 2 */
  piko_constant_state cs;
  int queue_cover;         /* QUEUE: PIKO_COVER_* for this job's scene class; 0 = the batch's */
} piko_raster_job;

typedef struct {
  int input_kind;          /* PIKO_INPUT_* for every job */
  int subpixel_bits;       /* S for every job */
  int max_jobs;            /* job-table capacity */
  long long max_tris;      /* initial per-batch triangle capacity (grows) */
  long long max_pixels;    /* initial per-batch pixel capacity (grows) */
  long long pair_capacity; /* 0 = automatic */
  int raster_mode;         /* PIKO_RASTER_* for every batch; 0 = AUTO */
  int part_pairs;          /* QUEUE: queue entries per part item; 0 = 65536
                              (whole tiles, no merge launch, at the batch's
                              tile sizes) */
  int waves_per_cu;        /* QUEUE: persistent waves per CU, at most what a CU
                              holds at once (three workgroups at 64-px tiles);
                              0 = 24 */
  int queue_cover;         /* QUEUE: PIKO_COVER_* of jobs that name none; 0 = AUTO (SPANS) */
} piko_batch_options;

/* Creates a batch pipe over `input_tris` resident triangles (pos/attr as for
 * piko_raster_allocate; on_device != 0: device pointers, not copied). */
int piko_batch_create(piko_raster_pipe **pipe, const float *pos, const void *attr, long long input_tris,
                      int on_device, const piko_batch_options *opt);
/* Enqueues one frame of every job (asynchronous; waits for the pipe's
 * previous batch first).  All jobs run through one kernel sequence: vertex
 * stage, binning and sort over the batch, one raster launch per bin size. */
int piko_batch_submit(piko_raster_pipe *pipe, const piko_raster_job *jobs, int njobs);
/* 1 if the last batch has finished, 0 if it is still running. */
int piko_batch_query(piko_raster_pipe *pipe);
/* Waits for the last batch (re-running it if its queues had to grow). */
int piko_batch_wait(piko_raster_pipe *pipe);
/* Framebuffers of job `job` of the last batch (W*H words / floats). */
int piko_batch_get_color(piko_raster_pipe *pipe, int job, uint32_t *color);
int piko_batch_get_depth(piko_raster_pipe *pipe, int job, float *depth);

/* ---- serving ------------------------------------------------------------ */

/* A render request on a resident scene: camera and screen in `cs`. */
typedef struct {
  int scene;
  int bin_size;
  piko_constant_state cs;
  double arrival_s;        /* seconds after the start of the trace */
} piko_request;

typedef struct {
  int input_kind;
  int subpixel_bits;
  int pipes;                    /* batch pipes (HIP streams) in flight, <= 16 */
  int max_jobs;                 /* requests per batch at most */
  long long batch_budget_bytes; /* device bytes one batch may use (admission);
                                   at most each pipe's share of nine tenths of
                                   the device memory free once the geometry is
                                   resident */
} piko_server_options;

typedef struct {
  long long requests, batches;
  double wall_s, mean_batch;
} piko_server_stats;

typedef struct piko_server piko_server;

/* Uploads (on_device == 0) or adopts the scenes' geometry once; scene k is
 * triangles [scene_base[k], scene_base[k] + scene_count[k]).  Each of the
 * `pipes` batch pipes is allocated here for the largest batch admission lets
 * through (max_jobs requests of the largest scene at 1920x1080, at most
 * batch_budget_bytes), so serving itself does not allocate; a batch beyond
 * that (larger screens) grows its pipe. */
int piko_server_create(piko_server **server, const float *pos, const void *attr, long long tris, int on_device,
                       const long long *scene_base, const int *scene_count, int nscenes,
                       const piko_server_options *opt);
/* A finished request: its id (from piko_server_submit), completion time in
 * seconds on the server's clock, and status (PIKO_OK, or the error its batch
 * met). */
typedef struct {
  long long id;
  double done_s;
  int status;
} piko_completion;

/* Online serving (new API above RasterPipe::prepare / run_single,
 * rasterPipelineFixPt/main.cpp:172-193).  Any thread may submit while batches
 * run: the request joins the next batch a free pipe forms (continuous
 * batching by a worker thread that owns the pipes; admission in arrival
 * order under the per-batch budget).  `color` (optional, host, W*H words)
 * receives the frame and must stay valid until the request completes;
 * `arrival_s` is ignored.  *id names the request in completions.
 * PIKO_ERR_BUSY if the request can never fit the batch budget. */
int piko_server_submit(piko_server *server, const piko_request *req, uint32_t *color, long long *id);
/* Moves up to `cap` completions to `out` and returns their number, waiting
 * up to wait_ms for one when none is ready and requests are outstanding. */
int piko_server_poll(piko_server *server, piko_completion *out, int cap, int wait_ms);
/* Replays `reqs` (sorted by arrival) in real time through submit / poll: each
 * request is submitted at its arrival time.  done_s[i] receives request i's
 * completion time (same clock as arrival_s).  keep_color (optional, per
 * request, may hold NULLs) receives framebuffers.  PIKO_ERR_BUSY if a
 * request can never fit the batch budget; PIKO_ERR_INVALID while submitted
 * requests are outstanding. */
int piko_server_replay(piko_server *server, const piko_request *reqs, int n, double *done_s,
                       uint32_t **keep_color, piko_server_stats *st);
/* Finishes the submitted requests, then frees the pipes and the geometry it
 * uploaded. */
int piko_server_destroy(piko_server *server);

/* ---- Reyes pipeline ------------------------------------------------------ */

/* The pipe pikoc generates for ReyesPipe (reyesPipe/reyesPipe.h:12-40):
 * Split (looping on itself) -> Dice -> Shade -> PikoScreen, here as a split
 * kernel that walks each patch's split tree on the device (the reference
 * loops on the host with a device-to-host read per level,
 * PTXBackend.cpp:962-980), a dice + sample kernel and a resolve.
 * The input is bicubic Bezier patches: 16 control points x (x,y,z,w) floats
 * per patch (piko_patch.CP, reyesPipe/pikoTypes.h:65-73).  The camera is
 * cs->viewProjMatrix and the screen cs->screenSizeX/Y; the shading constants
 * are the ones ShadeStage hard-codes (shade.pikostage:64-85). */
typedef struct {
  float threshold;     /* split until both screen extents are below this many
                          pixels (THRESHOLD_PIXELS, split.pikostage:12: 4) */
  int max_splits;      /* split depth cap, <= 20 (split.pikostage:284) */
  int dice;            /* micropolygon grid edge per leaf: 1, 2, 4, 8, 16 or
                          32 quads covering the patch; 0 = the reference's
                          (nu+1) x (nv+1) grid from the leaf's extents
                          (split.pikostage:307-316, dice.pikostage:176-177) */
  int spp;             /* samples per pixel: 1 (the pixel centre,
                          shade.pikostage:166-167) or 4 (2x2 strata centres) */
  float displacement;  /* displacement along the normal by
                          displacement*(x^2 - y^2) of the surface point; 0 = none */
  int roots_per_wave;  /* patches whose split trees one wave walks; 0 = 1 */
  int keep_samples;    /* 1: keep per-sample depth for piko_reyes_get_samples */
  int sampling;        /* PIKO_SAMPLE_*: how covered samples are composited */
  int tile;            /* tiled sampling's tile edge in pixels: 16 or 32; 0 = 16 */
} piko_reyes_options;

/* How Reyes composites samples.  Both give the same frame bit for bit.
 *   GLOBAL: the dice kernel takes a 64-bit atomicMin of (depth, order) per
 *           covered sample in a W*H*spp buffer; a resolve pass filters it.
 *   TILED:  the dice kernel keeps the micropolygons and lists each leaf in
 *           the screen tiles it touches (16 pixels, or 32 where a screen
 *           would have more than 8192 tiles; larger screens are refused);
 *           workgroups take the same minimum per tile in LDS and resolve
 *           its pixels.  A tile with many leaves is split into parts whose
 *           keys meet in the visibility buffer.
 *   AUTO:   GLOBAL (the faster of the two on the sphere configuration). */
enum { PIKO_SAMPLE_AUTO = 0, PIKO_SAMPLE_GLOBAL = 1, PIKO_SAMPLE_TILED = 2 };

typedef struct piko_reyes_pipe piko_reyes_pipe;

/* ReyesPipe::allocate (reyesPipe/main.cpp:225, PTXBackend.cpp:649-789):
 * patches [count*64] floats, uploaded once. */
int piko_reyes_allocate(piko_reyes_pipe **pipe, const piko_constant_state *cs, const float *patches, int count,
                        const piko_reyes_options *opt);
/* Same with the patches already in device memory (must outlive the pipe). */
int piko_reyes_allocate_device(piko_reyes_pipe **pipe, const piko_constant_state *cs, const float *d_patches,
                               int count, const piko_reyes_options *opt);
/* ReyesPipe::prepare: camera for the next frames (screen size fixed). */
int piko_reyes_prepare(piko_reyes_pipe *pipe, const piko_constant_state *cs);
/* ReyesPipe::run_single: one frame, synchronous. */
int piko_reyes_run_single(piko_reyes_pipe *pipe);
/* nframes frames back to back, one synchronisation at the end.  Work queues
 * that overflow grow and the frames re-run (PIKO_ERR_CAPACITY if they cannot). */
int piko_reyes_run_frames(piko_reyes_pipe *pipe, int nframes);
/* As piko_raster_profile_frames. */
int piko_reyes_profile_frames(piko_reyes_pipe *pipe, int nframes, float *kernel_ms, const char **names, int cap,
                              int *nk, float *frame_ms);
/* Resolved frame: W*H*4 floats RGBA (box filter over the samples; uncovered
 * samples hold the PikoScreen clear colour 0xff663313, stage.h:83-85), row 0
 * at the bottom. */
int piko_reyes_get_color(piko_reyes_pipe *pipe, float *rgba);
/* W*H floats: the nearest sample depth of each pixel (1.0 where empty). */
int piko_reyes_get_depth(piko_reyes_pipe *pipe, float *depth);
/* W*H*spp float bits of every sample's depth (needs keep_samples). */
int piko_reyes_get_samples(piko_reyes_pipe *pipe, int32_t *z);
/* Split leaves of the last frame in the order the split stage emits them
 * (level by level, in order within a level): cp [cap*64] floats and the
 * reference's dice grid nu/nv [cap] (split.pikostage:307-316).  Returns the
 * leaf count (>= 0) or a piko_status; writes at most cap leaves. */
long long piko_reyes_get_leaves(piko_reyes_pipe *pipe, float *cp, int *nu, int *nv, long long cap);

typedef struct {
  long long patches;
  long long leaves;          /* split leaves of the last frame */
  long long micropolygons;   /* diced quads of the last frame */
  int levels;                /* split levels used */
  long long device_bytes;
  int capacity_reruns;       /* frames re-run by the last run call because a
                                work queue grew */
  long long tile_pairs;      /* (tile, leaf) list entries of the last frame
                                (tiled sampling; 0 otherwise) */
} piko_reyes_stats;

int piko_reyes_get_stats(piko_reyes_pipe *pipe, piko_reyes_stats *st);
/* ReyesPipe::destroy. */
int piko_reyes_destroy(piko_reyes_pipe *pipe);

/* ---- runtime shim (generated-code layer) ------------------------------- */

typedef struct piko_rt_module *piko_module_t;
typedef struct piko_rt_function *piko_fn_t;
typedef struct piko_rt_stream *piko_stream_t;   /* NULL = default stream */
typedef struct piko_rt_event *piko_event_t;

int piko_rt_init(int device);                                     /* cuInit..cuCtxCreate, PTXBackend.cpp:112-130 */
int piko_rt_module_load(const char *path, piko_module_t *mod);    /* cuModuleLoad, PTXBackend.cpp:127 */
int piko_rt_module_unload(piko_module_t mod);
int piko_rt_get_function(piko_module_t mod, const char *name, piko_fn_t *fn); /* cuModuleGetFunction, :870 */
int piko_rt_get_global(piko_module_t mod, const char *name, void **dptr, size_t *bytes); /* cuModuleGetGlobal, :244-247 */
int piko_rt_malloc(void **dptr, size_t bytes);                    /* cuMemAlloc, :143 ff. */
int piko_rt_free(void *dptr);                                     /* cuMemFree, :294-320 */
int piko_rt_memcpy_htod(void *dst, const void *src, size_t bytes);/* cuMemcpyHtoD */
int piko_rt_memcpy_dtoh(void *dst, const void *src, size_t bytes);/* cuMemcpyDtoH, :971 */
int piko_rt_memcpy_dtod(void *dst, const void *src, size_t bytes);/* device-to-device copy */
int piko_rt_memset_d32(void *dst, uint32_t value, size_t count);
int piko_rt_launch(piko_fn_t fn, unsigned gx, unsigned bx, unsigned shmem, piko_stream_t s,
                   void **args);                                  /* cuLaunchKernel, :996 */
int piko_rt_sync(piko_stream_t s);                                /* cuCtxSynchronize, :251 */
int piko_rt_stream_create(piko_stream_t *s);
int piko_rt_stream_destroy(piko_stream_t s);
int piko_rt_event_create(piko_event_t *e);
int piko_rt_event_destroy(piko_event_t e);
int piko_rt_event_record(piko_event_t e, piko_stream_t s);
int piko_rt_event_elapsed(piko_event_t start, piko_event_t stop, float *ms);
int piko_rt_mem_info(size_t *free_bytes, size_t *total_bytes);
int piko_rt_shutdown(void);                                       /* cuCtxDestroy, :840-854 */
const char *piko_rt_error_string(int status);                     /* getCudaDrvErrorString */

#ifdef __cplusplus
}
#endif
#endif
