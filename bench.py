"""Throughput of the Piko raster pipeline on one MI355X.

Workload: the GPU-scale raster configuration -- a 2,000,000-triangle
height field (1000x1000 grid, 6-octave value noise from 16 lattice cells per
unit, amplitude 0.2, seed 2) rendered at 1920x1080 with a perspective camera
(60 deg) low over the terrain (mean depth complexity 3.94 over 99.8% of the
screen), per-vertex normals and one directional Lambert light, 8 subpixel
bits, FP32 depth, raster path and bin size swept (queue, tiled, direct,
binned; 8/16/32/64).  One step = one frame (vertex stage, binning, raster,
depth, Gouraud, framebuffer) with the geometry resident in HBM.

Prints one JSON line.  `value` is Mtri/s at the fastest bin size.  The same
line carries the batched form (256 uneven scenes as one batch -- a bin
launch, one persistent raster launch and a merge -- vs sequential plain
launches, autotuned per scene class) under "batch" and the
served form (resident scenes, Poisson request stream, continuous batching,
p50/p99 latency) under "serve", the served form at saturation against
sequential plain launches of the same requests under "serve_saturated", and
the Reyes pipeline (split, dice, sample, resolve of 4,096 Bezier patches at
720p, 4 spp) under "reyes".
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "piko-public_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
SERVE_BUDGET = 12 << 30  # device bytes one serving batch may use
# fragments (covered pixel samples before the depth test) and covered pixels
# of the 2M-triangle height field at 1080p, from the oracle's coverage
# counts (profiles/heightfield_depth_complexity_sweep.md)
HF_FRAGMENTS = 8159414
HF_COVERED = 2069833


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_info():
    """The CPU the baselines ran on: lscpu's model name (with its family /
    model / stepping numbers and base/max clock where lscpu reports them),
    nproc (the CPUs this process may use) and the logical CPU count."""
    model = None
    extra = {}
    keys = {"CPU family:": "family", "Model:": "model_number", "Stepping:": "stepping",
            "CPU max MHz:": "max_mhz", "Socket(s):": "sockets", "Thread(s) per core:": "threads_per_core"}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
            for k, name in keys.items():
                if line.startswith(k):
                    extra[name] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        nproc = len(os.sched_getaffinity(0))
    info = {"model": model or cpu_model(), "nproc": nproc, "logical_cpus": os.cpu_count()}
    info.update(extra)
    return info


def cpu_baseline(scene, T, S, budget_s=12.0):
    """The oracle (single-threaded restatement of the reference CPU backend)
    timed on this host on whole frames of the same workload."""
    from oracle import oracle
    oracle.build()
    times = []
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        oracle.frame(scene, S, T, threads=1)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all > budget_s or len(times) >= 10:
            break
    times.sort()
    med = times[len(times) // 2]
    h = host_info()
    return {"value": round(scene.n / med / 1e6, 4), "unit": "Mtri/s", "cores": 1, "kind": "port",
            "threads": 1, "host": h,
            "sample": "%d full frames of the same 2M-triangle 1080p scene (bin %d), median %.3f s/frame, "
                      "single-threaded oracle, host %s (nproc %d)" % (len(times), T, med, h["model"], h["nproc"])}


def cpu_reference_schedule(scene, T, S, budget_s=6.0):
    """The oracle on the reference CPU backend's schedule: the raster bins
    block-partitioned over host threads (CPUBackend.cpp:534-557, which
    hard-codes 64 threads), at 64 threads and at this host's CPU share
    (OMP_NUM_THREADS, else the core count).  The vertex stage and bin walk
    stay serial, as kernel0 is in the reference."""
    from oracle import oracle
    oracle.build()
    share = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
    runs = []
    for threads in sorted({64, share}, reverse=True):
        times = []
        t_all = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            oracle.frame(scene, S, T, threads=threads)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_all > budget_s or len(times) >= 10:
                break
        times.sort()
        med = times[len(times) // 2]
        runs.append({"threads": threads, "value": round(scene.n / med / 1e6, 4), "unit": "Mtri/s",
                     "frames": len(times), "median_s": round(med, 4)})
    return {"kind": "port", "bin_size": T, "runs": runs, "host": host_info(),
            "sample": "whole frames of the same scene; raster bins over host threads"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE measurements (profiles/traffic.json,
    each counter scaled by its reading on a copy of known size), or None."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f)[workload][kernel]
    except (OSError, KeyError, ValueError):
        return None, None
    return int(entry["bytes_per_launch"]), "profiles/traffic.json: %s" % entry["summary"]


def copy_calibration(mb):
    """A plain device-to-device copy of known size through the runtime shim:
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this run give the counters'
    reading for a known byte count."""
    import ctypes
    import piko_amd
    L = piko_amd.lib()
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if L.piko_rt_malloc(ctypes.byref(ptr), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("malloc failed")
    L.piko_rt_memset_d32.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_size_t]
    L.piko_rt_memcpy_dtod.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    L.piko_rt_memset_d32(a, 7, n // 4)
    for _ in range(5):
        L.piko_rt_memcpy_dtod(b, a, n)
    L.piko_rt_sync(None)
    L.piko_rt_free(a)
    L.piko_rt_free(b)
    print(json.dumps({"copy_bytes": n, "copies": 5}), flush=True)


def bench_batch(log, reps=5):
    """256 scenes from seeds 1000-1255 (uniform soups, log-uniform sizes and
    resolutions), autotuned; one batch vs sequential plain pipes.  The batch
    runs on the tiled path (tile size per scene class, tuned in the batch) and
    on the queue path (one persistent raster launch over every job's tile
    parts; tile size, waves per CU and part size tuned in the batch)."""
    import numpy as np

    import piko_amd
    from piko_amd import autotune, scenes

    spec = scenes.batch_spec()
    scs = [scenes.tris_uniform(seed, n, W, H) for seed, n, W, H in spec]
    ntris = sum(s.n for s in scs)
    npix = sum(s.W * s.H for s in scs)
    log("batch: %d scenes, %d triangles, %d pixels; autotuning" % (len(scs), ntris, npix))
    # plain pipes: the fastest mode and bin size per class
    per, table = autotune.autotune(scs, log=None, modes=(piko_amd.RASTER_DIRECT, piko_amd.RASTER_TILED,
                                                         piko_amd.RASTER_QUEUE))
    pipes = []
    for s, cfg in zip(scs, per):
        pipes.append(piko_amd.RasterPipe().allocate(s, cfg["bin_size"], 8, waves_per_cu=cfg["waves_per_cu"],
                                                    pair_capacity=int(cfg["queue_depth"] * s.n) + 64,
                                                    raster_mode=cfg["raster_mode"]))
    for p in pipes:
        p.run_frames(1)
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in pipes:
            p.run_single()
    t_seq = (time.perf_counter() - t0) / reps
    plain_cols = [p.pikoScreen.getData() for p in pipes]
    for p in pipes:
        p.destroy()
    pos, attr, parts = piko_amd.concat_scenes(scs)

    def run_batch(mode, tiles, covers=None, **kw):
        bp = piko_amd.BatchPipe(pos, attr, "pretransformed", max_jobs=len(scs), max_tris=ntris, max_pixels=npix,
                                raster_mode=mode, **kw)
        covers = covers or [0] * len(tiles)
        try:
            bp.submit([(n, base, T, piko_amd.ConstantState.for_scene(s), 0, 0.0, cv)
                       for (n, base, s), T, cv in zip(parts, tiles, covers)])
            bp.wait()
            bp.run_frames(1)
            t0 = time.perf_counter()
            bp.run_frames(reps)
            dt = (time.perf_counter() - t0) / reps
            kernels, _ = bp.profile_frames(3)
            exact = sum(1 for k in range(len(scs)) if np.array_equal(bp.color(k), plain_cols[k]))
        finally:
            bp.destroy()
        return dt, kernels, exact

    tiles, btable = autotune.tune_batch(scs, pos, attr, parts)
    t_batch, k_tiled, exact = run_batch(piko_amd.RASTER_TILED, tiles)
    qcfg, _ = autotune.tune_queue_batch(scs, pos, attr, parts)
    t_queue, k_queue, exact_q = run_batch(piko_amd.RASTER_QUEUE, qcfg["tiles"], qcfg["covers"],
                                          part_pairs=qcfg["part_pairs"], waves_per_cu=qcfg["waves_per_cu"])
    qcfg["tiles"] = {str(T): qcfg["tiles"].count(T) for T in sorted(set(qcfg["tiles"]))}
    qcfg["covers"] = {str(c): qcfg["covers"].count(c) for c in sorted(set(qcfg["covers"]))}
    out = {"workload": "256 uniform-soup scenes, seeds 1000-1255, N log-uniform [1K,500K], "
                       "pixels log-uniform [256^2,2048^2], aspect {1,4:3,16:9}, autotuned (plain pipes: raster "
                       "path and bin size per class; batch: tile size, waves per CU, part size and coverage mode, timed in "
                       "the batch)",
           "form": "queue path: bin launch + one persistent raster launch over every job's tile parts "
                   "(+ merge of split tiles), frames of one submitted job set",
           "scenes": len(scs), "triangles": ntris, "pixels": npix, "classes": len(table),
           "seq_ms_per_batch": round(t_seq * 1e3, 3), "batch_ms": round(t_queue * 1e3, 3),
           "seq_frames_per_s": round(len(scs) / t_seq, 1), "batch_frames_per_s": round(len(scs) / t_queue, 1),
           "speedup": round(t_seq / t_queue, 2), "batch_Mtri_per_s": round(ntris / t_queue / 1e6, 1),
           "plain_modes": {m: sum(1 for c in per if c["raster_mode"] == v)
                           for m, v in (("direct", piko_amd.RASTER_DIRECT), ("tiled", piko_amd.RASTER_TILED),
                                        ("queue", piko_amd.RASTER_QUEUE))},
           "tuned": qcfg,
           "batch_kernels_ms": [[n, round(m, 4)] for n, m in k_queue],
           "bit_exact_vs_sequential": exact_q == len(scs), "scenes_compared": len(scs), "scenes_equal": exact_q,
           "tiled": {"form": "tiled path: bin launch + one raster launch per tile-size class, one workgroup per "
                             "tile",
                     "batch_ms": round(t_batch * 1e3, 3), "speedup": round(t_seq / t_batch, 2),
                     "frames_per_s": round(len(scs) / t_batch, 1),
                     "batch_tile_sizes": {str(T): tiles.count(T) for T in sorted(set(tiles))},
                     "kernels_ms": [[n, round(m, 4)] for n, m in k_tiled],
                     "bit_exact_vs_sequential": exact == len(scs), "scenes_equal": exact}}
    log("batch: %s" % out)
    return out


def bench_batch_queue(log, T, waves, part_pairs, steps, warmup):
    """The 256-scene batch on the queue path alone, at fixed settings: one
    submit, `warmup` untimed frames, `steps` timed ones.  Each frame is one
    bin launch, one persistent raster launch and (when a tile is split) one
    merge launch; the submit adds the counting pass of the first frame."""
    import piko_amd
    from piko_amd import scenes

    spec = scenes.batch_spec()
    scs = [scenes.tris_uniform(seed, n, W, H) for seed, n, W, H in spec]
    ntris = sum(s.n for s in scs)
    npix = sum(s.W * s.H for s in scs)
    pos, attr, parts = piko_amd.concat_scenes(scs)
    log("batch-queue: %d scenes, %d triangles, %d pixels, T=%d, %d waves/CU, parts of %d" % (
        len(scs), ntris, npix, T, waves, part_pairs))
    with piko_amd.BatchPipe(pos, attr, "pretransformed", max_jobs=len(scs), max_tris=ntris, max_pixels=npix,
                            raster_mode=piko_amd.RASTER_QUEUE, part_pairs=part_pairs, waves_per_cu=waves) as bp:
        bp.submit([(n, base, T, piko_amd.ConstantState.for_scene(s)) for (n, base, s) in parts])
        bp.wait()
        bp.run_frames(max(warmup, 1))
        t0 = time.perf_counter()
        bp.run_frames(steps)
        dt = (time.perf_counter() - t0) / steps
    return {"metric": "batch frames/s (256 scenes)", "workload": "batch-queue", "scenes": len(scs),
            "triangles": ntris, "pixels": npix, "bin_size": T, "waves_per_cu": waves, "part_pairs": part_pairs,
            "steps": steps, "warmup": warmup, "batch_ms": round(dt * 1e3, 3),
            "frames_per_s": round(len(scs) / dt, 1)}


def bench_serve(log, ds, nreq, rate_hz=2000.0):
    """64 resident world-space scenes (soups and height fields, `ds`, built
    in device memory), Poisson requests with random cameras and
    540p/720p/1080p screens, served with continuous batching."""
    from piko_amd import serve
    gb = ds.nbytes / 1e9
    log("serve: 64 scenes, %.2f GB resident, %d requests at %.0f/s" % (gb, nreq, rate_hz))
    reqs, _ = serve.poisson_trace(nreq, rate_hz, 64, seed=11, bin_size=8)
    warm, _ = serve.poisson_trace(64, 1e6, 64, seed=12, bin_size=8)
    with serve.Server(ds, pipes=4, max_jobs=48, budget_bytes=SERVE_BUDGET) as srv:
        srv.replay(warm)
        _, st, _ = srv.replay(reqs)
    out = {"workload": "64 resident world-space scenes (triangle soups + height fields, %.1f GB), Poisson "
                       "%.0f req/s, camera radius U[2,5], FOV U[40,90], {540p,720p,1080p}" % (gb, rate_hz),
           "resident_gb": round(gb, 2), "triangles_per_scene": int(ds.count.mean()), "requests": nreq,
           "frames_per_s": round(st["frames_per_s"], 1), "p50_ms": round(st["p50_ms"], 3),
           "p99_ms": round(st["p99_ms"], 3), "batches": st["batches"], "mean_batch": round(st["mean_batch"], 2)}
    log("serve: %s" % out)
    return out


def bench_serve_saturated(log, ds, nreq=512):
    """The served form at saturation (every request has arrived before the
    first batch ends) against sequential plain launches of the same requests:
    one plain pipe per (scene, resolution) over the same resident geometry,
    prepare + run_single per request in arrival order."""
    import numpy as np

    import piko_amd
    from piko_amd import serve
    reqs, _ = serve.poisson_trace(nreq, 1e7, 64, seed=13, bin_size=8)
    log("serve (saturated): %d requests at once over 64 resident scenes (%.1f GB)" % (nreq, ds.nbytes / 1e9))
    with serve.Server(ds, pipes=4, max_jobs=48, budget_bytes=SERVE_BUDGET) as srv:
        # steady state: one untimed replay of the trace sizes every pipe's
        # buffers for it; then the median of three timed replays
        srv.replay(reqs)
        runs = sorted((srv.replay(reqs)[1] for _ in range(3)), key=lambda s: s["frames_per_s"])
        st = runs[1]
        # the frames to compare come from a second, untimed replay of the same
        # trace: reading framebuffers back inside the timed loop would stall it
        _, _, cols = srv.replay(reqs, keep=True)
    pipes = {}
    try:
        def pipe_of(r):
            key = (r.scene, r.cs.screenSizeX, r.cs.screenSizeY)
            if key not in pipes:
                b, n = int(ds.base[r.scene]), int(ds.count[r.scene])
                pipes[key] = piko_amd.RasterPipe().allocate_device(ds.pos_ptr + 36 * b, ds.nrm_ptr + 36 * b, n, r.cs,
                                                                   bin_size=8, raster_mode=piko_amd.RASTER_DIRECT)
            return pipes[key]
        for r in reqs:
            pipe_of(r)
        for p in pipes.values():
            p.run_single()
        passes = []
        for _ in range(3):  # the median of three passes, as for the served form
            t0 = time.perf_counter()
            for r in reqs:
                p = pipe_of(r)
                p.prepare(cs=r.cs)
                p.run_single()
            passes.append(time.perf_counter() - t0)
        t_seq = sorted(passes)[1]
        exact = 0
        for k, r in enumerate(reqs):  # every request against its plain launch
            p = pipe_of(r)
            p.prepare(cs=r.cs)
            p.run_single()
            exact += int(np.array_equal(p.pikoScreen.getData(), cols[k]))
    finally:
        for p in pipes.values():
            p.destroy()
    seq_fps = nreq / t_seq
    out = {"workload": "%d requests arriving at once (Poisson at 1e7 req/s) over 64 resident world-space scenes "
                       "(%.1f GB), camera radius U[2,5], FOV U[40,90], {540p,720p,1080p}" % (nreq, ds.nbytes / 1e9),
           "resident_gb": round(ds.nbytes / 1e9, 2),
           "requests": nreq, "served_frames_per_s": round(st["frames_per_s"], 1),
           "sequential_plain_frames_per_s": round(seq_fps, 1),
           "speedup": round(st["frames_per_s"] / seq_fps, 2), "p50_ms": round(st["p50_ms"], 3),
           "p99_ms": round(st["p99_ms"], 3), "batches": st["batches"], "mean_batch": round(st["mean_batch"], 2),
           "bit_exact_vs_plain": exact == nreq, "requests_compared": nreq, "requests_equal": exact}
    log("serve (saturated): %s" % out)
    return out


def bench_reyes(log, steps, warmup, cpu=True, sampling="auto", tile=0):
    """The Reyes configuration: 4,096 bicubic patches tiling a noisy sphere,
    split below 16 px, 8x8 dice, displacement, 4 samples, 1280x720 FP32.
    `sampling`: auto, global or tiled (piko_reyes_options.sampling)."""
    import piko_amd
    from piko_amd import scenes
    sc = scenes.sphere_patches()
    alg = scenes.algorithmic_bytes(sc)
    smode = {"auto": piko_amd.SAMPLE_AUTO, "global": piko_amd.SAMPLE_GLOBAL, "tiled": piko_amd.SAMPLE_TILED}[sampling]
    with piko_amd.ReyesPipe().allocate(sc, sampling=smode, tile=tile) as pipe:
        pipe.prepare()
        pipe.run_frames(max(warmup, 1))
        t0 = time.perf_counter()
        pipe.run_frames(steps)
        dt = (time.perf_counter() - t0) / steps
        kernels, frame_ms = pipe.profile_frames(max(5, min(steps, 20)))
        st = pipe.stats()
    dom = max(kernels, key=lambda k: k[1])
    out = {"workload": "reyes-sphere-4096-720p: 64x64 bicubic patches on a unit sphere, radial noise +-0.05 "
                       "(seed 3), split < 16 px, dice 8x8, displacement, 4 spp, 1280x720 FP32 RGBA",
           "sampling": "tiled" if st["tile_pairs"] > 0 else "global", "tile_pairs": st["tile_pairs"],
           "frames_per_s": round(1.0 / dt, 2), "ms_per_frame": round(dt * 1e3, 4),
           "event_frame_ms": round(frame_ms, 4), "patches": st["patches"], "leaves": st["leaves"],
           "micropolygons": st["micropolygons"], "split_levels": st["levels"],
           "Mmicropolygons_per_s": round(st["micropolygons"] / dt / 1e6, 1),
           "kernels_ms": [[n, round(m, 4)] for n, m in kernels],
           "algorithmic_bytes_per_frame": alg,
           "frame_GBs": round(alg / (frame_ms / 1e3) / 1e9, 2),
           "dominant_kernel": dom[0], "dominant_ms": round(dom[1], 4)}
    if cpu:
        from oracle import oracle
        oracle.build()
        times = []
        t_all = time.perf_counter()
        while len(times) < 5 and time.perf_counter() - t_all < 12.0:
            t0 = time.perf_counter()
            oracle.reyes_frame(sc)
            times.append(time.perf_counter() - t0)
        times.sort()
        med = times[len(times) // 2]
        out["cpu_baseline"] = {"value": round(1.0 / med, 4), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": "%d full frames, median %.3f s/frame, single-threaded oracle, host %s" % (
                                   len(times), med, cpu_model())}
    log("reyes: %s" % out)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bins", default="8,16,32,64", help="bin sizes of the binned raster mode")
    ap.add_argument("--modes", default="queue,tiled,direct,binned", help="raster modes to sweep")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="all", choices=["raster", "batch", "serve", "reyes", "all", "batch-queue"],
                    help="batch-queue: only the 256-scene batch on the queue path at --batch-queue settings "
                         "(for kernel traces: one bin, one raster and one merge launch per frame)")
    ap.add_argument("--batch-queue", default="64,24,65536",
                    help="tile size, waves per CU and part size of the batch-queue workload")
    ap.add_argument("--reyes-sampling", default="auto", choices=["auto", "global", "tiled"],
                    help="how the Reyes frame composites samples (auto = global)")
    ap.add_argument("--reyes-tile", type=int, default=0, help="tiled sampling's tile edge: 16 or 32 (0 = 16)")
    ap.add_argument("--serve-tris", type=int, default=8_700_000,
                    help="triangles per resident scene when serving (8.7M: ~40 GB over 64 scenes)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_mb:
        copy_calibration(args.copy_mb)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs a single MI355X (no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch-queue":
        T, w, pp = (int(x) for x in args.batch_queue.split(","))
        print(json.dumps(bench_batch_queue(log, T, w, pp, args.steps, args.warmup)), flush=True)
        return

    import piko_amd
    from piko_amd import scenes

    log("generating height field (2,000,000 triangles)")
    scene = scenes.heightfield()
    S = 8
    alg_bytes = scenes.algorithmic_bytes(scene)
    sweep = {}
    best = None
    runs = []
    modes = args.modes.split(",")
    if "tiled" in modes:  # tile edges of 16 and up (8 is rasterised in 16-pixel tiles)
        runs += [("tiled", int(t)) for t in args.bins.split(",") if int(t) >= 16]
    if "binned" in modes:
        runs += [("binned", int(t)) for t in args.bins.split(",")]
    if "direct" in modes:
        runs.append(("direct", 16))  # the bin size only shapes the on-demand queues
    if "queue" in modes:  # tiles of 32 and 64 (a 1080p screen at 8 or 16 is binned in 32-pixel tiles)
        runs += [("queue", int(t)) for t in args.bins.split(",") if int(t) >= 32]
    rmodes = {"binned": piko_amd.RASTER_BINNED, "direct": piko_amd.RASTER_DIRECT, "tiled": piko_amd.RASTER_TILED,
              "queue": piko_amd.RASTER_QUEUE}
    for mode, T in runs:
        rm = rmodes[mode]
        pipe = piko_amd.RasterPipe().allocate(scene, T, S, raster_mode=rm)
        pipe.prepare()
        pipe.run_frames(max(args.warmup, 1))
        t0 = time.perf_counter()
        pipe.run_frames(args.steps)  # synchronises at the end
        dt = (time.perf_counter() - t0) / args.steps
        kernels, frame_ms = pipe.profile_frames(max(5, min(args.steps, 20)))
        st = pipe.stats()
        key = "direct" if mode == "direct" else "%s-%d" % (mode, T)
        sweep[key] = {"ms_per_frame": round(dt * 1e3, 4), "event_frame_ms": round(frame_ms, 4),
                      "pairs": st["pairs"], "kernels_ms": [[n, round(m, 4)] for n, m in kernels]}
        log("%s: %.3f ms/frame (events %.3f ms), kernels %s" % (key, dt * 1e3, frame_ms, sweep[key]["kernels_ms"]))
        if best is None or dt < best[1]:
            best = (T, dt, kernels, frame_ms, st, mode)
        pipe.destroy()

    T, dt, kernels, frame_ms, st, mode = best
    # dominant kernel: the longest single launch of the frame (HIP events on
    # the pipe's stream); the raster kernel at every bin size measured so far
    dom, dom_ms = max(kernels, key=lambda k: k[1])
    dom_key = "%s<%d>" % (dom, T) if dom in ("k_raster", "k_tile_raster", "k_qraster") else dom
    if dom == "k_raster":
        # framebuffer written once (colour + depth), each emitted triangle's
        # setup (3 fixed-point xy pairs, 3 depths, 3 packed colours) read once,
        # each (triangle, bin) queue entry read once
        units = {"pixels": scene.W * scene.H, "emitted_triangles": st["emitted"], "queue_entries": st["pairs"]}
        alg_k = 8 * units["pixels"] + 48 * units["emitted_triangles"] + 4 * units["queue_entries"]
    elif dom in ("k_tile_bin", "k_vs_direct", "k_qbin"):
        # the input's vertex positions (3 x xyz f32) read once per triangle
        units = {"input_triangles": scene.n}
        alg_k = 36 * scene.n
    elif dom in ("k_tile_raster", "k_resolve", "k_qraster"):
        # the input's vertex attributes (normals 36 B, or packed colours 12 B)
        # read once per triangle, the framebuffer (colour + depth) written once
        b_attr = 36 if scene.kind == "world" else 12
        units = {"input_triangles": scene.n, "pixels": scene.W * scene.H}
        alg_k = b_attr * scene.n + 8 * scene.W * scene.H
    else:
        units = {"frame": 1}
        alg_k = alg_bytes
    achieved_k = alg_k / (dom_ms / 1e3) / 1e9
    traffic, traffic_src = measured_traffic(dom_key, "heightfield-2M-1080p")
    frame_s = frame_ms / 1e3
    achieved = alg_bytes / frame_s / 1e9
    out = {
        "metric": "Mtri/s + frames/s at 1080p, HBM GB/s vs peak",
        "value": round(scene.n / dt / 1e6, 3),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "frames_per_s": round(1.0 / dt, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64 edge functions + f32 shading",
        "data": "synthetic",
        "config": {"workload": "heightfield-2M-1080p raster (VS->bin->raster->depth<=->Gouraud)",
                   "triangles": scene.n, "W": scene.W, "H": scene.H, "subpixel_bits": S, "raster_mode": mode,
                   "bin_size": None if mode == "direct" else T,
                   # run_frames keeps two frames in flight on the queue path
                   # (two streams, two sets of frame buffers); the event times
                   # below are of one frame alone
                   "frames_in_flight": 2 if mode == "queue" else 1,
                   "emitted": st["emitted"], "pairs": st["pairs"],
                   "base_freq": scenes.HF_FREQ, "eye": list(scenes.HF_EYE), "target": list(scenes.HF_TARGET),
                   # the scene's fragment statistics (oracle coverage counts, recorded
                   # with the pose sweep in profiles/heightfield_depth_complexity_sweep.md)
                   "fragments": HF_FRAGMENTS, "covered_pixels": HF_COVERED,
                   "depth_complexity": round(HF_FRAGMENTS / HF_COVERED, 2),
                   "coverage": round(HF_COVERED / (scene.W * scene.H), 4),
                   "sweep": sweep},
        "roofline": {"bound": "hbm", "achieved": round(achieved_k, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_k / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom_key, "kernel_ms": round(dom_ms, 4),
                     "algorithmic_bytes_per_launch": alg_k, "units": units,
                     "traffic_source": traffic_src},
        "frame_roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "event_frame_ms": round(frame_ms, 4),
                           "launches": len(kernels), "algorithmic_bytes_per_frame": alg_bytes,
                           # the same bytes over the timed steps' wall clock per frame
                           "achieved_per_step": round(alg_bytes / dt / 1e9, 2),
                           "frac_per_step": round(alg_bytes / dt / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if not args.no_cpu:
        log("timing the CPU oracle on a bounded sample")
        out["cpu_baseline"] = cpu_baseline(scene, T, S)
        # the same oracle on the reference CPU backend's threaded schedule
        out["cpu_baseline"]["reference_schedule"] = cpu_reference_schedule(scene, T, S)
    if args.workload in ("batch", "all"):
        out["batch"] = bench_batch(log)
    if args.workload in ("serve", "all"):
        from piko_amd import serve
        log("building the resident scene set (64 scenes x %d triangles) in device memory" % args.serve_tris)
        ds = serve.DeviceScenes(64, args.serve_tris, log=log)
        try:
            out["serve"] = bench_serve(log, ds, args.serve_requests)
            out["serve_saturated"] = bench_serve_saturated(log, ds)
        finally:
            ds.free()
    if args.workload in ("reyes", "all"):
        out["reyes"] = bench_reyes(log, args.steps, args.warmup, cpu=not args.no_cpu, sampling=args.reyes_sampling,
                                   tile=args.reyes_tile)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
