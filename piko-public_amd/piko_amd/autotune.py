"""Per-scene-class autotuner for the raster pipeline.

A scene class is (log2 of the triangle count, log2 of the pixel-count edge).
For one representative scene per class, every candidate is timed on the GPU
with the plain pipe; the fastest wins.  Candidates are the bin (tile) size T
and, for the binned raster, the persistent raster waves per CU (the queue
path's persistent grid and part size are tuned for batches by
tune_queue_batch).  The queue
depth recorded for the class is the measured (triangle, bin) pairs per
triangle at the chosen T, with 25% headroom, so binned batches can size their
queues up front instead of growing them after an overflow.
"""
import json
import math
import time

import numpy as np

from . import RASTER_AUTO, RASTER_BINNED, RASTER_QUEUE, RasterPipe

BIN_SIZES = (8, 16, 32, 64)
WAVES_PER_CU = (8, 16, 24)
TILE_SIZES = (16, 32, 64)  # the tiled raster's tile edges (8 is drawn in 16-pixel tiles)


def candidates(raster_mode):
    """(bin sizes, waves per CU) worth timing for a raster mode: the waves only
    shape the binned raster's persistent grid."""
    if raster_mode == RASTER_BINNED:
        return BIN_SIZES, WAVES_PER_CU
    if raster_mode == RASTER_QUEUE:
        return (32, 64), (16,)
    return TILE_SIZES, (0,)


def scene_class(n, W, H):
    return (int(round(math.log2(max(n, 1)))), int(round(math.log2(max(W * H, 1)) / 2)))


def time_plain(scene, T, waves, S=8, frames=4, raster_mode=RASTER_AUTO):
    with RasterPipe().allocate(scene, T, S, waves_per_cu=waves, raster_mode=raster_mode) as pipe:
        pipe.prepare()
        pipe.run_frames(1)
        t0 = time.perf_counter()
        pipe.run_frames(frames)
        dt = (time.perf_counter() - t0) / frames
        st = pipe.stats()
    return dt, st["pairs"] / max(scene.n, 1)


def tune(scene, bin_sizes=None, waves=None, S=8, raster_mode=RASTER_AUTO, modes=None):
    """Times every candidate of `raster_mode` (or of each mode in `modes`) on
    `scene`.  Returns the fastest setting overall ("raster_mode", "bin_size",
    "waves_per_cu", "queue_depth", "ms") and, per mode, its fastest bin size
    ("best_bin_size": {mode: T})."""
    best = None
    per_mode = {}
    trials = []
    for mode in (modes or (raster_mode,)):
        cb, cw = candidates(mode)
        for T in bin_sizes or cb:
            for w in waves or cw:
                dt, qd = time_plain(scene, T, w, S, raster_mode=mode)
                trials.append((mode, T, w, dt))
                if mode not in per_mode or dt < per_mode[mode][1]:
                    per_mode[mode] = (T, dt)
                if best is None or dt < best[3]:
                    best = (mode, T, w, dt, qd)
    mode, T, w, dt, qd = best
    return {"raster_mode": mode, "bin_size": T, "waves_per_cu": w, "queue_depth": round(1.25 * qd + 0.05, 3),
            "ms": dt * 1e3, "best_bin_size": {m: v[0] for m, v in per_mode.items()}, "trials": trials}


def autotune(scene_list, S=8, log=None, raster_mode=RASTER_AUTO, modes=None):
    """Returns (per-scene settings list, class table).  The representative of
    a class is its scene with the median triangle count.  With `modes`, the
    raster mode is tuned too."""
    groups = {}
    for k, sc in enumerate(scene_list):
        groups.setdefault(scene_class(sc.n, sc.W, sc.H), []).append(k)
    table = {}
    for cls, members in sorted(groups.items()):
        members = sorted(members, key=lambda k: scene_list[k].n)
        rep = scene_list[members[len(members) // 2]]
        table[cls] = tune(rep, S=S, raster_mode=raster_mode, modes=modes)
        if log:
            log("class %s (%d scenes): mode %d T=%d waves/CU=%d queue depth %.2f (%.3f ms)" % (
                cls, len(members), table[cls]["raster_mode"], table[cls]["bin_size"], table[cls]["waves_per_cu"],
                table[cls]["queue_depth"], table[cls]["ms"]))
    per_scene = [table[scene_class(sc.n, sc.W, sc.H)] for sc in scene_list]
    return per_scene, table


def tune_batch(scene_list, pos, attr, parts, S=8, tile_sizes=TILE_SIZES, frames=3, log=None):
    """Tile size of the batched (tiled) form per scene class, timed in the
    batch itself: each class's scenes are run as one sub-batch at every
    candidate tile size and the fastest wins.  `pos`/`attr`/`parts` are
    piko_amd.concat_scenes(scene_list).  Returns (per-scene tile sizes,
    {class: (T, ms)})."""
    from . import BatchPipe, ConstantState, RASTER_TILED
    groups = {}
    for k, sc in enumerate(scene_list):
        groups.setdefault(scene_class(sc.n, sc.W, sc.H), []).append(k)
    kind = scene_list[0].kind if scene_list else "pretransformed"
    ntris = sum(sc.n for sc in scene_list)
    npix = sum(sc.W * sc.H for sc in scene_list)
    table = {}
    bp = BatchPipe(pos, attr, kind, subpixel_bits=S, max_jobs=max(len(scene_list), 1), max_tris=ntris,
                   max_pixels=npix, raster_mode=RASTER_TILED)
    try:
        for cls, members in sorted(groups.items()):
            best = None
            for T in tile_sizes:
                bp.submit([(parts[k][0], parts[k][1], T, ConstantState.for_scene(scene_list[k])) for k in members])
                bp.wait()
                bp.run_frames(1)
                t0 = time.perf_counter()
                bp.run_frames(frames)
                dt = (time.perf_counter() - t0) / frames
                if best is None or dt < best[1]:
                    best = (T, dt)
            table[cls] = (best[0], best[1] * 1e3)
            if log:
                log("batch class %s (%d scenes): T=%d (%.3f ms)" % (cls, len(members), best[0], best[1] * 1e3))
    finally:
        bp.destroy()
    per_scene = [table[scene_class(sc.n, sc.W, sc.H)][0] for sc in scene_list]
    return per_scene, table


def tune_queue_batch(scene_list, pos, attr, parts, S=8, tile_sizes=(32, 64), waves=(16, 24),
                     part_pairs=(2048, 65536), class_tiles=(16, 32, 64), class_covers=None, frames=3, log=None):
    """The queue path's batch settings, timed in the batch itself.  First one
    tile size for every job with the persistent grid's waves per CU and the
    queue entries per part item (the raster is one persistent launch over
    all of the batch's tile parts, so those two are per batch); then, per
    scene class (largest first), each other tile size in `class_tiles` and
    each other coverage mode in `class_covers` (COVER_WALK / COVER_SPANS;
    default both) for that class's jobs, kept when the whole batch gets at
    least 2% faster.  Returns ({"bin_size": the common tile size,
    "waves_per_cu", "part_pairs", "tiles": per-scene tile sizes, "covers":
    per-scene coverage modes (0: the batch's), "class_bin_size" and
    "class_cover": {class: value} where it differs, "ms"}, trials)."""
    from . import BatchPipe, ConstantState, RASTER_QUEUE, COVER_WALK, COVER_SPANS
    if class_covers is None:
        class_covers = (COVER_WALK, COVER_SPANS)
    kind = scene_list[0].kind if scene_list else "pretransformed"
    ntris = sum(sc.n for sc in scene_list)
    npix = sum(sc.W * sc.H for sc in scene_list)
    best, trials = None, []

    def timed(bp, tiles, covers=None):
        covers = covers or [0] * len(tiles)
        bp.submit([(n, base, T, ConstantState.for_scene(sc), 0, 0.0, cv)
                   for (n, base, sc), T, cv in zip(parts, tiles, covers)])
        bp.wait()
        bp.run_frames(1)
        t0 = time.perf_counter()
        bp.run_frames(frames)
        return (time.perf_counter() - t0) / frames

    for w in waves:
        for pp in part_pairs:
            bp = BatchPipe(pos, attr, kind, subpixel_bits=S, max_jobs=max(len(scene_list), 1), max_tris=ntris,
                           max_pixels=npix, raster_mode=RASTER_QUEUE, part_pairs=pp, waves_per_cu=w)
            try:
                for T in tile_sizes:
                    dt = timed(bp, [T] * len(scene_list))
                    trials.append((T, w, pp, dt * 1e3))
                    if log:
                        log("queue batch T=%d waves/CU=%d part pairs=%d: %.3f ms" % (T, w, pp, dt * 1e3))
                    if best is None or dt < best[3]:
                        best = (T, w, pp, dt)
            finally:
                bp.destroy()
    T, w, pp, dt = best
    tiles = [T] * len(scene_list)
    groups = {}
    for k, sc in enumerate(scene_list):
        groups.setdefault(scene_class(sc.n, sc.W, sc.H), []).append(k)
    order = sorted(groups, key=lambda c: -sum(scene_list[k].n * 48 + scene_list[k].W * scene_list[k].H * 8
                                              for k in groups[c]))
    class_T, class_cv = {}, {}
    covers = [0] * len(scene_list)
    if (class_tiles and len(groups) > 1) or class_covers:
        bp = BatchPipe(pos, attr, kind, subpixel_bits=S, max_jobs=max(len(scene_list), 1), max_tris=ntris,
                       max_pixels=npix, raster_mode=RASTER_QUEUE, part_pairs=pp, waves_per_cu=w)
        try:
            dt = timed(bp, tiles, covers)
            for cls in order:
                for alt in (class_tiles if len(groups) > 1 else ()):
                    if alt == tiles[groups[cls][0]]:
                        continue
                    trial = list(tiles)
                    for k in groups[cls]:
                        trial[k] = alt
                    dt2 = timed(bp, trial, covers)
                    trials.append((("class",) + tuple(cls), alt, w, pp, dt2 * 1e3))
                    if dt2 < 0.98 * dt:
                        tiles, dt = trial, dt2
                        class_T[cls] = alt
                        if log:
                            log("queue batch class %s: T=%d (%.3f ms)" % (cls, alt, dt * 1e3))
                for cv in class_covers:
                    trial = list(covers)
                    for k in groups[cls]:
                        trial[k] = cv
                    dt2 = timed(bp, tiles, trial)
                    trials.append((("class",) + tuple(cls), "cover", cv, w, pp, dt2 * 1e3))
                    if dt2 < 0.98 * dt:
                        covers, dt = trial, dt2
                        class_cv[cls] = cv
                        if log:
                            log("queue batch class %s: cover %d (%.3f ms)" % (cls, cv, dt * 1e3))
        finally:
            bp.destroy()
    return {"bin_size": T, "waves_per_cu": w, "part_pairs": pp, "tiles": tiles, "covers": covers,
            "class_bin_size": {"%d,%d" % c: t for c, t in class_T.items()},
            "class_cover": {"%d,%d" % c: v for c, v in class_cv.items()}, "ms": dt * 1e3}, trials


def save(table, path):
    with open(path, "w") as f:
        json.dump({"%d,%d" % k: {kk: (vv if kk != "best_bin_size" else {str(m): t for m, t in vv.items()})
                                 for kk, vv in v.items() if kk != "trials"} for k, v in table.items()}, f, indent=1)


__all__ = ["scene_class", "candidates", "tune", "autotune", "tune_batch", "tune_queue_batch", "save", "np"]
