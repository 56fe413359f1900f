"""Autotuner for the batched substep: kernel shape, workgroups per CU, substeps per graph.

The reference has nothing to tune: Warp launches one thread per element
(flatworld/femspringmanager.py:947-967).  Here three choices change the speed of
the same computation and none changes its result (every nodal sum has a fixed
order whatever the tiling or launch shape):

* kernel shape -- threads per workgroup, element slots per tile, owned nodes per
  tile (the compiled shapes of the substep kernel); bigger tiles evaluate fewer
  halo elements, smaller ones fit more workgroups per CU;
* workgroups per CU -- one workgroup per tile (the hardware schedules tiles), or
  a persistent grid of ``k`` workgroups per CU looping over tiles;
* substeps per graph -- how many substep launches one captured hipGraph replays.

``tune(specs)`` times every candidate on a size-stratified sample of the scenes
(the same scenes for every candidate, so the comparison is fair) and returns the
fastest; results are cached in a JSON file keyed by the workload signature (scene
size histogram, material mix, ground kinds, device), so a serving process tunes
once per kind of traffic.
"""

import json
import os
import time

import numpy as np

from . import _native

# (threads, element slots, owned nodes per tile) of the compiled kernel shapes
SHAPES = [
    (256, 1024, 512),
    (256, 768, 384),
    (512, 1024, 512),
    (256, 512, 256),
    (512, 2048, 1024),
    (512, 2048, 512),
    (1024, 2048, 1024),
]
GRID_PER_CU = [0, 4]  # 0: one workgroup per tile; 4: persistent grid
FRAMES = [30, 60, 120]

CACHE_ENV = "FLATWORLD_AMD_TUNE_CACHE"


def default_cache_path():
    base = os.environ.get(CACHE_ENV)
    if base:
        return base
    return os.path.join(os.path.expanduser("~"), ".cache", "flatworld_amd", "tune.json")


def signature(specs, device_name=""):
    """Workload key: log2 histogram of element counts, material and ground mix."""
    ne = np.array([s.n_elems for s in specs], dtype=np.float64)
    hist = np.bincount(np.clip(np.log2(np.maximum(ne, 1)).astype(int), 0, 24), minlength=25)
    frac = hist / max(len(specs), 1)
    mats = np.bincount([int(s.mat_type) for s in specs], minlength=4)[1:] / max(len(specs), 1)
    grounds = np.bincount([g.kind for s in specs for g in s.grounds], minlength=4)[1:]
    grounds = grounds / max(grounds.sum(), 1)
    parts = [
        "n%d" % int(round(np.log2(max(len(specs), 1)))),
        "h" + ",".join("%d" % int(round(10 * f)) for f in frac),
        "m" + ",".join("%d" % int(round(10 * f)) for f in mats),
        "g" + ",".join("%d" % int(round(10 * f)) for f in grounds),
        device_name,
    ]
    return "|".join(parts)


def stratified_sample(specs, max_elems, seed=0):
    """About ``max_elems`` triangles of scenes spread evenly over the size order.

    Systematic sampling of the size-sorted scene list keeps the shape of the
    size histogram (every size range contributes in proportion to its count).
    """
    total = sum(s.n_elems for s in specs)
    if total <= max_elems:
        return list(specs)
    order = np.argsort([s.n_elems for s in specs], kind="stable")
    n_pick = max(1, int(round(len(specs) * max_elems / total)))
    u = np.random.default_rng(seed).uniform()
    idx = np.minimum(((np.arange(n_pick) + u) * (len(specs) / n_pick)).astype(int), len(specs) - 1)
    return [specs[i] for i in order[idx]]


def _time_candidate(sample, block, slots, nodes, gpc, frames, reps, capacity):
    arena = _native.Arena(capacity_bytes=capacity, tile_nodes=nodes, tile_slots=slots, block=block, grid_per_cu=gpc)
    try:
        for s in sample:
            arena.add(s, track_disp=True)
        arena.step(1, -1.0, 0)  # warm-up substep of every scene
        arena.sync()
        out = {}
        for f in frames:
            _graph_time(arena, f, 1)  # capture and first replay
            out[f] = _graph_time(arena, f, reps) / (f * reps)
        return out, arena.info()
    finally:
        arena.close()


def _graph_time(arena, frame, reps):
    """Milliseconds for ``reps`` replays of a ``frame``-substep graph."""
    arena.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        arena.step_graph(frame, 0)
    arena.sync()
    return (time.perf_counter() - t0) * 1e3


def tune(specs, cache_path=None, max_elems=4_000_000, reps=3, shapes=None, grid_per_cu=None, frames=None,
         use_cache=True, log=None):
    """Pick the fastest (shape, workgroups per CU, frame) for ``specs``.

    Returns a dict with tile_nodes, tile_slots, block, grid_per_cu, frame and the
    measured table.  The measurement replays hipGraph frames on a sample of the
    scenes; the cache maps the workload signature to the winner.
    """
    shapes = shapes or SHAPES
    grid_per_cu = grid_per_cu if grid_per_cu is not None else GRID_PER_CU
    frames = frames or FRAMES
    key = signature(specs)
    path = cache_path or default_cache_path()
    if use_cache and os.path.exists(path):
        try:
            with open(path) as fh:
                cache = json.load(fh)
            if key in cache:
                hit = dict(cache[key])
                hit["cached"] = True
                return hit
        except (OSError, ValueError):
            pass
    sample = stratified_sample(specs, max_elems)
    e_sample = sum(s.n_elems for s in sample)
    # arena sized for the sample: under 256 B per triangle plus per-array alignment per scene
    capacity = (256 * e_sample + 16384 * len(sample) + (64 << 20)) // 256 * 256
    table = []
    best = None
    for block, slots, nodes in shapes:
        for gpc in grid_per_cu:
            t0 = time.perf_counter()
            per_step, info = _time_candidate(sample, block, slots, nodes, gpc, frames, reps, capacity)
            for f, ms in per_step.items():
                row = {"block": block, "tile_slots": slots, "tile_nodes": nodes, "grid_per_cu": gpc, "frame": f,
                       "ms_per_substep": ms, "elem_substeps_per_s": e_sample / (ms * 1e-3),
                       "tiles": info["n_tiles"], "tiles_per_cu": info["tiles_per_cu"]}
                table.append(row)
                if best is None or ms < best["ms_per_substep"]:
                    best = row
            if log:
                log(f"[tune] block={block} slots={slots} nodes={nodes} grid_per_cu={gpc}: "
                    + " ".join(f"f{f}={ms * 1e3:.1f}us" for f, ms in per_step.items())
                    + f" ({time.perf_counter() - t0:.1f}s)")
    result = {
        "tile_nodes": best["tile_nodes"],
        "tile_slots": best["tile_slots"],
        "block": best["block"],
        "grid_per_cu": best["grid_per_cu"],
        "frame": best["frame"],
        "sample_scenes": len(sample),
        "sample_triangles": e_sample,
        "signature": key,
        "table": table,
        "cached": False,
    }
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            cache = {}
            if os.path.exists(path):
                with open(path) as fh:
                    cache = json.load(fh)
            cache[key] = {k: v for k, v in result.items() if k not in ("table", "cached")}
            tmp = path + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(cache, fh, indent=1)
            os.replace(tmp, path)
        except (OSError, ValueError):
            pass
    return result


def arena_kwargs(result):
    """Arena / FemBatch keyword arguments of a tuning result."""
    return {k: result[k] for k in ("tile_nodes", "tile_slots", "block", "grid_per_cu")}


__all__ = ["tune", "signature", "stratified_sample", "arena_kwargs", "SHAPES", "GRID_PER_CU", "FRAMES"]
