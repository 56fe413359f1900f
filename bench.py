#!/usr/bin/env python3
"""Benchmark: explicit-FEM substep throughput on one MI355X.

Default workload: the batched mix -- 4096 independent scenes, grid side
log-uniform in [8, 128], Elastic / Neo-Hookean / J2 mix, one height field per
scene (24.08 M triangles, 12.40 M nodes).  A "step" is one substep of every
scene.  Inputs are resident in HBM before timing starts.  Before the timed
region the arena advances (untimed) until every block has reached its height
field, so the timed substeps run with ground contact active and J2 elements
yielding; the timed region is replays of one captured hipGraph whose length
divides --steps (at most the 60-substep frame).  config.timed_window states the
launch mode and the contact / yield fractions at the end of the window.

Printed JSON line:
  value            element-substeps/s of the batched, graph-replayed form
  per_scene        the same scenes advanced with one launch per scene per substep
                   (hbm_frac: algorithmic bytes per second over the HBM peak)
  roofline         algorithmic bytes per substep / average substep-kernel time,
                   measured live with HIP events on the launch stream; traffic =
                   HBM bytes per launch from the committed rocprof counter profile
                   of the same kernel shape and workload
  tuning           kernel shape and workgroups per CU chosen by the autotuner
                   (flatworld_amd.autotune) for this workload; its own pick of
                   substeps per graph is frame_tuned, while the timed frames hold the
                   60 substeps the 4096-scene mix is defined with (--frame)
  serving          a Poisson stream of jobs (grid side 16-256, random material and
                   ground, horizon 60-600 substeps) at 80% of the capacity measured
                   on a closed-loop burst: p50 / p99 request latency, throughput,
                   hbm_frac of the served work over the wall clock
  springs          4096 spring lattices (side log-uniform in [8, 128], one height field
                   each) batched the same way: spring-substeps/s, the spring substep's
                   hbm_frac, its own autotuned shape and a CPU baseline on a sample
  contact          2048 two-block loops with FEM-FEM contact on a plane: element-
                   substeps/s with the contact pass and the hash rebuilds included,
                   the contact + substep kernel time and a CPU baseline on a sample
  single_scene     the 1000x1000 Neo-Hookean sheet on a plane and the 512x512 J2 block on
                   256x256 voxels, each tuned, graph-replayed and timed the same way, with
                   its roofline fraction and the oracle on the same scene as CPU baseline
  cpu_baseline     the oracle (serial C restatement) on a bounded sample
roofline.frac uses the live kernel time (HIP events); roofline.frac_wall uses the
wall-clock time per substep of the timed region; roofline.copy_ceiling is the fastest
read + write stream the device sustains (a tuned copy), the practical ceiling beside
the 8 TB/s peak.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILES = {
    "batched": os.path.join(ROOT, "profiles", "substep_batched4096_traffic.json"),
    "springs": os.path.join(ROOT, "profiles", "substep_springs4096_traffic.json"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_specs(config, seed, n_scenes):
    from flatworld_amd import workloads as W

    if config == "batched":
        return W.batched_config(n_scenes=n_scenes, seed=seed)
    if config == "large":
        return [W.large_config(seed=seed)]
    if config == "j2voxel":
        return [W.j2_voxel_config(seed=seed)]
    if config == "small":
        return [W.small_config()]
    if config == "springs":
        return W.spring_batched_config(n_scenes=n_scenes, seed=seed)
    raise SystemExit(f"unknown config {config}")


def cpu_baseline(specs, budget_s=12.0, all_cores=True):
    """The oracle (serial C restatement, the checker) on a bounded sample of the same scenes.

    Timed twice: serially on one core, as Warp-CPU launches run, and with the
    sample's scenes spread over every host core (the scenes are independent and
    the C calls release the GIL), each for about ``budget_s / 2`` seconds.
    """
    from concurrent.futures import ThreadPoolExecutor

    from oracle.oracle import OracleScene

    order = np.random.Generator(np.random.PCG64(1)).permutation(len(specs))
    sample, elems = [], 0
    for i in order:
        sample.append(specs[i])
        elems += specs[i].n_elems
        if elems >= 400_000:
            break
    scenes = [OracleScene(s) for s in sample]
    for o in scenes:
        o.warmup()
        o.substep(o.spec.dt, o.spec.damping)  # first substep excluded from timing

    def serial(n_limit_s):
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < n_limit_s:
            for o in scenes:
                o.substep(o.spec.dt, o.spec.damping)
            n += 1
        return elems * n / (time.perf_counter() - t0), n

    one, n1 = serial(budget_s / 2)
    if not all_cores:
        return {"value": one, "unit": "element-substeps/s", "cores": 1, "kind": "port",
                "sample": f"{len(sample)} of {len(specs)} scenes ({elems} elements) x {n1} substeps, serial C oracle"}
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 16))
    groups = [scenes[k::cores] for k in range(cores)]

    def run_group(g, reps):
        for _ in range(reps):
            for o in g:
                o.substep(o.spec.dt, o.spec.damping)

    reps = max(1, n1)
    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda g: run_group(g, reps), groups))
        multi = elems * reps / (time.perf_counter() - t0)
    return {
        "value": one,
        "unit": "element-substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(sample)} of {len(specs)} scenes ({elems} elements) x {n1} substeps, serial C oracle",
        "all_cores": {"value": multi, "cores": cores,
                      "sample": f"same scenes x {reps} substeps, spread over {cores} threads"},
    }


def serving_bench(n_jobs, seed, arena_kw, frame, burst_jobs, tune_resident=24):
    """Capacity on a closed-loop burst, then a Poisson stream at 80% of it.

    With ``tune_resident`` > 0 the served arena's kernel shape and workgroups per CU
    come from the autotuner run on that many of the jobs: a served arena holds a few
    dozen jobs at a time, a different regime from the 4096-scene arena ``arena_kw``
    was tuned for (there, one workgroup per tile wins; here a persistent grid does).
    """
    from flatworld_amd import _native, autotune
    from flatworld_amd import workloads as W
    from flatworld_amd.server import Job, SimServer, poisson_arrivals

    rng = W.rng_for(seed + 7919)
    t = time.perf_counter()
    specs = [W.serving_job(rng) for _ in range(n_jobs)]  # the requests' payloads, built before the clock starts
    log(f"[bench] serving: {n_jobs} job payloads built in {time.perf_counter() - t:.1f}s")
    tuning = {k: arena_kw[k] for k in ("tile_nodes", "tile_slots", "block", "grid_per_cu")}
    tuning["source"] = "the batched arena's tuning"
    if tune_resident > 0:
        res = autotune.tune(specs[:tune_resident], frames=[frame], log=log)
        tuning = {k: res[k] for k in ("tile_nodes", "tile_slots", "block", "grid_per_cu")}
        tuning["source"] = f"autotuner on {min(tune_resident, n_jobs)} served jobs" + (" (cached)" if res.get("cached") else "")
        arena_kw = autotune.arena_kwargs(dict(res, frame=frame))
    log(f"[bench] serving: tuning {tuning}")

    # the tiling cache starts empty for the burst and again for the stream (the stream
    # re-serves the burst's payloads): only connectivities repeated within a run hit it
    nb = min(burst_jobs, n_jobs)
    _native.tiling_cache_clear()
    srv = SimServer(frame=frame, keep_results=False, **arena_kw)
    srv.run([Job(i, specs[i], 0.0) for i in range(nb)])
    burst = srv.stats()
    cache_burst = _native.tiling_cache_stats()
    srv.batch.close()
    capacity = burst["jobs"] / burst["wall_s"]
    rate = 0.8 * capacity
    log(f"[bench] serving: burst of {nb} jobs in {burst['wall_s']:.2f}s -> capacity {capacity:.1f} jobs/s")

    arrivals = poisson_arrivals(n_jobs, rate, rng)
    _native.tiling_cache_clear()
    srv = SimServer(frame=frame, keep_results=False, **arena_kw)
    srv.run([Job(i, specs[i], float(arrivals[i])) for i in range(n_jobs)])
    st = srv.stats()
    cache_stream = _native.tiling_cache_stats()
    srv.batch.close()
    # algorithmic bytes the served stream moved per second (each job: its per-substep bytes x horizon)
    done = {r.job_id for r in srv.results}
    served_bytes = sum(W.algorithmic_bytes(specs[i]) * specs[i].horizon for i in done)
    log(f"[bench] serving: {st['jobs']} jobs at {rate:.1f}/s, p50 {st['latency_p50_s'] * 1e3:.1f} ms, "
        f"p99 {st['latency_p99_s'] * 1e3:.1f} ms, {st['elem_substeps_per_s']:.3e} elem-substeps/s")
    return {
        "jobs": st["jobs"],
        "failed": st["failed"],
        "arrival_rate_jobs_per_s": rate,
        "capacity_jobs_per_s": capacity,
        "latency_p50_ms": st["latency_p50_s"] * 1e3,
        "latency_p99_ms": st["latency_p99_s"] * 1e3,
        "latency_mean_ms": st["latency_mean_s"] * 1e3,
        "elem_substeps_per_s": st["elem_substeps_per_s"],
        "hbm_frac": served_bytes / st["wall_s"] / 1e9 / HBM_PEAK_GBS,
        "burst_elem_substeps_per_s": burst["elem_substeps_per_s"],
        "gpu_busy_frac": st["gpu_busy_frac"],
        "frame_substeps": frame,
        "tuning": tuning,
        "tiling_cache": {"burst": cache_burst, "stream": cache_stream,
                         "note": "tilings reused by connectivity (jobs of one grid side share it); empty at the start of each"},
        "workload": "Poisson stream: grid side U{16..256}, Elastic/Neo-Hookean/J2, plane/height field/voxel ground, "
                    "horizon U{60..600}; latency = arrival -> final positions on the host",
    }


def traffic_from_profile(shape, config):
    """HBM bytes per substep launch from the committed counter profile of this shape, if any."""
    path = TRAFFIC_PROFILES.get(config.split("-")[0])
    try:
        with open(path) as fh:
            prof = json.load(fh)
    except (OSError, ValueError, TypeError):
        return None, None
    if prof.get("workload") != config or any(prof.get("shape", {}).get(k) != v for k, v in shape.items()):
        return None, None
    return prof.get("bytes_per_launch"), os.path.relpath(path, ROOT)


def tune_for(specs, args):
    """Kernel shape, workgroups per CU and substeps per graph for these scenes."""
    from flatworld_amd import autotune

    if args.tile_nodes > 0 or args.no_tune:
        return {"tile_nodes": args.tile_nodes or 512, "tile_slots": args.tile_slots, "block": args.block,
                "grid_per_cu": args.grid_per_cu, "frame": args.frame or 60, "source": "explicit"}
    t = time.perf_counter()
    res = autotune.tune(specs, log=log)
    tuning = {k: res[k] for k in ("tile_nodes", "tile_slots", "block", "grid_per_cu", "frame")}
    tuning["source"] = "autotuner (cached)" if res.get("cached") else "autotuner"
    tuning["tune_s"] = time.perf_counter() - t
    tuning["frame_tuned"] = res["frame"]
    if args.frame:
        tuning["frame"] = args.frame
    return tuning


def graph_length(steps, frame):
    """Substeps per replayed graph in the timed region: the longest length no larger than
    ``frame`` that divides ``steps``, so that the whole timed region is graph replays."""
    for g in range(min(frame, steps), 0, -1):
        if steps % g == 0:
            return g
    return 1


def timed_arena(specs, tuning, steps, warmup, no_graph=False, settle=0, window_stats=False):
    """Every scene resident in one arena and warmed up; ``settle`` substeps (untimed, in
    graph frames) carry the scenes past their approach; ``warmup`` more substeps and one
    replay capture the timed graph; then ``steps`` substeps are timed as replays of one
    graph of ``graph_length(steps, frame)`` substeps (plain launches with ``no_graph``).
    Returns the arena, the wall time of the timed region, the average substep-kernel
    duration from HIP events over the timed region on the arena stream, and the timed
    region's description."""
    from flatworld_amd import _native
    from flatworld_amd import autotune

    arena = _native.Arena(**autotune.arena_kwargs(tuning))
    ids = [arena.add(s, track_disp=True) for s in specs]
    arena.step(1, -1.0, 0)  # zero-dt warm-up of every scene
    arena.sync()
    frame = max(1, tuning["frame"])
    g = graph_length(steps, frame)

    def run(n, length):
        if no_graph:
            arena.step(n, -1.0, 0)
            return
        full, rest = divmod(n, length)
        for _ in range(full):
            arena.step_graph(length, 0)
        if rest:
            arena.step(rest, -1.0, 0)

    t = time.perf_counter()
    run(settle, frame)
    run(warmup, g)
    if not no_graph:
        arena.step_graph(g, 0)  # captures (if new) and warms the timed graph
    arena.sync()
    settle_s = time.perf_counter() - t
    # HIP events bracket every stepping call of the timed region on the arena stream:
    # their summed GPU time over the substeps is the average substep-kernel duration
    # (within a graph the launches run back to back; the bracket also holds the
    # one-thread clock kernel that ends each graph)
    arena.track_busy(True)
    busy0 = arena.busy()[0]
    t0 = time.perf_counter()
    run(steps, g)
    arena.sync()
    total_s = time.perf_counter() - t0
    kernel_ms = (arena.busy()[0] - busy0) / steps
    arena.track_busy(False)
    window = {
        "settle_substeps": settle,
        "warmup_substeps": warmup + (0 if no_graph else g),
        "launch": "plain launches, one per substep and element kind" if no_graph
        else f"hipGraph replay: {steps // g} graphs of {g} substeps",
        "settle_s": settle_s,
    }
    if window_stats:
        window.update(window_fractions(arena, ids, specs))
    # for comparison: events around each plain launch of the same kernel, after the window
    _, window["kernel_ms_per_launch_events"] = arena.time_steps(min(steps, 240), mode=0)
    return arena, total_s, kernel_ms, window


def window_fractions(arena, ids, specs):
    """Contact and yield state at the end of the timed region, from one more substep with
    diagnostics: scenes with a ground-contact node, contact-active boundary nodes, scenes
    and J2 elements yielding, scenes whose state is not finite."""
    from flatworld_amd import _native

    arena.step(1, -1.0, _native.FW_STEP_DIAG)
    arena.sync()
    sc = bn = nb = sy = ey = ej2 = bad = 0
    for sid, s in zip(ids, specs):
        c = arena.read(sid, _native.FIELD_CONTACT)
        hit = int(np.count_nonzero(c))
        sc += int(hit > 0)
        bn += hit
        nb += len(s.bnodes)
        if s.mat_type == 3:
            y = arena.read(sid, _native.FIELD_YIELD)
            ny = int(np.count_nonzero(y))
            sy += int(ny > 0)
            ey += ny
            ej2 += s.n_elems
        if not np.isfinite(arena.read(sid, _native.FIELD_X)).all():
            bad += 1
    return {
        "contact_scenes_frac": sc / len(specs),
        "contact_boundary_nodes_frac": bn / max(nb, 1),
        "j2_scenes_yielding_frac": sy / max(sum(1 for s in specs if s.mat_type == 3), 1),
        "j2_elements_yielding_frac": ey / max(ej2, 1),
        "nonfinite_scenes": bad,
    }


def landing_substeps(specs, g=9.8, margin=60):
    """Substeps until every scene standing over a height field has reached it: the
    largest free-fall time from the body's lowest node to the highest height sample
    under it, over the scene's dt, plus a margin.  Scenes on other grounds count 0."""
    worst = 0
    for s in specs:
        for gr in s.grounds:
            if gr.kind != 2:
                continue
            h = np.asarray(gr.heights, dtype=np.float64)
            xs = np.linspace(float(gr.hf_lb_x), float(gr.hf_ub_x), h.size)
            x0 = np.asarray(s.x0, dtype=np.float64)
            under = (xs >= x0[:, 0].min()) & (xs <= x0[:, 0].max())
            gap = x0[:, 1].min() - (h[under].max() if under.any() else h.max())
            if gap > 0:
                worst = max(worst, int(np.ceil(np.sqrt(2.0 * gap / g) / float(s.dt))))
    return worst + margin if worst else 0


def single_scene_leg(config, args):
    """One of the single-scene configurations (the 1000x1000 Neo-Hookean sheet on a plane,
    or the 512x512 J2 block on 256x256 voxels), timed the way the batched mix is: tuned,
    graph-replayed, kernel time from HIP events, contact and yield state of the window,
    and the oracle on the same scene as its CPU baseline."""
    from flatworld_amd import workloads as W

    t = time.perf_counter()
    specs = build_specs(config, args.seed, 1)
    E = specs[0].n_elems
    N = specs[0].n_nodes
    algo = float(W.algorithmic_bytes(specs[0]))
    tuning = tune_for(specs, args)
    # the J2 block falls 2 mm into the top voxel row in about 100 substeps; the sheet
    # presses on its plane from the first substep
    settle = 200 if config == "j2voxel" else 0
    arena, total_s, kernel_ms, window = timed_arena(specs, tuning, args.single_steps, args.warmup, args.no_graph,
                                                    settle=settle, window_stats=True)
    info = arena.info()
    arena.close()
    achieved = algo / (kernel_ms * 1e-3) / 1e9
    ms = total_s * 1e3 / args.single_steps
    log(f"[bench] {config}: E={E}, tiles={info['n_tiles']}, {args.single_steps} substeps in {total_s * 1e3:.1f} ms, "
        f"kernel {kernel_ms * 1e3:.1f} us, frac {achieved / HBM_PEAK_GBS:.3f} ({time.perf_counter() - t:.1f}s with setup)")
    out = {
        "value": E * args.single_steps / total_s,
        "unit": "element-substeps/s",
        "workload": {"large": "1000x1000 jittered grid, 2 M triangles, Neo-Hookean, plane y=0, CFL dt",
                     "j2voxel": "512x512 block, 524 k triangles, J2, falling at 2 m/s onto 256x256 value-noise "
                                "voxels"}[config],
        "elements": E,
        "nodes": N,
        "tiles": info["n_tiles"],
        "tiles_per_cu": info["tiles_per_cu"],
        "steps": args.single_steps,
        "ms_per_step": ms,
        "window": window,
        "tuning": tuning,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "frac_wall": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernel_ms": kernel_ms,
            "algo_bytes_per_substep": algo,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(specs, args.cpu_budget / 3.0, all_cores=False)
    return out


def spring_bench(args):
    """4096 batched spring lattices: throughput, HBM fraction of the spring substep, CPU baseline."""
    from flatworld_amd import workloads as W

    t = time.perf_counter()
    specs = W.spring_batched_config(n_scenes=args.spring_scenes, seed=args.seed)
    E = sum(s.n_elems for s in specs)
    N = sum(s.n_nodes for s in specs)
    algo = float(sum(W.algorithmic_bytes(s) for s in specs))
    tuning = tune_for(specs, args)
    arena, total_s, kernel_ms, window = timed_arena(specs, tuning, args.spring_steps, args.warmup, args.no_graph,
                                                    settle=landing_substeps(specs))
    info = arena.info()
    arena.close()
    achieved = algo / (kernel_ms * 1e-3) / 1e9
    log(f"[bench] springs: {len(specs)} lattices, springs={E}, nodes={N}, tiles={info['n_tiles']}, "
        f"{args.spring_steps} substeps in {total_s * 1e3:.1f} ms, kernel {kernel_ms * 1e3:.1f} us "
        f"({time.perf_counter() - t:.1f}s with setup)")
    out = {
        "value": E * args.spring_steps / total_s,
        "unit": "spring-substeps/s",
        "workload": f"springs-{len(specs)}: spring lattices (structural, shear, bending springs), side log-uniform "
                    "in [8,128], spacing 0.01, one height field each",
        "scenes": len(specs),
        "springs": E,
        "nodes": N,
        "ms_per_step": total_s * 1e3 / args.spring_steps,
        "kernel_ms": kernel_ms,
        "algo_bytes_per_substep": algo,
        "hbm_frac": achieved / HBM_PEAK_GBS,
        "achieved_GBps": achieved,
        "tuning": tuning,
        "window": window,
    }
    shape = {k: tuning[k] for k in ("tile_nodes", "tile_slots", "block")}
    out["traffic"], out["traffic_source"] = traffic_from_profile(shape, f"springs-{len(specs)}")
    if not args.no_cpu:
        cpu = cpu_baseline(specs, args.cpu_budget / 3.0)
        cpu["unit"] = "spring-substeps/s"
        out["cpu_baseline"] = cpu
    return out


def contact_bench(args):
    """Batched two-block loops with FEM-FEM contact: per substep the contact pass (work
    items of every loop, then one workgroup per loop for the lists) and the substep
    launch, replayed as 60-substep hipGraph frames; the hash of every loop is rebuilt in
    one launch every 1/500 s of simulated time, as each loop's schedule has it,
    scheduled on the device inside the frames."""
    from flatworld_amd import _native
    from flatworld_amd import workloads as W
    from flatworld_amd.flexcontact import REBUILD_INTERVAL

    t = time.perf_counter()
    groups = W.contact_batched_config(n_groups=args.contact_groups, seed=args.seed)
    specs = [s for sp, _ in groups for s in sp]
    E = sum(s.n_elems for s in specs)
    N = sum(s.n_nodes for s in specs)
    items = sum(g.n_items for _, g in groups)
    slots = sum(int(g.body_bnodes.shape[0]) for _, g in groups)
    dt = specs[0].dt
    every = max(1, int(np.ceil(REBUILD_INTERVAL / dt - 1e-9)))
    algo_sub = float(sum(W.algorithmic_bytes(s) for s in specs))
    algo_contact = float(items * W.CONTACT_ITEM_BYTES + slots * W.CONTACT_SLOT_BYTES)
    arena = _native.Arena(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0)
    for sp, g in groups:
        arena.flex_add(g, [arena.add(s, track_disp=True) for s in sp])
    arena.step(1, -1.0, 0)  # zero-dt warm-up of every scene
    arena.flex_rebuild(-1)
    arena.sync()

    elapsed = [0.0]
    frame = 60

    def run(n):
        # the reference's schedule: a rebuild precedes the substep at which 1/500 s
        # of simulated time has passed (flatworld/femspringmanager.py:1035-1039);
        # frames of `frame` substeps replay as hipGraphs, each with its rebuilds
        # scheduled on the device (plain launches with --no-graph)
        rebuild_at = []
        for k in range(n):
            elapsed[0] += dt
            if elapsed[0] >= REBUILD_INTERVAL:
                rebuild_at.append(k)
                elapsed[0] = 0.0
        for f0 in range(0, n, frame):
            m = min(frame, n - f0)
            for k in rebuild_at:
                if f0 <= k < f0 + m:
                    arena.flex_schedule_rebuild(-1, k - f0)
            if args.no_graph:
                arena.step(m, -1.0, 0)
            else:
                arena.step_graph(m, 0)

    # timed by the wall clock: the first contact_steps substeps after release
    # (approach and impact), then two rebuild intervals with the blocks in contact;
    # then the contact pass + substep kernels with HIP events over one interval.
    # The frame graphs are captured before the timed region.
    if not args.no_graph:
        for m in {frame, args.contact_steps % frame, (2 * every) % frame} - {0}:
            arena.prepare_graph(m, 0)
    t0 = time.perf_counter()
    run(args.contact_steps)
    arena.sync()
    total_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    run(2 * every)
    arena.sync()
    contact_s = time.perf_counter() - t0
    _, kernel_ms = arena.time_steps(every, mode=0)
    t1 = time.perf_counter()
    arena.flex_rebuild(-1)
    arena.sync()
    rebuild_ms = (time.perf_counter() - t1) * 1e3
    info = arena.info()
    arena.close()
    achieved = (algo_sub + algo_contact) / (kernel_ms * 1e-3) / 1e9  # in contact
    log(f"[bench] contact: {len(groups)} loops, elements={E}, nodes={N}, work items={items}, tiles={info['n_tiles']}, "
        f"{args.contact_steps} substeps in {total_s * 1e3:.1f} ms, in contact {contact_s * 1e6 / (2 * every):.1f} us/substep, "
        f"contact+substep kernels {kernel_ms * 1e3:.1f} us, "
        f"rebuild {rebuild_ms:.2f} ms every {every} substeps ({time.perf_counter() - t:.1f}s with setup)")
    out = {
        "value": E * args.contact_steps / total_s,
        "unit": "element-substeps/s",
        "workload": f"contact-{len(groups)}: two-block loops (Elastic / J2, side log-uniform in [8,32]) on a plane, "
                    "the upper block falling onto the lower; FEM-FEM contact with the spatial hash",
        "loops": len(groups),
        "elements": E,
        "nodes": N,
        "work_items": items,
        "dt": dt,
        "rebuild_every": every,
        "ms_per_step": total_s * 1e3 / args.contact_steps,
        "in_contact_ms_per_step": contact_s * 1e3 / (2 * every),
        "window": f"value, ms_per_step: substeps 1-{args.contact_steps} after release; in_contact_ms_per_step: the "
                  f"next {2 * every} (rebuilds included); kernel_ms: the {every} after those (HIP events)",
        "kernel_ms": kernel_ms,
        "rebuild_ms": rebuild_ms,
        "algo_bytes_per_substep": algo_sub + algo_contact,
        "algo_bytes_contact": algo_contact,
        "hbm_frac": achieved / HBM_PEAK_GBS,
        "achieved_GBps": achieved,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = contact_cpu_baseline(groups, args.cpu_budget / 3.0)
    return out


def contact_cpu_baseline(groups, budget_s):
    """The oracle (serial C, contact pass included) on a bounded sample of the loops."""
    from oracle.oracle import OracleLoop

    order = np.random.Generator(np.random.PCG64(1)).permutation(len(groups))
    sample, elems = [], 0
    for i in order:
        sample.append(groups[i])
        elems += sum(s.n_elems for s in groups[i][0])
        if elems >= 100_000:
            break
    loops = [OracleLoop(sp, g) for sp, g in sample]
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s:
        for o, (sp, _) in zip(loops, sample):
            o.substep(sp[0].dt)
        n += 1
    rate = elems * n / (time.perf_counter() - t0)
    return {"value": rate, "unit": "element-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{len(sample)} of {len(groups)} loops ({elems} elements) x {n} substeps, serial C oracle"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200, help="timed substeps (whole frames replay as hipGraphs)")
    ap.add_argument("--warmup", type=int, default=120)
    ap.add_argument("--frame", type=int, default=60,
                    help="substeps per hipGraph frame (the 4096-scene mix is defined with 60; 0: the autotuner's choice)")
    ap.add_argument("--config", default="batched", choices=["batched", "large", "j2voxel", "small", "springs"])
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tile-nodes", type=int, default=0, help="explicit kernel shape (skips the autotuner)")
    ap.add_argument("--tile-slots", type=int, default=1024)
    ap.add_argument("--block", type=int, default=256, help="threads per tile workgroup")
    ap.add_argument("--grid-per-cu", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="default kernel shape instead of the autotuner")
    ap.add_argument("--per-scene-steps", type=int, default=6)
    ap.add_argument("--serve-jobs", type=int, default=600, help="jobs in the served Poisson stream (0: skip)")
    ap.add_argument("--serve-burst", type=int, default=300, help="jobs of the closed-loop capacity burst")
    ap.add_argument("--serve-frame", type=int, default=60, help="substeps per hipGraph frame of the served stream")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of hipGraph frames")
    ap.add_argument("--settle", type=int, default=-1,
                    help="untimed substeps before the timed region (-1: until every block has reached its height field)")
    ap.add_argument("--single-steps", type=int, default=240,
                    help="timed substeps of the single-scene legs (the 1000x1000 sheet, the J2 block on voxels)")
    ap.add_argument("--no-single", action="store_true", help="skip the single-scene legs")
    ap.add_argument("--spring-scenes", type=int, default=4096, help="spring lattices of the spring leg (0: skip)")
    ap.add_argument("--spring-steps", type=int, default=600)
    ap.add_argument("--contact-groups", type=int, default=2048,
                    help="two-block loops of the FEM-FEM contact leg (0: skip)")
    ap.add_argument("--contact-steps", type=int, default=600)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_probe:
        from flatworld_amd import _native

        gbs, ms = _native.bandwidth_probe(0, args.copy_probe, 10)
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "read_write_GBps": gbs, "kernel_ms": ms}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU scaling is not part of it", "n_gpus": args.gpus}))
        sys.exit(2)

    from flatworld_amd import _native, autotune
    from flatworld_amd import workloads as W

    t_setup = time.perf_counter()
    specs = build_specs(args.config, args.seed, args.scenes)
    E = sum(s.n_elems for s in specs)
    N = sum(s.n_nodes for s in specs)
    algo = float(sum(W.algorithmic_bytes(s) for s in specs))

    # kernel shape, workgroups per CU, substeps per graph
    tuning = tune_for(specs, args)
    arena_kw = autotune.arena_kwargs(tuning)
    log(f"[bench] tuning: {tuning}")
    frame = max(1, tuning["frame"])

    # timed graph frames, then the live kernel duration: HIP events around each substep
    # kernel on the arena stream
    settle = landing_substeps(specs) if args.settle < 0 else args.settle
    arena, total_s, kernel_ms, window = timed_arena(specs, tuning, args.steps, args.warmup, args.no_graph,
                                                    settle=settle, window_stats=True)
    info = arena.info()
    value = E * args.steps / total_s
    log(f"[bench] window: {window}")
    achieved = algo / (kernel_ms * 1e-3) / 1e9
    log(f"[bench] {args.config}: {len(specs)} scenes, E={E}, N={N}, tiles={info['n_tiles']}, "
        f"halo slots={info['halo_slots']}, {time.perf_counter() - t_setup:.1f}s with setup; "
        f"{args.steps} substeps in {total_s * 1e3:.1f} ms")

    per_scene = None
    if len(specs) > 1 and args.per_scene_steps > 0:
        ps_total_ms, _ = arena.time_steps(args.per_scene_steps, mode=2)
        ps_value = E * args.per_scene_steps / (ps_total_ms * 1e-3)
        per_scene = {
            "value": ps_value,
            "unit": "element-substeps/s",
            "launches_per_substep": len(specs),
            "ms_per_step": ps_total_ms / args.per_scene_steps,
            "hbm_frac": algo / (ps_total_ms / args.per_scene_steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "batched_speedup": value / ps_value,
        }
        log(f"[bench] per-scene launches: {ps_total_ms / args.per_scene_steps:.2f} ms/substep")
    arena.close()

    serving = None
    if args.serve_jobs > 0 and args.config == "batched":
        serving = serving_bench(args.serve_jobs, args.seed, arena_kw, args.serve_frame, args.serve_burst)

    springs = None
    if args.spring_scenes > 0 and args.config == "batched":
        springs = spring_bench(args)

    contact = None
    if args.contact_groups > 0 and args.config == "batched":
        contact = contact_bench(args)

    single = None
    if args.config == "batched" and not args.no_single:
        single = {c: single_scene_leg(c, args) for c in ("large", "j2voxel")}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(specs, args.cpu_budget)
        log(f"[bench] cpu baseline {cpu['value']:.3e} elem-substeps/s")

    shape = {k: tuning[k] for k in ("tile_nodes", "tile_slots", "block")}
    traffic, traffic_src = traffic_from_profile(shape, f"{args.config}-{len(specs)}")
    ceiling = _native.bandwidth_ceiling(0, 1 << 30, 5)
    log(f"[bench] copy ceiling: {ceiling}")
    out = {
        "metric": "element-substeps/s, batched vs per-scene; p50/p99 request latency @1e-4 tol",
        "value": value,
        "unit": "element-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}-{len(specs)}",
            "scenes": len(specs),
            "elements": E,
            "nodes": N,
            "frame_substeps": frame,
            "seed": args.seed,
            "timed_window": window,
        },
        "tuning": tuning,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "frac_wall": algo / (total_s / args.steps) / 1e9 / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "traffic_GBps": traffic / (kernel_ms * 1e-3) / 1e9 if traffic else None,
            "kernel_ms": kernel_ms,
            "algo_bytes_per_substep": algo,
            # the fastest read + write stream this device sustains (tuned copy), the
            # practical ceiling beside the 8 TB/s peak
            "copy_ceiling": ceiling,
            "frac_of_copy_ceiling": achieved / ceiling["GBps"],
            "traffic_frac_of_copy_ceiling": traffic / (kernel_ms * 1e-3) / 1e9 / ceiling["GBps"] if traffic else None,
        },
        "per_scene": per_scene,
        "serving": serving,
        "springs": springs,
        "contact": contact,
        "single_scene": single,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
