#!/usr/bin/env python3
"""Benchmark: explicit-FEM substep throughput on one MI355X.

Default workload: the batched mix -- 4096 independent scenes, grid side
log-uniform in [8, 128], Elastic / Neo-Hookean / J2 mix, one height field per
scene (24.08 M triangles, 12.40 M nodes).  A "step" is one substep of every
scene; the timed region replays hipGraph-captured 60-substep frames.  Inputs
are resident in HBM before timing starts.

Printed JSON line:
  value            element-substeps/s of the batched, graph-replayed form
  per_scene        the same scenes advanced with one launch per scene per substep
  roofline         algorithmic bytes per substep / average substep-kernel time,
                   measured live with HIP events on the launch stream
  cpu_baseline     the oracle (serial C restatement) on a bounded sample
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
    raise SystemExit(f"unknown config {config}")


def cpu_baseline(specs, budget_s=12.0):
    """Serial oracle on a bounded sample of the same scenes (the checker, timed)."""
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
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s:
        for o in scenes:
            o.substep(o.spec.dt, o.spec.damping)
        n += 1
    dt = time.perf_counter() - t0
    return {
        "value": elems * n / dt,
        "unit": "element-substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(sample)} of {len(specs)} scenes ({elems} triangles) x {n} substeps, serial C oracle",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200, help="timed substeps (multiple of --frame is replayed as graphs)")
    ap.add_argument("--warmup", type=int, default=120)
    ap.add_argument("--frame", type=int, default=60, help="substeps per hipGraph frame")
    ap.add_argument("--config", default="batched", choices=["batched", "large", "j2voxel", "small"])
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tile-nodes", type=int, default=512)
    ap.add_argument("--tile-slots", type=int, default=1024)
    ap.add_argument("--block", type=int, default=256, help="threads per tile workgroup")
    ap.add_argument("--per-scene-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-graph", action="store_true", help="plain launches instead of hipGraph frames")
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

    from flatworld_amd import _native
    from flatworld_amd import workloads as W

    t_setup = time.perf_counter()
    specs = build_specs(args.config, args.seed, args.scenes)
    E = sum(s.n_elems for s in specs)
    N = sum(s.n_nodes for s in specs)
    algo = float(sum(W.algorithmic_bytes(s) for s in specs))
    arena = _native.Arena(tile_nodes=args.tile_nodes, tile_slots=args.tile_slots, block=args.block)
    for s in specs:
        arena.add(s, track_disp=True)
    arena.step(1, -1.0, 0)  # zero-dt warm-up of every scene
    arena.sync()
    info = arena.info()
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] {args.config}: {len(specs)} scenes, E={E}, N={N}, tiles={info['n_tiles']}, "
          f"halo slots={info['halo_slots']}, setup {t_setup:.1f}s", file=sys.stderr, flush=True)

    frame = max(1, args.frame)

    def run(n):
        if args.no_graph:
            arena.step(n, -1.0, 0)
            return
        full, rest = divmod(n, frame)
        for _ in range(full):
            arena.step_graph(frame, 0)
        if rest:
            arena.step(rest, -1.0, 0)

    run(args.warmup)
    arena.sync()
    t0 = time.perf_counter()
    run(args.steps)
    arena.sync()
    t1 = time.perf_counter()
    total_s = t1 - t0
    value = E * args.steps / total_s
    print(f"[bench] batched: {args.steps} substeps in {total_s*1e3:.1f} ms", file=sys.stderr, flush=True)

    # live kernel duration: HIP events around each substep kernel on the arena stream
    n_k = min(args.steps, 240)
    _, kernel_ms = arena.time_steps(n_k, mode=0)
    achieved = algo / (kernel_ms * 1e-3) / 1e9

    per_scene = None
    if len(specs) > 1 and args.per_scene_steps > 0:
        ps_total_ms, _ = arena.time_steps(args.per_scene_steps, mode=2)
        ps_value = E * args.per_scene_steps / (ps_total_ms * 1e-3)
        per_scene = {
            "value": ps_value,
            "unit": "element-substeps/s",
            "launches_per_substep": len(specs),
            "ms_per_step": ps_total_ms / args.per_scene_steps,
            "batched_speedup": value / ps_value,
        }
        print(f"[bench] per-scene launches: {ps_total_ms / args.per_scene_steps:.2f} ms/substep", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(specs, args.cpu_budget)
        print(f"[bench] cpu baseline {cpu['value']:.3e} elem-substeps/s", file=sys.stderr, flush=True)

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
            "workload": args.config,
            "scenes": len(specs),
            "triangles": E,
            "nodes": N,
            "frame_substeps": frame,
            "tile_nodes": args.tile_nodes,
            "tile_slots": args.tile_slots,
            "seed": args.seed,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_ms": kernel_ms,
            "algo_bytes_per_substep": algo,
        },
        "per_scene": per_scene,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
