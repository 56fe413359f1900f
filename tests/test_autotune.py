"""Autotuner host logic on the CPU: workload signature and the tuning sample."""

import numpy as np

from flatworld_amd import autotune as A
from flatworld_amd import workloads as W


def test_signature_is_stable_and_separates_mixes():
    a = W.batched_config(n_scenes=40, seed=3, side_lo=6, side_hi=30)
    b = W.batched_config(n_scenes=40, seed=3, side_lo=6, side_hi=30)
    assert A.signature(a) == A.signature(b)
    bigger = W.batched_config(n_scenes=40, seed=3, side_lo=40, side_hi=60)
    assert A.signature(a) != A.signature(bigger)


def test_sample_keeps_size_histogram():
    specs = W.batched_config(n_scenes=400, seed=1, side_lo=6, side_hi=60)
    total = sum(s.n_elems for s in specs)
    sample = A.stratified_sample(specs, total // 4)
    got = sum(s.n_elems for s in sample)
    assert 0.15 * total < got < 0.4 * total
    q = [10, 50, 90]
    full = np.percentile([s.n_elems for s in specs], q)
    part = np.percentile([s.n_elems for s in sample], q)
    assert np.all(np.abs(np.log(part / full)) < 0.35)
    assert A.stratified_sample(specs, 10 * total) == specs


def test_tuning_cache_roundtrip(tmp_path, monkeypatch):
    specs = W.batched_config(n_scenes=8, seed=2, side_lo=6, side_hi=12)
    path = tmp_path / "tune.json"
    calls = []

    def fake_candidate(sample, block, slots, nodes, gpc, frames, reps, capacity):
        calls.append((block, slots, gpc))
        # pretend the (512, 2048) persistent shape with 60-substep frames is fastest
        base = 1.0 if (block, slots, gpc) == (512, 2048, 4) else 2.0
        return {f: base + (0.0 if f == 60 else 0.1) for f in frames}, {"n_tiles": 1, "tiles_per_cu": 2}

    monkeypatch.setattr(A, "_time_candidate", fake_candidate)
    r = A.tune(specs, cache_path=str(path))
    assert (r["block"], r["tile_slots"], r["grid_per_cu"], r["frame"]) == (512, 2048, 4, 60)
    assert not r["cached"] and len(calls) == len(A.SHAPES) * len(A.GRID_PER_CU)
    r2 = A.tune(specs, cache_path=str(path))
    assert r2["cached"] and r2["tile_slots"] == 2048 and len(calls) == len(A.SHAPES) * len(A.GRID_PER_CU)
    assert A.arena_kwargs(r2) == {"tile_nodes": 1024, "tile_slots": 2048, "block": 512, "grid_per_cu": 4}
