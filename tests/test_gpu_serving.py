"""Served and batched forms on the GPU.

* every served job's final state equals the same job run alone (bit for bit);
* a job admitted mid-stream (while others are on the GPU) is also identical;
* an arena too small for all jobs serves them all, admitting as memory frees;
* a job whose state turns non-finite is reported failed while the others stay
  identical to their solo runs.
"""

import dataclasses

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import workloads as W
from flatworld_amd.batch import FemBatch
from flatworld_amd.server import Job, SimServer
from tests.parity import log

pytestmark = pytest.mark.gpu


def small_jobs(n, seed, side_lo=6, side_hi=24, h_lo=20, h_hi=140):
    rng = W.rng_for(seed)
    return [Job(i, W.serving_job(rng, side_lo, side_hi, h_lo, h_hi), 0.0) for i in range(n)]


def solo(spec):
    b = FemBatch(tile_nodes=256)
    sid = b.add(spec)
    b.step(1 + spec.horizon, graph=False)
    x = b.positions(sid)
    b.close()
    return x


def test_served_results_equal_solo_runs():
    jobs = small_jobs(14, seed=21)
    for k, j in enumerate(jobs):
        j.arrival = 0.002 * k  # staggered: later jobs join while earlier ones run
    log(f"[case] serve {len(jobs)} jobs, compare with solo runs")
    srv = SimServer(frame=20, tile_nodes=256)
    res = {r.job_id: r for r in srv.run(jobs)}
    assert len(res) == len(jobs)
    for j in jobs:
        assert res[j.job_id].ok
        assert np.array_equal(res[j.job_id].positions, solo(j.spec)), f"job {j.job_id}"


def test_admission_waits_for_memory():
    jobs = small_jobs(8, seed=22, side_lo=10, side_hi=12)
    probe = FemBatch(capacity_bytes=1 << 26, tile_nodes=256)
    fp = {j.job_id: probe.footprint(j.spec) for j in jobs}
    probe.close()
    cap = 2 * max(fp.values()) + 4096
    log(f"[case] admission into {cap} bytes, footprints {sorted(fp.values())}")
    srv = SimServer(frame=30, tile_nodes=256, capacity_bytes=cap)
    res = srv.run(jobs)
    assert len(res) == 8 and all(r.ok for r in res)
    # at every admission the resident footprints fit the arena, and the arena
    # was small enough that some admissions had to wait for a retirement
    spans = [(r.admitted, r.released, fp[r.job_id]) for r in res]
    most = 0
    for t, _, _ in spans:
        live = [b for a, f, b in spans if a <= t < f]
        assert sum(live) <= cap
        most = max(most, len(live))
    assert most < len(jobs)
    assert sum(fp.values()) > cap


def test_failed_job_is_isolated():
    jobs = small_jobs(4, seed=23)
    bad = jobs[1].spec
    # a Neo-Hookean body slammed into the ground at 200 m/s inverts elements -> log(J<0) = NaN
    mat = fw.HyperElastic(1e4, 0.3, 100.0)
    dom = W.grid_domain(8, (0.3, 0.1), 0.4, mat, [fw.Gravity([0.0, -9.8])], initials=[fw.InitialVel("ALL", [0.0, -200.0])])
    jobs[1] = Job(1, W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]), horizon=bad.horizon), 0.0)
    log("[case] failed job isolation")
    srv = SimServer(frame=20, tile_nodes=256)
    res = {r.job_id: r for r in srv.run(jobs)}
    assert not res[1].ok
    for k in (0, 2, 3):
        assert res[k].ok and np.array_equal(res[k].positions, solo(jobs[k].spec))


def test_prepared_scenes_equal_direct_adds():
    """Two-phase admission (host prepare on a worker thread, then add) gives the same bits as add()."""
    from concurrent.futures import ThreadPoolExecutor

    from flatworld_amd import _native

    jobs = small_jobs(6, seed=24)
    log("[case] prepared admission vs direct add")
    a = FemBatch(tile_nodes=256)
    b = FemBatch(tile_nodes=256)
    ia = [a.add(j.spec) for j in jobs]
    with ThreadPoolExecutor(3) as ex:
        preps = list(ex.map(lambda j: b.prepare(j.spec), jobs))
    ib = [b.add_prepared(p) for p in preps]
    a.step(61, graph=False)
    b.step(61, graph=False)
    for x, y in zip(ia, ib):
        assert np.array_equal(a.positions(x), b.positions(y))
        assert np.array_equal(a.velocities(x), b.velocities(y))
    other = _native.Prepared(jobs[0].spec, dict(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0))
    with pytest.raises(ValueError, match="another kernel shape"):
        b.add_prepared(other)
    a.close()
    b.close()


def test_reused_tiling_matches_oracle():
    """A scene admitted with a tiling fw_scene_prepare kept from another scene of the same
    connectivity (other positions, material and ground) gives the oracle's bits: the
    tiling is valid whatever the positions, and results do not depend on it."""
    from flatworld_amd import _native
    from oracle.oracle import OracleScene

    hf = W.sinusoid_heightfield(W.rng_for(3))
    plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    first = W.one_body_spec(W.grid_domain(40, (1.2, 0.5), 0.3, fw.Elastic(1e4, 0.3, 10.0), [fw.Gravity([0.0, -9.8])]),
                            plane, horizon=120)
    second = W.one_body_spec(W.grid_domain(40, (0.3, W.ground_top(hf)), 0.4, fw.MisesPlastic(2e4, 0.3, 10.0, 60.0, 1e2),
                                           [fw.Gravity([0.0, -9.8])], jitter=0.1, seed=4), hf, horizon=120)
    b = FemBatch(tile_nodes=128, tile_slots=1024)
    _native.tiling_cache_clear()
    p1 = b.prepare(first)
    p2 = b.prepare(second)
    st = _native.tiling_cache_stats()
    log(f"[case] reused tiling: cache {st}")
    assert (st["hits"], st["misses"]) == (1, 1)
    i1, i2 = b.add_prepared(p1), b.add_prepared(p2)
    assert b.arena.info()["n_tiles"] > 20
    b.step(121, graph=False)
    for sid, spec in ((i1, first), (i2, second)):
        o = OracleScene(spec)
        o.warmup()
        o.run(120)
        assert np.isfinite(o.x).all()
        assert np.array_equal(b.positions(sid), o.x) and np.array_equal(b.velocities(sid), o.v)
    b.close()


@pytest.mark.timeout(400)
def test_served_c5_stream_matches_oracle():
    """A served stream shaped like the serving configuration (grid side 16..256, horizon
    60..600 substeps, plane / height-field / voxel grounds chosen uniformly, Poisson arrivals
    while earlier jobs are on the GPU), with a stratified sample of the served jobs -- one per
    (ground kind, size third) and the largest -- checked bit for bit against the oracle."""
    from flatworld_amd.server import poisson_arrivals
    from oracle.oracle import OracleScene

    rng = W.rng_for(55)
    n = 36
    jobs = [Job(i, W.serving_job(rng), 0.0) for i in range(n)]
    for j, t in zip(jobs, poisson_arrivals(n, 300.0, rng)):
        j.arrival = float(t)
    sides = {j.job_id: int(round(np.sqrt(j.spec.n_elems / 2))) for j in jobs}
    log(f"[case] served C5-shaped stream: {n} jobs, sides {min(sides.values())}..{max(sides.values())}, "
        f"horizons {min(j.spec.horizon for j in jobs)}..{max(j.spec.horizon for j in jobs)}")
    srv = SimServer(frame=60)
    res = {r.job_id: r for r in srv.run(jobs)}
    assert len(res) == n and all(r.ok for r in res.values())
    # overlap: some job was admitted while another was still resident
    spans = sorted((r.admitted, r.finished) for r in res.values())
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1)), "jobs should overlap on the device"
    strata = {}
    for j in jobs:
        key = (j.spec.grounds[0].kind, min(2, (sides[j.job_id] - 16) * 3 // 241))
        strata.setdefault(key, j)
    sample = list(strata.values())
    largest = max(jobs, key=lambda j: j.spec.n_elems)
    if largest not in sample:
        sample.append(largest)
    assert {j.spec.grounds[0].kind for j in sample} == {1, 2, 3}
    touching = 0
    for j in sample:
        o = OracleScene(j.spec)
        o.warmup()
        o.run(j.spec.horizon)
        log(f"  job {j.job_id}: side {sides[j.job_id]}, ground {j.spec.grounds[0].kind}, horizon {j.spec.horizon}, "
            f"contact nodes at the end {int(o.contact_flag.sum())}")
        assert np.isfinite(o.x).all()
        assert np.array_equal(res[j.job_id].positions, o.x), f"job {j.job_id} positions"
        assert np.array_equal(res[j.job_id].velocities, o.v), f"job {j.job_id} velocities"
        touching += int(o.contact_flag.any())
    assert touching >= len(sample) // 2, "served jobs start on their grounds: most should be in contact"
