"""HIP substep vs the CPU oracle on identical seeded inputs.

Every material must agree bit for bit: both sides evaluate the same float32
operations in the same order without contraction, and the Neo-Hookean log is
one shared double-operation routine (correctly rounded in practice, pinned
against numpy in the oracle tests).  The Neo-Hookean case also reports the
tolerance figures (nodal forces within 1e-5 of the largest nodal force,
positions within 1e-4 of the minimum edge length) for a reader comparing a
build whose log differs.
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from flatworld_amd import workloads as W
from oracle.oracle import OracleScene
from tests.parity import FORCE_RTOL, POS_TOL_EDGE, GpuScene, compare, log

pytestmark = pytest.mark.gpu


def run_both(spec, chunks, exact, label, **arena_kw):
    log(f"[case] {label}: N={spec.n_nodes} E={spec.n_elems} mat={spec.mat_type} grounds={[g.kind for g in spec.grounds]}")
    g = GpuScene(spec, **arena_kw)
    o = OracleScene(spec)
    g.warmup()
    o.warmup()
    compare(g, o, exact, label + " warm-up") if exact else None
    out = []
    for n in chunks:
        g.run(n)
        o.run(n)
        out.append(compare(g, o, exact, label))
    return g, o, out


def test_small_block_bitexact():
    spec = W.small_config()
    run_both(spec, [1, 99, 500], True, "small 20x20 elastic, plane")


def test_neohookean_multitile_bitexact():
    dom = W.grid_domain(48, (0.0, 0.0), 1.0, fw.HyperElastic(1e5, 0.3, 1000.0), [fw.Gravity([0.0, -9.8])], jitter=0.1, seed=3)
    spec = W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]))
    g, o, _ = run_both(spec, [1, 99, 200], True, "nh 48x48 jittered, plane", tile_nodes=256)
    res = compare(g, o, False, "nh 48x48 tolerance figures")
    log(f"  {res}")
    assert res["force_rel"] < FORCE_RTOL
    assert res["pos_err_edges"] < POS_TOL_EDGE


def test_neohookean_heightfield_bitexact():
    rng = W.rng_for(17)
    dom = W.grid_domain(28, (0.3, 0.1), 0.4, fw.HyperElastic(4e4, 0.35, 200.0), [fw.Gravity([0.0, -9.8])],
                        initials=[fw.InitialVel("ALL", [0.3, -0.5])])
    spec = W.one_body_spec(dom, W.sinusoid_heightfield(rng))
    _, o, _ = run_both(spec, [1, 300, 300], True, "nh 28x28 on height field", tile_nodes=200, tile_slots=512)
    assert np.count_nonzero(o.contact_flag) > 0, "scenario should touch the height field"


def test_elastic_multitile_bitexact():
    dom = W.grid_domain(40, (0.2, 0.05), 0.5, fw.Elastic(5e4, 0.3, 1000.0), [fw.Gravity([0.0, -9.8])], jitter=0.05, seed=4)
    spec = W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.07], [0.1, 1.0]))
    run_both(spec, [1, 150], True, "elastic 40x40 multi-tile, tilted plane", tile_nodes=200)


def _j2_voxel_spec(side=32):
    rng = W.rng_for(7)
    mat = fw.MisesPlastic(1e5, 0.3, 300.0, 500.0, 1e3)
    dom = W.grid_domain(side, (0.25, 0.56), 0.5, mat, [fw.Gravity([0.0, -9.8])], initials=[fw.InitialVel("ALL", [0.0, -1.0])])
    ground = W.value_noise_voxels(rng, 64, 64, [0.0, 0.0], [1.0, 0.55], floor_rows=2)
    return W.one_body_spec(dom, ground)


def test_j2_voxel_single_tile_bitexact():
    # 30x30 cells: 961 nodes and 1800 elements fit one 1024-thread, 2048-slot tile
    g, o, _ = run_both(_j2_voxel_spec(30), [1, 100, 100, 100, 300], True, "j2 30x30 on voxels, one tile", tile_nodes=1024,
                       tile_slots=2048, block=1024)
    assert o.eqps.max() > 0.0, "scenario should yield"


def test_j2_voxel_bitexact():
    g, o, _ = run_both(_j2_voxel_spec(), [1, 100, 100, 100, 300], True, "j2 32x32 on voxels", tile_nodes=300)
    assert o.eqps.max() > 0.0, "scenario should yield"


def test_heightfield_bitexact():
    rng = W.rng_for(11)
    mat = fw.Elastic(3e4, 0.25, 1000.0)
    dom = W.grid_domain(24, (0.3, 0.12), 0.4, mat, [fw.Gravity([0.0, -9.8])])
    spec = W.one_body_spec(dom, W.sinusoid_heightfield(rng))
    run_both(spec, [1, 400], True, "elastic 24x24 on height field")


def test_batch_alone_equals_mixed():
    """A scene's result is the same alone and in a mix of uneven scenes."""
    specs = W.batched_config(n_scenes=12, seed=5, side_lo=6, side_hi=40)
    log("[case] batch alone vs mixed, 12 scenes")
    mixed = _native.Arena(tile_nodes=256)
    ids = [mixed.add(s) for s in specs]
    mixed.step(1, -1.0, _native.FW_STEP_NO_CONTACT)
    mixed.step(120, -1.0, _native.FW_STEP_DIAG)
    for k in (0, 5, 11):
        alone = GpuScene(specs[k], tile_nodes=256)
        alone.warmup()
        alone.arena.step(120, -1.0, _native.FW_STEP_DIAG)
        for fld in (_native.FIELD_X, _native.FIELD_V, _native.FIELD_FORCE):
            assert np.array_equal(mixed.read(ids[k], fld), alone.read(fld)), f"scene {k} field {fld}"


def test_per_scene_launches_equal_batched():
    specs = W.batched_config(n_scenes=6, seed=9, side_lo=6, side_hi=30)
    log("[case] per-scene launch pattern vs batched")
    a = _native.Arena(tile_nodes=256)
    b = _native.Arena(tile_nodes=256)
    ia = [a.add(s) for s in specs]
    ib = [b.add(s) for s in specs]
    a.step(61, -1.0, _native.FW_STEP_DIAG)
    b.step(61, -1.0, _native.FW_STEP_DIAG | _native.FW_STEP_PER_SCENE)
    for x, y in zip(ia, ib):
        assert np.array_equal(a.read(x, _native.FIELD_X), b.read(y, _native.FIELD_X))


def test_graph_replay_equals_plain():
    specs = W.batched_config(n_scenes=5, seed=13, side_lo=6, side_hi=30)
    log("[case] hipGraph replay vs plain launches")
    a = _native.Arena(tile_nodes=256)
    b = _native.Arena(tile_nodes=256)
    ia = [a.add(s) for s in specs]
    ib = [b.add(s) for s in specs]
    a.step(1, -1.0, 0)
    b.step(1, -1.0, 0)
    for rep in range(4):
        a.step(20, -1.0, 0)
        b.step_graph(20, 0)
        b.sync()
        for x, y in zip(ia, ib):
            assert np.array_equal(a.read(x, _native.FIELD_X), b.read(y, _native.FIELD_X)), f"replay {rep}"
