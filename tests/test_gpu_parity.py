"""HIP substep vs the CPU oracle on identical seeded inputs.

Every material must agree bit for bit: both sides evaluate the same float32
operations in the same order without contraction, and the Neo-Hookean log is
one shared double-operation routine (correctly rounded in practice, pinned
against numpy in the oracle tests).  Every comparison covers positions,
velocities, nodal forces, displacement U, contact-active flags, the boundary
AABB, Cauchy stress, Green strain, the plastic-yield flag and the J2 state
(tests/parity.py ``check_fields``).  The Neo-Hookean case also reports the
tolerance figures (nodal forces within 1e-5 of the largest nodal force,
positions within 1e-4 of the minimum edge length) for a reader comparing a
build whose log differs.
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from flatworld_amd import autotune
from flatworld_amd import workloads as W
from oracle.oracle import OracleScene
from tests.parity import FORCE_RTOL, POS_TOL_EDGE, GpuScene, check_fields, compare, log

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
        out.append(compare(g, o, exact, label + f" after {o.counter} substeps"))
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


# ---------------------------------------------------------------- full sizes of the benchmark configurations


@pytest.mark.timeout(400)
def test_large_sheet_full_size_bitexact():
    """1000x1000 jittered Neo-Hookean sheet (2 M triangles) on a plane, CFL dt, all 1000 substeps
    of the configuration, every field compared after 1, 100 and 1000 substeps."""
    spec = W.large_config()
    g, o, out = run_both(spec, [1, 99, 900], True, "large 1000x1000 nh, plane")
    res = compare(g, o, False, "large tolerance figures")
    log(f"  {res} {out}")
    assert res["force_rel"] < FORCE_RTOL and res["pos_err_edges"] < POS_TOL_EDGE
    assert all(c["contact"] > 0 for c in out), "jittered bottom row should touch the plane throughout"


@pytest.mark.timeout(400)
def test_j2_voxel_full_size_bitexact():
    """512x512 J2 block (524 k triangles) dropped at 2 m/s onto 256x256 value-noise voxels, all 2000
    substeps of the configuration; the block must be in voxel contact and yielding."""
    spec = W.j2_voxel_config()
    g, o, out = run_both(spec, [1, 199, 800, 1000], True, "j2 512x512 on 256x256 voxels")
    log(f"  contact / yield per chunk: {out}")
    assert all(c["contact"] > 0 and c["yielded"] > 0 for c in out[1:]), "block should touch the voxels and yield"
    assert o.eqps.max() > 0.0


@pytest.mark.timeout(400)
def test_batched_4096_sampled_scenes_bitexact():
    """The full 4096-scene arena of the benchmark, stepped as the benchmark steps it (hipGraph
    frames of 60 substeps) for 1200 substeps, long enough for most blocks to land on their
    height fields: 48 sampled scenes against the oracle, every field."""
    specs = W.batched_config(n_scenes=4096, seed=0)
    log(f"[case] batched 4096 scenes, {sum(s.n_elems for s in specs)} triangles, 20 graph frames")
    arena = _native.Arena()
    ids = [arena.add(s) for s in specs]
    arena.step(1, -1.0, 0)
    for _ in range(20):
        arena.step_graph(60, _native.FW_STEP_DIAG)
    arena.sync()
    pick = W.rng_for(99).choice(len(specs), size=48, replace=False)
    touched = plastic = 0
    for k in pick:
        o = OracleScene(specs[k])
        o.warmup()
        o.run(1200)
        c = check_fields(lambda f: arena.read(ids[k], f), o, f"scene {k}")
        touched += int(c["contact"] > 0)
        plastic += int(specs[k].mat_type == 3 and o.eqps.max() > 0.0)
    log(f"  sampled scenes in contact at the end: {touched}, J2 scenes with plastic strain: {plastic}")
    assert touched >= 10 and plastic >= 5, "the window should cover height-field contact and yielding"
    arena.close()


def test_every_kernel_shape_gives_identical_results():
    """Tiling, kernel shape and launch mode change no bit of the result."""
    specs = W.batched_config(n_scenes=10, seed=31, side_lo=6, side_hi=48)
    log("[case] all compiled shapes, one workgroup per tile and persistent")
    ref = None
    for block, slots, nodes in autotune.SHAPES:
        for gpc in (0, 4):
            a = _native.Arena(capacity_bytes=1 << 30, tile_nodes=nodes, tile_slots=slots, block=block, grid_per_cu=gpc)
            ids = [a.add(s) for s in specs]
            a.step(1, -1.0, 0)
            a.step(80, -1.0, _native.FW_STEP_DIAG)
            got = [(a.read(i, _native.FIELD_X), a.read(i, _native.FIELD_V), a.read(i, _native.FIELD_FORCE)) for i in ids]
            a.close()
            if ref is None:
                ref = got
                o = OracleScene(specs[3])
                o.warmup()
                o.run(80)
                assert np.array_equal(got[3][0], o.x) and np.array_equal(got[3][2], o.fext)
                continue
            for k, (x, v, f) in enumerate(got):
                assert np.array_equal(x, ref[k][0]) and np.array_equal(v, ref[k][1]) and np.array_equal(f, ref[k][2]), \
                    f"shape ({block},{slots},{nodes}) grid_per_cu={gpc} scene {k}"


# ---------------------------------------------------------------- edge cases


def _nan_equal(a, b):
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(np.where(na, 0.0, a), np.where(nb, 0.0, b))


def test_ragged_arena_bitexact():
    """One- and two-triangle scenes next to a multi-tile 64x64 scene, three materials, three ground kinds."""
    rng = W.rng_for(41)
    tri = fw.FemDomain(fw.Mesh(2, np.array([[0, 1, 2]], np.int32), np.array([[0.4, 0.0005], [0.6, 0.0005], [0.5, 0.2]], np.float32)),
                       fw.SolidProp(fw.Elastic(2e4, 0.2, 40.0)), [fw.Gravity([0.0, -9.8])],
                       initials=[fw.InitialVel("ALL", [0.0, -0.5])])
    hf = W.sinusoid_heightfield(rng)
    y0 = float(hf.height.max()) + 0.001
    two = fw.FemDomain(fw.Mesh(2, np.array([[0, 1, 3], [0, 3, 2]], np.int32),
                               np.array([[0.5, y0], [0.7, y0], [0.5, y0 + 0.2], [0.7, y0 + 0.2]], np.float32)),
                       fw.SolidProp(fw.HyperElastic(3e4, 0.3, 50.0)), [fw.Gravity([0.0, -9.8])],
                       initials=[fw.InitialVel("ALL", [0.0, -0.5])])
    big = W.grid_domain(64, (0.25, 0.105), 0.5, fw.MisesPlastic(1e5, 0.3, 300.0, 500.0, 1e3), [fw.Gravity([0.0, -9.8])],
                        initials=[fw.InitialVel("ALL", [0.0, -1.0])])
    mid = W.grid_domain(8, (0.45, 0.2), 0.1, fw.Elastic(1e4, 0.25, 100.0), [fw.Gravity([0.0, -9.8])])
    specs = [
        W.one_body_spec(tri, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])),
        W.one_body_spec(two, hf),
        W.one_body_spec(big, W.value_noise_voxels(rng, 64, 32, [0.0, 0.0], [1.0, 0.1], floor_rows=2)),
        W.one_body_spec(mid, fw.GroundDomain(2, [0.0, 0.19], [0.05, 1.0])),
    ]
    log("[case] ragged arena: 1, 2, 128, 8192 triangles")
    arena = _native.Arena(capacity_bytes=1 << 30, tile_nodes=512)
    ids = [arena.add(s) for s in specs]
    arena.step(1, -1.0, 0)
    arena.step(120, -1.0, _native.FW_STEP_DIAG)
    assert arena.info()["n_tiles"] > len(specs)
    for k, s in enumerate(specs):
        o = OracleScene(s)
        o.warmup()
        o.run(120)
        check_fields(lambda f: arena.read(ids[k], f), o, f"ragged scene {k}")
    arena.close()


def test_degenerate_element_and_isolated_node_match_oracle():
    """A zero-area element (repeated node) and a node no element references: NaN appears where the
    oracle's does.  The reference's wp.inverse of a singular matrix cannot be reproduced without Warp,
    so this pins the choice both sides make (adjugate times 1/det, no guard), not the reference."""
    conn, coords = fw.structured_grid(5, lower=(0.3, 0.2), extent=0.3)
    coords = np.concatenate([coords, [[0.9, 0.9]]]).astype(np.float32)
    conn = np.concatenate([conn, [[0, 1, 1]]]).astype(np.int32)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(2e4, 0.3, 100.0)), [fw.Gravity([0.0, -9.8])])
    spec = W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]), dt=1e-4)
    log("[case] degenerate element + isolated node")
    for tile_nodes, tile_slots, block in ((512, 1024, 256), (8, 512, 256)):
        g = GpuScene(spec, tile_nodes=tile_nodes, tile_slots=tile_slots, block=block)
        o = OracleScene(spec)
        g.warmup()
        o.warmup()
        g.run(20)
        o.run(20)
        for name, a, b in (("x", g.read(_native.FIELD_X), o.x), ("v", g.read(_native.FIELD_V), o.v)):
            assert _nan_equal(a, b), f"{name} (tile_nodes {tile_nodes})"
        assert np.isnan(o.x[-1]).any(), "the massless isolated node should be NaN, as in the reference"
        g.arena.close()


def _fan(n_tri, r=0.1, c=(0.5, 0.102)):
    ang = np.linspace(0.0, 2.0 * np.pi, n_tri, endpoint=False)
    coords = np.concatenate([[c], np.asarray(c) + r * np.stack([np.cos(ang), np.sin(ang)], 1)]).astype(np.float32)
    conn = np.array([[0, 1 + k, 1 + (k + 1) % n_tri] for k in range(n_tri)], dtype=np.int32)
    return conn, coords


@pytest.mark.parametrize("n_tri,shape", [(64, (256, 1024, 64)), (300, (256, 1024, 8)), (1000, (512, 2048, 64))])
def test_high_valence_fan_bitexact(n_tri, shape):
    """A disc of n_tri triangles around one node (n_tri element corners at the centre, beyond
    the 32 a corner word holds) falling onto a plane, cut into several tiles."""
    block, slots, nodes = shape
    conn, coords = _fan(n_tri)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(2e4, 0.3, 100.0)), [fw.Gravity([0.0, -9.8])],
                       initials=[fw.InitialVel("ALL", [0.0, -1.0])])
    spec = W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]))
    _, o, out = run_both(spec, [1, 200, 200], True, f"fan of {n_tri} triangles", tile_nodes=nodes, tile_slots=slots,
                         block=block)
    assert _native.tiling_check(spec, nodes, slots)["tiles"] > 1
