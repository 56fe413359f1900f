"""Spring-mass scenes and selective mass scaling on the GPU, bit for bit against the oracle.

Spring tiles share the triangle kernel: two corners per slot, the rest length in
the slot record, velocities staged in place of rest positions and pushed to the
neighbouring tiles' halo copies every substep (the damping term reads both ends'
velocities).  Every case here runs on several tiles as well as on one, and the
arena mixes spring and triangle scenes in one launch.
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from flatworld_amd import workloads as W
from oracle.oracle import OracleScene
from tests.parity import GpuScene, compare, log

pytestmark = pytest.mark.gpu

FRAME = 1.0 / 60.0


def _lattice_spec(side, ground="plane", k=200.0, c=0.3, m=0.01, spacing=0.02, y0=0.04, jitter=0.1, seed=0,
                  initials=(), bending=True):
    dom = W.spring_lattice(side, (0.5 - 0.5 * side * spacing, y0), spacing, [k, c, m], [fw.Gravity([0.0, -9.8])],
                           initials=initials, jitter=jitter, seed=seed, bending=bending)
    g = None
    if ground == "plane":
        g = fw.GroundDomain(2, [0.0, 0.0], [0.05, 1.0])
    elif ground == "hf":
        g = W.sinusoid_heightfield(W.rng_for(seed))
    return W.one_body_spec(dom, g)


def _run(spec, chunks, label, **arena_kw):
    """Warm-up, then the chunks of substeps, comparing after each; returns the scenes and
    the number of chunks whose last substep had a node in ground contact."""
    log(f"[case] {label}: N={spec.n_nodes} springs={spec.n_elems} grounds={[g.kind for g in spec.grounds]} {arena_kw}")
    g = GpuScene(spec, **arena_kw)
    o = OracleScene(spec)
    g.warmup()
    o.warmup()
    compare(g, o, True, label + " warm-up")
    touched = 0
    for n in chunks:
        g.run(n)
        o.run(n)
        compare(g, o, True, label)
        assert np.array_equal(g.read(_native.FIELD_CONTACT), o.contact_flag), f"{label}: contact flags differ"
        touched += int(np.count_nonzero(o.contact_flag) > 0)
    return g, o, touched


def test_spring_lattice_one_tile_plane():
    spec = _lattice_spec(10, "plane")
    g, o, touched = _run(spec, [1] + [50] * 12, "spring 10x10 lattice, one tile, plane", tile_nodes=512, tile_slots=1024)
    assert touched > 0, "lattice should reach the plane"
    assert np.array_equal(g.read(_native.FIELD_MASS), o.mass)


def test_spring_lattice_multitile_heightfield():
    spec = _lattice_spec(60, "hf", initials=[fw.InitialVel("ALL", [0.2, -0.3])], y0=0.06, seed=5)
    g, o, touched = _run(spec, [1] + [75] * 6, "spring 60x60 lattice, many tiles, height field", tile_nodes=256,
                         tile_slots=1024)
    stats = _native.tiling_check(spec, 256, 1024)
    assert stats["tiles"] > 10 and stats["halo_slots"] > 0
    assert touched > 0, "lattice should reach the height field"


@pytest.mark.parametrize("shape", [(256, 512, 256), (256, 768, 512), (256, 1024, 512), (512, 1024, 512),
                                   (512, 2048, 1024), (512, 2048, 512), (1024, 2048, 1024)])
def test_spring_every_kernel_shape(shape):
    block, slots, nodes = shape
    spec = _lattice_spec(36, "plane", jitter=0.2, seed=2)
    _run(spec, [1, 120], f"spring 36x36, shape {shape}", tile_nodes=nodes, tile_slots=slots, block=block)


def test_spring_velocity_write_reaches_halo_copies():
    """fw_scene_write of velocities also refreshes the halo velocity copies the next substep reads."""
    spec = _lattice_spec(40, "plane", seed=3)
    g, o, _ = _run(spec, [1, 50], "spring 40x40 before a velocity write", tile_nodes=128, tile_slots=1024)
    rng = np.random.default_rng(9)
    v = (o.v + rng.normal(0.0, 0.05, o.v.shape)).astype(np.float32)
    g.arena.write(g.sid, _native.FIELD_V, v)
    o.v[:] = v
    for n in (1, 60):
        g.run(n)
        o.run(n)
        compare(g, o, True, "spring 40x40 after a velocity write")


def test_mixed_arena_springs_and_triangles():
    """Spring and triangle scenes in one arena, one launch per substep, each bit-exact."""
    arena = _native.Arena(tile_nodes=256, tile_slots=1024)
    specs = [_lattice_spec(24, "plane", seed=1), W.small_config(), _lattice_spec(48, "hf", seed=4),
             W.one_body_spec(W.grid_domain(30, (0.3, 0.05), 0.4, fw.MisesPlastic(1e5, 0.3, 300.0, 500.0, 1e3),
                                           [fw.Gravity([0.0, -9.8])]), fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]))]
    log(f"[case] mixed arena: {[(s.mat_type, s.n_nodes, s.n_elems) for s in specs]}")
    gs = [GpuScene(s, arena=arena) for s in specs]
    os_ = [OracleScene(s) for s in specs]
    arena.step(1, 0.0, _native.FW_STEP_NO_CONTACT | _native.FW_STEP_DIAG, damping_override=0.0)
    for o in os_:
        o.warmup()
    for n in (1, 100, 200):
        arena.step(n, -1.0, _native.FW_STEP_DIAG)  # every scene with its own dt
        for o in os_:
            o.run(n)
        for g, o in zip(gs, os_):
            compare(g, o, True, f"mixed arena scene mat={g.spec.mat_type}")


# ------------------------------------------------------------------ mass scaling


def _fem_spec(side, mat, jitter=0.25, seed=0):
    dom = W.grid_domain(side, (0.3, 0.05), 0.4, mat, [fw.Gravity([0.0, -9.8])], jitter=jitter, seed=seed)
    return W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]))


@pytest.mark.parametrize("case", ["elastic-one-tile", "j2-multitile", "springs-multitile"])
def test_mass_scaling_bitexact(case):
    if case == "elastic-one-tile":
        spec, kw = _fem_spec(12, fw.Elastic(5e4, 0.3, 500.0)), dict(tile_nodes=512, tile_slots=1024)
    elif case == "j2-multitile":
        spec, kw = _fem_spec(40, fw.MisesPlastic(1e5, 0.3, 300.0, 500.0, 1e3), seed=3), dict(tile_nodes=200, tile_slots=1024)
    else:
        spec, kw = _lattice_spec(40, "plane", seed=7), dict(tile_nodes=128, tile_slots=1024)
    log(f"[case] mass scaling {case}: N={spec.n_nodes} E={spec.n_elems}")
    g = GpuScene(spec, **kw)
    o = OracleScene(spec)
    g.warmup()
    o.warmup()
    # a target that scales a good share of the elements (springs: 4 k dt^2 above the node mass)
    dt_t = (12.0 if spec.mat_type == 4 else 3.0) * spec.dt
    g.arena.mass_scaling(g.sid, dt_t)
    added = o.mass_scaling(dt_t)
    m = g.read(_native.FIELD_MASS)
    assert added > 0.0
    assert np.array_equal(m, o.mass), f"mass differs at {np.flatnonzero(m != o.mass)[:5]}"
    for n in (1, 150):  # heavier nodes only: the scene's own dt stays stable
        g.run(n)
        o.run(n)
        compare(g, o, True, f"mass-scaled {case}")


def test_mass_scaling_after_motion_uses_current_positions():
    """The scaling reads the positions of the substep it is enqueued behind, as the reference's kernel reads coords."""
    spec = _fem_spec(20, fw.HyperElastic(4e4, 0.3, 400.0), seed=4)
    g = GpuScene(spec, tile_nodes=128, tile_slots=1024)
    o = OracleScene(spec)
    g.warmup()
    o.warmup()
    g.run(77)
    o.run(77)
    g.arena.mass_scaling(g.sid, 2.5 * spec.dt)
    o.mass_scaling(2.5 * spec.dt)
    assert np.array_equal(g.read(_native.FIELD_MASS), o.mass)


# ------------------------------------------------------------------ drop-in API


def test_reference_spring_network_loop():
    """test2D/test_2Dspring_10000.py through ExplicitLoop: the reference's bounds, and bit-exact positions."""
    from tests.test_oracle_springs import _spring_grid_10000

    coords, conn = _spring_grid_10000()
    dom = fw.SpringMassDomain(2, coords + 0.25, conn, [10.0, 1.0, 1.0], [fw.Gravity([0, -100.0])], False)
    loop = fw.ExplicitLoop(0.0, [dom], useAdapativeDT=True)
    o = OracleScene(fw.pack_spring_domain(dom))
    o.warmup()
    log(f"[case] spring network loop, dt={loop.dt:.4f}")
    t = 0.0
    while t < 0.1:
        loop.advanceWithTime(FRAME)
        for sd in loop.substep_sequence(FRAME):
            o.substep(sd, loop.damping)
        t += FRAME
    fsm = loop.femSpringManager
    pos = fsm.coords.numpy()[: fsm.totalNodes.numpy()[0]]
    e2n = fsm.connectivity.numpy()[: fsm.totalElements.numpy()[0]]
    assert np.array_equal(e2n[:, :2], conn) and np.all(e2n[:, 2:] == -1)
    assert np.allclose(pos.min(0), [0.25, -0.5278], atol=1e-2) and np.allclose(pos.max(0), [0.85, 0.072], atol=1e-2)
    assert np.array_equal(pos, o.x)
    assert np.array_equal(fsm.elemVolume.numpy()[: conn.shape[0]], dom.restLength)


def test_reference_spring_bcs_loop():
    """test2D/test_2Dspring_bcs.py through ExplicitLoop: three spring domains, min = [0.0, -0.252]."""
    conn = np.array([[0, 1], [1, 2], [0, 2], [2, 3], [3, 1], [0, 3]], dtype=np.int32)
    coords = np.array([[0.5, 0.5], [0.7, 0.5], [0.5, 0.7], [0.7, 0.7]], dtype=np.float32)
    doms = [fw.SpringMassDomain(2, coords, conn, [100.0, 1.0, 1.0], [fw.Force([0, 1], [0, -1.0]), fw.Fixed([2, 3])], False),
            fw.SpringMassDomain(2, coords - 0.25, conn, [100.0, 1.0, 1.0], [fw.EnforceVel([0, 1], [0, -1.0]), fw.Fixed([2, 3])], False),
            fw.SpringMassDomain(2, coords - 0.5, conn, [100.0, 1.0, 1.0], [fw.EnforceAcc([0, 1], [0, -10.0]), fw.Fixed([2, 3])], False)]
    loop = fw.ExplicitLoop(0.0, doms, useAdapativeDT=True)
    os_ = [OracleScene(fw.pack_spring_domain(d)) for d in doms]
    for o in os_:
        o.warmup()
    t = 0.0
    while t < 0.2:
        loop.advanceWithTime(FRAME)
        for sd in loop.substep_sequence(FRAME):
            for o in os_:
                o.substep(sd, loop.damping)
        t += FRAME
    fsm = loop.femSpringManager
    pos = fsm.coords.numpy()[: fsm.totalNodes.numpy()[0]]
    assert np.allclose(pos.min(0), [0.0, -0.252], atol=1e-3)
    assert np.array_equal(pos, np.concatenate([o.x for o in os_]))


def test_spring_lattice_on_plane_loop():
    """A lattice dropped on a plane through ExplicitLoop: spring-plane pairs (penalty 100 k), contact stable time."""
    dom = W.spring_lattice(12, (0.38, 0.08), 0.02, [200.0, 0.3, 0.01], [fw.Gravity([0.0, -9.8])])
    g = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    loop = fw.ExplicitLoop(0.0, [g, dom], useAdapativeDT=True)
    c = loop.contacts[0]
    assert c.penalty == 20000.0 and loop.dt == min(1e-3, dom.stableTime, c.stableTime)
    o = OracleScene(fw.pack_spring_domain(dom, [(g, c)]))
    o.warmup()
    for _ in range(30):
        loop.advanceWithTime(FRAME)
        for sd in loop.substep_sequence(FRAME):
            o.substep(sd, loop.damping)
    pos = loop.femSpringManager.coords.numpy()[: dom.nnodes]
    assert np.array_equal(pos, o.x)
    assert pos[:, 1].min() > -0.01, "lattice fell through the plane"


def test_mass_scaling_loop():
    """ExplicitLoop(mass_scaling_dt=...): masses scaled after the warm-up, stable time raised to the target."""
    dom = W.grid_domain(16, (0.3, 0.1), 0.3, fw.Elastic(1e5, 0.3, 200.0), [fw.Gravity([0.0, -9.8])], jitter=0.2, seed=6)
    target = 3.0 * dom.stableTime
    loop = fw.ExplicitLoop(0.0, [dom], mass_scaling_dt=target, useAdapativeDT=True)
    assert loop.stableTime == target and loop.dt == target
    o = OracleScene(fw.pack_fem_domain(dom))
    o.warmup()
    o.mass_scaling(target)
    assert np.array_equal(loop.femSpringManager.mass.numpy()[: dom.nnodes], o.mass)
    for _ in range(10):
        loop.advanceWithTime(FRAME)
        for sd in loop.substep_sequence(FRAME):
            o.substep(sd, loop.damping)
    assert np.array_equal(loop.femSpringManager.coords.numpy()[: dom.nnodes], o.x)


def test_mixed_arena_graph_and_per_scene_launches_equal_plain():
    """Triangle tiles and spring tiles run as two launches per substep; graph-replayed frames
    and one launch per scene give the same bits as plain launches, and adding or removing
    the last spring scene re-captures the frame graph with the right launches."""
    specs = [_lattice_spec(30, "hf", seed=11), W.small_config(), _lattice_spec(16, "plane", seed=12)]
    arenas = {m: _native.Arena(tile_nodes=256, tile_slots=1024) for m in ("plain", "graph", "per_scene")}
    ids = {m: [a.add(s) for s in specs] for m, a in arenas.items()}
    log(f"[case] mixed arena launch modes: {[(s.mat_type, s.n_elems) for s in specs]}")
    for a in arenas.values():
        a.step(1, -1.0, 0)
    for _ in range(3):
        arenas["plain"].step(40, -1.0, 0)
        arenas["graph"].step_graph(40, 0)
        arenas["per_scene"].step(40, -1.0, _native.FW_STEP_PER_SCENE)
    for k in range(len(specs)):
        ref = arenas["plain"].read(ids["plain"][k], _native.FIELD_X)
        for m in ("graph", "per_scene"):
            assert np.array_equal(arenas[m].read(ids[m][k], _native.FIELD_X), ref), (m, k)
    # drop the spring scenes: frames hold the triangle launch only; add one back: both again
    for m in ("plain", "graph"):
        arenas[m].remove(ids[m][0])
        arenas[m].remove(ids[m][2])
    arenas["plain"].step(40, -1.0, 0)
    arenas["graph"].step_graph(40, 0)
    assert np.array_equal(arenas["graph"].read(ids["graph"][1], _native.FIELD_X),
                          arenas["plain"].read(ids["plain"][1], _native.FIELD_X))
    extra = _lattice_spec(20, "plane", seed=13)
    e_plain, e_graph = arenas["plain"].add(extra), arenas["graph"].add(extra)
    for _ in range(2):
        arenas["plain"].step(40, -1.0, 0)
        arenas["graph"].step_graph(40, 0)
    for sp, sg in ((ids["plain"][1], ids["graph"][1]), (e_plain, e_graph)):
        assert np.array_equal(arenas["graph"].read(sg, _native.FIELD_X), arenas["plain"].read(sp, _native.FIELD_X))
    o = OracleScene(extra)
    o.warmup()  # the first of the 80 launches after the add is the scene's warm-up
    o.run(79)
    assert np.array_equal(arenas["graph"].read(e_graph, _native.FIELD_X), o.x)


def test_served_spring_and_triangle_jobs_equal_solo_runs():
    """Continuous batching with spring lattices among triangle jobs: every served result equals its solo run."""
    import dataclasses

    from flatworld_amd.batch import FemBatch
    from flatworld_amd.server import Job, SimServer

    rng = W.rng_for(41)
    specs = []
    for k in range(10):
        if k % 2:
            specs.append(W.serving_job(rng, 6, 20, 20, 120))
        else:
            s = _lattice_spec(int(rng.integers(6, 30)), "plane" if k % 4 else "hf", seed=50 + k)
            specs.append(dataclasses.replace(s, horizon=int(rng.integers(20, 120))))
    jobs = [Job(i, s, 0.002 * i) for i, s in enumerate(specs)]
    log(f"[case] serve {len(jobs)} mixed spring / triangle jobs")
    srv = SimServer(frame=20, tile_nodes=256)
    res = {r.job_id: r for r in srv.run(jobs)}
    assert len(res) == len(jobs) and all(r.ok for r in res.values())
    for j in jobs:
        b = FemBatch(tile_nodes=256)
        sid = b.add(j.spec)
        b.step(1 + j.spec.horizon, graph=False)
        x = b.positions(sid)
        b.close()
        assert np.array_equal(res[j.job_id].positions, x), f"job {j.job_id} (material {j.spec.mat_type})"


def test_spring_4096_arena_sampled_scenes_bitexact():
    """The full 4096-lattice spring arena of the benchmark, in its tuned shape with graph frames:
    32 sampled lattices against the oracle after 60 substeps."""
    specs = W.spring_batched_config(n_scenes=4096, seed=0)
    log(f"[case] spring arena, 4096 lattices, {sum(s.n_elems for s in specs)} springs")
    arena = _native.Arena(tile_nodes=1024, tile_slots=2048, block=512)
    ids = [arena.add(s) for s in specs]
    arena.step(1, -1.0, 0)
    arena.step_graph(59, 0)
    arena.step(1, -1.0, _native.FW_STEP_DIAG)
    pick = W.rng_for(98).choice(len(specs), size=32, replace=False)
    for k in pick:
        o = OracleScene(specs[k])
        o.warmup()
        o.run(60)
        for fld, ref in ((_native.FIELD_X, o.x), (_native.FIELD_V, o.v), (_native.FIELD_FORCE, o.fext)):
            assert np.isfinite(ref).all()
            assert np.array_equal(arena.read(ids[k], fld), ref), f"lattice {k} field {fld}"
    arena.close()


def test_irregular_spring_network_bitexact():
    """A random network: edges of a jittered triangle mesh, a random subset dropped, plus
    duplicate springs and isolated nodes, so node degrees run from 0 to about 8 and the
    incidence ranks and halos are irregular; several tiles, plane contact."""
    rng = np.random.default_rng(77)
    conn_t, X = fw.structured_grid(40, lower=(0.2, 0.03), extent=0.6, jitter=0.3, seed=77)
    e = np.concatenate([conn_t[:, [0, 1]], conn_t[:, [1, 2]], conn_t[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    e = e[rng.random(e.shape[0]) < 0.8]
    e = np.concatenate([e, e[rng.choice(e.shape[0], 50, replace=False)]])  # duplicate springs
    e = e[rng.permutation(e.shape[0])].astype(np.int32)
    flip = rng.random(e.shape[0]) < 0.5  # either orientation
    e[flip] = e[flip][:, ::-1]
    X = np.concatenate([X, rng.uniform(0.2, 0.8, (5, 2)).astype(np.float32)])  # five isolated nodes
    dom = fw.SpringMassDomain(2, X, np.ascontiguousarray(e), [150.0, 0.2, 0.01], [fw.Gravity([0.0, -9.8])])
    spec = W.one_body_spec(dom, fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0]))
    assert np.bincount(e.reshape(-1), minlength=X.shape[0]).min() == 0
    _run(spec, [1, 100, 200], "irregular spring network", tile_nodes=128, tile_slots=1024)
