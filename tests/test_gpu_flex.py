"""FEM-FEM contact on the GPU against the oracle, bit for bit.

Loops run through ``ExplicitLoop.advanceWithTime``; the oracle (OracleLoop) gets
the same packed scenes, the same contact group and the same float64 dt sequence,
so positions, velocities and every work item's cached segment must agree
exactly.  Cases: the reference's own two-circle test
(test2D/test_2Dfem_contact.py), stacked boxes on a plane (the gated path, ties
between collinear edges, friction), three bodies of three materials on a height
field, bodies spread over many tiles, and tied pairs (ExplicitLoop ``ties``):
resolved once at the first substep, then held by penalty springs, alone and
beside sliding pairs.

FEM-FEM contact at the reference's adaptive step diverges as soon as both
bodies press on each other through two work items (the contact stable time
counts one penalty spring), in the oracle exactly as on the device, so the
multi-body cases below step at a fixed fraction of it.
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from oracle.oracle import OracleLoop
from tests.parity import log

pytestmark = pytest.mark.gpu

FRAME = 1.0 / 60.0


def ccw_grid(side, lower, extent, mat, bcs=(), initials=(), jitter=0.0, seed=0, friction=0.0):
    conn, coords = fw.structured_grid(side, lower=lower, extent=extent, jitter=jitter, seed=seed)
    conn = np.ascontiguousarray(conn[:, [0, 2, 1]])
    return fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(mat), list(bcs), initials=list(initials),
                        friction=friction)


def oracle_of(loop):
    m = loop.femSpringManager
    return OracleLoop(m._specs, m._flex)


def assert_same(loop, o, what):
    m = loop.femSpringManager
    for k, body in enumerate(o.bodies):
        x = m.domain_coords(k)
        v = m._arena.read(m._scene_ids[k], _native.FIELD_V)
        assert np.array_equal(x, body.x), f"{what}: positions of body {k} differ (max {np.abs(x - body.x).max():.3e})"
        assert np.array_equal(v, body.v), f"{what}: velocities of body {k} differ"
    if o.flex is not None:
        assert np.array_equal(m.flex_cache(), o.flex.cache), f"{what}: cached segments differ"
        if o.flex.group.n_tied:
            seg, state = m.flex_ties()
            so, sto = o.flex.ties
            assert np.array_equal(seg, so) and np.array_equal(state, sto), f"{what}: tie state differs"


def run_frames(loop, o, frames, check_every=1, what=""):
    contact = 0
    for f in range(frames):
        loop.advanceWithTime(FRAME)
        for sd in loop.substep_sequence(FRAME):
            o.substep(sd, loop.damping)
        contact += int((o.flex.cache >= 0).sum())
        if (f + 1) % check_every == 0 or f == frames - 1:
            assert_same(loop, o, f"{what} frame {f + 1}")
    return contact


def test_reference_two_circles_bitexact():
    """test2D/test_2Dfem_contact.py: two 16-segment circles under opposite gravity."""
    msh = fw.FEMesher(2)
    prop = fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0))
    d1 = fw.FemDomain(msh.createCircle([0.5, 0.5], 0.1), prop, [fw.Gravity([0, 10.0])])
    d2 = fw.FemDomain(msh.createCircle([0.5, 0.8], 0.1), prop, [fw.Gravity([0, -10.0])])
    loop = fw.ExplicitLoop(0.0, [d1, d2], useAdapativeDT=True)
    log(f"[case] two circles, dt={loop.dt:.3e}, items={loop.femSpringManager._flex.n_items}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 60, check_every=5, what="two circles")
    assert contact > 0


def test_hash_build_matches_oracle():
    mat = fw.Elastic(1e4, 0.3, 10.0)
    a = ccw_grid(8, (0.0, 0.0), 0.3, mat, [fw.Gravity([0, -5.0])], jitter=0.2, seed=1)
    b = ccw_grid(6, (0.05, 0.305), 0.2, mat, [fw.Gravity([0, -5.0])], jitter=0.2, seed=2)
    loop = fw.ExplicitLoop(0.0, [a, b], useAdapativeDT=True)
    o = oracle_of(loop)
    m = loop.femSpringManager
    for rnd in range(3):
        gd = m._arena.flex_read(m._flex_gid, _native.FLEX_GRID)
        go = o.flex.grid
        log(f"[case] hash round {rnd}: cells {gd['cn']}, used {gd['tc']}")
        assert np.array_equal(gd["glb"], go["glb"]) and np.array_equal(gd["gs"], go["gs"])
        assert np.array_equal(gd["cn"], go["cn"]) and gd["tc"] == go["tc"] and gd["exact"]
        cs = m._arena.flex_read(m._flex_gid, _native.FLEX_CELL_START)[: gd["tc"]]
        ce = m._arena.flex_read(m._flex_gid, _native.FLEX_CELL_END)[: gd["tc"]]
        srt = m._arena.flex_read(m._flex_gid, _native.FLEX_SORTED)
        assert np.array_equal(cs, o.flex.cell_start[: go["tc"]]) and np.array_equal(ce, o.flex.cell_end[: go["tc"]])
        assert np.array_equal(srt[: ce.max()], o.flex.sorted_seg[: ce.max()])
        run_frames(loop, o, 2, what=f"hash round {rnd}")  # several rebuilds per frame


def test_stacked_boxes_on_plane_bitexact():
    """A box dropped on a box resting on a plane: the BVH-gated path, aligned grids
    whose nodes sit exactly over edge ends (penetration ties), and friction."""
    mat = fw.Elastic(2e4, 0.3, 50.0)
    lower = ccw_grid(6, (0.2, 0.0), 0.3, mat, [fw.Gravity([0, -9.8])], friction=0.3)
    upper = ccw_grid(6, (0.2, 0.31), 0.3, mat, [fw.Gravity([0, -9.8])], friction=0.1)
    plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    loop = fw.ExplicitLoop(3.5e-4, [lower, upper, plane], useAdapativeDT=False, damping=0.5)  # half the stable step
    g = loop.femSpringManager._flex
    assert g.gated
    log(f"[case] stacked boxes on a plane, dt={loop.dt:.3e}, items={g.n_items}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 20, check_every=4, what="stacked boxes")
    assert contact > 0


def test_three_materials_on_height_field_bitexact():
    hf_h = (0.05 + 0.03 * np.sin(np.linspace(0.0, 6.0, 65))).astype(np.float32)
    hf = fw.HeightFieldDomain(2, hf_h, lb=[0.0, 0.0], ub=[1.0, 1.0])
    a = ccw_grid(8, (0.1, 0.1), 0.25, fw.Elastic(3e4, 0.3, 60.0), [fw.Gravity([0, -9.8])], jitter=0.15, seed=3)
    b = ccw_grid(7, (0.3, 0.37), 0.22, fw.HyperElastic(2e4, 0.35, 40.0), [fw.Gravity([0, -9.8])], jitter=0.15,
                 seed=4, initials=[fw.InitialVel("ALL", [-0.2, -0.5])])
    c = ccw_grid(6, (0.12, 0.62), 0.2, fw.MisesPlastic(2e4, 0.3, 50.0, 300.0, 1e3), [fw.Gravity([0, -9.8])],
                 jitter=0.1, seed=5, friction=0.2)
    loop = fw.ExplicitLoop(1e-4, [a, b, c, hf], useAdapativeDT=False)  # 0.38 x the stable step
    g = loop.femSpringManager._flex
    assert g.gated and g.n_pairs == 3
    log(f"[case] three materials on a height field, dt={loop.dt:.3e}, items={g.n_items}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 10, check_every=5, what="three materials")
    assert contact > 0


def test_bodies_over_many_tiles_bitexact():
    """40 x 40 and 30 x 30 bodies cut into 128-node tiles: internal node order
    differs from the caller's, halos cross the contact zone."""
    mat = fw.Elastic(1e4, 0.3, 20.0)
    a = ccw_grid(40, (0.0, 0.0), 0.4, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=7)
    b = ccw_grid(30, (0.1, 0.401), 0.3, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=8,
                 initials=[fw.InitialVel("ALL", [0.0, -0.8])])
    loop = fw.ExplicitLoop(3e-5, [a, b], useAdapativeDT=False,  # 0.3 x the stable step
                           arena_kwargs=dict(tile_nodes=128, tile_slots=1024))
    assert loop.femSpringManager.arena_info()["n_tiles"] > 10
    log(f"[case] many tiles, dt={loop.dt:.3e}, items={loop.femSpringManager._flex.n_items}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 3, check_every=1, what="many tiles")
    assert contact > 0


def test_group_refuses_scene_removal_and_per_scene_launches():
    mat = fw.Elastic(1e4, 0.3, 10.0)
    a = ccw_grid(4, (0.0, 0.0), 0.2, mat)
    b = ccw_grid(4, (0.0, 0.21), 0.2, mat)
    loop = fw.ExplicitLoop(0.0, [a, b], useAdapativeDT=True)
    m = loop.femSpringManager
    with pytest.raises(NotImplementedError, match="FW_STEP_PER_SCENE"):
        m._arena.step(2, -1.0, _native.FW_STEP_PER_SCENE)
    with pytest.raises(ValueError, match="remove the group first"):
        m._arena.remove(m._scene_ids[0])
    with pytest.raises(ValueError, match=">= 0"):
        m._arena.flex_schedule_rebuild(m._flex_gid, -1)
    for at in range(4):
        m._arena.flex_schedule_rebuild(m._flex_gid, 10 + at)
    m._arena.flex_schedule_rebuild(m._flex_gid, 11)  # already pending: nothing to add
    with pytest.raises(RuntimeError, match="pending rebuilds"):
        m._arena.flex_schedule_rebuild(m._flex_gid, 20)
    m._arena.step_graph(12)  # runs the first two; the other two stay pending
    m._arena.flex_schedule_rebuild(m._flex_gid, 5)
    m._arena.flex_schedule_rebuild(m._flex_gid, 6)
    with pytest.raises(RuntimeError, match="pending rebuilds"):
        m._arena.flex_schedule_rebuild(m._flex_gid, 7)
    m._arena.step_graph(12)
    m._arena.flex_remove(m._flex_gid)
    m._flex_gid = None
    m._arena.step_graph(4)
    m._arena.remove(m._scene_ids[0])
    m._arena.sync()


def test_graph_frames_with_scheduled_rebuilds_equal_plain_launches():
    """256 loops of the contact workload: 60-substep hipGraph frames with every
    hash rebuild scheduled on the device, against plain launches with the rebuilds
    enqueued between them (the reference's schedule, one rebuild every 147
    substeps), and against fw_step_noalloc with scheduled rebuilds: positions,
    velocities, contact flags and every work item's cache, bit for bit."""
    from flatworld_amd import workloads as W
    from flatworld_amd.flexcontact import REBUILD_INTERVAL

    groups = W.contact_batched_config(n_groups=256, seed=2)
    dt = groups[0][0][0].dt
    arenas, ids = [], []
    for _ in range(3):
        ar = _native.Arena(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0, capacity_bytes=1 << 30)
        sids, gids = [], []
        for sp, g in groups:
            s = [ar.add(x) for x in sp]
            sids.append(s)
            gids.append(ar.flex_add(g, s))
        ar.step(1, -1.0, 0)
        ar.flex_rebuild(-1)
        arenas.append(ar)
        ids.append((sids, gids))
    plain, graph, noalloc = arenas
    noalloc.step(0)
    n_total, frame = 660, 60
    rebuild_at, elapsed = [], 0.0
    for k in range(n_total):
        elapsed += dt
        if elapsed >= REBUILD_INTERVAL:
            rebuild_at.append(k)
            elapsed = 0.0
    assert len(rebuild_at) >= 4
    log(f"[case] 256 contact loops, rebuilds before substeps {rebuild_at}")
    pending = 0
    for k in range(n_total):
        if k in rebuild_at:
            if pending:
                plain.step(pending, -1.0, 0)
                pending = 0
            plain.flex_rebuild(-1)
        pending += 1
    plain.step(pending, -1.0, 0)
    for f0 in range(0, n_total, frame):
        for k in rebuild_at:
            if f0 <= k < f0 + frame:
                graph.flex_schedule_rebuild(-1, k - f0)
                noalloc.flex_schedule_rebuild(-1, k - f0)
        graph.step_graph(frame, _native.FW_STEP_DIAG)
        noalloc.step_noalloc(frame, -1.0, _native.FW_STEP_DIAG)
    for ar in arenas:
        ar.sync()
    # the plain run's last substep carried no DIAG: compare positions, velocities, caches
    cached = 0
    for gi in range(len(groups)):
        for m, ar in ((1, graph), (2, noalloc)):
            for j, sid in enumerate(ids[0][0][gi]):
                for fld in (_native.FIELD_X, _native.FIELD_V):
                    assert np.array_equal(ar.read(ids[m][0][gi][j], fld), plain.read(sid, fld)), (gi, m, j, fld)
            c = plain.flex_read(ids[0][1][gi], _native.FLEX_CACHE)
            assert np.array_equal(ar.flex_read(ids[m][1][gi], _native.FLEX_CACHE), c), (gi, m)
        cached += int((c >= 0).sum())
    assert cached > 100, "the loops should be in contact"
    for ar in arenas:
        ar.close()


def test_tie_resolution_needs_a_built_hash():
    """Tie resolution walks the group's spatial hash: before the first rebuild
    the call is refused instead of resolving nothing."""
    from flatworld_amd import workloads as W

    (sp, g), = W.contact_batched_config(n_groups=1, seed=4)
    g.item_tied = np.ones(g.n_items, np.int32)
    arena = _native.Arena(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0)
    gid = arena.flex_add(g, [arena.add(x) for x in sp])
    arena.step(1, -1.0, 0)
    with pytest.raises(ValueError, match="no spatial hash yet"):
        arena.flex_init_ties(gid)
    arena.flex_rebuild(gid)
    arena.flex_init_ties(gid)
    arena.step(4, -1.0, 0)
    arena.sync()


def test_contact_workload_sampled_loops_bitexact():
    """The bench's 2048-loop contact workload in one arena, 600 substeps on the
    reference's rebuild schedule (one launch rebuilds every loop's hash); 24
    sampled loops against the oracle: positions, velocities and caches."""
    from flatworld_amd import workloads as W
    from flatworld_amd.flexcontact import REBUILD_INTERVAL

    groups = W.contact_batched_config(n_groups=2048, seed=0)
    dt = groups[0][0][0].dt
    arena = _native.Arena(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0)
    sids, gids = [], []
    for sp, g in groups:
        s = [arena.add(x) for x in sp]
        sids.append(s)
        gids.append(arena.flex_add(g, s))
    arena.step(1, -1.0, 0)
    arena.flex_rebuild(-1)
    elapsed, pending = 0.0, 0
    for _ in range(600):
        elapsed += dt
        if elapsed >= REBUILD_INTERVAL:
            if pending:
                arena.step(pending, -1.0, 0)
                pending = 0
            arena.flex_rebuild(-1)
            elapsed = 0.0
        pending += 1
    arena.step(pending, -1.0, 0)
    arena.sync()
    sample = np.random.default_rng(5).choice(len(groups), 24, replace=False)
    hits = 0
    for gi in sample:
        sp, g = groups[gi]
        o = OracleLoop(sp, g)
        for _ in range(600):
            o.substep(dt)
        for k, body in enumerate(o.bodies):
            assert np.all(np.isfinite(body.x)), f"loop {gi}: oracle not finite"
            x = arena.read(sids[gi][k], _native.FIELD_X)
            v = arena.read(sids[gi][k], _native.FIELD_V)
            assert np.array_equal(x, body.x), f"loop {gi} body {k}: positions differ"
            assert np.array_equal(v, body.v), f"loop {gi} body {k}: velocities differ"
        cache = arena.flex_read(gids[gi], _native.FLEX_CACHE)
        assert np.array_equal(cache, o.flex.cache), f"loop {gi}: caches differ"
        hits += int((cache >= 0).sum())
    log(f"[case] contact workload: 24 of 2048 loops bit-exact after 600 substeps, {hits} cached segments")
    assert hits > 0
    arena.close()


def test_contact_workload_with_tied_loops_bitexact():
    """The 2048-loop contact workload with every third loop tied: tied and sliding
    loops share each contact launch, and each tied loop's ties resolve once, in a
    launch of its own before the first substep.  All hashes are rebuilt in one
    launch on the schedule (a tied loop's hash is not read after its ties
    resolve); 12 tied and 12 sliding loops against the oracle after 400 substeps:
    positions, velocities, caches and tie state."""
    from flatworld_amd import workloads as W
    from flatworld_amd.flexcontact import REBUILD_INTERVAL

    groups = W.contact_batched_config(n_groups=2048, seed=1)
    tied = list(range(0, len(groups), 3))
    for gi in tied:
        g = groups[gi][1]
        g.item_tied = np.ones(g.n_items, np.int32)
    dt = groups[0][0][0].dt
    arena = _native.Arena(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0)
    sids, gids = [], []
    for sp, g in groups:
        s = [arena.add(x) for x in sp]
        sids.append(s)
        gids.append(arena.flex_add(g, s))
    arena.step(1, -1.0, 0)
    arena.flex_rebuild(-1)
    elapsed, pending = 0.0, 0
    for n in range(400):
        elapsed += dt
        if elapsed >= REBUILD_INTERVAL:
            if pending:
                arena.step(pending, -1.0, 0)
                pending = 0
            arena.flex_rebuild(-1)
            elapsed = 0.0
        if n == 0:
            for gi in tied:
                arena.flex_init_ties(gids[gi])
        pending += 1
    arena.step(pending, -1.0, 0)
    arena.sync()
    rng = np.random.default_rng(9)
    untied = [gi for gi in range(len(groups)) if gi % 3]
    resolved = hits = 0
    for gi in list(rng.choice(tied, 12, replace=False)) + list(rng.choice(untied, 12, replace=False)):
        sp, g = groups[gi]
        o = OracleLoop(sp, g)
        for _ in range(400):
            o.substep(dt)
        for k, body in enumerate(o.bodies):
            assert np.all(np.isfinite(body.x)), f"loop {gi}: oracle not finite"
            assert np.array_equal(arena.read(sids[gi][k], _native.FIELD_X), body.x), f"loop {gi} body {k}: positions"
            assert np.array_equal(arena.read(sids[gi][k], _native.FIELD_V), body.v), f"loop {gi} body {k}: velocities"
        cache = arena.flex_read(gids[gi], _native.FLEX_CACHE)
        assert np.array_equal(cache, o.flex.cache), f"loop {gi}: caches differ"
        hits += int((cache >= 0).sum())
        if g.n_tied:
            seg = arena.flex_read(gids[gi], _native.FLEX_TIE_SEG)
            state = arena.flex_read(gids[gi], _native.FLEX_TIE_STATE)[:, :3]
            so, sto = o.flex.ties
            assert np.array_equal(seg, so) and np.array_equal(state, sto), f"loop {gi}: tie state differs"
            resolved += int((seg >= 0).sum())
    log(f"[case] contact workload with tied loops: 24 loops bit-exact after 400 substeps, {resolved} ties, "
        f"{hits} cached segments")
    assert resolved > 0 and hits > 0
    arena.close()


@pytest.mark.parametrize("caps", [dict(max_cells=40), dict(max_query=2), dict(max_pairs=60),
                                  dict(max_cells=25, max_query=3, max_pairs=40)])
def test_capped_hash_and_query_bitexact(caps):
    """Hash limits small enough to bind: a grid clipped to max_cells, pairs dropped
    past max_pairs (both make the walk repeat itself to deduplicate) and query
    lists cut at max_query (the serial walk decides)."""
    import dataclasses

    from flatworld_amd import workloads as W

    rng = np.random.default_rng(11)
    loops = [W.contact_pair_group(rng, 10, 16) for _ in range(6)]
    dt = min(s.dt for sp, _ in loops for s in sp)
    arena = _native.Arena(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0)
    ids = []
    for k, (sp, g) in enumerate(loops):
        g = dataclasses.replace(g, **caps)
        loops[k] = (sp, g)
        s = [arena.add(x) for x in sp]
        ids.append((s, arena.flex_add(g, s)))
    arena.step(1, -1.0, 0)
    arena.flex_rebuild(-1)
    oracles = [OracleLoop(sp, g) for sp, g in loops]
    elapsed, hits = 0.0, 0
    for step in range(700):
        elapsed += dt
        if elapsed >= loops[0][1].rebuild_interval:
            arena.flex_rebuild(-1)
            elapsed = 0.0
        arena.step(1, float(np.float32(dt)), 0)
        for o in oracles:
            o.substep(dt)
    arena.sync()
    for (sids, gid), o in zip(ids, oracles):
        grid = arena.flex_read(gid, _native.FLEX_GRID)
        assert grid["tc"] == o.flex.grid["tc"]
        for k, body in enumerate(o.bodies):
            assert np.all(np.isfinite(body.x))
            assert np.array_equal(arena.read(sids[k], _native.FIELD_X), body.x), f"{caps}: positions differ"
            assert np.array_equal(arena.read(sids[k], _native.FIELD_V), body.v), f"{caps}: velocities differ"
        cache = arena.flex_read(gid, _native.FLEX_CACHE)
        assert np.array_equal(cache, o.flex.cache), f"{caps}: caches differ"
        hits += int((cache >= 0).sum())
    log(f"[case] capped hash {caps}: 6 loops bit-exact after 700 substeps, {hits} cached segments")
    arena.close()


def test_contact_beside_a_quiet_body_and_springs_bitexact():
    """A loop holding two FEM bodies in contact, a FEM body outside every pair (its
    edges are in the hash, but no work item targets it) and a spring lattice, on a
    plane: the contact passes run beside the triangle and spring launches."""
    from flatworld_amd import workloads as W

    mat = fw.Elastic(2e4, 0.3, 50.0)
    lower = ccw_grid(8, (0.1, 0.0), 0.3, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=21)
    upper = ccw_grid(6, (0.15, 0.302), 0.2, fw.MisesPlastic(3e4, 0.3, 60.0, 200.0, 1e3), [fw.Gravity([0, -9.8])],
                     jitter=0.1, seed=22, initials=[fw.InitialVel("ALL", [0.05, -0.3])], friction=0.2)
    conn, coords = fw.structured_grid(5, lower=(0.32, 0.2), extent=0.15)
    quiet = fw.FemDomain(fw.Mesh(2, np.ascontiguousarray(conn[:, [0, 2, 1]]), coords), fw.SolidProp(mat),
                         [fw.Gravity([0, -9.8])], considerContact=False)
    lattice = W.spring_lattice(6, (0.6, 0.3), 0.02, [200.0, 0.3, 0.01], [fw.Gravity([0.0, -9.8])])
    lattice.considerContact = False
    plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    loop = fw.ExplicitLoop(5e-5, [lower, upper, quiet, lattice, plane], useAdapativeDT=False)
    g = loop.femSpringManager._flex
    assert g.n_bodies == 3 and g.n_pairs == 1 and g.gated
    log(f"[case] contact beside a quiet body and springs, items={g.n_items}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 6, check_every=2, what="quiet body and springs")
    assert contact > 0


def test_tied_blocks_bitexact():
    """Two jittered blocks 2 mm apart, tied, the second launched away: the ties
    resolve once at the first substep, every substep applies the springs, and the
    hash stops refreshing once no sliding item is left."""
    mat = fw.Elastic(1e4, 0.3, 10.0)
    a = ccw_grid(6, (0.0, 0.0), 0.2, mat, jitter=0.2, seed=0)
    b = ccw_grid(5, (0.202, 0.03), 0.18, mat, jitter=0.2, seed=7, initials=[fw.InitialVel("ALL", [0.3, 0.1])])
    loop = fw.ExplicitLoop(1.5e-4, [a, b], ties=[(0, 1)], useAdapativeDT=False)  # 0.35 x the stable step
    m = loop.femSpringManager
    g = m._flex
    assert g.n_tied == g.n_items and not g.gated
    log(f"[case] tied blocks, dt={loop.dt:.3e}, items={g.n_items}")
    o = oracle_of(loop)
    run_frames(loop, o, 2, check_every=1, what="tied blocks")
    seg, _ = m.flex_ties()
    assert (seg >= 0).sum() > 10 and (seg < 0).sum() > 0
    gd, go = m._arena.flex_read(m._flex_gid, _native.FLEX_GRID), o.flex.grid
    assert np.array_equal(gd["glb"], go["glb"]) and np.array_equal(gd["gs"], go["gs"])
    for body in o.bodies:
        assert np.all(np.isfinite(body.x))


def test_tied_and_sliding_pairs_on_plane_bitexact():
    """Two blocks tied side by side on a plane and a third dropped across the
    joint: one group of tied and sliding pairs on the plane-gated path; the joint's
    nodes take their tie springs before their contacts."""
    mat = fw.Elastic(2e4, 0.3, 50.0)
    left = ccw_grid(6, (0.1, 0.0), 0.3, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=31)
    right = ccw_grid(6, (0.4, 0.0), 0.3, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=32)
    top = ccw_grid(5, (0.3, 0.301), 0.2, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=33, friction=0.2,
                   initials=[fw.InitialVel("ALL", [0.0, -0.5])])
    plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    loop = fw.ExplicitLoop(1e-4, [left, right, top, plane], ties=[(1, 0)], useAdapativeDT=False)
    g = loop.femSpringManager._flex
    assert g.gated and g.n_pairs == 3 and 0 < g.n_tied < g.n_items
    log(f"[case] tied and sliding pairs on a plane, dt={loop.dt:.3e}, items={g.n_items}, tied={g.n_tied}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 2, check_every=1, what="tied and sliding")
    assert contact > 0
    for body in o.bodies:
        assert np.all(np.isfinite(body.x))


def test_spring_fem_pairs_gate_fem_fem_contact_bitexact():
    """A spring lattice in contact range of two FEM blocks: the FEM-spring pairs
    apply no force (the reference never runs its unbatched contacts) but lower the
    stable time and turn the BVH gate on, so the blocks' FEM-FEM pair is gated
    with no ground in the loop."""
    from flatworld_amd import workloads as W
    from flatworld_amd.femcontact import ContactSpringFlex

    mat = fw.Elastic(2e4, 0.3, 50.0)
    lower = ccw_grid(6, (0.1, 0.0), 0.3, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=41)
    upper = ccw_grid(5, (0.15, 0.31), 0.2, mat, [fw.Gravity([0, -9.8])], jitter=0.1, seed=42,
                     initials=[fw.InitialVel("ALL", [0.0, -0.8])])
    lattice = W.spring_lattice(6, (0.2, 0.1), 0.02, [200.0, 0.3, 0.01], [fw.Gravity([0.0, -9.8])])
    with pytest.warns(UserWarning, match="FEM-spring"):
        loop = fw.ExplicitLoop(1e-4, [lower, upper, lattice], useAdapativeDT=False)
    fs = [c for c in loop.contacts if isinstance(c, ContactSpringFlex)]
    assert len(fs) == 2 and loop.stableTime <= min(c.stableTime for c in fs)
    g = loop.femSpringManager._flex
    assert g.gated and g.n_pairs == 1
    log(f"[case] FEM-spring pairs gate a FEM-FEM pair, items={g.n_items}")
    o = oracle_of(loop)
    contact = run_frames(loop, o, 3, check_every=1, what="FEM-spring gate")
    assert contact > 0


def test_group_of_300_bodies_bitexact():
    """One FEM-FEM group of 300 bodies (75 stacks of four small blocks on a plane,
    each block paired with the one above it): more bodies than eight bits of a
    packed hash entry could name; the query walk tells the target body by its
    segment range."""
    from flatworld_amd import workloads as W
    from flatworld_amd.definitions import ContactType
    from flatworld_amd.femcontact import ContactFlexFlex, contactHelper
    from flatworld_amd.flexcontact import build_flex_group
    from flatworld_amd.scene import add_flex_record, pack_domain

    mat = fw.Elastic(2e4, 0.3, 100.0)
    plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    bodies, pairs = [], []
    for col in range(75):
        x0 = 0.02 * col
        for lvl in range(4):
            b = ccw_grid(2, (x0 + 0.002 * (lvl % 2), 0.0161 * lvl), 0.016, mat, [fw.Gravity([0, -9.8])],
                         initials=[fw.InitialVel("ALL", [0.0, -0.05 * lvl])], friction=0.1 * (lvl % 2))
            if lvl:
                pairs.append(ContactFlexFlex(b, bodies[-1], ContactType.FLEXFLEX))
            bodies.append(b)
    group = build_flex_group(bodies, pairs, True, sum(b.nnodes for b in bodies))
    assert group.n_bodies == 300 and group.n_pairs == 225
    grounds = [(plane, contactHelper(plane, b, plane.type | b.type)) for b in bodies]
    dt = 0.2 * min([1e-3] + [b.stableTime for b in bodies] + [p.stableTime for p in pairs] + [c.stableTime for _, c in grounds])
    specs = [add_flex_record(pack_domain(b, [grounds[k]], dt=dt), group.flex_cap(k)) for k, b in enumerate(bodies)]
    arena = _native.Arena(tile_nodes=256, tile_slots=512, block=256, grid_per_cu=0)
    sids = [arena.add(s) for s in specs]
    gid = arena.flex_add(group, sids)
    log(f"[case] 300-body group, dt={dt:.3e}, items={group.n_items}")
    arena.step(1, -1.0, 0)  # warm-up
    arena.flex_rebuild(gid)
    o = OracleLoop(specs, group)
    elapsed = 0.0
    for k in range(300):
        elapsed += dt
        if elapsed >= group.rebuild_interval:
            arena.flex_rebuild(gid)
            elapsed = 0.0
        arena.step(1, -1.0, 0)
        o.substep(dt)
    hits = int((o.flex.cache >= 0).sum())
    assert all(np.isfinite(b.x).all() for b in o.bodies)
    assert np.array_equal(arena.flex_read(gid, _native.FLEX_CACHE), o.flex.cache)
    for k, (sid, body) in enumerate(zip(sids, o.bodies)):
        x = arena.read(sid, _native.FIELD_X)
        assert np.array_equal(x, body.x), f"300-body group: body {k} differs (max {np.abs(x - body.x).max():.3e})"
    assert hits > 0
    arena.close()
