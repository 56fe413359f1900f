"""FEM bodies against kinematic rigid primitives on the GPU, bit for bit against the oracle.

The rigid record (FW_GROUND_RIGID) follows _batched_rigid_prim_kernel
(flatworld/mixedcontact.py:1025-1117): balls, boxes and capsules, rotated, with
friction against the rigid's surface velocity, tied pairs, and rigids moved between
calls (fw_scene_set_rigid on the device, the same pose on the oracle).  Every check
covers positions, velocities, nodal forces, U, contact flags, AABB, stress, strain,
yield flags and J2 state (tests/parity.py check_fields).
"""

import copy

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from flatworld_amd import workloads as W
from flatworld_amd.femcontact import contactHelper
from flatworld_amd.scene import rigid_spec, update_rigid_spec
from oracle.oracle import OracleScene
from tests.parity import GpuScene, check_fields, log

pytestmark = pytest.mark.gpu


def _scene(dom, grounds, dt=None):
    """SceneSpec of ``dom`` with ordered (ground or rigid) contacts, at the loop's stable dt."""
    pairs = [(g, contactHelper(g, dom, g.type | dom.type)) for g in grounds]
    st = min([1.0, 1e-3, dom.stableTime] + [c.stableTime for _, c in pairs])
    return fw.pack_fem_domain(dom, pairs, dt=st if dt is None else dt)


def _run(spec, chunks, label, moves=None, **arena_kw):
    """Warm-up, then chunks of substeps; ``moves(k)`` (optional) moves the rigids before
    chunk k and returns the record indices to refresh.  Returns (gpu, oracle, contacts)."""
    log(f"[case] {label}: N={spec.n_nodes} E={spec.n_elems} grounds={[g.kind for g in spec.grounds]} {arena_kw}")
    g = GpuScene(spec, **arena_kw)
    o = OracleScene(spec)
    g.warmup()
    o.warmup()
    check_fields(g.read, o, label + " warm-up")
    contacts = []
    for k, n in enumerate(chunks):
        if moves is not None:
            for gi in moves(k):
                rec = spec.grounds[gi]
                update_rigid_spec(rec)
                g.arena.set_rigid(g.sid, gi, rec)
                o.set_rigid(gi, rec)
        g.run(n)
        o.run(n)
        contacts.append(check_fields(g.read, o, f"{label} chunk {k}")["contact"])
    return g, o, contacts


def test_block_falls_on_ball():
    dom = W.grid_domain(16, (0.4, 0.31), 0.2, fw.Elastic(2e4, 0.3, 200.0), [fw.Gravity([0.0, -9.8])],
                        initials=[fw.InitialVel("ALL", [0.0, -0.3])])
    ball = fw.RigidBodyDomain(fw.BallRigid(2, [0.5, 0.2], 0.1, 1.0))
    spec = _scene(dom, [ball])
    _, o, contacts = _run(spec, [1, 300, 300, 300], "elastic block onto a ball")
    assert max(contacts) > 0, "the block should reach the ball"


def test_j2_block_slides_on_rotated_box_with_friction():
    dom = W.grid_domain(24, (0.3, 0.33), 0.2, fw.MisesPlastic(5e4, 0.3, 300.0, 250.0, 500.0), [fw.Gravity([0.0, -9.8])],
                        initials=[fw.InitialVel("ALL", [0.2, -0.5])])
    box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.5, 0.2], [0.8, 0.2], 0.25, 1.0), friction=0.4)
    spec = _scene(dom, [box])
    _, o, contacts = _run(spec, [1, 250, 250, 250], "J2 block on a rotated box, friction 0.4", tile_nodes=128)
    assert max(contacts) > 0 and o.eqps.max() > 0.0, "the block should touch the box and yield"


def test_moving_spinning_capsule_pushes_block_on_plane():
    dom = W.grid_domain(20, (0.4, 0.0), 0.2, fw.HyperElastic(3e4, 0.3, 300.0), [fw.Gravity([0.0, -9.8])])
    plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
    cap = fw.RigidBodyDomain(fw.CapsuleRigid(2, [0.25, 0.05], [0.35, 0.12], 0.0, 0.03, 1.0), friction=0.3)
    cap.set_motion([0.6, 0.0], 2.0)
    spec = _scene(dom, [plane, cap])
    assert [g.kind for g in spec.grounds] == [1, 5]
    dt = spec.dt

    def moves(k):
        if k > 1:
            cap.advance(40 * dt)  # the rigid moved over the previous 40-substep chunk
        return [1]

    _, o, contacts = _run(spec, [1, 40] + [40] * 10, "capsule moving into a block on a plane", moves=moves,
                          tile_nodes=200)
    assert max(contacts) > 0


def test_spring_lattice_falls_on_ball_and_rotated_box():
    """Spring-rigid pairs (flatworld/femcontact.py:243-253): every lattice node is a work
    item of the rigid-primitive contact; a jittered lattice with bending springs drops
    onto a ball and a rotated box with friction, on one tile and on several."""
    dom = W.spring_lattice(24, (0.38, 0.30), 0.01, [300.0, 0.4, 0.01], [fw.Gravity([0.0, -9.8])],
                           initials=[fw.InitialVel("ALL", [0.1, -0.4])], jitter=0.1, seed=3)
    ball = fw.RigidBodyDomain(fw.BallRigid(2, [0.45, 0.2], 0.08, 1.0))
    box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.62, 0.18], [0.3, 0.06], -0.3, 1.0), friction=0.3)
    pairs = [(g, contactHelper(g, dom, g.type | dom.type)) for g in (ball, box)]
    assert all(type(c).__name__ == "ContactSpringRigid" for _, c in pairs)
    dt = min([1e-3, dom.stableTime] + [c.stableTime for _, c in pairs])
    spec = fw.pack_domain(dom, pairs, dt=dt)
    assert [g.kind for g in spec.grounds] == [5, 5]
    for kw in ({}, dict(tile_nodes=128)):
        # the lattice bounces: compare every 40 substeps so that the checks see its contacts
        _, o, contacts = _run(spec, [1] + [40] * 30, f"spring lattice onto a ball and a box {kw}", **kw)
        assert sum(c > 0 for c in contacts) >= 5, "the lattice should reach the rigids"


def test_tied_block_hangs_from_box():
    dom = W.grid_domain(12, (0.4, 0.1995), 0.2, fw.Elastic(2e4, 0.3, 100.0), [fw.Gravity([0.0, -9.8])])
    box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.5, 0.1], [0.6, 0.2], 0.0, 1.0))
    c = contactHelper(box, dom, box.type | dom.type)
    c.tied = True
    rec = rigid_spec(box, c)
    assert rec.rigid_tied == 1
    spec = _scene(dom, [])
    spec.grounds = [rec]
    spec.dt = min(dom.stableTime, c.stableTime, 1e-3)
    _, o, contacts = _run(spec, [1, 200, 200], "block tied to a box")
    assert max(contacts) > 0


def test_eight_grounds_on_one_body():
    """Six planes around a block, a ball and a box: eight contact records in one scene."""
    dom = W.grid_domain(14, (0.35, 0.05), 0.3, fw.Elastic(3e4, 0.3, 200.0), [fw.Gravity([0.0, -9.8])],
                        initials=[fw.InitialVel("ALL", [0.4, -0.2])])
    planes = [fw.GroundDomain(2, p, n) for p, n in (([0.0, 0.0], [0.0, 1.0]), ([0.3, 0.0], [1.0, 0.1]),
                                                    ([0.7, 0.0], [-1.0, 0.2]), ([0.0, 0.02], [0.1, 1.0]),
                                                    ([0.0, 0.01], [-0.1, 1.0]), ([0.66, 0.0], [-1.0, 0.0]))]
    ball = fw.RigidBodyDomain(fw.BallRigid(2, [0.5, 0.42], 0.06, 1.0))
    box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.36, 0.3], [0.05, 0.3], 0.1, 1.0), friction=0.2)
    spec = _scene(dom, planes + [ball, box])
    assert len(spec.grounds) == 8
    _, o, contacts = _run(spec, [1, 300, 300], "eight contact records", tile_nodes=128)
    assert max(contacts) > 0
    bad = copy.deepcopy(spec)
    bad.grounds = bad.grounds + [bad.grounds[0]]
    with pytest.raises(NotImplementedError):
        _native.Arena().add(bad)


def test_explicitloop_with_kinematic_ball():
    """The drop-in API: a FEM block on a plane, a ball moving down onto it; the oracle
    follows the same float64 dt sequence and the same ball poses."""
    def build():
        dom = W.grid_domain(10, (0.4, 0.0), 0.2, fw.Elastic(2e4, 0.3, 100.0), [fw.Gravity([0.0, -9.8])])
        plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
        ball = fw.RigidBodyDomain(fw.BallRigid(2, [0.5, 0.32], 0.1, 1.0), friction=0.2)
        ball.set_motion([0.0, -0.5], 1.0)
        return dom, plane, ball

    dom, plane, ball = build()
    loop = fw.ExplicitLoop(0.0, [plane, dom, ball])
    odom, oplane, oball = build()
    c_pl = contactHelper(oplane, odom, oplane.type | odom.type)
    c_b = contactHelper(oball, odom, oball.type | odom.type)
    spec = fw.pack_fem_domain(odom, [(oplane, c_pl), (oball, c_b)])
    o = OracleScene(spec)
    o.warmup()
    log(f"[case] ExplicitLoop with a kinematic ball, dt={loop.dt:.3e}")
    touched = 0
    for frame in range(8):
        loop.advanceWithTime(1.0 / 240.0)
        for sd in loop.substep_sequence(1.0 / 240.0):
            update_rigid_spec(spec.grounds[1])
            o.set_rigid(1, spec.grounds[1])
            o.substep(sd, loop.damping)
            oball.advance(sd)
        pos = loop.femSpringManager.coords.numpy()[: dom.nnodes]
        assert np.array_equal(pos, o.x), f"frame {frame}"
        touched += int(o.contact_flag.any())
    assert touched > 0


def test_explicitloop_spring_network_with_kinematic_box():
    """The drop-in API with a spring network: a lattice on a plane, a box sliding along
    the plane into it with friction; the oracle follows the same float64 dt sequence
    and the same box poses."""
    def build():
        dom = W.spring_lattice(12, (0.45, 0.001), 0.01, [200.0, 0.3, 0.01], [fw.Gravity([0.0, -9.8])])
        plane = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])
        box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.398, 0.05], [0.1, 0.1], 0.0, 1.0), friction=0.3)
        box.set_motion([0.6, 0.0], 0.0)
        return dom, plane, box

    dom, plane, box = build()
    loop = fw.ExplicitLoop(0.0, [plane, dom, box])
    odom, oplane, obox = build()
    c_pl = contactHelper(oplane, odom, oplane.type | odom.type)
    c_b = contactHelper(obox, odom, obox.type | odom.type)
    spec = fw.pack_domain(odom, [(oplane, c_pl), (obox, c_b)])
    o = OracleScene(spec)
    o.warmup()
    log(f"[case] ExplicitLoop: spring network and a kinematic box, dt={loop.dt:.3e}")
    for frame in range(10):
        loop.advanceWithTime(1.0 / 240.0)
        for sd in loop.substep_sequence(1.0 / 240.0):
            update_rigid_spec(spec.grounds[1])
            o.set_rigid(1, spec.grounds[1])
            o.substep(sd, loop.damping)
            obox.advance(sd)
        pos = loop.femSpringManager.coords.numpy()[: dom.nnodes]
        assert np.array_equal(pos, o.x), f"frame {frame}"
    assert o.x[:, 0].min() > 0.46, "the box should have pushed the lattice"


def test_rigid_records_under_graph_per_scene_and_noalloc_launches():
    """Rigid-bearing scenes beside plain ones: hipGraph replay, one launch per scene and
    fw_step_noalloc give the bits of plain launches (the rigid pass runs in each)."""
    specs = W.batched_config(n_scenes=4, seed=21, side_lo=6, side_hi=20)
    for k, (rig, fr) in enumerate(((fw.BallRigid(2, [0.5, 0.0], 0.08, 1.0), 0.0),
                                  (fw.BoxRigid(2, [0.5, -0.02], [0.3, 0.1], 0.2, 1.0), 0.3))):
        dom = W.grid_domain(12, (0.4, 0.085), 0.2, fw.Elastic(2e4, 0.3, 100.0), [fw.Gravity([0.0, -9.8])],
                            initials=[fw.InitialVel("ALL", [0.1, -0.5])])
        specs.append(_scene(dom, [fw.RigidBodyDomain(rig, friction=fr)]))
    log("[case] rigid records: plain vs graph vs per-scene vs noalloc")
    arenas = [_native.Arena(tile_nodes=128, capacity_bytes=256 << 20) for _ in range(4)]
    ids = [[a.add(s) for s in specs] for a in arenas]
    for a in arenas:
        a.step(1, -1.0, 0)
    a0, a1, a2, a3 = arenas
    a3.step(0)
    for rep in range(5):
        a0.step(60, -1.0, _native.FW_STEP_DIAG)
        a1.step_graph(60, _native.FW_STEP_DIAG)
        a2.step(60, -1.0, _native.FW_STEP_DIAG | _native.FW_STEP_PER_SCENE)
        a3.step_noalloc(60, -1.0, _native.FW_STEP_DIAG)
        for a in arenas:
            a.sync()
        for j in range(len(specs)):
            for fld in (_native.FIELD_X, _native.FIELD_V, _native.FIELD_FORCE, _native.FIELD_CONTACT):
                ref = a0.read(ids[0][j], fld)
                for m in (1, 2, 3):
                    assert np.array_equal(arenas[m].read(ids[m][j], fld), ref), f"rep {rep} scene {j} field {fld} mode {m}"
    o = OracleScene(specs[-1])
    o.warmup()
    o.run(300)
    assert o.contact_flag.any()
    check_fields(lambda f: a0.read(ids[0][-1], f), o, "rigid scene in the mixed arena")
    for a in arenas:
        a.close()
