"""Pins of the oracle's FEM-vs-rigid primitive contact (flatworld/mixedcontact.py:1025-1117,
flatworld/contact_detection.py:147-191, flatworld/utils.py:34-145, 235-302).

Every check is a closed form, a float64 brute force or an invariance, so that a
dropped term, a wrong sign, a transposed rotation or a wrong face choice fails:
* ball: penetration = distance - radius, normal radial, contact point on the sphere;
* box: inside points take the nearest face (normal of that face, penetration its
  signed distance), outside points the distance to the clamped point; rotating the
  box and the point together changes nothing (to float rounding); the AABB of an
  unrotated box is centre +- extent / 2, and a quarter turn swaps the extents;
* capsule: penetration = float64 distance to the axis segment - radius;
* force: none outside the rigid's AABB or above the 1e-4 skin; the normal force is
  k (1e-4 - pen) along the normal; friction opposes the relative tangential velocity
  (including the rigid's rotation) with magnitude mu |normal force|; a tied node is
  pulled to its contact point;
* a block resting on a wide static box behaves as on a plane 1e-4 above the box top.
"""

import math

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import workloads as W
from flatworld_amd.femcontact import contactHelper
from flatworld_amd.rigid import primitive_aabb, rotation_matrix
from flatworld_amd.scene import GroundSpec, rigid_spec, update_rigid_spec
from oracle.oracle import OracleScene, ground_force, rigid_aabb, rigid_detect

K = 1.0e4


def _rec(rigid, friction=0.0, velocity=(0.0, 0.0), omega=0.0, tied=False, theta=None, center=None):
    d = fw.RigidBodyDomain(rigid, friction=friction)
    d.set_motion(velocity, omega)
    d.set_pose(center, theta)
    c = type("C", (), {"penalty": K, "tied": tied})()
    return rigid_spec(d, c)


def test_ball_closed_form():
    g = _rec(fw.BallRigid(2, [0.5, 0.2], 0.1, 1.0))
    for ang, depth in ((0.3, 0.01), (2.0, 0.03), (-1.2, 0.002)):
        p = np.array([0.5 + (0.1 - depth) * math.cos(ang), 0.2 + (0.1 - depth) * math.sin(ang)], np.float32)
        pen, n, cp = rigid_detect(g, p)
        dist = np.hypot(*(p.astype(np.float64) - [0.5, 0.2]))
        assert abs(pen - (dist - 0.1)) < 1e-6
        assert np.allclose(n, [math.cos(ang), math.sin(ang)], atol=1e-6)
        assert np.allclose(cp, [0.5 + 0.1 * math.cos(ang), 0.2 + 0.1 * math.sin(ang)], atol=1e-6)
        f, hit = ground_force(g, p, [0.0, 0.0], 1e-4)
        assert hit
        assert np.allclose(f, n.astype(np.float64) * K * (1e-4 - float(pen)), rtol=1e-5)


def test_box_faces_corners_and_rotation():
    box = fw.BoxRigid(2, [0.5, 0.2], [0.4, 0.2], 0.0, 1.0)
    g = _rec(box)
    # inside, nearest the top face
    pen, n, cp = rigid_detect(g, [0.55, 0.29])
    assert np.isclose(pen, 0.29 - 0.3, atol=1e-7) and np.array_equal(n, [0.0, 1.0])
    assert np.allclose(cp, [0.55, 0.3], atol=1e-7)
    # inside, nearest the left face
    pen, n, _ = rigid_detect(g, [0.31, 0.21])
    assert np.isclose(pen, 0.3 - 0.31, atol=1e-7) and np.array_equal(n, [-1.0, 0.0])
    # outside a corner: distance to the corner, normal along the offset
    pen, n, cp = rigid_detect(g, [0.73, 0.34])
    assert np.isclose(pen, np.hypot(0.03, 0.04), atol=1e-6) and np.allclose(n, [0.6, 0.8], atol=1e-5)
    assert np.array_equal(cp, np.float32([0.73, 0.34]))  # no contact point update outside
    # rotating box and point together about the centre
    rng = np.random.default_rng(3)
    for a in (0.3, 1.1, -2.5):
        gr = _rec(box, theta=a)
        R = rotation_matrix(a).astype(np.float64)
        for _ in range(20):
            q = rng.uniform([-0.25, -0.15], [0.25, 0.15])
            p0 = np.float32(q + [0.5, 0.2])
            p1 = np.float32(R @ q + [0.5, 0.2])
            pen0, n0, _ = rigid_detect(g, p0)
            pen1, n1, _ = rigid_detect(gr, p1)
            assert abs(float(pen0) - float(pen1)) < 2e-6
            assert np.allclose(R @ n0.astype(np.float64), n1, atol=2e-5)


def test_capsule_against_float64_segment_distance():
    cap = fw.CapsuleRigid(2, [0.3, 0.2], [0.7, 0.3], 0.2, 0.05, 1.0)
    g = _rec(cap, theta=0.2)
    R = rotation_matrix(0.2).astype(np.float64)
    c = np.array([0.5, 0.25])
    lc = c + R @ (np.array([0.3, 0.2]) - c)
    uc = 2 * c - lc
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform([0.2, 0.1], [0.8, 0.4]).astype(np.float32)
        ab = uc - lc
        t = np.clip(np.dot(p - lc, ab) / np.dot(ab, ab), 0.0, 1.0)
        d = np.hypot(*(p - (lc + t * ab)))
        pen, n, _ = rigid_detect(g, p)
        assert abs(float(pen) - (d - 0.05)) < 2e-6
        assert np.allclose(n, (p - (lc + t * ab)) / d, atol=1e-4)


def test_aabb_closed_forms():
    c = np.float32([0.5, 0.2])
    lo, hi = primitive_aabb(fw.RigidType.BOX, c, rotation_matrix(0.0), [0.4, 0.2], 0.0)
    assert np.allclose(lo, [0.3, 0.1], atol=1e-7) and np.allclose(hi, [0.7, 0.3], atol=1e-7)
    lo, hi = primitive_aabb(fw.RigidType.BOX, c, rotation_matrix(math.pi / 2), [0.4, 0.2], 0.0)
    assert np.allclose(lo, [0.4, 0.0], atol=1e-6) and np.allclose(hi, [0.6, 0.4], atol=1e-6)
    lo, hi = primitive_aabb(fw.RigidType.BALL, c, rotation_matrix(1.0), [0.0, 0.0], 0.1)
    assert np.array_equal(lo, c - np.float32(0.1)) and np.array_equal(hi, c + np.float32(0.1))
    # capsule with a horizontal axis of half-length 0.2 and radius 0.05
    lo, hi = primitive_aabb(fw.RigidType.CAPSULE, c, rotation_matrix(0.0), [-0.2, 0.0], 0.05)
    assert np.allclose(lo, [0.25, 0.15], atol=1e-7) and np.allclose(hi, [0.75, 0.25], atol=1e-7)
    # the host packing and the oracle compute the same float32 AABB
    for rig, a in ((fw.BoxRigid(2, [0.5, 0.2], [0.3, 0.1], 0.7, 1.0), 0.7),
                   (fw.CapsuleRigid(2, [0.4, 0.2], [0.6, 0.25], 0.3, 0.05, 1.0), 0.3)):
        r = fw.RigidBodyDomain(rig).record()
        lb, ub = rigid_aabb(r["rigid_type"], r["rigid_center"], r["rigid_rot"], r["rigid_prim"], r["rigid_radius"])
        assert np.array_equal(lb, r["aabb_lo"]) and np.array_equal(ub, r["aabb_hi"])


def test_force_skin_aabb_friction_and_tie():
    ball = fw.BallRigid(2, [0.5, 0.2], 0.1, 1.0)
    g = _rec(ball)
    # outside the rigid's AABB: nothing, even though within the margin
    f, hit = ground_force(g, [0.5, 0.305], [0.0, 0.0], 1e-4)
    assert not hit and not f.any()
    # the contact starts 1e-4 outside the surface (penetration - 1e-4 < 0), inside the AABB
    s2 = math.sqrt(0.5)
    f, hit = ground_force(g, [0.5 + (0.1 + 5e-5) * s2, 0.2 + (0.1 + 5e-5) * s2], [0.0, 0.0], 1e-4)
    assert hit and np.allclose(f, K * 5e-5 * s2, rtol=2e-2)
    f, hit = ground_force(g, [0.5 + (0.1 + 2e-4) * s2, 0.2 + (0.1 + 2e-4) * s2], [0.0, 0.0], 1e-4)
    assert not hit
    # friction: opposes the tangential relative velocity, magnitude mu |N|
    gf = _rec(ball, friction=0.4)
    p = np.float32([0.5, 0.29])
    f, hit = ground_force(gf, p, [0.3, -0.1], 1e-4)
    fn, _ = ground_force(g, p, [0.3, -0.1], 1e-4)
    ft = f.astype(np.float64) - fn
    assert hit and ft[0] < 0.0 and abs(ft[1]) < 1e-3 * abs(ft[0])
    assert np.isclose(np.hypot(*ft), 0.4 * np.hypot(*fn), rtol=1e-5)
    # a rigid spinning so that its surface moves with the node: no friction
    gs = _rec(ball, friction=0.4, omega=-3.0)  # surface velocity at the top: (+0.3, 0)
    f, _ = ground_force(gs, p, [0.3, 0.0], 1e-4)
    assert np.allclose(f, fn, atol=1e-6 * np.abs(fn).max())
    # tied: pulled to the contact point while within the margin
    gt = _rec(ball, tied=True)
    pen, n, cp = rigid_detect(gt, p)
    f, hit = ground_force(gt, p, [0.0, 0.0], 1e-4)
    assert hit and np.allclose(f, K * (cp.astype(np.float64) - p), rtol=1e-6)


def test_block_on_wide_box_matches_plane():
    """A block pressed into a wide static box top at y = 0 moves as on a plane at y = 1e-4
    (the box's 1e-4 skin), as long as its nodes stay below the box top (the box's AABB
    ends there): a heavy block pushed down at 0.2 m/s."""
    dom = W.grid_domain(6, (0.3, 0.0), 0.2, fw.Elastic(2e4, 0.3, 2000.0), [fw.Gravity([0.0, -9.8])],
                        initials=[fw.InitialVel("ALL", [0.0, -0.2])])
    box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.4, -0.5], [4.0, 1.0], 0.0, 1.0))
    c = contactHelper(box, dom, box.type | dom.type)
    spec_box = W.one_body_spec(dom, None, dt=2e-5)
    spec_box.grounds = [rigid_spec(box, c)]
    plane = fw.GroundDomain(2, [0.0, 1e-4], [0.0, 1.0])
    spec_pl = W.one_body_spec(dom, plane, dt=2e-5)
    assert np.float32(spec_pl.grounds[0].penalty) == np.float32(spec_box.grounds[0].penalty)
    a, b = OracleScene(spec_box), OracleScene(spec_pl)
    for o in (a, b):
        o.warmup()
        o.run(300)
    assert a.contact_flag.any() and b.contact_flag.any()
    assert a.x[spec_box.bnodes, 1].max() > 0.1 and a.x[spec_box.bnodes][:, 1][a.contact_flag[spec_box.bnodes] > 0].max() <= 0.0
    h = 0.2 / 6
    assert np.abs(a.x - b.x).max() < 1e-4 * h


def test_spring_network_on_wide_box_matches_plane():
    """Spring-rigid pairs (flatworld/femcontact.py:243-253): penalty 100 x spring and
    stable time 0.5 sqrt(mass / k), as against a plane; every spring node is a work
    item.  A lattice pushed down into a wide static box top at y = 0 moves as on a
    plane at y = 1e-4 while its touching nodes stay below the box top."""
    dom = W.spring_lattice(5, (0.3, 0.0), 0.02, [200.0, 0.3, 0.01], [fw.Gravity([0.0, -9.8])],
                           initials=[fw.InitialVel("ALL", [0.0, -0.2])])
    box = fw.RigidBodyDomain(fw.BoxRigid(2, [0.4, -0.5], [4.0, 1.0], 0.0, 1.0))
    plane = fw.GroundDomain(2, [0.0, 1e-4], [0.0, 1.0])
    c_box = contactHelper(box, dom, box.type | dom.type)
    c_pl = contactHelper(plane, dom, plane.type | dom.type)
    assert type(c_box).__name__ == "ContactSpringRigid" and c_box.domain1 is dom
    assert c_box.penalty == 100.0 * 200.0 == c_pl.penalty
    assert c_box.stableTime == c_pl.stableTime == 0.5 * math.sqrt(0.01 / 2e4)
    spec_box = fw.pack_domain(dom, [(box, c_box)], dt=2e-5)
    spec_pl = fw.pack_domain(dom, [(plane, c_pl)], dt=2e-5)
    assert spec_box.bnodes.shape[0] == dom.nnodes
    a, b = OracleScene(spec_box), OracleScene(spec_pl)
    for o in (a, b):
        o.warmup()
        o.run(200)
    assert a.contact_flag.sum() >= 6 and b.contact_flag.any()
    assert a.x[a.contact_flag > 0, 1].max() <= 0.0
    assert np.abs(a.x - b.x).max() < 1e-4 * 0.02
