"""Pins of the oracle's FEM-vs-rigid-mesh contact and of the host's mesh hash
(flatworld/mixedcontact.py:519-588, 1119-1565; flatworld/rigidmanager.py:4095-4193,
4581-4641; flatworld/spatialmanager.py:200-330).

Every check is a closed form, a brute force or an invariance:
* placement: the boundary nodes at a pose are the rest nodes turned and moved as a
  rigid body (float64 comparison), and the AABB bounds them;
* hash: every registered edge sits in exactly the cells its widened box covers, each
  cell's entries ascend in registration order, and every edge of a mesh is
  registered by an unfiltered build;
* a single node d below a flat top edge gets k d along the edge's outward normal and
  nothing else; a node above it, or outside the rigid's AABB, gets nothing;
* friction: the first contact step has no history and no friction; once the contact
  point slides by s along the edge, the friction is k s along the slide (the sign of
  f_trial = prev_force + k_t slip, mixedcontact.py:1352-1372), capped at mu |N|;
* a tied node is pulled to its projection on the nearest edge with k (cp - p),
  whatever its side of the edge;
* a node inside the rigid's AABB whose query cells hold no edge of the rigid is
  skipped until the skip epochs reset.
"""

import ctypes
import math

import numpy as np

import flatworld_amd as fw
from flatworld_amd.femcontact import contactHelper
from flatworld_amd.rigid import KinematicRigids, RigidMeshSet
from flatworld_amd.scene import pack_domain
from oracle.oracle import OracleScene


def slab(lower, width, height, side=4, angle=0.0):
    conn, coords = fw.structured_grid(side, lower=lower, extent=width)
    conn = np.ascontiguousarray(conn[:, [0, 2, 1]])
    coords = coords.copy()
    coords[:, 1] = lower[1] + (coords[:, 1] - lower[1]) * (height / width)
    return fw.MeshRigid(2, fw.Mesh(2, conn, coords), [angle], 1.0)


def triangle_scene(pts, rigid, friction=0.0, tied=False, subcycle=2):
    """One undeformed triangle (all three nodes boundary nodes) against a mesh rigid;
    returns (scene spec, its mesh record, the mesh set, the contact)."""
    conn = np.array([[0, 1, 2]], np.int32)
    body = fw.FemDomain(fw.Mesh(2, conn, np.asarray(pts, np.float32)), fw.SolidProp(fw.Elastic(1e4, 0.3, 100.0)), [])
    rd = fw.RigidBodyDomain(rigid, friction=friction)
    c = contactHelper(body, rd, body.type | rd.type)
    c.tied = tied
    ms = RigidMeshSet([rd], [rd])
    c.mesh_set, c.mesh_did, c.mesh_subcycle = ms, 0, subcycle
    spec = pack_domain(body, [(rd, c)], dt=1e-4)
    ms.populate()  # every edge registered
    return spec, spec.grounds[0], ms, c, rd


def step(o, g, ms, counter_reset=-1):
    """One zero-dt substep; returns the contact force per node: the nodal force less the
    same state's nodal force without contact (the element's rounding-level stress)."""
    g.aabb_lo, g.aabb_hi = ms.aabbs[0]
    g.mesh_reset_at = counter_reset
    o.set_rigid(0, g)
    x = o.x.copy()
    o._lib.ora_substep(ctypes.byref(o._s), 0.0, 0.0, o.counter, 0)
    internal = o.fext.astype(np.float64)
    o.x[:] = x
    o.substep(0.0)
    return o.fext.astype(np.float64) - internal


def test_placement_is_a_rigid_motion():
    r = slab((0.1, 0.1), 0.6, 0.1)
    d = fw.RigidBodyDomain(r)
    w0, lb0, ub0 = d.mesh_world()
    d.set_pose([0.7, 0.4], 0.6)
    w1, lb1, ub1 = d.mesh_world()
    c, s = math.cos(0.6), math.sin(0.6)
    expect = (r.local.astype(np.float64) @ np.array([[c, s], [-s, c]])) + [0.7, 0.4]
    assert np.allclose(w1, expect, atol=2e-7)
    assert np.allclose(w0, r.local.astype(np.float64) + r.origin, atol=2e-7)
    assert np.all(w1 >= lb1) and np.all(w1 <= ub1) and np.any(w1 == lb1) and np.any(w1 == ub1)


def test_hash_brute_force():
    a = fw.RigidBodyDomain(slab((0.1, 0.1), 0.6, 0.1, side=6))
    b = fw.RigidBodyDomain(fw.MeshRigid(2, fw.FEMesher(2).createCircle([0.5, 0.5], 0.1), [0.0], 1.0))
    b.set_motion([0.3, -0.2], 1.0)
    ms = RigidMeshSet([a, b], [a, b])
    ms.max_v = ms.max_speed()
    ms.populate()
    assert ms.reg.shape[0] == ms.n_elems  # unfiltered: every edge
    assert sorted(ms.reg[:, 0].tolist()) == [0] * 24 + [1] * (ms.n_elems - 24)
    f = np.float32
    buf = f(2.0 * ms.max_v / 500.0)
    nx, ny = (int(v) for v in ms.cn)
    seen = {}
    for c in range(ms.tc):
        ent = ms.sorted[ms.cell[c, 0]:ms.cell[c, 1]].tolist()
        assert ent == sorted(ent)  # registration order inside a cell
        for ei in ent:
            seen.setdefault(ei, set()).add(c)
    for ei, (rid, e) in enumerate(ms.reg):
        lb = np.minimum(ms.coords[ms.conn[e, 0]], ms.coords[ms.conn[e, 1]]) - buf
        ub = np.maximum(ms.coords[ms.conn[e, 0]], ms.coords[ms.conn[e, 1]]) + buf
        lx, ly = (np.floor((lb - ms.glb) / ms.gs)).astype(int)
        ux, uy = (np.floor((ub - ms.glb) / ms.gs)).astype(int)
        cells = {ix + iy * nx for ix in range(max(lx, 0), min(ux, nx - 1) + 1) for iy in range(max(ly, 0), min(uy, ny - 1) + 1)}
        assert seen.get(ei, set()) == {c for c in cells if c < ms.tc}


def test_penetration_force_closed_form():
    r = slab((0.0, 0.0), 1.0, 0.2)  # top edge at y = 0.2, outward normal +y
    for d in (1e-3, 4e-3):
        spec, g, ms, c, _ = triangle_scene([[0.5, 0.2 - d], [0.55, 0.26], [0.45, 0.26]], r)
        o = OracleScene(spec)
        o.warmup()
        f = step(o, g, ms)
        k = float(np.float32(c.penalty))
        assert abs(f[0, 1] - k * d) < 1e-4 * k * d and abs(f[0, 0]) < 1e-6 * k * d
        assert np.all(f[1:] == 0.0)  # above the edge: no force
    # outside the rigid's AABB: nothing, whatever the penetration of the infinite edge line
    spec, g, ms, c, _ = triangle_scene([[1.2, 0.19], [1.25, 0.26], [1.15, 0.26]], r)
    o = OracleScene(spec)
    o.warmup()
    assert np.all(step(o, g, ms) == 0.0)


def test_friction_history():
    r = slab((0.0, 0.0), 1.0, 0.2)
    d, mu = 2e-3, 0.5
    spec, g, ms, c, rd = triangle_scene([[0.6, 0.2 - d], [0.65, 0.26], [0.55, 0.26]], r, friction=mu)
    o = OracleScene(spec)
    o.warmup()
    k = float(np.float32(c.penalty))
    f = step(o, g, ms)
    assert abs(f[0, 0]) < 1e-9  # no history: no friction
    st = o.mesh_state[0]
    assert st.pvalid[0] == 1
    # move the node along the edge (the cached edge still holds it): the contact point
    # slides with it
    for s in (1e-5, 3e-3):
        o2 = OracleScene(spec)
        o2.warmup()
        step(o2, g, ms)
        o2.x[:, 0] += np.float32(s)
        f2 = step(o2, g, ms)
        expect = min(k * s, mu * k * d)
        assert abs(f2[0, 0] - expect) < 2e-3 * abs(expect)


def test_tie_pulls_to_projection():
    r = slab((0.0, 0.0), 1.0, 0.2)
    for dy in (-1e-3, 1e-3):
        spec, g, ms, c, _ = triangle_scene([[0.5, 0.2 + dy], [0.55, 0.26], [0.45, 0.26]], r, tied=True)
        o = OracleScene(spec)
        o.warmup()
        f = step(o, g, ms)
        k = float(np.float32(c.penalty))
        st = o.mesh_state[0]
        assert st.tres[0] == 1 and st.tres[1] == 0 and st.tres[2] == 0
        assert abs(f[0, 1] - k * (-dy)) < 1e-3 * k * abs(dy) and abs(f[0, 0]) < 1e-6
        # the same pull on later steps, whatever the gate
        assert np.allclose(step(o, g, ms), f)


def test_skip_epoch_until_reset():
    r = slab((0.0, 0.0), 1.0, 0.2, angle=0.6)  # turned: the AABB has empty corners
    rd = fw.RigidBodyDomain(r)
    w, lb, ub = rd.mesh_world()
    corner = [float(lb[0]) + 0.02, float(ub[1]) - 0.02]  # inside the AABB, far from every edge
    spec, g, ms, c, _ = triangle_scene([corner, [corner[0] + 0.01, corner[1] + 0.002], [corner[0], corner[1] + 0.01]], r)
    o = OracleScene(spec)
    o.warmup()
    assert np.all(step(o, g, ms) == 0.0)
    st = o.mesh_state[0]
    assert st.skip[0] == 1
    step(o, g, ms)
    assert st.skip[0] == 1  # stays skipped
    f = step(o, g, ms, counter_reset=o.counter)  # the reset before this step: queried again, skipped again
    assert st.skip[0] == 1 and np.all(f == 0.0)
    # node 0 pushed 0.4 mm into the highest edge (inside its margin, 2% of its box
    # diagonal): still skipped, so no force, until a reset
    a, b = w[r.edges[:, 0]], w[r.edges[:, 1]]
    top = int(np.argmax(a[:, 1] + b[:, 1]))
    mid = 0.5 * (a[top].astype(np.float64) + b[top])
    t = (b[top] - a[top]).astype(np.float64)
    n_out = np.array([t[1], -t[0]]) / np.hypot(*t)
    p0 = mid - 4e-4 * n_out
    o.x[:] = (o.x - o.x[0] + p0).astype(np.float32)
    o.counter += 1  # off the full-query subcycle's phase does not matter: the cache is empty
    assert np.all(step(o, g, ms) == 0.0) and st.skip[0] == 1
    o.counter += o.counter % 2  # a full-query step
    f = step(o, g, ms, counter_reset=o.counter)
    k = float(np.float32(c.penalty))
    assert np.allclose(f[0], k * 4e-4 * n_out, rtol=5e-3) and st.skip[0] == -1


def test_kinematic_schedule():
    """The skip reset and box refresh after every N2-th substep, N2 = max(1, (1/500) // dt)."""
    ms = RigidMeshSet([fw.RigidBodyDomain(slab((0.0, 0.0), 1.0, 0.2))], [])
    ms.all_rigids = ms.meshes
    kin = KinematicRigids(ms.meshes, ms)
    resets = []
    for k in range(25):
        resets.append(kin.before_substep())
        kin.after_substep(2.5e-4)  # N2 = 8
    assert [r for r in resets if r >= 0] == [8, 16, 24]
