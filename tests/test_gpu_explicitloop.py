"""The drop-in API on the GPU: the reference's own scenario tests, run through
``ExplicitLoop.advanceWithTime`` and ``femSpringManager.coords.numpy()``, and the
same loops compared bit for bit with the oracle driven by the same float64 dt
sequence (flatworld/explicitloop.py:558-589).
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd.scene import pack_fem_domain
from oracle.oracle import OracleScene
from tests.parity import log

pytestmark = pytest.mark.gpu

FRAME = 1.0 / 60.0


def oracle_for(loop, domains_with_grounds):
    """One oracle scene per FEM domain of ``loop``, warmed up as setGlobalAABB does."""
    out = []
    for dom, grounds in domains_with_grounds:
        o = OracleScene(pack_fem_domain(dom, grounds))
        o.warmup()
        out.append(o)
    return out


def oracle_frame(loop, oracles, frame_dt=FRAME):
    for sd in loop.substep_sequence(frame_dt):
        for o in oracles:
            o.substep(sd, loop.damping)


def square_block():
    conn = np.array([[0, 1, 3], [0, 3, 2]], dtype=np.int32)
    coords = np.array([[0.5, 0.5], [0.7, 0.5], [0.5, 0.7], [0.7, 0.7]], dtype=np.float32)
    return conn, coords


def test_reference_free_fall_two_triangles():
    """test2D/test_2Dfem_elastic.py:14-40: node 0 crosses y = 0 at t = 1.0 +- one frame."""
    conn, coords = square_block()
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0)), [fw.Gravity([0, -1.0])])
    loop = fw.ExplicitLoop(0.0, [dom], useAdapativeDT=True)
    oracles = oracle_for(loop, [(dom, [])])
    log(f"[case] two-triangle free fall, dt={loop.dt:.3e}")
    t = 0.0
    for _ in range(200):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
        t += FRAME
        pos = loop.femSpringManager.coords.numpy()
        assert np.array_equal(pos[:4], oracles[0].x)
        if pos[0, 1] < 0.0:
            break
    assert np.isclose(t, 1.0, atol=FRAME)


def test_reference_free_fall_100x100_grid():
    """test2D/test_2Dfem_elastic10000.py:14-53 (its grid numbering): node 0 crosses y = 0 at t = sqrt(0.6)."""
    conn, coords = fw.structured_grid(100, lower=(0.0, 0.0), extent=0.6)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords + np.float32(0.3)), fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0)),
                       [fw.Gravity([0, -1.0])])
    loop = fw.ExplicitLoop(0.0, [dom], useAdapativeDT=True)
    log(f"[case] 100x100 free fall, dt={loop.dt:.3e}")
    t = 0.0
    for _ in range(120):
        loop.advanceWithTime(FRAME)
        t += FRAME
        if loop.femSpringManager.coords.numpy()[0, 1] < 0.0:
            break
    assert np.isclose(t, np.sqrt(0.6), atol=FRAME)


def test_reference_two_domains():
    """test2D/test_2Dfem2domain_elastic.py:32-67: two one-triangle domains in one loop, same timing."""
    prop = fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0))
    d1 = fw.FemDomain(fw.Mesh(2, np.array([[0, 2, 1]], np.int32), np.array([[0.5, 0.5], [0.5, 0.7], [0.7, 0.7]], np.float32)),
                      prop, [fw.Gravity([0, -1.0])], considerContact=False)
    d2 = fw.FemDomain(fw.Mesh(2, np.array([[0, 1, 2]], np.int32), np.array([[0.5, 0.5], [0.7, 0.5], [0.7, 0.7]], np.float32)),
                      prop, [fw.Gravity([0, -1.0])], considerContact=False)
    loop = fw.ExplicitLoop(0.0, [d1, d2], useAdapativeDT=True)
    oracles = oracle_for(loop, [(d1, []), (d2, [])])
    log("[case] two domains")
    t = 0.0
    for _ in range(200):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
        t += FRAME
        pos = loop.femSpringManager.coords.numpy()
        assert np.array_equal(pos[:3], oracles[0].x) and np.array_equal(pos[3:6], oracles[1].x)
        if pos[0, 1] < 0.0:
            break
    assert np.isclose(t, 1.0, atol=FRAME)


def test_reference_bcs_three_domains():
    """test2D/test_2Dfem_bcs.py: Force + Fixed, EnforceVel + Fixed, EnforceAcc + Fixed, with damping."""
    conn, coords = square_block()
    prop = fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0))
    doms = [
        fw.FemDomain(fw.Mesh(2, conn, coords), prop, [fw.Force([0, 1], [0, -100.0]), fw.Fixed([2, 3])], False),
        fw.FemDomain(fw.Mesh(2, conn, coords + 0.25), prop, [fw.EnforceVel([0, 1], [0, -1.0]), fw.Fixed([2, 3])], False),
        fw.FemDomain(fw.Mesh(2, conn, coords - 0.25), prop, [fw.EnforceAcc([0, 1], [0, -10.0]), fw.Fixed([2, 3])], False),
    ]
    loop = fw.ExplicitLoop(0.0, doms, damping=0.5, useAdapativeDT=True)
    oracles = oracle_for(loop, [(d, []) for d in doms])
    log("[case] BC kinds, three domains, damping 0.5")
    for _ in range(30):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
    pos = loop.femSpringManager.coords.numpy()
    vel = loop.femSpringManager.V.numpy()
    for k, o in enumerate(oracles):
        assert np.array_equal(pos[4 * k: 4 * k + 4], o.x) and np.array_equal(vel[4 * k: 4 * k + 4], o.v)
    assert np.array_equal(vel[4:6], [[0.0, -1.0], [0.0, -1.0]])  # EnforceVel holds
    assert np.all(vel[[2, 3, 6, 7, 10, 11]] == 0.0)  # Fixed nodes never move


def test_reference_initial_velocity():
    """test2D/test_2Dfem_ics.py: InitialVel on all nodes and on a subset."""
    conn, coords = square_block()
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0)),
                       [fw.Gravity([0, -1.0])], initials=[fw.InitialVel("ALL", [1.0, 0.5]), fw.InitialVel([3], [0.0, 2.0])])
    loop = fw.ExplicitLoop(0.0, [dom], useAdapativeDT=True)
    oracles = oracle_for(loop, [(dom, [])])
    for _ in range(20):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
    assert np.array_equal(loop.femSpringManager.coords.numpy()[:4], oracles[0].x)
    assert np.array_equal(loop.femSpringManager.U.numpy()[:4], oracles[0].u)


def _bowl(nx=513, cx=0.5, r=0.5, zc=0.7):
    xs = np.linspace(0.0, 1.0, nx, dtype=np.float32)
    inside = r * r - (xs - cx) ** 2
    return np.where(inside >= 0.0, zc - np.sqrt(np.maximum(inside, 0.0)), zc).astype(np.float32)


def test_reference_heightfield_bowl():
    """test2D/test_2Dfem_heightfield_contact.py:27-66: a disc dropped into a bowl stays above 0.05."""
    mesh = fw.FEMesher(2).createCircle([0.5, 0.65], 0.1)
    fem = fw.FemDomain(mesh, fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0)), [fw.Gravity([0.0, -10.0])],
                       considerContact=True)
    hf = fw.HeightFieldDomain(2, _bowl(), lb=[0.0, 0.0], ub=[1.0, 1.0], considerContact=True)
    loop = fw.ExplicitLoop(0.0, [hf, fem], useAdapativeDT=True)
    oracles = oracle_for(loop, [(fem, [(hf, loop.contacts[0])])])
    log(f"[case] disc in height-field bowl, dt={loop.dt:.3e}")
    rebound = False
    for _ in range(60):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
        # the disc bounces in the bowl (elastic penalty, no friction): an upward mean
        # velocity means the height field pushed it
        rebound |= bool(loop.femSpringManager.V.numpy()[: mesh.numNodes, 1].mean() > 0.0)
    coords = loop.femSpringManager.coords.numpy()[: mesh.numNodes]
    assert coords[:, 1].min() > 0.05, "FEM object fell through heightfield"
    assert np.array_equal(coords, oracles[0].x)
    assert rebound


def _voxel_pit(nx=12, ny=8):
    occ = np.zeros((nx, ny), dtype=np.int32)
    cx, r, max_h = nx * 0.5, nx * 0.45, int(ny * 0.55)
    for i in range(nx):
        inside = r * r - (i - cx) ** 2
        h = int(max_h * (1.3 - inside / (r * r))) if inside > 0 else max_h
        if h > 0:
            occ[i, :h] = 1
    return fw.VoxelGridDomain(d=2, nx=nx, ny=ny, lb=[0.0, 0.0], ub=[1.0, 1.0], considerContact=True, occupancy_np=occ)


def test_reference_voxel_pit():
    """test2D/test_2Dfem_voxel_contact.py:34-70: a disc dropped into a voxel pit stays above 0.02."""
    mesh = fw.FEMesher(2).createCircle([0.5, 0.75], 0.07)
    fem = fw.FemDomain(mesh, fw.SolidProp(fw.Elastic(E=2e4, nu=0.25, rho=50.0)), [fw.Gravity([0.0, -9.8])],
                       considerContact=True)
    vox = _voxel_pit()
    loop = fw.ExplicitLoop(0.0, [vox, fem], useAdapativeDT=True)
    oracles = oracle_for(loop, [(fem, [(vox, loop.contacts[0])])])
    log(f"[case] disc in voxel pit, dt={loop.dt:.3e}")
    for _ in range(60):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
    coords = loop.femSpringManager.coords.numpy()[: mesh.numNodes]
    assert coords[:, 1].min() > 0.02, "FEM object fell through voxel pit"
    assert np.array_equal(coords, oracles[0].x)


def test_reference_disc_on_plane():
    """test2D/test_2Dfemanalytic_contact.py / test_2Dfemcontact_types.py: a disc resting on the plane y = 0.3."""
    mesh = fw.FEMesher(2).createCircle([0.5, 0.65], 0.1)
    fem = fw.FemDomain(mesh, fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0)), [fw.Gravity([0.0, -10.0])])
    ground = fw.GroundDomain(2, [0.0, 0.3], [0.0, 1.0])
    loop = fw.ExplicitLoop(0.0, [ground, fem], useAdapativeDT=True)
    oracles = oracle_for(loop, [(fem, [(ground, loop.contacts[0])])])
    for _ in range(60):
        loop.advanceWithTime(FRAME)
        oracle_frame(loop, oracles)
    coords = loop.femSpringManager.coords.numpy()[: mesh.numNodes]
    assert coords[:, 1].min() >= 0.2
    assert np.array_equal(coords, oracles[0].x)
