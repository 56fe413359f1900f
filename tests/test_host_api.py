"""Host-side logic and the C ABI surface (no GPU needed).

* Mesh topology and lengths vs the dict restatement of flatworld/mesh.py:53-112
  and the committed fixtures (bit-exact);
* BC / IC packing vs a straight restatement of
  flatworld/femspringmanager.py:660-704;
* stable-time and penalty scalars (flatworld/femdomain.py:24,
  flatworld/femcontact.py:94-141);
* libflatworld_fem.so loads and exports every function declared in
  include/flatworld_fem.h; the host tiler's invariants hold on regular and
  irregular meshes; argument errors map to the reference's exception types.
"""

import dataclasses
import json
import os
import re

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from flatworld_amd import workloads as W
from flatworld_amd.definitions import ATYPE, FORCETYPE, GRAVITY, UTYPE, VTYPE
from oracle.oracle import boundary_edges_dict, char_lengths_loop

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _topology_meshes():
    from tests.golden.make_golden import circle_mesh, grid100_mesh, quad_mesh

    return {"quad": quad_mesh(), "circle": circle_mesh(), "grid100": grid100_mesh()}


def test_mesh_topology_matches_fixture():
    topo = json.load(open(os.path.join(GOLDEN, "topology.json")))
    for name, (conn, coords) in _topology_meshes().items():
        m = fw.Mesh(2, conn, coords)
        assert m.boundaryElements.tolist() == topo[name]["boundary_edges"], name
        assert m.boundaryNodes.tolist() == topo[name]["boundary_nodes"], name
        assert m.numBoundNodes == len(topo[name]["boundary_nodes"])


@pytest.mark.parametrize("seed", range(8))
def test_mesh_boundary_random_orientations(seed):
    """Random grids with shuffled elements and flipped orientations: vectorized == dict order."""
    rng = np.random.default_rng(seed)
    s = int(rng.integers(2, 9))
    conn, coords = fw.structured_grid(s, jitter=0.2, seed=seed)
    conn = conn[rng.permutation(conn.shape[0])]
    flip = rng.random(conn.shape[0]) < 0.5
    conn[flip] = conn[flip][:, [0, 2, 1]]
    rot = rng.integers(0, 3, conn.shape[0])
    conn = np.stack([np.roll(c, r) for c, r in zip(conn, rot)])
    # drop some elements to create holes and notches
    keep = rng.random(conn.shape[0]) > 0.2
    conn = np.ascontiguousarray(conn[keep])
    m = fw.Mesh(2, conn, coords)
    ref = boundary_edges_dict(conn)
    assert np.array_equal(m.boundaryElements, ref)
    assert np.array_equal(m.boundaryNodes, np.array(sorted(set(ref.reshape(-1).tolist()))))


def test_char_lengths_bitexact():
    for name, (conn, coords) in _topology_meshes().items():
        if name == "grid100":
            conn = conn[:400]
        m = fw.Mesh(2, conn, coords)
        ref = char_lengths_loop(conn, coords)
        assert np.array_equal(m._collect_edge_lengths(), ref)
        assert m.charLength == float(np.min(ref))
        assert m.charAverageLength == float(np.mean(ref))


def _process_conditions_loop(domain):
    """Straight restatement of flatworld/femspringmanager.py:660-704 for one domain."""
    n = domain.nnodes
    bc = np.zeros(n, np.int32)
    bcv = np.zeros((n, 2), np.float32)
    V = np.zeros((n, 2), np.float32)
    for b in domain.bcs:
        t, nodes, values = b.processData()
        if values is None:
            values = [0.0] * 2
        if nodes is None:
            nodes = range(n)
        for i in nodes:
            bc[int(i)] |= int(t)
            bcv[int(i), 0] = float(values[0])
            bcv[int(i), 1] = float(values[1])
    for ic in domain.initials:
        if ic.is_all_nodes:
            V[:, 0], V[:, 1] = float(ic.vel[0]), float(ic.vel[1])
        else:
            for i in ic.nds:
                V[int(i)] = [float(ic.vel[0]), float(ic.vel[1])]
    return bc, bcv, V


def test_bc_packing_matches_restatement():
    conn, coords = fw.structured_grid(4)
    bcs = [fw.Gravity([0.0, -9.8]), fw.Force([0, 3, 7], [1.0, 2.0]), fw.Fixed([7, 8]), fw.EnforceVel([8, 9], [0.5, -0.5]),
           fw.EnforceAcc([10], [0.0, 3.0])]
    ics = [fw.InitialVel("ALL", [0.1, 0.2]), fw.InitialVel([3, 4], [-1.0, 0.0])]
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), bcs, initials=ics)
    spec = fw.pack_fem_domain(dom)
    bc, bcv, V = _process_conditions_loop(dom)
    assert np.array_equal(spec.bc, bc)
    assert np.array_equal(spec.bcv, bcv)
    assert np.array_equal(spec.v0, V)
    flags, table = spec.node_flags()
    assert np.array_equal(table[flags >> 16], bcv)
    assert np.array_equal(flags & 0x1FF, bc.astype(np.uint32))
    bn = np.zeros(spec.n_nodes, bool)
    bn[spec.bnodes] = True
    assert np.array_equal((flags & 0x8000) != 0, bn)
    assert bc[7] == FORCETYPE | GRAVITY | UTYPE and bc[8] == GRAVITY | UTYPE | VTYPE and bc[10] == GRAVITY | ATYPE


def test_bc_table_keeps_signed_zeros_distinct():
    """The BC value table is keyed by bit pattern: a node whose value is -0.0 keeps -0.0."""
    conn, coords = fw.structured_grid(3)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)),
                       [fw.Gravity([0.0, -9.8]), fw.Force([1, 2], [-0.0, 1.0]), fw.Force([5], [0.0, 1.0])])
    spec = fw.pack_fem_domain(dom)
    flags, table = spec.node_flags()
    got = table[flags >> 16]
    assert np.array_equal(got.view(np.uint32), spec.bcv.astype(np.float32).view(np.uint32))
    assert np.signbit(got[1, 0]) and not np.signbit(got[5, 0])
    assert table.shape[0] == 3
    # one distinct row: the single-row table
    dom1 = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [fw.Gravity([0.0, -9.8])])
    f1, t1 = fw.pack_fem_domain(dom1).node_flags()
    assert t1.shape == (1, 2) and np.all((f1 >> 16) == 0) and t1[0, 1] == np.float32(-9.8)


def test_stable_time_and_penalty_scalars():
    conn, coords = fw.structured_grid(10, extent=0.4)
    mat = fw.Elastic(5e4, 0.35, 900.0)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(mat), [])
    L = dom.mesh.charLength
    assert dom.stableTime == 0.4 * L / (5e4 / 900.0) ** 0.5
    g = fw.GroundDomain(2, [0, 0], [0, 1])
    spec = W.one_body_spec(dom, g)
    K = 5e4 / (3.0 * (1.0 - 2.0 * 0.35))
    assert spec.grounds[0].penalty == np.float32(2.0 * K)
    m_node = 900.0 * (L * L * 0.5) / 3.0
    dtc = 0.9 * (m_node / (2.0 * K)) ** 0.5
    assert spec.dt == min(1e-3, dom.stableTime, dtc)
    # nu is clamped at 0.4999 in the bulk modulus
    from flatworld_amd.femcontact import ContactBase

    assert ContactBase._bulk_modulus(fw.Elastic(1.0, 0.6, 1.0)) == 1.0 / (3.0 * (1.0 - 2.0 * 0.4999))


def test_ground_aabbs():
    g = fw.GroundDomain(2, [0.0, 0.3], [0.0, 2.0])
    lo, hi = g.getBBox()
    assert np.allclose(lo, [-10000.0, 0.2]) and np.allclose(hi, [10000.0, 0.4])
    hf = fw.HeightFieldDomain(2, np.array([0.1, 0.3, 0.2], np.float32), [0.0, 0.0], [2.0, 1.0])
    lo, hi = hf.getBBox()
    assert np.allclose(lo, [0.0, 0.0]) and np.allclose(hi, [2.0, 0.4])


# ------------------------------------------------------------------ C ABI


def test_library_exports_declared_symbols():
    header = open(os.path.join(ROOT, "include", "flatworld_fem.h")).read()
    declared = set(re.findall(r"\b(fw_[a-z_]+)\s*\(", header))
    assert declared, "no declarations found"
    L = _native.load()
    for name in sorted(declared):
        assert hasattr(L, name), name
        assert name in _native.SIGNATURES, f"{name} not bound in _native.SIGNATURES"
    assert L.fw_abi_version() == 8


@pytest.mark.parametrize("tile_nodes,tile_slots", [(512, 1024), (64, 1024), (1024, 2048), (7, 1024)])
def test_tiler_invariants_grids(tile_nodes, tile_slots):
    for side in (3, 17, 40):
        dom = W.grid_domain(side, (0, 0), 1.0, fw.Elastic(1e4, 0.3, 1.0), [], jitter=0.1, seed=side)
        st = _native.tiling_check(W.one_body_spec(dom, None, dt=1e-4), tile_nodes, tile_slots)
        assert st["csr_entries"] == 3 * dom.nelements
        assert st["max_slots"] <= tile_slots


def test_tiler_invariants_irregular():
    """Shuffled numbering, a fan node of degree 16 and a degenerate (repeated-node) element."""
    rng = np.random.default_rng(3)
    conn, coords = fw.structured_grid(12, jitter=0.2, seed=3)
    perm = rng.permutation(coords.shape[0])
    inv = np.argsort(perm)
    coords = coords[perm]
    conn = inv[conn].astype(np.int32)[rng.permutation(conn.shape[0])]
    conn = np.vstack([conn, [[0, 0, 5]]]).astype(np.int32)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [])
    for tn in (16, 100, 1000):
        st = _native.tiling_check(W.one_body_spec(dom, None, dt=1e-4), tn, 1024)
        assert st["csr_entries"] == 3 * conn.shape[0]
    m = fw.FEMesher(2).createCircle([0.5, 0.5], 0.1)
    dom = fw.FemDomain(m, fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [])
    assert _native.tiling_check(W.one_body_spec(dom, None, dt=1e-4), 3, 1024)["tiles"] >= 6


def test_abi_errors_map_to_exceptions():
    conn, coords = fw.structured_grid(3)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [])
    spec = W.one_body_spec(dom, None, dt=1e-4)
    spec.conn = spec.conn.copy()
    spec.conn[0, 0] = 999
    with pytest.raises(ValueError, match="outside"):
        _native.tiling_check(spec)
    spec = W.one_body_spec(dom, None, dt=1e-4)
    spec.mat_type = 7
    with pytest.raises(ValueError, match="material"):
        _native.tiling_check(spec)


def test_explicitloop_rejects_out_of_scope_domains():
    from flatworld_amd.explicitloop import ExplicitLoop

    conn, coords = fw.structured_grid(2)
    a = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [])
    b = fw.FemDomain(fw.Mesh(2, conn, coords + 2.0), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [])
    # FEM-FEM orderings whose reference behaviour reads another domain's entry
    ground = fw.GroundDomain(2, [0.0, -1.0], [0.0, 1.0])
    with pytest.raises(NotImplementedError, match="FEM domains listed first"):
        ExplicitLoop(0.0, [ground, a, b])
    quiet = fw.SpringMassDomain(2, coords + 4.0, np.array([[0, 1], [1, 2]], np.int32), considerContact=False)
    with pytest.raises(NotImplementedError, match="spring-mass domain listed before"):
        ExplicitLoop(0.0, [quiet, a, b])


def test_prepare_is_host_only_and_thread_safe():
    """fw_scene_prepare needs no GPU; concurrent preparations agree with a serial one."""
    from concurrent.futures import ThreadPoolExecutor

    from flatworld_amd import workloads as W

    rng = W.rng_for(4)
    specs = [W.serving_job(rng, 8, 40, 60, 90) for _ in range(12)]
    tune = dict(tile_nodes=256, tile_slots=512, block=256, grid_per_cu=0)
    serial = [_native.Prepared(s, tune).footprint for s in specs]
    with ThreadPoolExecutor(6) as ex:
        par = list(ex.map(lambda s: _native.Prepared(s, tune).footprint, specs))
    assert serial == par and all(f > 0 for f in serial)
    with pytest.raises(ValueError, match="unsupported"):
        _native.Prepared(specs[0], dict(tile_nodes=256, tile_slots=777, block=256, grid_per_cu=0))
    bad = dataclasses.replace(specs[0], conn=specs[0].conn.copy())
    bad.conn[0, 0] = bad.n_nodes + 5
    with pytest.raises(ValueError, match="outside"):
        _native.Prepared(bad, tune)


def _fan(n_tri, r=0.1):
    """n_tri triangles around one centre node (the centre has n_tri element corners)."""
    ang = np.linspace(0.0, 2.0 * np.pi, n_tri, endpoint=False)
    coords = np.concatenate([[[0.5, 0.5]], 0.5 + r * np.stack([np.cos(ang), np.sin(ang)], 1)]).astype(np.float32)
    conn = np.array([[0, 1 + k, 1 + (k + 1) % n_tri] for k in range(n_tri)], dtype=np.int32)
    return conn, coords


def test_node_degree_limit_is_reported():
    """A triangle node takes up to 1024 element corners (its rank spread over the slot
    record's corner word and the upper half of .y); fans of 33, 64 and 1024 triangles
    tile with every invariant checked; 1025 is refused with a clear error.  A spring
    node stays limited to 32 (the record's .y holds the rest length)."""
    for n_tri, slots in ((33, 1024), (64, 1024), (1024, 2048)):
        conn, coords = _fan(n_tri)
        dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [fw.Gravity([0.0, -9.8])])
        st = _native.tiling_check(fw.pack_fem_domain(dom), 64, slots)
        assert st["csr_entries"] == 3 * n_tri
    conn, coords = _fan(1025)
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [fw.Gravity([0.0, -9.8])])
    with pytest.raises(NotImplementedError, match="at most 1024"):
        _native.Prepared(fw.pack_fem_domain(dom), dict(tile_nodes=1024, tile_slots=2048, block=512, grid_per_cu=0))
    xy = np.concatenate([[[0.5, 0.5]], 0.5 + 0.1 * np.stack([np.cos(np.arange(33)), np.sin(np.arange(33))], 1)])
    star = np.stack([np.zeros(33, np.int64), np.arange(1, 34)], 1).astype(np.int32)
    sdom = fw.SpringMassDomain(2, xy.astype(np.float32), star, [100.0, 0.1, 0.01], [fw.Gravity([0.0, -9.8])])
    with pytest.raises(NotImplementedError, match="at most 32"):
        _native.Prepared(fw.pack_domain(sdom), dict(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0))


def test_isolated_node_and_degenerate_element_tile():
    """A node that no element references and a zero-area element still give a valid tiling."""
    conn, coords = fw.structured_grid(5)
    coords = np.concatenate([coords, [[2.0, 2.0]]]).astype(np.float32)  # isolated node
    conn = np.concatenate([conn, [[0, 1, 1]]]).astype(np.int32)  # repeated node: zero area
    dom = fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(fw.Elastic(1e4, 0.3, 1.0)), [fw.Gravity([0.0, -9.8])])
    spec = fw.pack_fem_domain(dom)
    for tile_nodes, tile_slots in ((512, 1024), (8, 512)):
        stats = _native.tiling_check(spec, tile_nodes, tile_slots)
        assert stats["tiles"] >= 1 and stats["csr_entries"] == 3 * spec.n_elems


# ------------------------------------------------------------------ spring-mass domains


def _spring_net(n, seed=0, jitter=0.2):
    """A spring network over an n x n node lattice: edges, diagonals and a few bending springs."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    ij = np.array([(i, j) for i in range(n) for j in range(n)])
    X = (ij * h + rng.uniform(-jitter * h, jitter * h, ij.shape)).astype(np.float32)
    idx = lambda i, j: i * n + j  # noqa: E731
    conn = []
    for i in range(n):
        for j in range(n):
            for di, dj in ((1, 0), (0, 1), (1, 1), (1, -1), (2, 0)):
                if 0 <= i + di < n and 0 <= j + dj < n:
                    conn.append((idx(i, j), idx(i + di, j + dj)))
    return X, np.array(conn, np.int32)


def test_spring_rest_lengths_equal_reference_loop():
    """Vectorized rest lengths == the reference's per-spring np.linalg.norm loop (flatworld/springmass.py:46-52)."""
    from flatworld_amd.springmass import spring_rest_lengths

    for seed in range(3):
        X, conn = _spring_net(20, seed)
        X = X * np.float32(3.7) - np.float32(1.3)
        loop = np.zeros(conn.shape[0], np.float32)
        for i in range(conn.shape[0]):
            loop[i] = np.linalg.norm(X[conn[i, 0]] - X[conn[i, 1]])
        assert np.array_equal(spring_rest_lengths(X, conn), loop)
    dom = fw.SpringMassDomain(2, X, conn, [10.0, 1.0, 2.0])
    assert dom.stableTime == 0.5 * np.sqrt(2.0 / 10.0) and dom.massInv == 0.5


def test_spring_packing():
    X, conn = _spring_net(6)
    dom = fw.SpringMassDomain(2, X, conn, [80.0, 0.5, 0.25], [fw.Gravity([0.0, -9.8]), fw.Fixed([0, 1])],
                              initials=[fw.InitialVel("ALL", [0.1, 0.0])])
    spec = fw.pack_spring_domain(dom)
    assert spec.mat_type == 4 and spec.conn.shape == (conn.shape[0], 2)
    assert spec.E == np.float32(80.0) and spec.rho == np.float32(0.25) and spec.spring_c == np.float32(0.5)
    assert np.array_equal(spec.bnodes, np.arange(X.shape[0]))  # every node is a contact work item
    flags, _ = spec.node_flags()
    assert np.all(flags & fw.definitions.NODE_BOUNDARY_BIT)
    assert np.array_equal(spec.rest_len, dom.restLength) and np.all(spec.v0[:, 0] == np.float32(0.1))
    d, keep = _native.scene_desc(spec)
    assert d.mat_type == 4 and bool(d.rest_len) and abs(d.spring_c - 0.5) < 1e-7


def test_spring_contact_factory():
    from flatworld_amd.femcontact import ContactSpringAnalytical, ContactSpringHeightField, contactHelper

    X, conn = _spring_net(4)
    sp = fw.SpringMassDomain(2, X, conn, [30.0, 0.0, 0.2])
    g = fw.GroundDomain(2, [0, 0], [0, 1])
    hf = fw.HeightFieldDomain(2, np.zeros(8, np.float32), [0.0, 0.0], [1.0, 1.0])
    vox = fw.VoxelGridDomain(2, 4, 4, [0.0, 0.0], [1.0, 1.0])
    c = contactHelper(g, sp, g.type | sp.type)
    assert isinstance(c, ContactSpringAnalytical) and c.domain1 is sp and c.penalty == 3000.0
    assert c.stableTime == 0.5 * (0.2 / 3000.0) ** 0.5
    assert isinstance(contactHelper(sp, hf, sp.type | hf.type), ContactSpringHeightField)
    # the reference's factory has no spring-voxel or spring-spring class: no contact
    assert contactHelper(sp, vox, sp.type | vox.type) is None
    assert contactHelper(sp, sp, sp.type) is None


def test_spring_fem_contact_scalars():
    """flatworld/femcontact.py:200-213: springs first whatever the argument order,
    penalty 100 min(spring, 2 K), the FEM body's contact stable time."""
    from flatworld_amd.femcontact import ContactSpringFlex, contactHelper

    X, conn = _spring_net(4)
    sp = fw.SpringMassDomain(2, X, conn, [30.0, 0.0, 0.2])
    c2, x2 = fw.structured_grid(3, extent=0.3)
    fem = fw.FemDomain(fw.Mesh(2, c2, x2), fw.SolidProp(fw.Elastic(1e4, 0.3, 50.0)), [])
    for a, b in ((fem, sp), (sp, fem)):
        c = contactHelper(a, b, a.type | b.type)
        assert isinstance(c, ContactSpringFlex) and c.domain1 is sp and c.domain2 is fem
        assert c.penalty == min(30.0, 2.0 * 1e4 / (3.0 * (1.0 - 0.6))) * 100.0 == 3000.0
        L = fem.mesh.charLength
        assert c.stableTime == 0.9 * ((50.0 * (L * L * 0.5) / 3.0) / 3000.0) ** 0.5
    stiff = fw.SpringMassDomain(2, X, conn, [1e5, 0.0, 0.2])
    assert contactHelper(stiff, fem, stiff.type | fem.type).penalty == 2.0 * 1e4 / (3.0 * (1.0 - 0.6)) * 100.0


@pytest.mark.parametrize("shape", [(256, 512), (512, 1024), (1024, 2048)])
def test_spring_tiling_invariants(shape):
    """The host tiler on spring networks: one tile and many (halo velocity copies pushed once each)."""
    tile_nodes, tile_slots = shape
    for n in (5, 40):
        X, conn = _spring_net(n, seed=n)
        spec = fw.pack_spring_domain(fw.SpringMassDomain(2, X, conn, [10.0, 1.0, 1.0]))
        st = _native.tiling_check(spec, tile_nodes, tile_slots)
        assert st["csr_entries"] == 2 * conn.shape[0]
        if n == 40:
            assert st["tiles"] > 1 and st["halo_slots"] > 0


def test_spring_scene_validation():
    X, conn = _spring_net(4)
    spec = fw.pack_spring_domain(fw.SpringMassDomain(2, X, conn, [10.0, 1.0, 1.0]))
    with pytest.raises(ValueError, match="rest_len"):
        _native.tiling_check(dataclasses.replace(spec, rest_len=spec.rest_len[:-1]))
    bad = dataclasses.replace(spec, conn=spec.conn.copy())
    bad.conn[3, 1] = X.shape[0] + 2
    with pytest.raises(ValueError, match="outside"):
        _native.tiling_check(bad)


def test_tie_entries_parse_as_the_reference(capsys):
    """flatworld/explicitloop.py:400-439: pairs of domain indices, negative ones from
    the end, stored as (low, high); anything else is counted and ignored."""
    import numpy as np

    from flatworld_amd.explicitloop import ExplicitLoop

    loop = ExplicitLoop.__new__(ExplicitLoop)
    loop.domains = [object()] * 4
    loop.ties = [(0, 1), (-1, 1), ("a", 1), (2, 2), (0, 7), [3, np.int64(0)], (1, 2, 3), {"domain1": 0}]
    loop._tie_pairs = set()
    loop._refresh_tie_pairs()
    assert loop._tie_pairs == {(0, 1), (1, 3), (0, 3)}
    assert loop._is_tie_pair(3, 0) and loop._is_tie_pair(1, 0) and not loop._is_tie_pair(1, 2)
    assert "5 tie entries could not be resolved" in capsys.readouterr().out


def test_tied_contact_flags_its_work_items():
    import numpy as np

    import flatworld_amd as fw
    from flatworld_amd.definitions import ContactType
    from flatworld_amd.femcontact import ContactFlexFlex
    from flatworld_amd.flexcontact import build_flex_group

    mat = fw.Elastic(1e4, 0.3, 1.0)
    b = [fw.FemDomain(fw.Mesh(2, *fw.structured_grid(2, lower=(2.0 * k, 0.0), extent=1.0)), fw.SolidProp(mat), [])
         for k in range(3)]
    cs = [ContactFlexFlex(b[1], b[0], ContactType.FLEXFLEX), ContactFlexFlex(b[2], b[1], ContactType.FLEXFLEX)]
    cs[1].tied = True
    g = build_flex_group(b, cs, False, sum(x.nnodes for x in b))
    assert g.n_tied == int(np.sum(g.item_pair == 1)) > 0
    assert np.array_equal(g.tied_flags(), (g.item_pair == 1).astype(np.int32))


def test_tiling_cache_reuses_by_connectivity():
    """fw_scene_prepare reuses a tiling for a scene with the same connectivity (and
    kernel shape) whatever its positions and material; another connectivity, another
    shape, or springs with other rest lengths build their own."""
    ground = fw.GroundDomain(2, [0.0, 0.0], [0.0, 1.0])

    def block(side, lower, mat):
        return W.one_body_spec(W.grid_domain(side, lower, 0.4, mat, [fw.Gravity([0.0, -9.8])]), ground, horizon=10)

    tun = dict(tile_nodes=128, tile_slots=512, block=256, grid_per_cu=0)
    _native.tiling_cache_clear()
    a = _native.Prepared(block(40, (0.3, 0.0), fw.Elastic(1e4, 0.3, 10.0)), tun)
    b = _native.Prepared(block(40, (1.7, 0.2), fw.MisesPlastic(2e4, 0.3, 10.0, 50.0, 1e2)), tun)
    st = _native.tiling_cache_stats()
    assert (st["hits"], st["misses"], st["entries"]) == (1, 1, 1) and st["bytes"] > 0
    c = _native.Prepared(block(40, (0.9, 0.5), fw.Elastic(3e4, 0.25, 20.0)), tun)
    assert a.footprint == c.footprint
    _native.Prepared(block(41, (0.3, 0.0), fw.Elastic(1e4, 0.3, 10.0)), tun)
    _native.Prepared(block(40, (0.3, 0.0), fw.Elastic(1e4, 0.3, 10.0)), dict(tun, tile_nodes=256))
    st = _native.tiling_cache_stats()
    assert (st["hits"], st["misses"], st["entries"]) == (2, 3, 3)
    lat = [W.spring_lattice(12, (0.1, 0.1), h, [200.0, 0.3, 0.01], [fw.Gravity([0.0, -9.8])]) for h in (0.01, 0.01, 0.02)]
    for d in lat:
        _native.Prepared(W.one_body_spec(d, ground, horizon=10), dict(tile_nodes=64, tile_slots=512, block=256, grid_per_cu=0))
    st = _native.tiling_cache_stats()
    assert (st["hits"], st["misses"]) == (3, 5)  # the second lattice hits, the third (other rest lengths) misses
    _native.tiling_cache_clear()
    assert _native.tiling_cache_stats() == {"hits": 0, "misses": 0, "bytes": 0, "entries": 0}
