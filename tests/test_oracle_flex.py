"""Pins of the FEM-FEM contact oracle (oracle/flex_oracle.c), CPU only.

The reference holds no recorded output for its FEM-FEM path
(test2D/test_2Dfem_contact.py asserts nothing), so each piece is pinned by what
the mathematics fixes:

* pointToEdgeContact (flatworld/contact_detection.py:15-61): closed forms on
  axis-aligned and rotated edges, the outward normal of a counter-clockwise
  boundary, the t-range and the degenerate edge;
* the spatial hash (flatworld/spatialmanager.py): every segment sits in exactly
  the cells of its box, cells list ascending segments, and a buffered query
  returns every segment of the target body whose box meets the query box, once;
* the contact pass (flatworld/mixedcontact.py:886-1025): with an empty cache the
  chosen segment is the brute-force best over all of the target's segments; a
  flat penetration gives force n k d on the node and -force * weights on the
  edge's nodes; every pass's contributions sum to zero (action and reaction);
  the cache keeps a segment while the node stays in its margin box;
* tied pairs (flatworld/mixedcontact.py:687-811): each tied item resolves to the
  brute-force nearest edge with the weights of its projection and a gap equal to
  minus its penetration; at the state it was resolved in the tie pulls with no
  force, and a rigid shift d of the node's body against the edge's body gives
  -k d on the node (+k d when the edge's body moves); each node lists its tie
  contributions before its sliding-contact ones; tied items never slide;
* whole loops: two free bodies colliding conserve momentum; the reference's
  two-circle scenario runs 1.5 s with contact and without blowing up.
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd.definitions import ContactType
from flatworld_amd.femcontact import ContactFlexFlex
from flatworld_amd.flexcontact import build_flex_group
from flatworld_amd.scene import GROUND_FLEX, add_flex_record, pack_domain
from oracle.oracle import OracleLoop, point_edge


# ------------------------------------------------------------------ point vs edge


def test_point_edge_closed_forms():
    # edge along +x: outward normal of a counter-clockwise boundary points to -y
    ins, pen, n, w = point_edge([0.25, -0.5], [0.0, 0.0], [1.0, 0.0])
    assert ins and pen == np.float32(0.5) and np.array_equal(n, [0.0, -1.0]) and np.array_equal(w, [0.75, 0.25])
    ins, pen, n, w = point_edge([0.5, 0.125], [0.0, 0.0], [1.0, 0.0])
    assert ins and pen == np.float32(-0.125) and np.array_equal(w, [0.5, 0.5])
    # reversed edge flips the normal and the weights
    ins, pen, n, w = point_edge([0.5, 0.125], [1.0, 0.0], [0.0, 0.0])
    assert ins and pen == np.float32(0.125) and np.array_equal(n, [0.0, 1.0])
    # a 3-4-5 edge: distance along the normal and the projection parameter
    ins, pen, n, w = point_edge([0.0, 5.0], [0.0, 0.0], [3.0, 4.0])
    assert ins
    assert abs(float(pen) - (-3.0)) < 1e-6  # n = (0.8, -0.6)
    assert abs(float(w[1]) - 0.8) < 1e-6 and abs(float(w[0] + w[1]) - 1.0) < 1e-6


def test_point_edge_outside_and_degenerate():
    ins, pen, n, w = point_edge([1.5, 0.1], [0.0, 0.0], [1.0, 0.0])
    assert not ins and pen == np.float32(1e9) and np.array_equal(w, [0.0, 0.0])
    assert np.array_equal(n, [0.0, -1.0])  # the normal is set before the range test
    ins, pen, n, w = point_edge([-1e-3, 0.1], [0.0, 0.0], [1.0, 0.0])
    assert not ins
    # the end points are inside (0 <= t <= 1)
    assert point_edge([0.0, 0.1], [0.0, 0.0], [1.0, 0.0])[0] and point_edge([1.0, 0.1], [0.0, 0.0], [1.0, 0.0])[0]
    ins, pen, n, w = point_edge([0.0, 0.1], [0.3, 0.3], [0.3, 0.3])
    assert not ins and pen == np.float32(1e9) and np.array_equal(n, [0.0, 0.0])


# ------------------------------------------------------------------ fixtures


def _grid_body(side, lower, extent, mat, bcs=(), initials=(), jitter=0.0, seed=0, friction=0.0):
    """A grid body with counter-clockwise triangles, so that its boundary edges run
    counter-clockwise and their normals (dy, -dx) point outwards."""
    conn, coords = fw.structured_grid(side, lower=lower, extent=extent, jitter=jitter, seed=seed)
    conn = np.ascontiguousarray(conn[:, [0, 2, 1]])
    return fw.FemDomain(fw.Mesh(2, conn, coords), fw.SolidProp(mat), list(bcs), initials=list(initials),
                        friction=friction)


def _loop(bodies, gated=False, tied=()):
    """OracleLoop of FEM bodies with every pair in contact, in the reference's pair
    order; ``tied`` lists the (i, j) body pairs whose contact is tied."""
    contacts = []
    tied = {frozenset(t) for t in tied}
    for i in range(len(bodies)):
        for j in range(i):
            c = ContactFlexFlex(bodies[i], bodies[j], ContactType.FLEXFLEX)
            c.tied = frozenset((i, j)) in tied
            contacts.append(c)
    g = build_flex_group(bodies, contacts, gated, sum(b.nnodes for b in bodies))
    specs = [add_flex_record(pack_domain(b, []), g.flex_cap(k)) for k, b in enumerate(bodies)]
    dt = min([1.0] + [b.stableTime for b in bodies] + [c.stableTime for c in contacts])
    return OracleLoop(specs, g), g, dt


def _seg_xy(loop, g, s):
    b = int(np.searchsorted(g.body_seg_off, s, side="right") - 1)
    a, c = g.seg_node[s]
    x = loop.flex.bodies[b].x
    return b, x[a], x[c]


# ------------------------------------------------------------------ spatial hash


def test_hash_cells_hold_each_segment_in_its_box_cells():
    mat = fw.Elastic(1e4, 0.3, 10.0)
    bodies = [_grid_body(6, (0.0, 0.0), 0.3, mat, jitter=0.2, seed=1),
              _grid_body(5, (0.25, 0.1), 0.2, mat, jitter=0.2, seed=2)]
    o, g, _ = _loop(bodies)
    F = o.flex
    gr = F.grid
    nx, ny = gr["cn"]
    assert gr["tc"] == nx * ny
    expect = {c: [] for c in range(gr["tc"])}
    for s in range(g.n_segs):
        _, a, c = _seg_xy(o, g, s)
        lb, ub = np.minimum(a, c), np.maximum(a, c)
        rng = [int(np.floor(np.float32(v - o_) / np.float32(h))) for v, o_, h in
               ((lb[0], gr["glb"][0], gr["gs"][0]), (lb[1], gr["glb"][1], gr["gs"][1]),
                (ub[0], gr["glb"][0], gr["gs"][0]), (ub[1], gr["glb"][1], gr["gs"][1]))]
        lx, ly = min(max(rng[0], 0), nx - 1), min(max(rng[1], 0), ny - 1)
        ux, uy = min(max(rng[2], 0), nx - 1), min(max(rng[3], 0), ny - 1)
        for ix in range(lx, ux + 1):
            for iy in range(ly, uy + 1):
                expect[ix + iy * nx].append(s)
    for cid, segs in expect.items():
        got = F.sorted_seg[F.cell_start[cid]: F.cell_end[cid]].tolist()
        assert got == segs, cid  # ascending, since segments are visited in order
    assert gr["n_pairs"] == sum(len(v) for v in expect.values())
    # bounds: the union of the bodies' boundary boxes widened by 1% of its larger side
    lo = np.minimum(o.bodies[0].aabb[:2], o.bodies[1].aabb[:2]).astype(np.float64)
    hi = np.maximum(o.bodies[0].aabb[2:], o.bodies[1].aabb[2:]).astype(np.float64)
    pad = 0.01 * (hi - lo).max()
    assert np.allclose(gr["glb"], lo - pad, atol=1e-6)
    # cells: at least 1.5 x the smallest segment-box diagonal, as few as tile the bounds
    diag = min(float(np.hypot(*(_seg_xy(o, g, s)[2] - _seg_xy(o, g, s)[1]))) for s in range(g.n_segs))
    span = (hi - lo) + 2 * pad
    for k, cells in enumerate((nx, ny)):
        assert cells == max(1, int(np.ceil(span[k] / (1.5 * diag + 1e-6) - 1e-6)))
        assert abs(float(gr["gs"][k]) * cells - span[k]) < 1e-5


@pytest.mark.parametrize("seed", range(4))
def test_hash_query_is_conservative_and_unique(seed):
    rng = np.random.default_rng(seed)
    mat = fw.Elastic(1e4, 0.3, 10.0)
    bodies = [_grid_body(5, (0.0, 0.0), 0.25, mat, jitter=0.3, seed=seed),
              _grid_body(4, (0.2, 0.05), 0.2, mat, jitter=0.3, seed=seed + 10),
              _grid_body(3, (0.05, 0.22), 0.15, mat, jitter=0.3, seed=seed + 20)]
    o, g, _ = _loop(bodies)
    for _ in range(200):
        p = rng.uniform(-0.05, 0.45, 2).astype(np.float32)
        buf = np.float32(rng.uniform(0.005, 0.08))
        target = int(rng.integers(0, 3))
        got = o.flex.query(p, buf, target)
        assert len(set(got.tolist())) == len(got)
        lo, hi = p - buf, p + buf
        for s in range(g.body_seg_off[target], g.body_seg_off[target + 1]):
            _, a, c = _seg_xy(o, g, s)
            if np.all(np.minimum(a, c) <= hi) and np.all(np.maximum(a, c) >= lo):
                assert s in got, (s, p, buf)
        for s in got:
            assert g.body_seg_off[target] <= s < g.body_seg_off[target + 1]


# ------------------------------------------------------------------ contact pass


def _brute_best(o, g, w):
    """Best segment of the item's target over all its segments (cache empty)."""
    b, nid, t = int(g.item_body[w]), int(g.item_node[w]), int(g.item_target[w])
    p = o.flex.bodies[b].x[nid]
    gr = o.flex.grid
    qbuf = np.float32(max(max(1e-4, float(max(gr["gs"]))), 0.016))
    best, best_s = np.float32(1e9), -1
    for s in range(g.body_seg_off[t], g.body_seg_off[t + 1]):
        _, a, c = _seg_xy(o, g, s)
        ins, pen, n, wts = point_edge(p, a, c)
        if not (ins and abs(pen) < qbuf):
            continue
        if pen < best and pen < np.float32(1e-4):
            best, best_s = pen, s
    return best_s


@pytest.mark.parametrize("seed", range(3))
def test_first_pass_picks_the_brute_force_segment(seed):
    """Two jittered blocks overlapping by 0.015: with an empty cache every work
    item's segment is the best one over all of the target's edges."""
    mat = fw.Elastic(1e4, 0.3, 10.0)
    a = _grid_body(6, (0.0, 0.0), 0.2, mat, jitter=0.3, seed=seed)
    b = _grid_body(6, (0.185, 0.03), 0.2, mat, jitter=0.3, seed=seed + 5)  # 0.015 overlap
    o, g, _ = _loop([a, b])
    o.flex.contacts(1)
    hits = 0
    for w in range(g.n_items):
        exp = _brute_best(o, g, w)
        assert o.flex.cache[w] == exp, w
        hits += exp >= 0
    assert hits > 5


def test_flat_penetration_force_closed_form():
    """One node 0.01 inside a long flat edge: force n k d on it, the reaction split
    by the projection parameter over the edge's two nodes."""
    mat = fw.Elastic(1e4, 0.3, 10.0)
    top = fw.FemDomain(fw.Mesh(2, np.array([[0, 1, 2]], np.int32),
                               np.array([[0.3, 0.49], [0.5, 0.7], [0.2, 0.7]], np.float32)), fw.SolidProp(mat), [])
    bottom = fw.FemDomain(fw.Mesh(2, np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                                  np.array([[-1.0, 0.0], [2.0, 0.0], [2.0, 0.5], [-1.0, 0.5]], np.float32)),
                          fw.SolidProp(mat), [])
    o, g, _ = _loop([top, bottom])
    o.flex.contacts(1)
    k = np.float32(g.item_penalty[0])
    off0, f0 = o.flex.contributions(0)
    assert off0[1] - off0[0] == 1 and off0[-1] == 1  # only node 0 of the top body
    pen = np.float32(0.49) - np.float32(0.5)  # edge (2.0, 0.5) -> (-1.0, 0.5): outward normal (0, 1)
    tf = f0[off0[0]]
    assert tf[0] == 0.0 and tf[1] == np.float32(np.float32(-k) * pen)
    off1, f1 = o.flex.contributions(1)
    assert off1[-1] == 2 and off1[3] - off1[2] == 1 and off1[4] - off1[3] == 1
    t = (2.0 - 0.3) / 3.0  # projection parameter from the edge's first node
    r2, r3 = f1[off1[2]], f1[off1[3]]
    assert abs(float(r2[1]) + float(tf[1]) * (1.0 - t)) < 1e-5 * float(tf[1])
    assert abs(float(r3[1]) + float(tf[1]) * t) < 1e-5 * float(tf[1])


@pytest.mark.parametrize("friction", [0.0, 0.4])
def test_contributions_sum_to_zero(friction):
    mat = fw.Elastic(1e4, 0.3, 10.0)
    a = _grid_body(7, (0.0, 0.0), 0.2, mat, jitter=0.2, seed=3, friction=friction,
                   initials=[fw.InitialVel("ALL", [0.3, -0.2])])
    b = _grid_body(7, (0.19, 0.02), 0.2, mat, jitter=0.2, seed=4, friction=friction)
    o, g, _ = _loop([a, b])
    o.flex.contacts(1)
    tot = np.zeros(2, np.float64)
    scale = 0.0
    n = 0
    for k in range(2):
        off, f = o.flex.contributions(k)
        tot += f.astype(np.float64).sum(axis=0)
        scale = max(scale, float(np.abs(f).max(initial=0.0)))
        n += f.shape[0]
    assert n > 10
    assert np.all(np.abs(tot) <= 4 * n * scale * 2.0 ** -23), tot


def test_cache_keeps_its_segment_inside_the_margin_box():
    """Once a node has a cached edge it keeps it while it stays inside the edge's
    box widened by the margin, even where a fresh query would pick a deeper edge
    (flatworld/mixedcontact.py:942-967, 969-1001)."""
    mat = fw.Elastic(1e4, 0.3, 10.0)
    bottom = _grid_body(2, (0.0, 0.0), 1.0, mat)  # the unit square, edges of length 0.5
    probe = fw.FemDomain(fw.Mesh(2, np.array([[0, 1, 2]], np.int32),
                                 np.array([[0.6, 0.999], [0.7, 1.2], [0.5, 1.2]], np.float32)), fw.SolidProp(mat), [])
    o, g, _ = _loop([probe, bottom])
    w0 = [w for w in range(g.n_items) if g.item_body[w] == 0 and g.item_node[w] == 0][0]
    o.flex.contacts(1)
    top_edge = int(o.flex.cache[w0])
    _, a, c = _seg_xy(o, g, top_edge)
    assert a[1] == 1.0 and c[1] == 1.0 and min(a[0], c[0]) == 0.5  # the top edge over x in [0.5, 1]
    # into the corner: the right edge (x = 1) is now 0.02 deep, the top edge 0.002
    o.flex.bodies[0].x[0] = [0.98, 0.998]
    o.flex.contacts(2)
    assert o.flex.cache[w0] == top_edge
    off, f = o.flex.contributions(0)
    assert abs(float(f[off[0], 1]) - 0.002 * float(g.item_penalty[w0])) < 1e-3 * float(g.item_penalty[w0]) * 0.002
    # a fresh query takes the deepest edge
    o.flex.cache[w0] = -1
    o.flex.contacts(3)
    _, a, c = _seg_xy(o, g, int(o.flex.cache[w0]))
    assert a[0] == 1.0 and c[0] == 1.0


# ------------------------------------------------------------------ whole loops


def test_free_collision_conserves_momentum():
    mat = fw.Elastic(5e3, 0.3, 20.0)
    a = _grid_body(6, (0.0, 0.0), 0.2, mat, initials=[fw.InitialVel("ALL", [0.4, 0.05])])
    b = _grid_body(5, (0.22, 0.02), 0.16, fw.Elastic(8e3, 0.25, 15.0), initials=[fw.InitialVel("ALL", [-0.2, 0.0])])
    o, g, dt = _loop([a, b])
    def momentum():
        return sum((bd.mass[:, None].astype(np.float64) * bd.v.astype(np.float64)).sum(axis=0) for bd in o.bodies)
    p0 = momentum()
    contact = 0
    for _ in range(600):
        o.substep(dt)
        contact += int(any(o.flex.contributions(k)[0][-1] > 0 for k in range(2)))
    p1 = momentum()
    assert contact >= 5
    assert np.all(np.isfinite(p1))
    assert np.allclose(p1, p0, rtol=0, atol=1e-4 * np.abs(p0).max())
    # they separated again: body a moves left of b's left edge
    assert o.bodies[0].v[:, 0].mean() < o.bodies[1].v[:, 0].mean()


def test_reference_two_circle_scenario_runs():
    """test2D/test_2Dfem_contact.py: two circles, gravity towards each other, 1.5 s."""
    msh = fw.FEMesher(2)
    prop = fw.SolidProp(fw.Elastic(E=2e4, nu=0.2, rho=40.0))
    d1 = fw.FemDomain(msh.createCircle([0.5, 0.5], 0.1), prop, [fw.Gravity([0, 10.0])])
    d2 = fw.FemDomain(msh.createCircle([0.5, 0.8], 0.1), prop, [fw.Gravity([0, -10.0])])
    o, g, dt = _loop([d1, d2])
    assert g.n_items == 32 and list(g.item_body[:16]) == [0] * 16 and list(g.item_target[:16]) == [1] * 16
    n = int(1.5 / dt)
    touched = 0
    for i in range(n):
        o.substep(dt)
        touched += int(o.flex.contributions(0)[0][-1] > 0)
    assert touched > 10
    for b in o.bodies:
        assert np.all(np.isfinite(b.x)) and np.abs(b.x).max() < 10.0


def test_work_list_order_and_scalars():
    """flatworld/mixedcontact.py:1879-1911: per pair, domain2's boundary nodes against
    domain1, then domain1's against domain2; penalty min of both sides, friction max."""
    m1, m2, m3 = fw.Elastic(1e4, 0.3, 1.0), fw.Elastic(3e4, 0.45, 1.0), fw.Elastic(2e4, 0.2, 1.0)
    b = [_grid_body(2, (0, 0), 1.0, m1, friction=0.1), _grid_body(3, (2, 0), 1.0, m2, friction=0.3),
         _grid_body(2, (4, 0), 1.0, m3)]
    cs = [ContactFlexFlex(b[1], b[0], ContactType.FLEXFLEX), ContactFlexFlex(b[2], b[0], ContactType.FLEXFLEX),
          ContactFlexFlex(b[2], b[1], ContactType.FLEXFLEX)]
    g = build_flex_group(b, cs, True, sum(x.nnodes for x in b))
    rows = []
    for q, c in enumerate(cs):
        d1, d2 = b.index(c.domain1), b.index(c.domain2)
        for src, dst in ((d2, d1), (d1, d2)):
            for nd in b[src].mesh.boundaryNodes:
                rows.append((src, int(nd), dst, q))
    assert list(zip(g.item_body.tolist(), g.item_node.tolist(), g.item_target.tolist(), g.item_pair.tolist())) == rows
    K = lambda m: m.E / (3.0 * (1.0 - 2.0 * min(m.nu, 0.4999)))  # noqa: E731
    assert g.item_penalty[0] == np.float32(min(2 * K(m1), 2 * K(m2)))
    assert g.item_friction[0] == np.float32(0.3)
    assert g.pair_body.tolist() == [[1, 0], [2, 0], [2, 1]]
    assert g.max_pairs == 9 * max(g.n_segs + 1024, max(sum(x.nnodes for x in b) + 1024, 2048))
    spec = add_flex_record(pack_domain(b[0], []), g.flex_cap(0))
    assert [gr.kind for gr in spec.grounds] == [GROUND_FLEX]
    assert g.flex_cap(0) == int(np.sum(g.item_body == 0) + 2 * np.sum(g.item_target == 0))


# ------------------------------------------------------------------ tied pairs


def _tie_blocks(seed, gap=0.002, tied=((1, 0),)):
    mat = fw.Elastic(1e4, 0.3, 10.0)
    a = _grid_body(6, (0.0, 0.0), 0.2, mat, jitter=0.2, seed=seed)
    b = _grid_body(5, (0.2 + gap, 0.03), 0.18, mat, jitter=0.2, seed=seed + 7)
    return _loop([a, b], tied=tied)


@pytest.mark.parametrize("seed", range(3))
def test_tie_init_resolves_the_nearest_edge(seed):
    o, g, _ = _tie_blocks(seed)
    assert g.n_tied == g.n_items
    o.flex.ties_init()
    seg, state = o.flex.ties
    qbuf = np.float32(max(max(1e-4, float(max(o.flex.grid["gs"]))), 0.001))
    resolved = 0
    for w in range(g.n_items):
        b, nid, t = int(g.item_body[w]), int(g.item_node[w]), int(g.item_target[w])
        p = o.flex.bodies[b].x[nid]
        cands = []
        for s in range(g.body_seg_off[t], g.body_seg_off[t + 1]):
            _, a, c = _seg_xy(o, g, s)
            ins, pen, n, wts = point_edge(p, a, c)
            if ins and abs(pen) < qbuf:
                cands.append((abs(float(pen)), s, pen, wts))
        if not cands:
            assert seg[w] == -1 and np.all(state[w] == 0.0), w
            continue
        resolved += 1
        best = min(x[0] for x in cands)
        hit = [x for x in cands if x[1] == seg[w]]
        assert hit and hit[0][0] == best, w
        _, _, pen, wts = hit[0]
        assert np.array_equal(state[w, :2], wts)
        # the tie normal is the contact normal turned by pi: gap = -penetration
        assert abs(float(state[w, 2]) + float(pen)) <= 2e-7 * (1.0 + abs(float(pen))) + 1e-7, w
    assert resolved > 10


@pytest.mark.parametrize("seed", range(2))
def test_tie_force_is_a_spring_to_the_resolved_point(seed):
    o, g, _ = _tie_blocks(seed)
    o.flex.ties_init()
    seg, _ = o.flex.ties
    k = float(g.item_penalty[0])
    L = 0.4  # coordinates are below 0.4: rounding of a point is about L * 2^-24
    o.flex.contacts(1)
    for b in range(2):
        off, f = o.flex.contributions(b)
        assert off[-1] > 0
        assert np.abs(f).max() <= k * 8 * L * 2.0 ** -24
    # shift body 0 rigidly: its tied nodes feel -k d, body 1's tied nodes +k d
    d = np.array([1.5e-3, -1e-3], np.float32)
    o.flex.bodies[0].x[:] += d
    o.flex.contacts(2)
    for w in range(g.n_items):
        if seg[w] < 0:
            continue
        b = int(g.item_body[w])
        off, f = o.flex.contributions(b)
        nid = int(g.item_node[w])
        # the node's own entry: its first one that is not a reaction; items of this
        # pair are the node's only contributions, so the entry is unique per item
        own = f[off[nid]:off[nid + 1]]
        exp = -k * d if b == 0 else k * d
        assert np.any(np.all(np.abs(own - exp) <= 1e-3 * k * float(np.abs(d).max()), axis=1)), w
    tot = sum(o.flex.contributions(b)[1].astype(np.float64).sum(axis=0) for b in range(2))
    assert np.all(np.abs(tot) <= 1e-5 * k * float(np.abs(d).max()) * g.n_items)


def test_tied_items_never_slide():
    """A tied pair overlapping by more than the sliding contact would accept still
    gives only tie springs, and the cache stays empty."""
    o, g, _ = _tie_blocks(0, gap=-0.01)
    o.flex.ties_init()
    o.flex.contacts(1)
    assert np.all(o.flex.cache == -1)
    tot = max(float(np.abs(o.flex.contributions(b)[1]).max(initial=0.0)) for b in range(2))
    assert tot <= float(g.item_penalty[0]) * 8 * 0.4 * 2.0 ** -24
    # the same overlap untied slides: contacts push the bodies apart
    o2, g2, _ = _tie_blocks(0, gap=-0.01, tied=())
    o2.flex.contacts(1)
    assert np.any(o2.flex.cache >= 0)
    assert float(np.abs(o2.flex.contributions(0)[1]).max()) > 1e-3 * float(g2.item_penalty[0]) * 0.01


def test_ties_come_before_contacts_in_each_node():
    """Body 1 is tied to body 0 on its left side and pressed by body 2 on its top:
    its top-left nodes receive both; every node lists the tie entries (shifted
    body 0: forces along x) before the contact entries (flat top: forces along y)."""
    mat = fw.Elastic(1e4, 0.3, 10.0)
    b0 = _grid_body(4, (0.0, 0.0), 0.2, mat)
    b1 = _grid_body(4, (0.2, 0.0), 0.2, mat)
    b2 = _grid_body(4, (0.19, 0.195), 0.2, mat)  # 0.005 into body 1's top
    o, g, _ = _loop([b0, b1, b2], tied=[(1, 0)])
    assert 0 < g.n_tied < g.n_items
    o.flex.ties_init()
    o.flex.bodies[0].x[:] += np.array([1e-3, 0.0], np.float32)
    o.flex.contacts(1)
    off, f = o.flex.contributions(1)
    both = 0
    for n in range(len(off) - 1):
        e = f[off[n]:off[n + 1]]
        kinds = [0 if abs(x[0]) > abs(x[1]) else 1 for x in e if np.any(x != 0.0)]
        assert kinds == sorted(kinds), (n, e)
        both += (0 in kinds and 1 in kinds)
    assert both >= 1


def test_tied_free_bodies_stay_together():
    """Two free bodies side by side, one launched away from the other: tied, the
    interface stays closed and the total momentum is kept; untied, it opens.

    The reference's tie is not conservative: its query buffer (at least the cell
    size) also ties nodes a cell away from the interface with a nonzero normal
    gap, and the force on such a node and its reactions on the edge form a couple
    (gap n x F) under shear, so the tied pair's oscillation grows.  The check stays
    within the first 80 half-steps, before that growth dominates."""
    mat = fw.Elastic(5e3, 0.3, 20.0)
    out = {}
    for tied in ([(1, 0)], []):
        a = _grid_body(5, (0.0, 0.0), 0.2, mat)
        b = _grid_body(5, (0.2, 0.0), 0.2, mat, initials=[fw.InitialVel("ALL", [0.5, 0.0])])
        o, g, dt = _loop([a, b], tied=tied)
        def momentum():
            return sum((bd.mass[:, None].astype(np.float64) * bd.v.astype(np.float64)).sum(axis=0) for bd in o.bodies)
        p0 = momentum()
        # coincident sides: each node is tied both ways, which doubles the stiffness
        # the contact's stable step assumes, so the loop runs at half of it
        for _ in range(80):
            o.substep(0.5 * dt)
        assert o.ties_done == bool(tied)
        p1 = momentum()
        assert np.all(np.isfinite(p1))
        assert np.allclose(p1, p0, rtol=0, atol=1e-6 * np.abs(p0).max())
        out[bool(tied)] = o.bodies[1].x[:, 0].min() - o.bodies[0].x[:, 0].max()
    assert abs(out[True]) < 1e-3 and out[False] > 5e-3, out
