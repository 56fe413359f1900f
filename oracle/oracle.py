"""Python side of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
import this module, as the checker or the timed CPU baseline.  The product
package (flatworld_amd) never imports it.

Contents
* ``OracleScene``: ctypes driver of oracle/liboracle.so (fem_oracle.c), the
  serial float32 restatement of the substep, fed the same ``SceneSpec`` arrays
  the device receives.
* ``OracleFlexGroup`` / ``OracleLoop``: the FEM-FEM contact of a loop
  (flex_oracle.c) and a multi-body loop driver with the reference's hash
  schedule.
* ``boundary_edges_dict``: a straightforward restatement of
  flatworld/mesh.py:53-71 (a dict over element edges), used to pin the
  vectorized topology of flatworld_amd.mesh bit for bit.
"""

import ctypes
import os
import subprocess
from collections import defaultdict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class OraGround(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("penalty", ctypes.c_float),
        ("aabb_lo", ctypes.c_float * 2),
        ("aabb_hi", ctypes.c_float * 2),
        ("point", ctypes.c_float * 2),
        ("normal", ctypes.c_float * 2),
        ("heights", _f32p),
        ("hf_nx", ctypes.c_int32),
        ("hf_lb_x", ctypes.c_float),
        ("hf_ub_x", ctypes.c_float),
        ("reverse", ctypes.c_int32),
        ("occ", _i32p),
        ("vox_nx", ctypes.c_int32),
        ("vox_ny", ctypes.c_int32),
        ("vox_lb", ctypes.c_float * 2),
        ("vox_dx", ctypes.c_float),
        ("vox_dz", ctypes.c_float),
        ("flex_off", _i32p),
        ("flex_f", _f32p),
        ("rigid_type", ctypes.c_int32),
        ("rigid_center", ctypes.c_float * 2),
        ("rigid_rot", ctypes.c_float * 4),
        ("rigid_prim", ctypes.c_float * 2),
        ("rigid_radius", ctypes.c_float),
        ("rigid_vel", ctypes.c_float * 2),
        ("rigid_omega", ctypes.c_float),
        ("rigid_friction", ctypes.c_float),
        ("rigid_tied", ctypes.c_int32),
        ("mesh_coords", _f32p),
        ("mesh_conn", _i32p),
        ("mesh_elb", _f32p),
        ("mesh_eub", _f32p),
        ("mesh_margin", _f32p),
        ("mesh_reg", _i32p),
        ("mesh_cell", _i32p),
        ("mesh_sorted", _i32p),
        ("mesh_glb", ctypes.c_float * 2),
        ("mesh_gs", ctypes.c_float * 2),
        ("mesh_cn", ctypes.c_int32 * 2),
        ("mesh_tc", ctypes.c_int32),
        ("mesh_did", ctypes.c_int32),
        ("mesh_subcycle", ctypes.c_int32),
        ("mesh_reset_at", ctypes.c_int32),
        ("st_cache", _i32p),
        ("st_skip", _i32p),
        ("st_pelem", _i32p),
        ("st_pvalid", _i32p),
        ("st_pw", _f32p),
        ("st_pf", _f32p),
        ("st_ppen", _f32p),
        ("st_tres", _i32p),
        ("st_tw", _f32p),
    ]


class OraScene(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32),
        ("n_elems", ctypes.c_int32),
        ("n_bnodes", ctypes.c_int32),
        ("n_grounds", ctypes.c_int32),
        ("x", _f32p),
        ("v", _f32p),
        ("u", _f32p),
        ("fext", _f32p),
        ("mass", _f32p),
        ("conn", _i32p),
        ("bc", _i32p),
        ("bcv", _f32p),
        ("bnodes", _i32p),
        ("jinv", _f32p),
        ("area", _f32p),
        ("mat_type", ctypes.c_int32),
        ("E", ctypes.c_float),
        ("nu", ctypes.c_float),
        ("rho", ctypes.c_float),
        ("sigma_y", ctypes.c_float),
        ("H", ctypes.c_float),
        ("eps_p", _f32p),
        ("eqps", _f32p),
        ("stress", _f32p),
        ("strain", _f32p),
        ("yield_flag", _u8p),
        ("contact_flag", _u8p),
        ("aabb", ctypes.c_float * 4),
        ("grounds", ctypes.POINTER(OraGround)),
        ("rest_len", _f32p),
        ("spring_c", ctypes.c_float),
    ]


class OraFlex(ctypes.Structure):
    _fields_ = [
        ("n_bodies", ctypes.c_int32),
        ("n_segs", ctypes.c_int32),
        ("body_seg_off", _i32p),
        ("seg_body", _i32p),
        ("seg_node", _i32p),
        ("n_items", ctypes.c_int32),
        ("item_body", _i32p),
        ("item_node", _i32p),
        ("item_target", _i32p),
        ("item_pair", _i32p),
        ("item_penalty", _f32p),
        ("item_friction", _f32p),
        ("cache", _i32p),
        ("n_pairs", ctypes.c_int32),
        ("pair_body", _i32p),
        ("gated", ctypes.c_int32),
        ("item_tied", _i32p),
        ("tie_elem", _i32p),
        ("tie_w", _f32p),
        ("tie_gap", _f32p),
        ("max_cells", ctypes.c_int32),
        ("max_query", ctypes.c_int32),
        ("max_pairs", ctypes.c_int32),
        ("glb", ctypes.c_float * 2),
        ("gub", ctypes.c_float * 2),
        ("gs", ctypes.c_float * 2),
        ("cn", ctypes.c_int32 * 2),
        ("total_cells", ctypes.c_int32),
        ("n_hash_pairs", ctypes.c_int32),
        ("cell_start", _i32p),
        ("cell_end", _i32p),
        ("sorted_seg", _i32p),
        ("c_off", ctypes.POINTER(_i32p)),
        ("c_f", ctypes.POINTER(_f32p)),
    ]


_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_prepare.argtypes = [ctypes.POINTER(OraScene)]
        L.ora_substep.argtypes = [ctypes.POINTER(OraScene), ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_int32]
        L.ora_stress2d.argtypes = [ctypes.c_int32, ctypes.c_float, ctypes.c_float, _f32p, _f32p, _f32p]
        L.ora_mises2d.argtypes = [ctypes.c_float] * 4 + [_f32p, _f32p, ctypes.c_float, _f32p, _f32p, _f32p, _f32p]
        L.ora_mises2d.restype = ctypes.c_int32
        L.ora_ground_force.argtypes = [ctypes.POINTER(OraGround), _f32p, _f32p, ctypes.c_float, _f32p]
        L.ora_ground_force.restype = ctypes.c_int32
        L.ora_gate.argtypes = [_f32p, _f32p, ctypes.POINTER(OraGround), ctypes.c_int32]
        L.ora_gate.restype = ctypes.c_int32
        L.ora_update_bbox.argtypes = [ctypes.POINTER(OraScene)]
        L.ora_spring_force.argtypes = [ctypes.POINTER(OraScene), ctypes.c_int32, _f32p]
        L.ora_mass_scaling.argtypes = [ctypes.POINTER(OraScene), ctypes.c_float]
        L.ora_mass_scaling.restype = ctypes.c_float
        L.ora_rigid_detect.argtypes = [ctypes.POINTER(OraGround), _f32p, _f32p, _f32p]
        L.ora_rigid_detect.restype = ctypes.c_float
        L.ora_rigid_aabb.argtypes = [ctypes.c_int32, _f32p, _f32p, _f32p, ctypes.c_float, _f32p, _f32p]
        L.ora_log_f32.argtypes = [ctypes.c_float]
        L.ora_log_f32.restype = ctypes.c_float
        _scenes = ctypes.POINTER(ctypes.POINTER(OraScene))
        L.ora_point_edge.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.ora_point_edge.restype = ctypes.c_int32
        L.ora_flex_hash_build.argtypes = [ctypes.POINTER(OraFlex), _scenes]
        L.ora_flex_query.argtypes = [ctypes.POINTER(OraFlex), _f32p, ctypes.c_float, ctypes.c_int32, _i32p]
        L.ora_flex_query.restype = ctypes.c_int32
        L.ora_flex_contacts.argtypes = [ctypes.POINTER(OraFlex), _scenes, ctypes.c_int32]
        L.ora_flex_ties_init.argtypes = [ctypes.POINTER(OraFlex), _scenes]
        L.ora_gate_pair.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.c_int32]
        L.ora_gate_pair.restype = ctypes.c_int32
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def make_ground(gs, keep):
    """OraGround from a flatworld_amd.scene.GroundSpec; ``keep`` retains numpy buffers."""
    g = OraGround()
    g.kind = int(gs.kind)
    g.penalty = float(gs.penalty)
    g.aabb_lo[:] = [float(v) for v in gs.aabb_lo]
    g.aabb_hi[:] = [float(v) for v in gs.aabb_hi]
    if gs.kind == 1:
        g.point[:] = [float(v) for v in gs.point]
        g.normal[:] = [float(v) for v in gs.normal]
    elif gs.kind == 2:
        h = np.ascontiguousarray(gs.heights, dtype=np.float32)
        keep.append(h)
        g.heights = _p(h, _f32p)
        g.hf_nx = h.shape[0]
        g.hf_lb_x = float(gs.hf_lb_x)
        g.hf_ub_x = float(gs.hf_ub_x)
        g.reverse = int(gs.reverse)
    elif gs.kind == 3:
        occ = np.ascontiguousarray(gs.occ, dtype=np.int32)
        keep.append(occ)
        g.occ = _p(occ, _i32p)
        g.vox_nx, g.vox_ny = occ.shape
        g.vox_lb[:] = [float(v) for v in gs.vox_lb]
        g.vox_dx = float(gs.vox_dx)
        g.vox_dz = float(gs.vox_dz)
    elif gs.kind == 5:
        set_rigid(g, gs)
    # kind 4 (FEM-FEM contributions): linked to its group's lists by OracleFlexGroup
    return g


RIGID_MESH = 16


class MeshItemState:
    """Per-work-item state of one mesh rigid record (the arrays the reference's mesh
    kernels keep per item): initial values as _build_batched_contacts sets them
    (flatworld/mixedcontact.py:2052-2093)."""

    def __init__(self, n):
        self.cache = np.full((n, 2), -1, np.int32)
        self.skip = np.full(n, -1, np.int32)
        self.pelem = np.full(n, -1, np.int32)
        self.pvalid = np.zeros(n, np.int32)
        self.pw = np.zeros((n, 2), np.float32)
        self.pf = np.zeros((n, 2), np.float32)
        self.ppen = np.zeros(n, np.float32)
        self.tres = np.zeros(n, np.int32)
        self.tw = np.zeros((n, 2), np.float32)

    def attach(self, g):
        g.st_cache, g.st_skip, g.st_pelem = _p(self.cache, _i32p), _p(self.skip, _i32p), _p(self.pelem, _i32p)
        g.st_pvalid, g.st_tres = _p(self.pvalid, _i32p), _p(self.tres, _i32p)
        g.st_pw, g.st_pf, g.st_ppen, g.st_tw = (_p(a, _f32p) for a in (self.pw, self.pf, self.ppen, self.tw))


def set_mesh(g, gs, keep):
    """Point an OraGround at the current arrays of a mesh record's RigidMeshSet."""
    m = gs.mesh_set
    arrs = dict(coords=np.ascontiguousarray(m.coords, np.float32), conn=np.ascontiguousarray(m.conn, np.int32),
                elb=np.ascontiguousarray(m.elb, np.float32), eub=np.ascontiguousarray(m.eub, np.float32),
                margin=np.ascontiguousarray(m.margin, np.float32),
                reg=np.ascontiguousarray(m.reg if m.reg.size else np.zeros((1, 2), np.int32), np.int32),
                cell=np.ascontiguousarray(m.cell if m.cell.size else np.zeros((1, 2), np.int32), np.int32),
                sorted=np.ascontiguousarray(m.sorted if m.sorted.size else np.zeros(1, np.int32), np.int32))
    keep.append(arrs)
    g.mesh_coords, g.mesh_elb, g.mesh_eub, g.mesh_margin = (_p(arrs[k], _f32p) for k in ("coords", "elb", "eub", "margin"))
    g.mesh_conn, g.mesh_reg, g.mesh_cell, g.mesh_sorted = (_p(arrs[k], _i32p) for k in ("conn", "reg", "cell", "sorted"))
    g.mesh_glb[:] = [float(v) for v in m.glb]
    g.mesh_gs[:] = [float(v) for v in m.gs]
    g.mesh_cn[:] = [int(v) for v in m.cn]
    g.mesh_tc = int(m.tc)
    g.mesh_did = int(gs.mesh_did)
    g.mesh_subcycle = int(gs.mesh_subcycle)
    g.mesh_reset_at = int(gs.mesh_reset_at)


def set_rigid(g, gs):
    """Copy a rigid record's pose (flatworld_amd.scene.GroundSpec of kind 5) into an OraGround."""
    g.aabb_lo[:] = [float(v) for v in gs.aabb_lo]
    g.aabb_hi[:] = [float(v) for v in gs.aabb_hi]
    g.rigid_type = int(gs.rigid_type)
    g.rigid_center[:] = [float(v) for v in gs.rigid_center]
    g.rigid_rot[:] = [float(v) for v in np.asarray(gs.rigid_rot, np.float32).reshape(4)]
    g.rigid_prim[:] = [float(v) for v in gs.rigid_prim]
    g.rigid_radius = float(gs.rigid_radius)
    g.rigid_vel[:] = [float(v) for v in gs.rigid_vel]
    g.rigid_omega = float(gs.rigid_omega)
    g.rigid_friction = float(gs.rigid_friction)
    g.rigid_tied = int(gs.rigid_tied)


class OracleScene:
    """Serial CPU run of one packed scene (one FEM or spring-mass body + grounds)."""

    def __init__(self, spec):
        L = lib()
        self.spec = spec
        N, E = spec.n_nodes, spec.n_elems
        self.x = np.ascontiguousarray(spec.x0, dtype=np.float32).copy()
        self.v = np.ascontiguousarray(spec.v0, dtype=np.float32).copy()
        self.u = np.zeros((N, 2), np.float32)
        self.fext = np.zeros((N, 2), np.float32)
        self.mass = np.zeros(N, np.float32)
        self.conn = np.ascontiguousarray(spec.conn, dtype=np.int32)
        rl = getattr(spec, "rest_len", None)
        self.rest_len = np.ascontiguousarray(rl if rl is not None else np.zeros(E), dtype=np.float32)
        self.bc = np.ascontiguousarray(spec.bc, dtype=np.int32)
        self.bcv = np.ascontiguousarray(spec.bcv, dtype=np.float32)
        self.bnodes = np.ascontiguousarray(spec.bnodes, dtype=np.int32)
        self.jinv = np.zeros((E, 4), np.float32)
        self.area = np.zeros(E, np.float32)
        self.eps_p = np.zeros((E, 3), np.float32)
        self.eqps = np.zeros(E, np.float32)
        self.stress = np.zeros((E, 3), np.float32)
        self.strain = np.zeros((E, 3), np.float32)
        self.yield_flag = np.zeros(E, np.uint8)
        self.contact_flag = np.zeros(N, np.uint8)
        self._keep = []
        gs = [make_ground(g, self._keep) for g in spec.grounds]
        self._grounds = (OraGround * max(len(gs), 1))(*gs) if gs else (OraGround * 1)()
        # mesh rigid records: the mesh set's arrays and per-item state
        self.mesh_state = {}
        for gi, g in enumerate(spec.grounds):
            if g.kind == 5 and int(g.rigid_type) == RIGID_MESH:
                st = MeshItemState(self.bnodes.shape[0])
                st.attach(self._grounds[gi])
                set_mesh(self._grounds[gi], g, self._keep)
                self.mesh_state[gi] = st
        s = OraScene()
        s.n_nodes, s.n_elems, s.n_bnodes, s.n_grounds = N, E, self.bnodes.shape[0], len(gs)
        s.x, s.v, s.u, s.fext, s.mass = (_p(a, _f32p) for a in (self.x, self.v, self.u, self.fext, self.mass))
        s.conn, s.bc, s.bnodes = _p(self.conn, _i32p), _p(self.bc, _i32p), _p(self.bnodes, _i32p)
        s.bcv, s.jinv, s.area = _p(self.bcv, _f32p), _p(self.jinv, _f32p), _p(self.area, _f32p)
        s.mat_type = int(spec.mat_type)
        s.E, s.nu, s.rho = float(spec.E), float(spec.nu), float(spec.rho)
        s.sigma_y, s.H = float(spec.sigma_y), float(spec.H)
        s.eps_p, s.eqps = _p(self.eps_p, _f32p), _p(self.eqps, _f32p)
        s.stress, s.strain = _p(self.stress, _f32p), _p(self.strain, _f32p)
        s.yield_flag, s.contact_flag = _p(self.yield_flag, _u8p), _p(self.contact_flag, _u8p)
        s.grounds = ctypes.cast(self._grounds, ctypes.POINTER(OraGround))
        s.rest_len = _p(self.rest_len, _f32p)
        s.spring_c = float(getattr(spec, "spring_c", 0.0))
        self._s = s
        self._lib = L
        self.counter = 0
        L.ora_prepare(ctypes.byref(s))

    @property
    def aabb(self):
        return np.array(self._s.aabb[:], dtype=np.float32)

    def set_rigid(self, ground_index, gs):
        """Move rigid record ``ground_index`` to the pose of GroundSpec ``gs`` (a mesh
        record also takes its mesh set's current positions and hash)."""
        assert self._grounds[ground_index].kind == 5
        set_rigid(self._grounds[ground_index], gs)
        if int(gs.rigid_type) == RIGID_MESH:
            self._keep = [k for k in self._keep if not isinstance(k, dict)]
            for gi, g in enumerate(self.spec.grounds):
                if g.kind == 5 and int(g.rigid_type) == RIGID_MESH:
                    set_mesh(self._grounds[gi], g, self._keep)

    def warmup(self):
        """Zero-dt femStep run by setGlobalAABB (flatworld/femspringmanager.py:872-879)."""
        self._lib.ora_substep(ctypes.byref(self._s), 0.0, 0.0, 0, 0)

    def mass_scaling(self, dt_target):
        """FemSpringManager.applyMassScaling (flatworld/femspringmanager.py:141-201, 899-924), which
        ExplicitLoop runs right after the warm-up (flatworld/explicitloop.py:164-170); returns the
        added mass as the reference accumulates it."""
        return float(self._lib.ora_mass_scaling(ctypes.byref(self._s), float(np.float32(dt_target))))

    def spring_force(self, e):
        """Force of spring e on its first node, with a minus sign (flatworld/femspringmanager.py:283-312)."""
        f = np.zeros(2, np.float32)
        self._lib.ora_spring_force(ctypes.byref(self._s), int(e), _p(f, _f32p))
        return f

    def substep(self, dt, damping=0.0):
        """One ExplicitLoop.advance(); ``dt`` is rounded to float32 like a Warp float argument."""
        self._lib.ora_substep(ctypes.byref(self._s), float(np.float32(dt)), float(np.float32(damping)), self.counter, 1)
        self.counter += 1

    def run(self, n, dt=None, damping=None):
        dt = self.spec.dt if dt is None else dt
        damping = self.spec.damping if damping is None else damping
        for _ in range(int(n)):
            self.substep(dt, damping)


def point_edge(p, a, b):
    """pointToEdgeContact: (is_inside, penetration, normal, weights)."""
    arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (p, a, b)]
    pen = np.zeros(1, np.float32)
    n = np.zeros(2, np.float32)
    w = np.zeros(2, np.float32)
    ins = lib().ora_point_edge(*(_p(v, _f32p) for v in arrs), _p(pen, _f32p), _p(n, _f32p), _p(w, _f32p))
    return bool(ins), pen[0], n, w


class OracleFlexGroup:
    """Serial CPU run of one FEM-FEM contact group over OracleScene bodies.

    ``bodies`` are the loop's FEM bodies in manager order (OracleScene, whose
    specs already carry the FEM-FEM ground record where the body is in a pair)
    and ``group`` the ``flatworld_amd.flexcontact.FlexGroupSpec``."""

    def __init__(self, bodies, group):
        L = lib()
        self._lib = L
        self.bodies = bodies
        self.group = g = group
        B = g.n_bodies
        self._arr = keep = {}
        keep["seg_off"] = np.ascontiguousarray(g.body_seg_off, np.int32)
        keep["seg_body"] = np.repeat(np.arange(B, dtype=np.int32), np.diff(g.body_seg_off)).astype(np.int32)
        keep["seg_node"] = np.ascontiguousarray(g.seg_node, np.int32).reshape(-1)
        for k in ("item_body", "item_node", "item_target", "item_pair"):
            keep[k] = np.ascontiguousarray(getattr(g, k), np.int32)
        keep["item_penalty"] = np.ascontiguousarray(g.item_penalty, np.float32)
        keep["item_friction"] = np.ascontiguousarray(g.item_friction, np.float32)
        self.cache = np.full(g.n_items, -1, np.int32)
        keep["item_tied"] = np.ascontiguousarray(g.tied_flags(), np.int32)
        self.tie_elem = np.full(g.n_items, -1, np.int32)
        self.tie_w = np.zeros(2 * max(g.n_items, 1), np.float32)
        self.tie_gap = np.zeros(max(g.n_items, 1), np.float32)
        keep["pair_body"] = np.ascontiguousarray(g.pair_body, np.int32).reshape(-1)
        self.cell_start = np.zeros(g.max_cells, np.int32)
        self.cell_end = np.zeros(g.max_cells, np.int32)
        self.sorted_seg = np.zeros(g.max_pairs, np.int32)
        self.c_off = [np.zeros(b.spec.n_nodes + 1, np.int32) for b in bodies]
        self.c_f = [np.zeros(2 * max(g.flex_cap(k), 1), np.float32) for k in range(B)]
        F = OraFlex()
        F.n_bodies, F.n_segs = B, g.n_segs
        F.body_seg_off, F.seg_body, F.seg_node = (_p(keep[k], _i32p) for k in ("seg_off", "seg_body", "seg_node"))
        F.n_items = g.n_items
        F.item_body, F.item_node, F.item_target, F.item_pair = (
            _p(keep[k], _i32p) for k in ("item_body", "item_node", "item_target", "item_pair"))
        F.item_penalty, F.item_friction = _p(keep["item_penalty"], _f32p), _p(keep["item_friction"], _f32p)
        F.cache = _p(self.cache, _i32p)
        F.n_pairs = g.n_pairs
        F.pair_body = _p(keep["pair_body"], _i32p)
        F.gated = 1 if g.gated else 0
        F.item_tied = _p(keep["item_tied"], _i32p)
        F.tie_elem, F.tie_w, F.tie_gap = _p(self.tie_elem, _i32p), _p(self.tie_w, _f32p), _p(self.tie_gap, _f32p)
        F.max_cells, F.max_query, F.max_pairs = g.max_cells, g.max_query, g.max_pairs
        F.cell_start, F.cell_end, F.sorted_seg = (_p(a, _i32p) for a in (self.cell_start, self.cell_end, self.sorted_seg))
        self._offp = (_i32p * B)(*[_p(a, _i32p) for a in self.c_off])
        self._fp = (_f32p * B)(*[_p(a, _f32p) for a in self.c_f])
        F.c_off = ctypes.cast(self._offp, ctypes.POINTER(_i32p))
        F.c_f = ctypes.cast(self._fp, ctypes.POINTER(_f32p))
        self._F = F
        self._scenes = (ctypes.POINTER(OraScene) * B)(*[ctypes.pointer(b._s) for b in bodies])
        # link each body's FEM-FEM ground record to its contribution lists
        for k, body in enumerate(bodies):
            for gi in range(body._s.n_grounds):
                if body._grounds[gi].kind == 4:
                    body._grounds[gi].flex_off = _p(self.c_off[k], _i32p)
                    body._grounds[gi].flex_f = _p(self.c_f[k], _f32p)

    def rebuild(self):
        self._lib.ora_flex_hash_build(ctypes.byref(self._F), self._scenes)

    def ties_init(self):
        self._lib.ora_flex_ties_init(ctypes.byref(self._F), self._scenes)

    @property
    def ties(self):
        """Resolved tie edge (-1: none) and (weight 0, weight 1, normal gap) per work item."""
        W = self.group.n_items
        state = np.stack([self.tie_w[0:2 * W:2], self.tie_w[1:2 * W:2], self.tie_gap[:W]], axis=1)
        return self.tie_elem.copy(), state.astype(np.float32)

    def contacts(self, counter):
        self._lib.ora_flex_contacts(ctypes.byref(self._F), self._scenes, int(counter))

    def query(self, p, buffer, target):
        out = np.zeros(self.group.max_query, np.int32)
        pp = np.ascontiguousarray(p, dtype=np.float32)
        n = self._lib.ora_flex_query(ctypes.byref(self._F), _p(pp, _f32p), float(np.float32(buffer)), int(target),
                                     _p(out, _i32p))
        return out[:n].copy()

    @property
    def grid(self):
        F = self._F
        return dict(glb=np.array(F.glb[:], np.float32), gs=np.array(F.gs[:], np.float32),
                    cn=np.array(F.cn[:], np.int32), tc=int(F.total_cells), n_pairs=int(F.n_hash_pairs))

    def contributions(self, b):
        """Node CSR and float2 entries of body b from the last contact pass."""
        off = self.c_off[b]
        return off.copy(), self.c_f[b][: 2 * off[-1]].reshape(-1, 2).copy()


class OracleLoop:
    """Serial CPU run of an ExplicitLoop over several bodies: the zero-dt warm-up,
    the initial hash build, then per substep the hash schedule, the one-time tie
    resolution at the first substep, the FEM-FEM pass and every body's substep,
    with the loop's counter driving the gates (flatworld/explicitloop.py:383-385,
    473-555; femspringmanager.py:1035-1039; mixedcontact.py:1668-1695)."""

    def __init__(self, specs, group=None):
        self.bodies = [OracleScene(sp) for sp in specs]
        self.group = group
        self.flex = None
        if group is not None:
            fem = [b for b in self.bodies if b.spec.mat_type != 4]
            self.flex = OracleFlexGroup(fem, group)
        self.counter = 0
        self.elapsed = 0.0
        self.ties_done = False
        for b in self.bodies:
            b.warmup()
        if self.flex is not None:
            self.flex.rebuild()

    def substep(self, dt, damping=0.0):
        if self.flex is not None:
            g = self.group
            if g.n_tied < g.n_items or not self.ties_done:
                self.elapsed += float(dt)
                if self.elapsed >= g.rebuild_interval:
                    self.flex.rebuild()
                    self.elapsed = 0.0
            if g.n_tied > 0 and not self.ties_done:
                self.flex.ties_init()
                self.ties_done = True
            self.flex.contacts(self.counter)
        d32, c32 = float(np.float32(dt)), float(np.float32(damping))
        for b in self.bodies:
            b._lib.ora_substep(ctypes.byref(b._s), d32, c32, self.counter, 1)
        self.counter += 1

    def run(self, n, dt, damping=0.0):
        for _ in range(int(n)):
            self.substep(dt, damping)


def stress2d(mat_type, E, nu, F):
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float32).reshape(4))
    s = np.zeros(3, np.float32)
    e = np.zeros(3, np.float32)
    lib().ora_stress2d(int(mat_type), float(np.float32(E)), float(np.float32(nu)), _p(F, _f32p), _p(s, _f32p), _p(e, _f32p))
    return s, e


def log_f32(x):
    """The oracle's float32 natural log (the one the Neo-Hookean stress uses)."""
    return np.float32(lib().ora_log_f32(float(np.float32(x))))


def mises2d(E, nu, sy, H, F, eps_p_old, eqps_old):
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float32).reshape(4))
    ep = np.ascontiguousarray(np.asarray(eps_p_old, dtype=np.float32).reshape(3))
    s = np.zeros(3, np.float32)
    epn = np.zeros(3, np.float32)
    eqn = np.zeros(1, np.float32)
    st = np.zeros(3, np.float32)
    y = lib().ora_mises2d(
        float(np.float32(E)), float(np.float32(nu)), float(np.float32(sy)), float(np.float32(H)),
        _p(F, _f32p), _p(ep, _f32p), float(np.float32(eqps_old)),
        _p(s, _f32p), _p(epn, _f32p), _p(eqn, _f32p), _p(st, _f32p),
    )
    return s, epn, float(eqn[0]), st, bool(y)


def ground_force(gs, p, vel, dt):
    keep = []
    g = make_ground(gs, keep)
    p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(2))
    v = np.ascontiguousarray(np.asarray(vel, dtype=np.float32).reshape(2))
    f = np.zeros(2, np.float32)
    hit = lib().ora_ground_force(ctypes.byref(g), _p(p, _f32p), _p(v, _f32p), float(np.float32(dt)), _p(f, _f32p))
    return f, bool(hit)


def rigid_detect(gs, p):
    """detectPointToPrimitive of the oracle: (penetration, normal, contact point)."""
    keep = []
    g = make_ground(gs, keep)
    pp = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(2))
    n = np.zeros(2, np.float32)
    cp = np.zeros(2, np.float32)
    pen = lib().ora_rigid_detect(ctypes.byref(g), _p(pp, _f32p), _p(n, _f32p), _p(cp, _f32p))
    return np.float32(pen), n, cp


def rigid_aabb(rtype, center, rot, prim, radius):
    """The oracle's getBallBBox / getBoxBBox / getCapsuleBBox."""
    arrs = [np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1)) for v in (center, rot, prim)]
    lb = np.zeros(2, np.float32)
    ub = np.zeros(2, np.float32)
    lib().ora_rigid_aabb(int(rtype), *(_p(a, _f32p) for a in arrs), float(np.float32(radius)), _p(lb, _f32p),
                         _p(ub, _f32p))
    return lb, ub


def gate(fem_lb, fem_ub, gs, counter):
    keep = []
    g = make_ground(gs, keep)
    lb = np.ascontiguousarray(np.asarray(fem_lb, dtype=np.float32).reshape(2))
    ub = np.ascontiguousarray(np.asarray(fem_ub, dtype=np.float32).reshape(2))
    return bool(lib().ora_gate(_p(lb, _f32p), _p(ub, _f32p), ctypes.byref(g), int(counter)))


def boundary_edges_dict(conns):
    """Straight restatement of flatworld/mesh.py:53-71 (dict insertion order)."""
    edge_count = defaultdict(int)
    for cell in conns:
        for i in range(3):
            v1, v2 = int(cell[i]), int(cell[(i + 1) % 3])
            if (v2, v1) in edge_count:
                edge_count[(v2, v1)] += 1
            else:
                edge_count[(v1, v2)] += 1
    return np.array([e for e, c in edge_count.items() if c == 1])


def char_lengths_loop(conns, coords):
    """Per-edge float32 lengths with np.linalg.norm (flatworld/mesh.py:90-112)."""
    out = []
    for c in conns:
        for ia, ib in ((0, 1), (1, 2), (2, 0)):
            out.append(np.linalg.norm(coords[c[ia]] - coords[c[ib]]))
    return np.asarray(out, dtype=np.float32)
