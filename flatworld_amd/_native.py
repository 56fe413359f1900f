"""ctypes binding of libflatworld_fem.so (include/flatworld_fem.h).

The library is the only compute path of this package.  Loading fails loudly
when it is missing or when no GPU is visible; there is no CPU fallback.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflatworld_fem.so")

FW_OK = 0
FW_E_INVALID = -1
FW_E_CAPACITY = -2
FW_E_HIP = -3
FW_E_NOT_FOUND = -4
FW_E_UNSUPPORTED = -5

FW_STEP_DIAG = 1
FW_STEP_NO_CONTACT = 2
FW_STEP_PER_SCENE = 4

FIELD_X, FIELD_V, FIELD_U, FIELD_FORCE, FIELD_MASS = 1, 2, 3, 4, 5
FIELD_STRESS, FIELD_STRAIN, FIELD_EPS_P, FIELD_EQPS = 6, 7, 8, 9
FIELD_AREA, FIELD_JINV, FIELD_YIELD, FIELD_CONTACT, FIELD_AABB = 10, 11, 12, 13, 14
FLEX_CACHE, FLEX_GRID, FLEX_CELL_START, FLEX_CELL_END, FLEX_SORTED = 1, 2, 3, 4, 5
FLEX_TIE_SEG, FLEX_TIE_STATE = 6, 7
GROUND_FLEX = 4
GROUND_RIGID = 5

# (per_node, width, dtype)
FIELD_SHAPES = {
    FIELD_X: (True, 2, np.float32),
    FIELD_V: (True, 2, np.float32),
    FIELD_U: (True, 2, np.float32),
    FIELD_FORCE: (True, 2, np.float32),
    FIELD_MASS: (True, 1, np.float32),
    FIELD_CONTACT: (True, 1, np.uint8),
    FIELD_STRESS: (False, 3, np.float32),
    FIELD_STRAIN: (False, 3, np.float32),
    FIELD_EPS_P: (False, 3, np.float32),
    FIELD_EQPS: (False, 1, np.float32),
    FIELD_AREA: (False, 1, np.float32),
    FIELD_JINV: (False, 4, np.float32),
    FIELD_YIELD: (False, 1, np.uint8),
}

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class GroundDesc(ctypes.Structure):
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
        ("flex_cap", ctypes.c_int32),
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
        ("mesh_nodes", ctypes.c_int32),
        ("mesh_edges", ctypes.c_int32),
        ("mesh_nreg", ctypes.c_int32),
        ("mesh_tc", ctypes.c_int32),
        ("mesh_npairs", ctypes.c_int32),
        ("mesh_cap_reg", ctypes.c_int32),
        ("mesh_cap_cells", ctypes.c_int32),
        ("mesh_cap_pairs", ctypes.c_int32),
        ("mesh_glb", ctypes.c_float * 2),
        ("mesh_gs", ctypes.c_float * 2),
        ("mesh_cn", ctypes.c_int32 * 2),
        ("mesh_did", ctypes.c_int32),
        ("mesh_subcycle", ctypes.c_int32),
        ("mesh_reset_at", ctypes.c_int32),
        ("mesh_update", ctypes.c_int32),
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32),
        ("n_elems", ctypes.c_int32),
        ("n_grounds", ctypes.c_int32),
        ("n_bc_values", ctypes.c_int32),
        ("x0", _f32p),
        ("v0", _f32p),
        ("conn", _i32p),
        ("flags", _u32p),
        ("bc_values", _f32p),
        ("mat_type", ctypes.c_int32),
        ("E", ctypes.c_float),
        ("nu", ctypes.c_float),
        ("rho", ctypes.c_float),
        ("sigma_y", ctypes.c_float),
        ("H", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("damping", ctypes.c_float),
        ("horizon", ctypes.c_int32),
        ("track_disp", ctypes.c_int32),
        ("grounds", ctypes.POINTER(GroundDesc)),
        ("rest_len", _f32p),
        ("spring_c", ctypes.c_float),
    ]


class FlexDesc(ctypes.Structure):
    _fields_ = [
        ("n_bodies", ctypes.c_int32),
        ("scene_ids", _i32p),
        ("body_seg_off", _i32p),
        ("seg_node", _i32p),
        ("n_segs", ctypes.c_int32),
        ("n_items", ctypes.c_int32),
        ("item_body", _i32p),
        ("item_node", _i32p),
        ("item_target", _i32p),
        ("item_pair", _i32p),
        ("item_penalty", _f32p),
        ("item_friction", _f32p),
        ("n_pairs", ctypes.c_int32),
        ("pair_body", _i32p),
        ("body_bnodes_off", _i32p),
        ("body_bnodes", _i32p),
        ("gated", ctypes.c_int32),
        ("max_cells", ctypes.c_int32),
        ("max_query", ctypes.c_int32),
        ("max_pairs", ctypes.c_int32),
        ("item_tied", _i32p),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("tile_nodes", ctypes.c_int32),
        ("tile_slots", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("grid_per_cu", ctypes.c_int32),
    ]


class ArenaInfo(ctypes.Structure):
    _fields_ = [
        ("capacity_bytes", ctypes.c_int64),
        ("used_bytes", ctypes.c_int64),
        ("n_scenes", ctypes.c_int32),
        ("n_tiles", ctypes.c_int32),
        ("n_nodes", ctypes.c_int64),
        ("n_elems", ctypes.c_int64),
        ("n_elems_j2", ctypes.c_int64),
        ("global_step", ctypes.c_int64),
        ("algo_bytes_per_substep", ctypes.c_double),
        ("halo_slots", ctypes.c_int64),
        ("n_cu", ctypes.c_int32),
        ("tiles_per_cu", ctypes.c_int32),
    ]


# exported symbols and their signatures (restype, argtypes)
SIGNATURES = {
    "fw_last_error": (ctypes.c_char_p, []),
    "fw_abi_version": (ctypes.c_int32, []),
    "fw_arena_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_void_p)]),
    "fw_arena_destroy": (ctypes.c_int32, [ctypes.c_void_p]),
    "fw_arena_get_info": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ArenaInfo)]),
    "fw_scene_footprint": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_int64)]),
    "fw_scene_add": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_int32)]),
    "fw_scene_remove": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_scene_retire": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "fw_retire_wait": (ctypes.c_int32, [ctypes.c_void_p]),
    "fw_scene_progress": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "fw_step": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_uint32]),
    "fw_scene_reset_gate": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_step_graph": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32]),
    "fw_step_graph_prepare": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32]),
    "fw_step_noalloc": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_uint32]),
    "fw_sync": (ctypes.c_int32, [ctypes.c_void_p]),
    "fw_arena_track_busy": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_arena_busy": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "fw_wait_launches": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_scene_read": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]),
    "fw_scene_write": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]),
    "fw_tiling_check": (ctypes.c_int32, [ctypes.POINTER(SceneDesc), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "fw_tiling_cache_stats": (ctypes.c_int32, [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "fw_tiling_cache_clear": (ctypes.c_int32, []),
    "fw_bandwidth_probe": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "fw_bandwidth_ceiling": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_int32)]),
    "fw_time_steps": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "fw_scene_prepare": (ctypes.c_int32, [ctypes.POINTER(SceneDesc), ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_void_p)]),
    "fw_prepared_footprint": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "fw_scene_add_prepared": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "fw_prepared_free": (ctypes.c_int32, [ctypes.c_void_p]),
    "fw_prepared_stage": (ctypes.c_int32, [ctypes.c_void_p]),
    "fw_scene_mass_scaling": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float]),
    "fw_scene_set_rigid": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(GroundDesc)]),
    "fw_flex_add": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(FlexDesc), ctypes.POINTER(ctypes.c_int32)]),
    "fw_flex_rebuild": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_flex_init_ties": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_flex_schedule_rebuild": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "fw_flex_remove": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "fw_flex_read": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load the shared library (no GPU needed just to load it)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m flatworld_amd.build` (hipcc, gfx950). "
            "flatworld_amd has no CPU fallback."
        )
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    """Map a C status to the reference's exception types."""
    if rc == FW_OK:
        return
    msg = load().fw_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == FW_E_INVALID:
        raise ValueError(msg)
    if rc == FW_E_CAPACITY:
        raise RuntimeError(msg)
    if rc == FW_E_NOT_FOUND:
        raise KeyError(msg)
    if rc == FW_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise NativeError(msg)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def fill_rigid(d, g):
    """The rigid_* fields of a GroundDesc from a rigid GroundSpec (kind 5)."""
    d.rigid_type = int(g.rigid_type)
    d.rigid_center[:] = [float(v) for v in np.asarray(g.rigid_center, np.float32)]
    d.rigid_rot[:] = [float(v) for v in np.asarray(g.rigid_rot, np.float32).reshape(4)]
    d.rigid_prim[:] = [float(v) for v in np.asarray(g.rigid_prim, np.float32)]
    d.rigid_radius = float(g.rigid_radius)
    d.rigid_vel[:] = [float(v) for v in np.asarray(g.rigid_vel, np.float32)]
    d.rigid_omega = float(g.rigid_omega)
    d.rigid_friction = float(g.rigid_friction)
    d.rigid_tied = int(g.rigid_tied)


RIGID_MESH = 16
MESH_UPDATE_COORDS, MESH_UPDATE_EDGES, MESH_UPDATE_HASH = 1, 2, 4


def fill_mesh(d, g, keep, update=MESH_UPDATE_COORDS | MESH_UPDATE_EDGES | MESH_UPDATE_HASH):
    """The mesh_* fields of a GroundDesc from a mesh rigid GroundSpec: the current
    arrays of its RigidMeshSet and the capacities the set reserves."""
    m = g.mesh_set
    arrs = [np.ascontiguousarray(m.coords, np.float32), np.ascontiguousarray(m.conn, np.int32),
            np.ascontiguousarray(m.elb, np.float32), np.ascontiguousarray(m.eub, np.float32),
            np.ascontiguousarray(m.margin, np.float32),
            np.ascontiguousarray(m.reg if m.reg.size else np.zeros((1, 2), np.int32), np.int32),
            np.ascontiguousarray(m.cell if m.cell.size else np.zeros((1, 2), np.int32), np.int32),
            np.ascontiguousarray(m.sorted if m.sorted.size else np.zeros(1, np.int32), np.int32)]
    keep.extend(arrs)
    d.mesh_coords, d.mesh_elb, d.mesh_eub, d.mesh_margin = (_ptr(arrs[k], _f32p) for k in (0, 2, 3, 4))
    d.mesh_conn, d.mesh_reg, d.mesh_cell, d.mesh_sorted = (_ptr(arrs[k], _i32p) for k in (1, 5, 6, 7))
    d.mesh_nodes, d.mesh_edges = int(m.n_nodes), int(m.n_elems)
    d.mesh_nreg, d.mesh_tc, d.mesh_npairs = int(m.reg.shape[0]), int(m.tc), int(m.sorted.shape[0])
    d.mesh_cap_reg, d.mesh_cap_cells, d.mesh_cap_pairs = int(m.n_elems), int(m.max_cells), int(m.max_pairs)
    d.mesh_glb[:] = [float(v) for v in m.glb]
    d.mesh_gs[:] = [float(v) for v in m.gs]
    d.mesh_cn[:] = [int(v) for v in m.cn]
    d.mesh_did, d.mesh_subcycle, d.mesh_reset_at = int(g.mesh_did), int(g.mesh_subcycle), int(g.mesh_reset_at)
    d.mesh_update = int(update)


def scene_desc(spec, track_disp=True):
    """Build the C description of a SceneSpec; returns (desc, keepalive list)."""
    keep = []
    x0 = np.ascontiguousarray(spec.x0, dtype=np.float32)
    v0 = np.ascontiguousarray(spec.v0, dtype=np.float32)
    conn = np.ascontiguousarray(spec.conn, dtype=np.int32)
    flags, table = spec.node_flags()
    keep += [x0, v0, conn, flags, table]
    gs = []
    for g in spec.grounds:
        d = GroundDesc()
        d.kind = int(g.kind)
        d.penalty = float(g.penalty)
        d.aabb_lo[:] = [float(v) for v in g.aabb_lo]
        d.aabb_hi[:] = [float(v) for v in g.aabb_hi]
        if g.kind == 1:
            d.point[:] = [float(v) for v in g.point]
            d.normal[:] = [float(v) for v in g.normal]
        elif g.kind == 2:
            h = np.ascontiguousarray(g.heights, dtype=np.float32)
            keep.append(h)
            d.heights = _ptr(h, _f32p)
            d.hf_nx = int(h.shape[0])
            d.hf_lb_x = float(g.hf_lb_x)
            d.hf_ub_x = float(g.hf_ub_x)
            d.reverse = int(g.reverse)
        elif g.kind == 3:
            occ = np.ascontiguousarray(g.occ, dtype=np.int32)
            keep.append(occ)
            d.occ = _ptr(occ, _i32p)
            d.vox_nx, d.vox_ny = int(occ.shape[0]), int(occ.shape[1])
            d.vox_lb[:] = [float(v) for v in g.vox_lb]
            d.vox_dx = float(g.vox_dx)
            d.vox_dz = float(g.vox_dz)
        elif g.kind == GROUND_FLEX:
            d.flex_cap = int(g.flex_cap)
        elif g.kind == GROUND_RIGID:
            fill_rigid(d, g)
            if int(g.rigid_type) == RIGID_MESH:
                fill_mesh(d, g, keep)
        gs.append(d)
    garr = (GroundDesc * max(len(gs), 1))(*gs)
    keep.append(garr)
    sd = SceneDesc()
    sd.n_nodes = spec.n_nodes
    sd.n_elems = spec.n_elems
    sd.n_grounds = len(gs)
    sd.n_bc_values = int(table.shape[0])
    sd.x0 = _ptr(x0, _f32p)
    sd.v0 = _ptr(v0, _f32p)
    sd.conn = _ptr(conn, _i32p)
    sd.flags = _ptr(flags, _u32p)
    sd.bc_values = _ptr(table, _f32p)
    sd.mat_type = int(spec.mat_type)
    sd.E, sd.nu, sd.rho = float(spec.E), float(spec.nu), float(spec.rho)
    sd.sigma_y, sd.H = float(spec.sigma_y), float(spec.H)
    sd.dt = float(np.float32(spec.dt))
    sd.damping = float(np.float32(spec.damping))
    sd.horizon = int(spec.horizon)
    sd.track_disp = 1 if track_disp else 0
    sd.grounds = ctypes.cast(garr, ctypes.POINTER(GroundDesc))
    if getattr(spec, "rest_len", None) is not None:
        rl = np.ascontiguousarray(spec.rest_len, dtype=np.float32)
        if rl.shape != (spec.n_elems,):
            raise ValueError(f"rest_len must hold one length per spring ({spec.n_elems}), got shape {rl.shape}")
        keep.append(rl)
        sd.rest_len = _ptr(rl, _f32p)
        sd.spring_c = float(np.float32(spec.spring_c))
    return sd, keep


def tiling_check(spec, tile_nodes=512, tile_slots=1024):
    """Run the host-side tiler on a scene and verify it (no GPU)."""
    d, keep = scene_desc(spec)
    stats = (ctypes.c_int64 * 4)()
    check(load().fw_tiling_check(ctypes.byref(d), int(tile_nodes), int(tile_slots), stats), "fw_tiling_check")
    return dict(tiles=stats[0], halo_slots=stats[1], max_slots=stats[2], csr_entries=stats[3])




def tiling_cache_stats():
    """Tilings reused by connectivity in fw_scene_prepare: hits, misses, bytes kept, entries."""
    h, m, b, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
    check(load().fw_tiling_cache_stats(ctypes.byref(h), ctypes.byref(m), ctypes.byref(b), ctypes.byref(e)),
          "fw_tiling_cache_stats")
    return {"hits": h.value, "misses": m.value, "bytes": b.value, "entries": e.value}


def tiling_cache_clear():
    check(load().fw_tiling_cache_clear(), "fw_tiling_cache_clear")


def bandwidth_ceiling(device=0, nbytes=1 << 30, reps=10):
    """Fastest streaming copy over loads in flight per thread and grid size:
    dict(GBps, kernel_ms, grid_per_cu, ilp); GBps counts bytes read + written."""
    g, k = ctypes.c_double(), ctypes.c_double()
    gp, il = ctypes.c_int32(), ctypes.c_int32()
    check(load().fw_bandwidth_ceiling(int(device), int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(k),
                                      ctypes.byref(gp), ctypes.byref(il)), "fw_bandwidth_ceiling")
    return dict(GBps=g.value, kernel_ms=k.value, grid_per_cu=gp.value, ilp=il.value)


def bandwidth_probe(device=0, nbytes=1 << 30, reps=20):
    """Plain copy kernel: (read+write GB/s, average kernel ms)."""
    g = ctypes.c_double()
    k = ctypes.c_double()
    check(load().fw_bandwidth_probe(int(device), int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(k)), "fw_bandwidth_probe")
    return g.value, k.value


class Prepared:
    """A scene tiled and packed on the host for one kernel shape (fw_scene_prepare).

    Building one needs no GPU and releases the GIL inside the C call, so servers
    prepare arriving scenes on worker threads while the device runs.
    """

    def __init__(self, spec, tuning=None, track_disp=True):
        L = load()
        self._L = L
        t = Tuning(**{k: int(v) for k, v in (tuning or {}).items() if k in ("tile_nodes", "tile_slots", "block", "grid_per_cu")}) \
            if tuning else Tuning(512, 1024, 256, 0)
        d, keep = scene_desc(spec, track_disp)
        h = ctypes.c_void_p()
        check(L.fw_scene_prepare(ctypes.byref(d), ctypes.byref(t), ctypes.byref(h)), "fw_scene_prepare")
        self.h = h
        self.spec = spec
        b = ctypes.c_int64()
        check(L.fw_prepared_footprint(self.h, ctypes.byref(b)), "fw_prepared_footprint")
        self.footprint = int(b.value)

    def stage(self):
        """Copy the packed image into page-locked memory now (worker thread), so that
        admission only queues its upload."""
        check(self._L.fw_prepared_stage(self.h), "fw_prepared_stage")
        return self

    def close(self):
        if self.h:
            self._L.fw_prepared_free(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Arena:
    """Owner of one fw_arena: a packed HBM region holding many scenes."""

    def __init__(self, device=0, capacity_bytes=0, tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0):
        L = load()
        self._L = L
        t = Tuning(tile_nodes, tile_slots, block, grid_per_cu)
        h = ctypes.c_void_p()
        check(L.fw_arena_create(int(device), int(capacity_bytes), ctypes.byref(t), ctypes.byref(h)), "fw_arena_create")
        self.h = h
        self.tuning = dict(tile_nodes=tile_nodes, tile_slots=tile_slots, block=block, grid_per_cu=grid_per_cu)
        self._shapes = {}
        self._flex = {}
        self._retiring = []

    def close(self):
        if self.h:
            self._L.fw_arena_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def free_bytes(self):
        """Arena bytes not held by scenes or groups."""
        i = ArenaInfo()
        check(self._L.fw_arena_get_info(self.h, ctypes.byref(i)))
        return int(i.capacity_bytes - i.used_bytes)

    def info(self):
        i = ArenaInfo()
        check(self._L.fw_arena_get_info(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in ArenaInfo._fields_}

    def footprint(self, spec, track_disp=True):
        d, keep = scene_desc(spec, track_disp)
        b = ctypes.c_int64()
        check(self._L.fw_scene_footprint(self.h, ctypes.byref(d), ctypes.byref(b)), "fw_scene_footprint")
        return int(b.value)

    def add(self, spec, track_disp=True):
        d, keep = scene_desc(spec, track_disp)
        sid = ctypes.c_int32()
        check(self._L.fw_scene_add(self.h, ctypes.byref(d), ctypes.byref(sid)), "fw_scene_add")
        self._shapes[sid.value] = (spec.n_nodes, spec.n_elems)
        return sid.value

    def prepare(self, spec, track_disp=True):
        """Host half of add() for this arena's kernel shape; thread-safe, no GPU work."""
        return Prepared(spec, self.tuning, track_disp)

    def add_prepared(self, prep):
        """Device half of add(): copy a Prepared scene in (it can be added once)."""
        sid = ctypes.c_int32()
        check(self._L.fw_scene_add_prepared(self.h, prep.h, ctypes.byref(sid)), "fw_scene_add_prepared")
        self._shapes[sid.value] = (prep.spec.n_nodes, prep.spec.n_elems)
        prep.close()
        return sid.value

    def remove(self, sid):
        check(self._L.fw_scene_remove(self.h, int(sid)), "fw_scene_remove")
        self._shapes.pop(sid, None)

    def retire(self, sid, velocities=True):
        """Remove a scene, its final positions (and velocities) queued for the host behind
        the substeps launched so far; the returned arrays are filled by retire_wait()."""
        n = self._shapes[sid][0]
        x = np.empty((n, 2), np.float32)
        v = np.empty((n, 2), np.float32) if velocities else None
        check(self._L.fw_scene_retire(self.h, int(sid), x.ctypes.data,
                                      v.ctypes.data if v is not None else None), "fw_scene_retire")
        self._shapes.pop(sid, None)
        self._retiring.append((x, v))  # kept alive until the copies land
        return x, v

    def wait_launches(self, keep=0):
        """Block until at most ``keep`` tracked stepping calls are unfinished (busy tracking on)."""
        check(self._L.fw_wait_launches(self.h, int(keep)), "fw_wait_launches")

    def retire_wait(self):
        """Wait for the retirements queued since the last call and fill their arrays."""
        check(self._L.fw_retire_wait(self.h), "fw_retire_wait")
        self._retiring.clear()

    def progress(self, sid):
        d = ctypes.c_int64()
        f = ctypes.c_int32()
        check(self._L.fw_scene_progress(self.h, int(sid), ctypes.byref(d), ctypes.byref(f)))
        return int(d.value), bool(f.value)

    def step(self, n, dt_override=-1.0, flags=0, damping_override=-1.0):
        check(self._L.fw_step(self.h, int(n), float(dt_override), float(damping_override), int(flags)), "fw_step")

    def step_noalloc(self, n, dt_override=-1.0, flags=0, damping_override=-1.0):
        """step() through fw_step_noalloc: no allocation; the FEM-FEM contact scratch lives in
        a fixed per-GPU array.  After adding or removing scenes or groups, call step(0) first."""
        check(self._L.fw_step_noalloc(self.h, int(n), float(dt_override), float(damping_override), int(flags)),
              "fw_step_noalloc")

    def reset_gate(self, sid):
        check(self._L.fw_scene_reset_gate(self.h, int(sid)), "fw_scene_reset_gate")

    def mass_scaling(self, sid, dt_target):
        """Selective mass scaling of one scene at its current positions (asynchronous)."""
        check(self._L.fw_scene_mass_scaling(self.h, int(sid), float(np.float32(dt_target))), "fw_scene_mass_scaling")

    def set_rigid(self, sid, ground_index, g, mesh_update=MESH_UPDATE_COORDS | MESH_UPDATE_EDGES | MESH_UPDATE_HASH):
        """Move rigid record ``ground_index`` of scene ``sid`` to the pose of GroundSpec ``g``
        (stream-ordered: later substeps see it); a mesh record also takes the parts of
        its mesh set that ``mesh_update`` selects, and its header."""
        d = GroundDesc()
        d.kind = int(g.kind)
        d.penalty = float(g.penalty)
        d.aabb_lo[:] = [float(v) for v in g.aabb_lo]
        d.aabb_hi[:] = [float(v) for v in g.aabb_hi]
        fill_rigid(d, g)
        keep = []
        if int(g.rigid_type) == RIGID_MESH:
            fill_mesh(d, g, keep, mesh_update)
        check(self._L.fw_scene_set_rigid(self.h, int(sid), int(ground_index), ctypes.byref(d)), "fw_scene_set_rigid")

    def step_graph(self, n, flags=0):
        check(self._L.fw_step_graph(self.h, int(n), int(flags)), "fw_step_graph")

    def prepare_graph(self, n, flags=0):
        """Capture the graph step_graph(n, flags) replays, without running it."""
        check(self._L.fw_step_graph_prepare(self.h, int(n), int(flags)), "fw_step_graph_prepare")

    def sync(self):
        check(self._L.fw_sync(self.h), "fw_sync")

    def track_busy(self, enable=True):
        """Bracket every stepping call with HIP events (GPU-busy accounting)."""
        check(self._L.fw_arena_track_busy(self.h, 1 if enable else 0), "fw_arena_track_busy")

    def busy(self):
        """(GPU milliseconds of the bracketed stepping calls, number of calls) since tracking began."""
        ms = ctypes.c_double()
        calls = ctypes.c_int64()
        check(self._L.fw_arena_busy(self.h, ctypes.byref(ms), ctypes.byref(calls)), "fw_arena_busy")
        return ms.value, int(calls.value)

    def read(self, sid, field):
        if field == FIELD_AABB:
            out = np.zeros(4, np.float32)
            check(self._L.fw_scene_read(self.h, int(sid), field, out.ctypes.data, out.nbytes), "fw_scene_read")
            return out
        n_nodes, n_elems = self._shapes[sid]
        per_node, width, dt = FIELD_SHAPES[field]
        count = n_nodes if per_node else n_elems
        out = np.zeros((count, width) if width > 1 else (count,), dtype=dt)
        check(self._L.fw_scene_read(self.h, int(sid), int(field), out.ctypes.data, out.nbytes), "fw_scene_read")
        return out

    def write(self, sid, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        check(self._L.fw_scene_write(self.h, int(sid), int(field), a.ctypes.data, a.nbytes), "fw_scene_write")

    def flex_add(self, group, scene_ids):
        """Register a FEM-FEM contact group (a ``flexcontact.FlexGroupSpec``) over the given scenes."""
        g = group
        arrs = dict(
            scene_ids=np.ascontiguousarray(scene_ids, dtype=np.int32),
            body_seg_off=np.ascontiguousarray(g.body_seg_off, dtype=np.int32),
            seg_node=np.ascontiguousarray(g.seg_node, dtype=np.int32).reshape(-1),
            item_body=np.ascontiguousarray(g.item_body, dtype=np.int32),
            item_node=np.ascontiguousarray(g.item_node, dtype=np.int32),
            item_target=np.ascontiguousarray(g.item_target, dtype=np.int32),
            item_pair=np.ascontiguousarray(g.item_pair, dtype=np.int32),
            item_penalty=np.ascontiguousarray(g.item_penalty, dtype=np.float32),
            item_friction=np.ascontiguousarray(g.item_friction, dtype=np.float32),
            pair_body=np.ascontiguousarray(g.pair_body, dtype=np.int32).reshape(-1),
            body_bnodes_off=np.ascontiguousarray(g.body_bnodes_off, dtype=np.int32),
            body_bnodes=np.ascontiguousarray(g.body_bnodes, dtype=np.int32),
            item_tied=np.ascontiguousarray(g.tied_flags(), dtype=np.int32),
        )
        d = FlexDesc()
        d.n_bodies = int(len(scene_ids))
        d.n_segs = int(g.n_segs)
        d.n_items = int(g.n_items)
        d.n_pairs = int(g.n_pairs)
        for k, a in arrs.items():
            setattr(d, k, _ptr(a, _f32p if a.dtype == np.float32 else _i32p))
        d.gated = 1 if g.gated else 0
        d.max_cells, d.max_query, d.max_pairs = int(g.max_cells), int(g.max_query), int(g.max_pairs)
        gid = ctypes.c_int32()
        check(self._L.fw_flex_add(self.h, ctypes.byref(d), ctypes.byref(gid)), "fw_flex_add")
        self._flex[gid.value] = g
        return gid.value

    def flex_rebuild(self, gid):
        check(self._L.fw_flex_rebuild(self.h, int(gid)), "fw_flex_rebuild")

    def flex_schedule_rebuild(self, gid, at):
        """Rebuild group ``gid``'s hash (-1: every group) on the device before substep ``at``
        (0: the next one) of the following steps, plain or graph-replayed."""
        check(self._L.fw_flex_schedule_rebuild(self.h, int(gid), int(at)), "fw_flex_schedule_rebuild")

    def flex_init_ties(self, gid):
        check(self._L.fw_flex_init_ties(self.h, int(gid)), "fw_flex_init_ties")

    def flex_remove(self, gid):
        check(self._L.fw_flex_remove(self.h, int(gid)), "fw_flex_remove")
        self._flex.pop(gid, None)

    def flex_read(self, gid, field):
        g = self._flex[gid]
        if field == FLEX_GRID:
            raw = np.zeros(8, np.int32)
            check(self._L.fw_flex_read(self.h, int(gid), field, raw.ctypes.data, raw.nbytes), "fw_flex_read")
            return dict(glb=raw[:2].view(np.float32).copy(), gs=raw[2:4].view(np.float32).copy(),
                        cn=raw[4:6].copy(), tc=int(raw[6]), exact=bool(raw[7]))
        if field == FLEX_TIE_STATE:
            out = np.zeros((g.n_items, 4), np.float32)
            check(self._L.fw_flex_read(self.h, int(gid), field, out.ctypes.data, out.nbytes), "fw_flex_read")
            return out
        count = {FLEX_CACHE: g.n_items, FLEX_CELL_START: g.max_cells, FLEX_CELL_END: g.max_cells,
                 FLEX_SORTED: g.max_pairs, FLEX_TIE_SEG: g.n_items}[field]
        out = np.zeros(count, np.int32)
        check(self._L.fw_flex_read(self.h, int(gid), int(field), out.ctypes.data, out.nbytes), "fw_flex_read")
        return out

    def time_steps(self, n, mode=0):
        tot = ctypes.c_double()
        ker = ctypes.c_double()
        check(self._L.fw_time_steps(self.h, int(n), int(mode), ctypes.byref(tot), ctypes.byref(ker)), "fw_time_steps")
        return tot.value, ker.value
