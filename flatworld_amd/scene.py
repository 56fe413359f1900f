"""Packed scene description: the flat arrays handed across the C-ABI.

A *scene* is one FEM body (one ``FemDomain``) or one spring network (one
``SpringMassDomain``) with the ordered list of grounds it touches.  The packing
mirrors what the reference manager and contact builder upload to the device:

* node/element arrays and per-element material
  (flatworld/femspringmanager.py:706-813; springs 814-870),
* BC bits ORed per node with the last-written value vector, then initial
  velocities (flatworld/femspringmanager.py:660-704),
* one ground record per (FEM, ground) contact pair with the float32 penalty,
  plane point/normal, height samples or occupancy grid, and the ground AABB the
  activation gate compares against (flatworld/mixedcontact.py:1794-1856,
  2131-2261; flatworld/explicitloop.py:198-216).

Every float that reaches the device is rounded to float32 here exactly once, in
the same place the reference rounds it.
"""

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .definitions import (
    MAX_BC_VALUES,
    NODE_BC_MASK,
    NODE_BCV_SHIFT,
    NODE_BOUNDARY_BIT,
    DomainType,
    MaterialType,
)

GROUND_PLANE = 1
GROUND_HEIGHTFIELD = 2
GROUND_VOXEL = 3
# the scene's FEM-FEM contact contributions (FW_GROUND_FLEX): filled by its
# contact group before each substep, placed after the plane records
GROUND_FLEX = 4
# a kinematic rigid primitive (FW_GROUND_RIGID): placed after the FEM-FEM record
# and before the height-field and voxel records, where the reference launches its
# FEM-rigid kernel (flatworld/mixedcontact.py:1697-1714)
GROUND_RIGID = 5

# material code of a spring-mass scene (FW_MAT_SPRING of include/flatworld_fem.h);
# the reference has no material type for springs: their elements carry matE and
# matDamping only (flatworld/femspringmanager.py:859-862)
MAT_SPRING = 4


@dataclass
class GroundSpec:
    """One (FEM body, ground) contact pair."""

    kind: int
    penalty: np.float32
    aabb_lo: np.ndarray  # float32[2], the ground AABB as stored for the gate
    aabb_hi: np.ndarray
    point: Optional[np.ndarray] = None  # plane
    normal: Optional[np.ndarray] = None
    heights: Optional[np.ndarray] = None  # height field
    hf_lb_x: np.float32 = np.float32(0.0)
    hf_ub_x: np.float32 = np.float32(0.0)
    reverse: int = 0
    occ: Optional[np.ndarray] = None  # voxel grid int32[nx, ny]
    vox_lb: Optional[np.ndarray] = None
    vox_dx: np.float32 = np.float32(0.0)
    vox_dz: np.float32 = np.float32(0.0)
    flex_cap: int = 0  # FEM-FEM record: contributions per substep
    # rigid record (flatworld_amd.rigid.RigidBodyDomain.record): the pose fields
    rigid_type: int = 0
    rigid_center: Optional[np.ndarray] = None
    rigid_rot: Optional[np.ndarray] = None  # float32[2, 2]
    rigid_prim: Optional[np.ndarray] = None
    rigid_radius: np.float32 = np.float32(0.0)
    rigid_vel: Optional[np.ndarray] = None
    rigid_omega: np.float32 = np.float32(0.0)
    rigid_friction: np.float32 = np.float32(0.0)
    rigid_tied: int = 0
    rigid_domain: object = None  # the RigidBodyDomain the record follows
    # mesh rigid record: the loop's RigidMeshSet (positions, edges, spatial hash), the
    # record's rigid id in it, substeps between full hash queries, and the loop step
    # before which the skip epochs reset (-1: none pending)
    mesh_set: object = None
    mesh_did: int = 0
    mesh_subcycle: int = 2
    mesh_reset_at: int = -1


@dataclass
class SceneSpec:
    """Flat arrays of one FEM body (or spring network) and its ground contacts.

    Spring scenes (``mat_type == MAT_SPRING``): ``conn`` is int32[E, 2], ``E`` the
    stiffness, ``rho`` the mass of every node, ``spring_c`` the damping coefficient
    and ``rest_len`` the float32[E] rest lengths; every node is a boundary node.
    """

    x0: np.ndarray  # float32[N, 2]
    v0: np.ndarray  # float32[N, 2]
    conn: np.ndarray  # int32[E, 3] (triangles) or int32[E, 2] (springs)
    bc: np.ndarray  # int32[N] BC bits
    bcv: np.ndarray  # float32[N, 2] last-written BC value
    bnodes: np.ndarray  # int32[Nb] sorted boundary nodes
    mat_type: int
    E: np.float32
    nu: np.float32
    rho: np.float32
    sigma_y: np.float32
    H: np.float32
    grounds: List[GroundSpec] = field(default_factory=list)
    dt: float = 0.0  # host float64; rounded to float32 on upload
    damping: float = 0.0
    horizon: int = 0  # substeps to run when served; 0 = unbounded
    rest_len: Optional[np.ndarray] = None  # springs: float32[E]
    spring_c: np.float32 = np.float32(0.0)  # springs: damping coefficient

    @property
    def n_nodes(self):
        return int(self.x0.shape[0])

    @property
    def n_elems(self):
        return int(self.conn.shape[0])

    def node_flags(self):
        """Device flag word per node and the BC value table it indexes.

        bits 0..8: BC bits; bit 15: boundary node; bits 16..31: row of the
        returned float32[K, 2] value table holding that node's BC vector.
        """
        bcv = np.ascontiguousarray(self.bcv, dtype=np.float32).reshape(-1, 2)
        # rows are deduplicated by their bit pattern (one uint64 per row), so -0.0 and
        # +0.0 stay distinct exactly as in the reference's per-node bc_values
        keys = bcv.view(np.uint64).reshape(-1)
        if keys.size == 0 or np.all(keys == keys[0]):
            table = bcv[:1] if keys.size else np.zeros((1, 2), np.float32)
            inv = np.zeros(keys.size, np.int64)
        else:
            uk, first, inv = np.unique(keys, return_index=True, return_inverse=True)
            table = bcv[first]
            inv = np.asarray(inv).reshape(-1)
        if table.shape[0] > MAX_BC_VALUES:
            raise ValueError(f"scene has {table.shape[0]} distinct BC vectors; at most {MAX_BC_VALUES} are supported")
        flags = (self.bc.astype(np.int64) & NODE_BC_MASK) | (inv.astype(np.int64) << NODE_BCV_SHIFT)
        flags[self.bnodes] |= NODE_BOUNDARY_BIT
        return flags.astype(np.uint32), np.ascontiguousarray(table, dtype=np.float32)


def pack_bcs(domain):
    """Per-node BC bits, values and initial velocity (flatworld/femspringmanager.py:660-704)."""
    n = domain.nnodes
    bc = np.zeros(n, dtype=np.int32)
    bcv = np.zeros((n, 2), dtype=np.float32)
    v0 = np.zeros((n, 2), dtype=np.float32)
    for b in domain.bcs:
        type_, nodes, values = b.processData()
        if values is None:
            values = [0.0] * domain.d
        idx = np.arange(n) if nodes is None else np.asarray(nodes, dtype=np.int64).reshape(-1)
        bc[idx] |= int(type_)
        bcv[idx, 0] = np.float32(float(values[0]))
        bcv[idx, 1] = np.float32(float(values[1]))
    for ic in domain.initials:
        vx, vy = float(ic.vel[0]), float(ic.vel[1])
        if getattr(ic, "is_all_nodes", False):
            v0[:, 0] = vx
            v0[:, 1] = vy
        else:
            nds = np.asarray(ic.nds, dtype=np.int64).reshape(-1)
            v0[nds, 0] = vx
            v0[nds, 1] = vy
    return bc, bcv, v0


def ground_spec(ground, contact):
    """Ground record for one FEM-vs-ground pair (flatworld/mixedcontact.py:1794-1856, 2131-2261)."""
    lo, hi = ground.getBBox()
    lo = np.asarray(lo, dtype=np.float32).reshape(2).copy()
    hi = np.asarray(hi, dtype=np.float32).reshape(2).copy()
    pen = np.float32(float(contact.penalty))
    if ground.type == DomainType.ANALYTICAL:
        return GroundSpec(
            GROUND_PLANE,
            pen,
            lo,
            hi,
            point=np.asarray(ground.point, dtype=np.float32).reshape(2).copy(),
            normal=np.asarray(ground.normal, dtype=np.float32).reshape(2).copy(),
        )
    if ground.type == DomainType.HEIGHTFIELD:
        return GroundSpec(
            GROUND_HEIGHTFIELD,
            pen,
            lo,
            hi,
            heights=np.ascontiguousarray(ground.height, dtype=np.float32),
            hf_lb_x=np.float32(float(ground.lb[0])),
            hf_ub_x=np.float32(float(ground.ub[0])),
            reverse=1 if ground.reverse else 0,
        )
    if ground.type == DomainType.VOXELMAP:
        return GroundSpec(
            GROUND_VOXEL,
            pen,
            lo,
            hi,
            occ=np.ascontiguousarray(ground.occ, dtype=np.int32),
            vox_lb=np.asarray(ground.lb, dtype=np.float32).reshape(2).copy(),
            vox_dx=np.float32(float(ground.dx)),
            vox_dz=np.float32(float(ground.dz)),
        )
    if ground.type == DomainType.RIGID:
        return rigid_spec(ground, contact)
    raise ValueError(f"not a ground domain: type {ground.type}")


def rigid_spec(rigid, contact):
    """Contact record of one FEM-vs-rigid pair at the rigid's current pose
    (flatworld/mixedcontact.py:1960-2043: penalty, the rigid's friction, tied flag)."""
    g = GroundSpec(GROUND_RIGID, np.float32(float(contact.penalty)), np.zeros(2, np.float32), np.zeros(2, np.float32),
                   rigid_friction=np.float32(float(rigid.friction)), rigid_tied=1 if getattr(contact, "tied", False) else 0,
                   rigid_domain=rigid, mesh_set=getattr(contact, "mesh_set", None),
                   mesh_did=int(getattr(contact, "mesh_did", 0)), mesh_subcycle=int(getattr(contact, "mesh_subcycle", 2)))
    update_rigid_spec(g)
    return g


def update_rigid_spec(g):
    """Refresh a rigid record from its domain's current pose and velocity."""
    for k, v in g.rigid_domain.record().items():
        setattr(g, k, v)
    return g


def pack_fem_domain(domain, grounds=(), dt=0.0, damping=0.0, horizon=0):
    """SceneSpec for one FemDomain and its (ground, contact) pairs in work-list order."""
    mesh = domain.mesh
    if mesh.d != 2:
        raise ValueError(f"only d=2 meshes are supported (got {mesh.d})")
    mat = domain.prop.getMaterial()
    x0 = np.ascontiguousarray(np.asarray(mesh.coords, dtype=np.float32)[:, :2])
    conn = np.ascontiguousarray(np.asarray(mesh.connectivity, dtype=np.int32)[:, :3])
    bc, bcv, v0 = pack_bcs(domain)
    bn = np.asarray(mesh.boundaryNodes).astype(np.int32) if mesh.numBoundNodes else np.zeros(0, np.int32)
    return SceneSpec(
        x0=x0,
        v0=v0,
        conn=conn,
        bc=bc,
        bcv=bcv,
        bnodes=np.ascontiguousarray(bn),
        mat_type=int(mat.type),
        E=np.float32(float(mat.E)),
        nu=np.float32(float(mat.nu)),
        rho=np.float32(float(mat.rho)),
        sigma_y=np.float32(float(getattr(mat, "sigma_y", 0.0))),
        H=np.float32(float(getattr(mat, "H", 0.0))),
        grounds=[ground_spec(g, c) for g, c in grounds],
        dt=float(dt),
        damping=float(damping),
        horizon=int(horizon),
    )


def pack_spring_domain(domain, grounds=(), dt=0.0, damping=0.0, horizon=0):
    """SceneSpec for one SpringMassDomain (flatworld/femspringmanager.py:814-870).

    Every node takes the domain mass, every spring the domain stiffness and damping
    coefficient; every node is a ground-contact work item and counts for the AABB
    (flatworld/mixedcontact.py:1815-1829, 2159-2167; femspringmanager.py:397-408).
    """
    x0 = np.ascontiguousarray(np.asarray(domain.coords, dtype=np.float32)[:, :2])
    conn = np.ascontiguousarray(np.asarray(domain.connectivity, dtype=np.int32)[:, :2])
    bc, bcv, v0 = pack_bcs(domain)
    return SceneSpec(
        x0=x0,
        v0=v0,
        conn=conn,
        bc=bc,
        bcv=bcv,
        bnodes=np.arange(domain.nnodes, dtype=np.int32),
        mat_type=MAT_SPRING,
        E=np.float32(float(domain.spring)),
        nu=np.float32(0.0),
        rho=np.float32(float(domain.mass)),
        sigma_y=np.float32(0.0),
        H=np.float32(0.0),
        grounds=[ground_spec(g, c) for g, c in grounds],
        dt=float(dt),
        damping=float(damping),
        horizon=int(horizon),
        rest_len=np.ascontiguousarray(domain.restLength, dtype=np.float32),
        spring_c=np.float32(float(domain.damping)),
    )


def pack_domain(domain, grounds=(), dt=0.0, damping=0.0, horizon=0):
    """SceneSpec of a FEM or spring-mass domain."""
    if domain.type == DomainType.SPRINGMASS:
        return pack_spring_domain(domain, grounds, dt, damping, horizon)
    return pack_fem_domain(domain, grounds, dt, damping, horizon)


def add_flex_record(spec, cap):
    """Insert the FEM-FEM contact record where the reference launches that kernel:
    after the plane kernel, before the rigid, height-field and voxel kernels
    (flatworld/mixedcontact.py:1654-1754)."""
    at = sum(1 for g in spec.grounds if g.kind == GROUND_PLANE)
    rec = GroundSpec(GROUND_FLEX, np.float32(0.0), np.zeros(2, np.float32), np.zeros(2, np.float32), flex_cap=int(cap))
    spec.grounds.insert(at, rec)
    return spec


def scene_stable_time(domain, contacts, has_plane):
    """Loop stable time for one FEM body (flatworld/explicitloop.py:145-170, 312-318)."""
    st = 1.0
    if has_plane:
        st = min(1.0 / 1000.0, st)
    st = min(st, min(domain.stableTime, 1.0))
    for c in contacts:
        st = min(st, c.stableTime)
    return st


MATERIAL_NAMES = {MaterialType.ELASTIC: "elastic", MaterialType.NEOHOOKEAN: "neohookean", MaterialType.MISES: "j2",
                  MAT_SPRING: "spring"}
