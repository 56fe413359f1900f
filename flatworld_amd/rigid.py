"""Rigid primitives as kinematic contact partners of FEM bodies.

The reference's rigid bodies (flatworld/rigid.py, flatworld/rigiddomain.py) are
integrated by its RigidManager (flatworld/rigidmanager.py: gravity, PGS contact
between rigids, joints).  That solver is outside this package.  Here a rigid
primitive -- ball, box or capsule -- is *kinematic*: its pose (centre, angle) and
velocity are set by the caller (``set_pose`` / ``set_motion``, or advanced by
``advance`` at constant velocity), and FEM bodies and spring networks touch it
through the reference's FEM-rigid / spring-rigid penalty contact
(flatworld/mixedcontact.py:1025-1117), which runs on the device as a contact
record of the FEM or spring scene.  The reaction on the
rigid (the impulses the reference accumulates for its solver) is not computed.

A MeshRigid (a rigid polygon given by a triangle Mesh) touches FEM bodies
through the reference's mesh kernel (_batched_rigid_mesh_kernel_sh and the mesh
tie kernels, flatworld/mixedcontact.py:1119-1565): boundary nodes against the
rigid's boundary edges, found through the RigidManager's spatial hash of those
edges.  ``RigidMeshSet`` keeps that hash on the host, as the RigidManager does
(flatworld/rigidmanager.py:4065-4193, flatworld/spatialmanager.py), and places
the kinematic mesh at its pose (updateMeshCoords, rigidmanager.py:4581-4641).

The device reads, per rigid record, the quantities the reference's kernel reads
from the RigidManager arrays: centre (rigidParams[i, 0]), primary parameters
(rigidParams[i, 1]: box extents; capsule axis end lc relative to the centre,
flatworld/rigidmanager.py:2257-2279), radius, the cached rotation matrix
[[cos a, -sin a], [sin a, cos a]] (rigidmanager.py:247, utils.py:6-11), linear and
angular velocity (friction) and the rigid's AABB at the pose (getBallBBox /
getBoxBBox / getCapsuleBBox, utils.py:235-302, as _write_primitive_aabb stores it,
rigidmanager.py:765-793).  Every one is rounded to float32 here once, with the
AABB evaluated in float32 operation by operation as those functions do.
"""

import math

import numpy as np

from .definitions import (
    COLLISION_CATEGORY_ORDINARY_RIGID,
    COLLISION_MASK_ALL,
    DomainType,
    RigidType,
)

f32 = np.float32


def rotation_matrix(angle):
    """cal2DRotationMat (flatworld/utils.py:6-11) of a float32 angle: cos and sin in
    double, rounded to float32 once (Warp's float32 cos / sin are not reproduced)."""
    a = float(f32(angle))
    c, s = f32(math.cos(a)), f32(math.sin(a))
    return np.array([[c, -s], [s, c]], dtype=np.float32)


def _mv(r, v0, v1):
    """mat22 @ vec2 in float32: (r_i0 v0) + (r_i1 v1)."""
    return f32(r[0, 0] * v0 + r[0, 1] * v1), f32(r[1, 0] * v0 + r[1, 1] * v1)


def primitive_aabb(rtype, center, rot, prim, radius):
    """getBallBBox / getBoxBBox / getCapsuleBBox (flatworld/utils.py:235-302) in float32."""
    c0, c1 = f32(center[0]), f32(center[1])
    r = np.asarray(rot, dtype=np.float32).reshape(2, 2)
    e0, e1 = f32(prim[0]), f32(prim[1])
    rad = f32(radius)
    if rtype == RigidType.BALL:
        return np.array([c0 - rad, c1 - rad], np.float32), np.array([c0 + rad, c1 + rad], np.float32)
    if rtype == RigidType.BOX:
        xs, ys = [], []
        for k, (s0, s1) in enumerate(((1, 1), (1, -1), (1, 1), (-1, 1))):
            fac = f32(-0.5) if k == 0 else f32(0.5)
            m0, m1 = _mv(r, f32(s0) * e0, f32(s1) * e1)
            xs.append(f32(c0 + f32(fac * m0)))
            ys.append(f32(c1 + f32(fac * m1)))
        lo = np.array([min(min(xs[0], xs[1]), min(xs[2], xs[3])), min(min(ys[0], ys[1]), min(ys[2], ys[3]))], np.float32)
        hi = np.array([max(max(xs[0], xs[1]), max(xs[2], xs[3])), max(max(ys[0], ys[1]), max(ys[2], ys[3]))], np.float32)
        return lo, hi
    if rtype == RigidType.CAPSULE:
        alen = f32(np.sqrt(f32(f32(f32(0.0) + e0 * e0) + e1 * e1)))
        d0, d1 = _mv(r, f32(e0 / alen), f32(e1 / alen))
        p00, p01 = f32(c0 - f32(d0 * alen)), f32(c1 - f32(d1 * alen))
        p10, p11 = f32(c0 + f32(d0 * alen)), f32(c1 + f32(d1 * alen))
        perp0 = f32(rad * f32(np.sqrt(max(f32(0.0), f32(f32(1.0) - f32(d0 * d0))))))
        perp1 = f32(rad * f32(np.sqrt(max(f32(0.0), f32(f32(1.0) - f32(d1 * d1))))))
        ax0, ax1 = f32(rad * abs(d0)), f32(rad * abs(d1))
        lo = np.array([f32(f32(min(p00, p10) - perp0) - ax0), f32(f32(min(p01, p11) - perp1) - ax1)], np.float32)
        hi = np.array([f32(f32(max(p00, p10) + perp0) + ax0), f32(f32(max(p01, p11) + perp1) + ax1)], np.float32)
        return lo, hi
    raise NotImplementedError(f"rigid type {rtype:#b} has no primitive AABB (ball, box and capsule do; a mesh rigid's comes from its placed nodes)")


def _angle(angle, d):
    if angle is None:
        return 0.0
    a = np.asarray(angle, dtype=np.float64).reshape(-1)
    return float(a[0]) if a.size else 0.0


class RigidBase:
    """Common part of the primitives (flatworld/rigid.py:20-50)."""

    def __init__(self, d):
        if d != 2:
            raise ValueError(f"rigid primitives: only d=2 is supported (got d={d})")
        self.d = d
        self.numNodes = 1
        self.mesh = None

    def getRadius(self):
        return 0.0

    def getPrimary(self):
        return np.zeros((self.d,), dtype=np.float32)

    def getRefPoint(self):
        return self.origin


class BallRigid(RigidBase):
    """BallRigid(d, origin, radius, mass, angle=None, inertia=None) (flatworld/rigid.py:52-84)."""

    def __init__(self, d, origin, radius, mass, angle=None, inertia=None):
        super().__init__(d)
        self.origin = np.asarray(origin, dtype=np.float32)
        self.radius = radius
        self.mass = mass
        self.rtype = RigidType.BALL
        self.angle = _angle(angle, d)
        self.inertia_body = inertia if inertia is not None else 0.5 * mass * radius ** 2

    def getRadius(self):
        return self.radius


class BoxRigid(RigidBase):
    """BoxRigid(d, origin, ext, angle, mass, inertia=None) (flatworld/rigid.py:86-134): ``ext``
    is the box's full extent."""

    def __init__(self, d, origin, ext, angle, mass, inertia=None):
        super().__init__(d)
        self.ext = np.asarray(ext, dtype=np.float32)
        self.origin = np.asarray(origin, dtype=np.float32)
        self.angle = _angle(angle, d)
        self.mass = mass
        self.rtype = RigidType.BOX
        w, h = float(self.ext[0]), float(self.ext[1])
        self.inertia_body = inertia if inertia is not None else (1.0 / 12.0) * mass * (w * w + h * h)

    def getPrimary(self):
        return self.ext


class CapsuleRigid(RigidBase):
    """CapsuleRigid(d, lc, uc, angle, radius, mass, inertia=None) (flatworld/rigid.py:136-193):
    the axis from lc to uc, centre at their midpoint; the primary parameter is lc
    relative to the centre (flatworld/rigidmanager.py:2275-2279)."""

    def __init__(self, d, lc, uc, angle, radius, mass, inertia=None):
        super().__init__(d)
        self.lc = np.asarray(lc, dtype=np.float32)
        self.uc = np.asarray(uc, dtype=np.float32)
        self.radius = radius
        self.mass = mass
        self.rtype = RigidType.CAPSULE
        self.origin = 0.5 * (self.lc + self.uc)
        self.angle = _angle(angle, d)
        axis_len = float(np.linalg.norm(self.uc - self.lc))
        if axis_len <= 0.0:
            raise ValueError("CapsuleRigid: lc and uc must differ")
        self.inertia_body = inertia if inertia is not None else (1.0 / 12.0) * mass * axis_len ** 2 + 0.5 * mass * radius ** 2

    def getPrimary(self):
        return self.lc

    def getRadius(self):
        return self.radius


class MeshRigid(RigidBase):
    """MeshRigid(d, mesh, angle, mass, origin=None, inertia=None, transform=None)
    (flatworld/rigid.py:195-227): a rigid body given by a triangle mesh, touched
    through its boundary edges.  The reference point is ``origin`` or the mesh's
    node centroid; the boundary nodes are kept relative to it (the RigidManager's
    geometry pool, flatworld/rigidmanager.py:2014-2044) and placed at the pose by
    ``RigidBodyDomain.mesh_world``."""

    def __init__(self, d, mesh, angle, mass, origin=None, inertia=None, transform=None):
        if transform is not None:
            raise NotImplementedError("MeshRigid: a scaling / offset transform is not supported")
        if mesh.d != d:
            raise AssertionError("MeshRigid: mesh dimension differs from d")
        super().__init__(d)
        self.mesh = mesh
        self.mass = mass
        if origin is not None:
            self.origin = np.asarray(origin, dtype=np.float32)
        else:
            self.origin = np.asarray(mesh.getCenterPoint(), dtype=np.float32)
        self.rtype = RigidType.MESH
        self.angle = _angle(angle, d)
        coords = np.asarray(mesh.coords, dtype=np.float32)
        self.minSize = float(np.min(np.max(coords, axis=0) - np.min(coords, axis=0)))
        self.inertia_body = inertia
        bn = np.asarray(mesh.boundaryNodes, dtype=np.int64).reshape(-1)
        if bn.size == 0:
            raise ValueError("MeshRigid: the mesh has no boundary")
        # boundary node offsets from the reference point, in boundary-node order
        # (coord - ref, float32), and the boundary edges in local boundary indices
        self.local = (coords[bn] - self.origin.astype(np.float32)).astype(np.float32)
        where = {int(g): k for k, g in enumerate(bn)}
        be = np.asarray(mesh.boundaryElements, dtype=np.int64).reshape(-1, 2)
        self.edges = np.array([[where[int(a)], where[int(b)]] for a, b in be], dtype=np.int32).reshape(-1, 2)
        self.numNodes = int(bn.size)


class RigidBodyDomain:
    """RigidBodyDomain(rigid, bcs=[], considerContact=True, ...) (flatworld/rigiddomain.py:28-70),
    kinematic: the pose and velocity are the caller's (see the module docstring)."""

    def __init__(self, rigid, bcs=[], considerContact=True, considerGroundContact=True, initials=[], friction=0.0,
                 restitution=0.9, name="", visual_mesh=None, collision_shapes=None,
                 category_bits=COLLISION_CATEGORY_ORDINARY_RIGID, collide_bits=COLLISION_MASK_ALL):
        if rigid.rtype not in (RigidType.BALL, RigidType.BOX, RigidType.CAPSULE, RigidType.MESH):
            raise NotImplementedError(f"rigid type {rigid.rtype:#b} is not part of this package")
        self.rigid = rigid
        self.d = rigid.d
        self.bcs = bcs
        self.considerContact = considerContact
        self.considerGroundContact = considerGroundContact
        self.type = DomainType.RIGID
        self.ref = rigid.getRefPoint()
        self.nnodes = rigid.numNodes
        self.nelements = 0
        self.initials = initials
        self.friction = float(friction)
        self.restitution = float(restitution)
        self.name = name
        self.category_bits = category_bits
        self.collide_bits = collide_bits
        # kinematic state, in double on the host
        self.center = np.asarray(rigid.getRefPoint(), dtype=np.float64).copy()
        self.theta = float(rigid.angle)
        self.velocity = np.zeros(2, np.float64)
        self.omega = 0.0
        for ic in initials:
            self.velocity[:] = [float(ic.vel[0]), float(ic.vel[1])]
        prim = np.asarray(rigid.getPrimary(), dtype=np.float32)
        self.is_mesh = rigid.rtype == RigidType.MESH
        if rigid.rtype == RigidType.CAPSULE:
            prim = (np.asarray(rigid.lc, np.float32) - np.asarray(rigid.getRefPoint(), np.float32)).astype(np.float32)
        self.prim = prim.astype(np.float32)

    # ------------------------------------------------------------ kinematics
    def set_pose(self, center=None, theta=None):
        if center is not None:
            self.center = np.asarray(center, dtype=np.float64).reshape(2).copy()
        if theta is not None:
            self.theta = float(theta)

    def set_motion(self, velocity=None, omega=None):
        if velocity is not None:
            self.velocity = np.asarray(velocity, dtype=np.float64).reshape(2).copy()
        if omega is not None:
            self.omega = float(omega)

    def advance(self, dt):
        """Constant-velocity motion over dt (the caller's kinematics, not the reference's solver)."""
        self.center = self.center + self.velocity * float(dt)
        self.theta = self.theta + self.omega * float(dt)

    # ------------------------------------------------------------ mesh placement
    def mesh_world(self):
        """World positions of a mesh rigid's boundary nodes at the pose and their AABB:
        center + offset + R @ (local * scale) in float32 numpy with R from the float64
        cosine and sine of the float32 angle (updateMeshCoords,
        flatworld/rigidmanager.py:4581-4641; scale 1, offset 0)."""
        center = self.center.astype(np.float32)
        offset = np.zeros(2, np.float32)
        angle = float(np.float32(self.theta))
        c, s = float(np.cos(angle)), float(np.sin(angle))
        R = np.array([[c, -s], [s, c]], dtype=np.float32)
        scale = np.ones(2, np.float32)
        world = np.empty_like(self.rigid.local)
        lb = np.array([1e30, 1e30], dtype=np.float32)
        ub = np.array([-1e30, -1e30], dtype=np.float32)
        for i, lr in enumerate(self.rigid.local):
            local = np.array([float(lr[0]) * float(scale[0]), float(lr[1]) * float(scale[1])], dtype=np.float32)
            w = center + offset + R @ local
            world[i] = w
            lb = np.minimum(lb, w)
            ub = np.maximum(ub, w)
        return world, lb, ub

    def max_boundary_speed(self):
        """Upper bound of the boundary speed, |v| + |omega| r_max in float32, with r_max
        the primitive's bounding radius or the farthest boundary node of a mesh at its
        pose (get_max_linear_velocity, flatworld/rigidmanager.py:4666-4738)."""
        v = self.velocity.astype(np.float32).reshape(1, 2)
        center_speed = np.linalg.norm(v, axis=1)
        omega_speed = np.abs(np.array([self.omega], dtype=np.float32))
        t = self.rigid.rtype
        if t == RigidType.BALL:
            r = np.array([self.rigid.radius], dtype=np.float32)
        elif t == RigidType.BOX:
            r = (0.5 * np.linalg.norm(self.prim.reshape(1, 2), axis=1)).astype(np.float32)
        elif t == RigidType.CAPSULE:
            r = (np.linalg.norm(self.prim.reshape(1, 2), axis=1) + np.float32(self.rigid.radius)).astype(np.float32)
        else:
            world, _, _ = self.mesh_world()
            r = np.array([np.max(np.linalg.norm(world - self.center.astype(np.float32), axis=1))], dtype=np.float32)
        return float((center_speed + omega_speed * r).max())

    # ------------------------------------------------------------ device record
    def record(self):
        """float32 pose fields of the device contact record, and the AABB at the pose."""
        rot = rotation_matrix(self.theta)
        c = self.center.astype(np.float32)
        if self.is_mesh:
            _, lo, hi = self.mesh_world()
            return dict(rigid_type=int(self.rigid.rtype), rigid_center=c, rigid_rot=rot,
                        rigid_prim=np.zeros(2, np.float32), rigid_radius=f32(0.0),
                        rigid_vel=self.velocity.astype(np.float32), rigid_omega=f32(self.omega), aabb_lo=lo, aabb_hi=hi)
        lo, hi = primitive_aabb(self.rigid.rtype, c, rot, self.prim, float(self.rigid.getRadius()))
        return dict(rigid_type=int(self.rigid.rtype), rigid_center=c, rigid_rot=rot, rigid_prim=self.prim.copy(),
                    rigid_radius=f32(float(self.rigid.getRadius())), rigid_vel=self.velocity.astype(np.float32),
                    rigid_omega=f32(self.omega), aabb_lo=lo, aabb_hi=hi)

    def getBBox(self):
        r = self.record()
        return r["aabb_lo"], r["aabb_hi"]


# ---------------------------------------------------------------------- mesh rigids

RIGID_SH_REBUILD_INTERVAL = 1.0 / 500.0  # RigidManager._sh_mesh_rebuild_interval (rigidmanager.py:1790)
MESH_CONTACT_MARGIN_RATIO = 0.02  # RigidManager.meshRigidContactMarginRatio (rigidmanager.py:1500)


class RigidMeshSet:
    """The mesh rigids of one loop, placed at their poses, and the RigidManager's
    spatial hash over their boundary edges.

    Numbering follows the RigidManager: boundary nodes and edges of the meshes one
    after another (meshBoundaryNodeOffset / meshBoundaryElementOffset), edges as
    pairs of node indices; rigid ids are positions in the loop's list of rigid
    domains.  The hash is sized as the RigidManager sizes it
    (flatworld/rigidmanager.py:1797-1815: max(edges + 4096, 10000) elements, at
    most 100000 cells, nine pairs per element) and filled by
    populate_spatial_hash_filtered (rigidmanager.py:4095-4193) with
    SpatialHashManager's build (flatworld/spatialmanager.py:200-330, 612-700).

    Schedule (kinematic rigids; the reference's RigidManager.substep and
    maybe_rebuild_spatial_hash, rigidmanager.py:2102-2197, 4065-4091, with every
    loop substep a rigid step): ``initial_build`` fills the hash at the loop's start
    with the elements that meet the FEM contact nodes' AABB (explicitloop.py:386-396);
    after every substep ``rigid_step(dt)`` accumulates time and rebuilds, with every
    element, once 1/500 s has passed or the fastest boundary point may have moved
    half a cell.  ``refresh_elem_aabbs`` recomputes the edge boxes from the current
    positions (update_mesh_element_aabbs, rigidmanager.py:4195-4221) and keeps each
    edge's margin base: the reference writes ratio x (ub - |lb|) there, a vector
    minus a scalar that Warp's Python scope does not evaluate as a float.
    """

    def __init__(self, mesh_domains, all_rigids):
        self.meshes = list(mesh_domains)
        self.all_rigids = list(all_rigids)
        node_off, elem_off, conn = [0], [0], []
        for r in self.meshes:
            conn.append(r.rigid.edges.astype(np.int64) + node_off[-1])
            node_off.append(node_off[-1] + r.rigid.local.shape[0])
            elem_off.append(elem_off[-1] + r.rigid.edges.shape[0])
        self.node_off, self.elem_off = node_off, elem_off
        self.conn = np.ascontiguousarray(np.concatenate(conn).astype(np.int32)).reshape(-1, 2)
        self.n_nodes, self.n_elems = node_off[-1], elem_off[-1]
        self.max_elements = max(self.n_elems + 4096, 10000)
        self.max_cells = min(self.max_elements, 100000)
        self.max_pairs = 9 * self.max_elements
        self.coords = np.zeros((self.n_nodes, 2), np.float32)
        self.aabbs = [None] * len(self.meshes)
        self.elb = np.zeros((self.n_elems, 2), np.float32)
        self.eub = np.zeros((self.n_elems, 2), np.float32)
        self.margin = np.zeros(self.n_elems, np.float32)
        # RigidManager schedule state
        self.contact_margin = np.float32(0.0)  # _sh_contact_margin
        self.max_v = 0.0                       # _sh_mesh_max_v
        self.elapsed = 0.0                     # _sh_mesh_elapsed
        self.cell_size = 0.0                   # _sh_mesh_cell_size
        self.needs_rebuild = True
        # bumped when the positions, the edge boxes / margins or the hash change
        self.v_coords = self.v_edges = self.v_hash = 0
        self.place()
        self._empty_hash()

    def did(self, rigid_domain):
        """A rigid's id: its position among the loop's rigid domains (by identity, so
        that a deep copy of the set and its domains keeps working)."""
        return next(k for k, r in enumerate(self.all_rigids) if r is rigid_domain)

    # ------------------------------------------------------------ placement
    def place(self):
        """World boundary positions and AABB of every mesh at its current pose."""
        for k, r in enumerate(self.meshes):
            world, lb, ub = r.mesh_world()
            self.coords[self.node_off[k]:self.node_off[k + 1]] = world
            self.aabbs[k] = (lb, ub)
        self.v_coords += 1

    def refresh_elem_aabbs(self):
        """Edge AABBs from the current positions (update_mesh_element_aabbs)."""
        a = self.coords[self.conn[:, 0]]
        b = self.coords[self.conn[:, 1]]
        self.elb = np.minimum(np.minimum(np.full_like(a, 1e30), a), b).astype(np.float32)
        self.eub = np.maximum(np.maximum(np.full_like(a, -1e30), a), b).astype(np.float32)
        self.v_edges += 1

    # ------------------------------------------------------------ hash
    def _empty_hash(self):
        self.glb = np.zeros(2, np.float32)
        self.gs = np.ones(2, np.float32)
        self.cn = np.ones(2, np.int32)
        self.tc = 0
        self.reg = np.zeros((0, 2), np.int32)
        self.cell = np.zeros((0, 2), np.int32)
        self.sorted = np.zeros(0, np.int32)

    def populate(self, fem_lb=None, fem_ub=None):
        """populate_spatial_hash_filtered (rigidmanager.py:4095-4193) and build()."""
        if fem_lb is None or fem_ub is None:
            fem_lb = np.array([-1e9, -1e9], np.float32)
            fem_ub = np.array([1e9, 1e9], np.float32)
        fem_lb = np.asarray(fem_lb, np.float32).reshape(-1)
        fem_ub = np.asarray(fem_ub, np.float32).reshape(-1)
        contact_margin = float(self.contact_margin)
        velocity_buffer = 2.0 * self.max_v * RIGID_SH_REBUILD_INTERVAL
        global_lb = np.array([1e9, 1e9], dtype=np.float32)
        global_ub = np.array([-1e9, -1e9], dtype=np.float32)
        for k in range(len(self.meshes)):
            rlb, rub = self.aabbs[k]
            ilb = np.maximum(rlb, fem_lb) - contact_margin
            iub = np.minimum(rub, fem_ub) + contact_margin
            if np.all(ilb <= iub):
                global_lb = np.minimum(global_lb, ilb)
                global_ub = np.maximum(global_ub, iub)
        reg = []  # (rigid id, edge, lb, ub) in insertion order
        if global_lb[0] <= global_ub[0]:
            expand = float(np.max(global_ub - global_lb)) * 0.05 + contact_margin
            global_lb = global_lb - expand
            global_ub = global_ub + expand
            for k, r in enumerate(self.meshes):
                rlb, rub = self.aabbs[k]
                ilb = np.maximum(rlb, fem_lb) - contact_margin
                iub = np.minimum(rub, fem_ub) + contact_margin
                has_isect = bool(np.all(ilb <= iub))
                rid = self.did(r)
                for e in range(self.elem_off[k], self.elem_off[k + 1]):
                    lb = np.array([1e30, 1e30], dtype=np.float32)
                    ub = np.array([-1e30, -1e30], dtype=np.float32)
                    for j in range(2):
                        c = self.coords[self.conn[e, j]]
                        lb = np.minimum(lb, c[:2])
                        ub = np.maximum(ub, c[:2])
                    self.elb[e] = lb
                    self.eub[e] = ub
                    self.margin[e] = MESH_CONTACT_MARGIN_RATIO * float(np.linalg.norm(ub - lb)) + velocity_buffer
                    if has_isect and bool(np.all(ub >= ilb) and np.all(lb <= iub)):
                        reg.append((rid, e, lb, ub))
        else:
            global_lb = np.array([0.0, 0.0], dtype=np.float32)
            global_ub = np.array([1.0, 1.0], dtype=np.float32)
        self._build(reg, global_lb.astype(np.float32), global_ub.astype(np.float32), velocity_buffer)
        self.contact_margin = np.float32(contact_margin + velocity_buffer)
        self.v_edges += 1
        self.v_hash += 1

    def _build(self, reg, glb, gub, buffer):
        """SpatialHashManager.add_element / set_bounds / build in float32: element boxes
        widened by the buffer, cell size 1.5 x the smallest box diagonal (at least the
        extent over sqrt(max_cells), the correctly rounded square root standing for
        pow(max_cells, 0.5)), pairs in element order, a cell's entries in pair order."""
        f = np.float32
        buf = f(buffer)
        n = min(len(reg), self.max_elements)
        boxes = np.zeros((n, 4), np.float32)
        est = f(1e30)
        for i in range(n):
            _, _, lb, ub = reg[i]
            b = np.array([lb[0] - buf, lb[1] - buf, ub[0] + buf, ub[1] + buf], dtype=np.float32)
            boxes[i] = b
            dx, dy = f(b[2] - b[0]), f(b[3] - b[1])
            diag = f(np.sqrt(f(f(f(0.0) + f(dx * dx)) + f(dy * dy))))
            est = diag if diag < est else est
        cell = f(est * f(1.5))
        ex, ey = f(gub[0] - glb[0]), f(gub[1] - glb[1])
        mpd = max(f(np.sqrt(f(self.max_cells))), f(1.0))
        cell = max(cell, f(ex / mpd))
        cell = max(cell, f(ey / mpd))
        cs = f(cell + f(1e-6))
        nx = int(max(np.ceil(f(ex / cs)), f(1.0)))
        ny = int(max(np.ceil(f(ey / cs)), f(1.0)))
        self.glb = glb.copy()
        self.cn = np.array([nx, ny], np.int32)
        self.gs = np.array([f(ex / f(nx)), f(ey / f(ny))], np.float32)
        self.tc = int(min(nx * ny, self.max_cells))
        self.reg = np.array([(rid, e) for rid, e, _, _ in reg[:n]], dtype=np.int32).reshape(-1, 2)
        tc = self.tc
        cells, elems = [], []
        if n:
            cnx, cny = nx - 1, ny - 1
            npairs = 0
            for i in range(n):
                b = boxes[i]
                lx = int(np.floor(f(f(b[0] - glb[0]) / self.gs[0])))
                ly = int(np.floor(f(f(b[1] - glb[1]) / self.gs[1])))
                ux = int(np.floor(f(f(b[2] - glb[0]) / self.gs[0])))
                uy = int(np.floor(f(f(b[3] - glb[1]) / self.gs[1])))
                lx, ly = min(max(lx, 0), cnx), min(max(ly, 0), cny)
                ux, uy = min(max(ux, 0), cnx), min(max(uy, 0), cny)
                for ix in range(lx, ux + 1):
                    for iy in range(ly, uy + 1):
                        cid = ix + iy * nx
                        if cid < tc:
                            if npairs < self.max_pairs:
                                cells.append(cid)
                                elems.append(i)
                            npairs += 1
        cells = np.asarray(cells, np.int64)
        order = np.argsort(cells, kind="stable")
        self.sorted = np.asarray(elems, np.int32)[order] if len(elems) else np.zeros(0, np.int32)
        counts = np.bincount(cells, minlength=tc) if len(cells) else np.zeros(tc, np.int64)
        start = np.concatenate([[0], np.cumsum(counts)[:-1]]) if tc else np.zeros(0, np.int64)
        self.cell = np.stack([start, start + counts], axis=1).astype(np.int32) if tc else np.zeros((0, 2), np.int32)

    # ------------------------------------------------------------ schedule
    def fem_contact_aabb(self, node_sets):
        """AABB of the FEM nodes in contact with any mesh (_compute_fem_contact_aabb,
        mixedcontact.py:591-618): float32 min / max over (positions, node ids) pairs."""
        lb = np.array([1e9, 1e9], np.float32)
        ub = np.array([-1e9, -1e9], np.float32)
        for x, nodes in node_sets:
            if len(nodes):
                p = np.asarray(x, np.float32)[np.asarray(nodes, np.int64)]
                lb = np.minimum(lb, p.min(axis=0))
                ub = np.maximum(ub, p.max(axis=0))
        return lb, ub

    def initial_build(self, fem_lb, fem_ub):
        self.populate(fem_lb, fem_ub)

    def aabb_of(self, rigid_domain):
        return self.aabbs[next(k for k, r in enumerate(self.meshes) if r is rigid_domain)]

    def max_speed(self):
        return max((r.max_boundary_speed() for r in self.all_rigids), default=0.0)

    def rigid_step(self, dt):
        """After a substep: the kinematic rigids have moved; update positions, the
        rebuild triggers and the query margin, and rebuild when due (unfiltered, as
        maybe_rebuild_spatial_hash runs it with rigid-rigid contact on).  Returns
        True when the hash was rebuilt."""
        self.place()
        self.elapsed += float(dt)
        if self.elapsed >= RIGID_SH_REBUILD_INTERVAL:
            self.needs_rebuild = True
        if not self.needs_rebuild:
            if self.max_v * self.elapsed > 0.5 * self.cell_size:
                self.needs_rebuild = True
        self.contact_margin = np.float32(self.max_v * self.elapsed)
        if not self.needs_rebuild:
            return False
        self.populate()
        self.cell_size = float(float(np.max(self.gs)))
        self.max_v = self.max_speed()
        self.contact_margin = np.float32(self.cell_size)
        self.elapsed = 0.0
        self.needs_rebuild = False
        return True


class KinematicRigids:
    """Per-substep host schedule of a loop's kinematic rigids: the moving ones advance
    at constant velocity after every substep, and the mesh set follows the
    RigidManager's schedule around it (flatworld/explicitloop.py:526-553):

    * before substep ``counter`` (0 at the loop's first substep): the skip epochs of
      the mesh items reset and the edge boxes are refreshed after every N2-th
      substep, N2 = max(1, (1/500 s) // dt of that substep); the edge boxes also
      before the first one;
    * after it: the rigid step (``RigidMeshSet.rigid_step``).
    """

    def __init__(self, rigid_domains, mesh_set=None):
        self.rigids = list(rigid_domains)
        self.mesh_set = mesh_set
        self.counter = 0
        self.prev_dt = None

    def before_substep(self):
        """The loop step before which the skip epochs reset (``counter``), or -1."""
        c = self.counter
        reset = c >= 1 and c % max(1.0, RIGID_SH_REBUILD_INTERVAL // float(self.prev_dt)) == 0
        if self.mesh_set is not None and (c == 0 or reset):
            self.mesh_set.refresh_elem_aabbs()
        return c if reset else -1

    def after_substep(self, dt):
        for r in self.rigids:
            if r.omega != 0.0 or np.any(r.velocity != 0.0):
                r.advance(dt)
        if self.mesh_set is not None:
            self.mesh_set.rigid_step(dt)
        self.counter += 1
        self.prev_dt = float(dt)
