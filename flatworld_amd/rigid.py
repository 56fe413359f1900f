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
    raise NotImplementedError(f"rigid type {rtype:#b}: only ball, box and capsule primitives are supported")


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


class RigidBodyDomain:
    """RigidBodyDomain(rigid, bcs=[], considerContact=True, ...) (flatworld/rigiddomain.py:28-70),
    kinematic: the pose and velocity are the caller's (see the module docstring)."""

    def __init__(self, rigid, bcs=[], considerContact=True, considerGroundContact=True, initials=[], friction=0.0,
                 restitution=0.9, name="", visual_mesh=None, collision_shapes=None,
                 category_bits=COLLISION_CATEGORY_ORDINARY_RIGID, collide_bits=COLLISION_MASK_ALL):
        if rigid.rtype not in (RigidType.BALL, RigidType.BOX, RigidType.CAPSULE):
            raise NotImplementedError("mesh rigids are not part of this package")
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

    # ------------------------------------------------------------ device record
    def record(self):
        """float32 pose fields of the device contact record, and the AABB at the pose."""
        rot = rotation_matrix(self.theta)
        c = self.center.astype(np.float32)
        lo, hi = primitive_aabb(self.rigid.rtype, c, rot, self.prim, float(self.rigid.getRadius()))
        return dict(rigid_type=int(self.rigid.rtype), rigid_center=c, rigid_rot=rot, rigid_prim=self.prim.copy(),
                    rigid_radius=f32(float(self.rigid.getRadius())), rigid_vel=self.velocity.astype(np.float32),
                    rigid_omega=f32(self.omega), aabb_lo=lo, aabb_hi=hi)

    def getBBox(self):
        r = self.record()
        return r["aabb_lo"], r["aabb_hi"]
