"""Scene-description domains: FEM body and the three ground kinds.

Host-side records with the reference's constructors and derived scalars:

* ``FemDomain``          -- flatworld/femdomain.py:8-50
* ``GroundDomain``       -- flatworld/grounddomain.py:18-70 (plane)
* ``HeightFieldDomain``  -- flatworld/heightfielddomain.py:18-57, 190-194
* ``VoxelGridDomain``    -- flatworld/voxeldomain.py:9-60, 194-195

All float arithmetic that feeds the device (normals, AABBs, voxel spacing) is
written with the same numpy float32 operations as the reference so that the
uploaded values agree bit for bit.
"""

import numpy as np

from .definitions import (
    COLLISION_CATEGORY_FEM,
    COLLISION_CATEGORY_GROUND,
    COLLISION_MASK_ALL,
    DomainType,
)


class DomainBase:
    """Common base (flatworld/numericaldomain.py:5-11)."""


class FemDomain(DomainBase):
    """FEM body (flatworld/femdomain.py:8-25).

    ``stableTime = 0.4 * charLength / sqrt(E / rho)`` in Python float64.
    """

    def __init__(self, mesh, prop, bcs=[], considerContact=True, considerGroundContact=True, initials=[], friction=0.0):
        self.mesh = mesh
        self.d = mesh.d
        self.prop = prop
        self.bcs = bcs
        self.friction = friction
        self.initials = initials
        self.nnodes = self.mesh.numNodes
        self.nelements = self.mesh.numElements
        self.considerContact = considerContact
        self.considerGroundContact = considerGroundContact
        self.type = DomainType.FEM
        self.category_bits = COLLISION_CATEGORY_FEM
        self.collide_bits = COLLISION_MASK_ALL
        mat = self.prop.mat
        self.stableTime = 0.4 * self.mesh.charLength / (mat.E / mat.rho) ** 0.5
        self.femManager = None
        self.domainIdx = None

    def attach(self, femManager, domainIdx):
        self.femManager = femManager
        self.domainIdx = domainIdx

    def getBBox(self):
        """Boundary-node AABB held by the manager (flatworld/femdomain.py:27-33)."""
        return self.femManager.domain_aabb(self.domainIdx)

    def getNumDofs(self):
        return self.d * self.mesh.numNodes

    def getNumNodes(self):
        return self.mesh.numNodes

    def getNumElements(self):
        return self.mesh.numElements

    def getCurrentCoords(self):
        """Current node positions of this domain (flatworld/femdomain.py:47-50)."""
        return self.femManager.domain_coords(self.domainIdx)


def _to_np(x):
    if hasattr(x, "numpy") and callable(getattr(x, "numpy")):
        return x.numpy()
    return np.asarray(x)


class GroundDomain(DomainBase):
    """Analytical plane through ``point`` with unit ``normal`` (flatworld/grounddomain.py:19-34)."""

    def __init__(self, d, point, normal, bcs=[], considerContact=True, considerGroundContact=True, initials=[]):
        self.d = d
        self.point = np.asarray(point, dtype=np.float32)
        n = np.asarray(normal, dtype=np.float32)
        nlen = float(np.linalg.norm(n))
        self.normal = n / max(nlen, 1e-12)
        self.nnodes = 1
        self.nelements = 0
        self.type = DomainType.ANALYTICAL
        self.considerContact = considerContact
        self.considerGroundContact = considerGroundContact
        self.initials = initials
        self.bcs = bcs
        self.category_bits = COLLISION_CATEGORY_GROUND
        self.collide_bits = COLLISION_MASK_ALL

    def getCurrentRefPoint(self):
        return self.point.copy()

    def getCurrentNormal(self):
        return self.normal.copy()

    def getBBox(self):
        """Plane AABB: +-10000 along the tangent, +-0.1 along the normal (flatworld/grounddomain.py:54-70)."""
        buffer = 0.1
        large_span = 10000.0
        p = self.point
        tangent = np.array([-self.normal[1], self.normal[0]], dtype=np.float32)
        lo_raw = p - tangent * large_span - self.normal * buffer
        hi_raw = p + tangent * large_span + self.normal * buffer
        lo = np.minimum(lo_raw, hi_raw).astype(np.float32)
        hi = np.maximum(lo_raw, hi_raw).astype(np.float32)
        return lo, hi


class HeightFieldDomain(DomainBase):
    """2D height field z = h(x) sampled at ``nx`` points on [lb.x, ub.x] (flatworld/heightfielddomain.py:18-53)."""

    def __init__(self, d, height_np, lb, ub, considerContact=True, reverse=False):
        if d != 2:
            raise ValueError("HeightFieldDomain: only d=2 is supported on this path")
        self.d = d
        self.type = DomainType.HEIGHTFIELD
        self.considerContact = considerContact
        self.considerGroundContact = considerContact
        self.initials = []
        self.bcs = []
        self.reverse = reverse
        height_np = np.asarray(height_np, dtype=np.float32)
        self.maxHeight = float(np.max(height_np)) + 0.1
        self.minHeight = float(np.min(height_np)) - 0.1
        assert height_np.ndim == 1, "2D height field expects 1D numpy array of shape (nx,)"
        self.nx = int(height_np.shape[0])
        self.ny = 1
        self.height = height_np.copy()
        self.point = np.array([0.0, -1e6], dtype=np.float32)
        self.normal = np.array([0.0, 1.0], dtype=np.float32)
        self.nnodes = 1
        self.nelements = 0
        self.lb = np.asarray(lb, dtype=np.float32)
        self.ub = np.asarray(ub, dtype=np.float32)
        self.category_bits = COLLISION_CATEGORY_GROUND
        self.collide_bits = COLLISION_MASK_ALL

    def sample_height_2d(self, x):
        """Host linear interpolation (flatworld/heightfielddomain.py:61-68)."""
        u = (x - self.lb[0]) / (self.ub[0] - self.lb[0])
        u = float(np.clip(u, 0.0, 1.0))
        s = u * (self.nx - 1)
        i0 = int(np.floor(s))
        i1 = min(i0 + 1, self.nx - 1)
        t = s - float(i0)
        return float(self.height[i0] + (self.height[i1] - self.height[i0]) * t)

    def getBBox(self):
        """[lb.x, min h - 0.1] x [ub.x, max h + 0.1] (flatworld/heightfielddomain.py:190-194)."""
        return (
            np.array([self.lb[0], self.minHeight], dtype=np.float32),
            np.array([self.ub[0], self.maxHeight], dtype=np.float32),
        )


class VoxelGridDomain(DomainBase):
    """Axis-aligned 2D occupancy grid ``occ[i, j]`` with i along x (flatworld/voxeldomain.py:9-60)."""

    def __init__(self, d, nx, ny, lb, ub, nz=None, considerContact=True, occupancy_np=None):
        assert d in (2, 3), "VoxelGridDomain currently supports d=2 or 3"
        if d != 2:
            raise ValueError("VoxelGridDomain: only d=2 is supported on this path")
        self.d = d
        self.type = DomainType.VOXELMAP
        self.considerContact = considerContact
        self.considerGroundContact = considerContact
        self.initials = []
        self.bcs = []
        self.nx = int(nx)
        self.ny = int(ny)
        self.lb = np.asarray(lb, dtype=np.float32)
        self.ub = np.asarray(ub, dtype=np.float32)
        self.dx = (self.ub[0] - self.lb[0]) / max(self.nx, 1)
        self.dz = (self.ub[1] - self.lb[1]) / max(self.ny, 1)
        self.occ = np.zeros((self.nx, self.ny), dtype=np.int32)
        if occupancy_np is not None:
            assert occupancy_np.shape == (self.nx, self.ny)
            self.occ[:] = occupancy_np.astype(np.int32)
        self.nnodes = 1
        self.nelements = 0
        self.category_bits = COLLISION_CATEGORY_GROUND
        self.collide_bits = COLLISION_MASK_ALL

    def getBBox(self):
        """Grid extent (flatworld/voxeldomain.py:194-195)."""
        return self.lb, self.ub
