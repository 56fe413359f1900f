"""Spring-mass domain (flatworld/springmass.py:1-80).

A network of two-node linear springs with one mass per node.  The constructor
keeps the reference's signature and derived values:

* ``prop = [spring, damping, mass]``: stiffness ``k`` of every spring, damping
  coefficient ``c`` along each spring, and the mass of every node;
* ``restLength[e] = |x_a - x_b|`` of the initial coordinates, float32;
* ``stableTime = 0.5 * sqrt(mass / spring)`` in Python float64.

On the device a spring-mass domain is one scene of the arena (material code 4):
its springs go through the same tiled substep kernel as triangles, two corners
per slot, with the rest length carried in the slot record.
"""

import numpy as np

from .definitions import COLLISION_CATEGORY_FEM, COLLISION_MASK_ALL, DomainType
from .domains import DomainBase


def spring_rest_lengths(coords, conns):
    """float32 ``|x_a - x_b|`` per spring (flatworld/springmass.py:46-52).

    The reference takes ``np.linalg.norm`` of each float32 difference vector in a
    Python loop; ``sqrt(dx*dx + dy*dy)`` in float32 gives the same values (pinned
    against that loop in the tests) in one vectorized pass.
    """
    coords = np.asarray(coords, dtype=np.float32)
    conns = np.asarray(conns)
    if conns.shape[0] == 0:
        return np.zeros(0, dtype=np.float32)
    d = coords[conns[:, 0]] - coords[conns[:, 1]]
    dx = d[:, 0]
    dy = d[:, 1]
    return np.sqrt(dx * dx + dy * dy).astype(np.float32)


class SpringMassDomain(DomainBase):
    """SpringMassDomain(d, coords, conns, prop=[100.0, 1.0, 1.0], bcs=[], considerContact=True,
    considerGroundContact=True, initials=[], friction=0.0) as in flatworld/springmass.py:6-43."""

    def __init__(
        self,
        d,
        coords,
        conns,
        prop=[100.0, 1.0, 1.0],
        bcs=[],
        considerContact=True,
        considerGroundContact=True,
        initials=[],
        friction=0.0,
    ):
        if d != 2:
            raise ValueError(f"SpringMassDomain: only d=2 is supported (got d={d})")
        self.d = d
        self.prop = prop
        self.bcs = bcs
        self.initials = initials
        self.friction = friction
        self.considerContact = considerContact
        self.considerGroundContact = considerGroundContact
        self.type = DomainType.SPRINGMASS
        self.spring, self.damping, self.mass = prop
        self.massInv = 1.0 / self.mass
        self.coords = np.array(coords, dtype=np.float32)
        self.connectivity = np.array(conns, dtype=np.int32)
        if self.connectivity.ndim != 2 or self.connectivity.shape[1] != 2:
            raise ValueError("SpringMassDomain: conns must be an (E, 2) array of node pairs")
        self.nnodes = coords.shape[0]
        self.nelements = conns.shape[0]
        self.restLength = spring_rest_lengths(self.coords[:, :2], self.connectivity)
        self.stableTime = 0.5 * np.sqrt(self.mass / self.spring)  # 2.0 * 0.25, 0.25 is the safety scale factor
        self.category_bits = COLLISION_CATEGORY_FEM
        self.collide_bits = COLLISION_MASK_ALL
        self.femManager = None
        self.domainIdx = None

    def attach(self, femManager, domainIdx):
        self.femManager = femManager
        self.domainIdx = domainIdx

    def getCurrentCoords(self):
        """Current node positions (flatworld/springmass.py:58-61)."""
        return self.femManager.domain_coords(self.domainIdx)

    def getBBox(self):
        """AABB of the initial coordinates, as the reference returns (flatworld/springmass.py:63-71)."""
        aabb = self.getBBoxKernel2D()
        return aabb[0], aabb[1]

    def getBBoxKernel2D(self):
        aabb = np.zeros((2, 2), dtype=np.float32)
        aabb[0, :] = np.min(self.coords, axis=0)
        aabb[1, :] = np.max(self.coords, axis=0)
        return aabb
