"""Boundary and initial conditions (flatworld/bcs.py:14-59, flatworld/initialconditions.py:31-50).

``processData`` returns ``(bit, nodes, values)``; the manager ORs the bits into
each node's flag word and overwrites its value vector, so the last BC listed
for a node decides the value (flatworld/femspringmanager.py:660-685).
"""

import numpy as np

from .definitions import (
    ATYPE,
    FORCETYPE,
    GRAVITY,
    UTYPE,
    VTYPE,
    BoundaryConditionType,
)


class BcBase:
    pass


class Gravity(BcBase):
    """Body acceleration ``g`` on every node of the domain (flatworld/bcs.py:14-20)."""

    def __init__(self, g):
        self.g = g
        self.type = BoundaryConditionType.GRAVITY

    def processData(self):
        return GRAVITY, None, self.g


class Force(BcBase):
    """Constant nodal force (flatworld/bcs.py:23-30)."""

    def __init__(self, nds, force):
        self.nds = nds
        self.force = force
        self.type = BoundaryConditionType.FORCE

    def processData(self):
        return FORCETYPE, self.nds, self.force


class EnforceAcc(BcBase):
    """Prescribed acceleration: f = m * a (flatworld/bcs.py:33-40)."""

    def __init__(self, nds, acc):
        self.nds = nds
        self.acc = acc
        self.type = BoundaryConditionType.ENFORCE_ACCELERATION

    def processData(self):
        return ATYPE, self.nds, self.acc


class EnforceVel(BcBase):
    """Prescribed velocity (flatworld/bcs.py:43-50)."""

    def __init__(self, nds, vel):
        self.nds = nds
        self.vel = vel
        self.type = BoundaryConditionType.ENFORCE_VELOCITY

    def processData(self):
        return VTYPE, self.nds, self.vel


class Fixed(BcBase):
    """Zero velocity and zero force (flatworld/bcs.py:53-59)."""

    def __init__(self, nds):
        self.nds = nds
        self.type = BoundaryConditionType.FIXED

    def processData(self):
        return UTYPE, self.nds, None


class InitialVel:
    """Initial nodal velocity, ``nds`` a list or "ALL" (flatworld/initialconditions.py:31-50)."""

    def __init__(self, nds, vel):
        if isinstance(nds, str) and nds == "ALL":
            self.is_all_nodes = True
            self.nds = None
        else:
            self.is_all_nodes = False
            self.nds = np.asarray(nds, dtype=np.int32)
        self.vel = np.asarray(vel, dtype=np.float32)
        self.type = VTYPE
