"""flatworld_amd -- FlatWorld's explicit-FEM substep, native to MI355X (gfx950).

Drop-in for the reference's FEM- and spring-vs-ground path: the same ``Mesh /
FemDomain / SpringMassDomain / SolidProp / ExplicitLoop.advanceWithTime`` API
(flatworld/__init__.py), with
the substep running as hand-written HIP kernels behind the C ABI in
include/flatworld_fem.h.  ``FemBatch`` and ``SimServer`` add the batched and
served forms: thousands of uneven scenes per launch, hipGraph frames,
HBM-footprint admission and continuous batching.
"""

from .bcs import EnforceAcc, EnforceVel, Fixed, Force, Gravity, InitialVel
from .definitions import *  # noqa: F401,F403
from .domains import FemDomain, GroundDomain, HeightFieldDomain, VoxelGridDomain
from .femesher import FEMesher, structured_grid
from .materials import Elastic, HyperElastic, MisesPlastic, SolidProp
from .mesh import Mesh
from .rigid import BallRigid, BoxRigid, CapsuleRigid, MeshRigid, RigidBodyDomain
from .scene import GroundSpec, SceneSpec, pack_domain, pack_fem_domain, pack_spring_domain
from .springmass import SpringMassDomain


def __getattr__(name):
    # device-facing classes load the HIP library on first use only
    if name == "ExplicitLoop":
        from .explicitloop import ExplicitLoop

        return ExplicitLoop
    if name == "FemSpringManager":
        from .femspringmanager import FemSpringManager

        return FemSpringManager
    if name == "FemBatch":
        from .batch import FemBatch

        return FemBatch
    if name == "SimServer":
        from .server import SimServer

        return SimServer
    raise AttributeError(name)


__version__ = "0.1.0"
