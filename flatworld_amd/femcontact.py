"""Contact pair records: penalty stiffness and contact stable time.

Host scalars only, following flatworld/femcontact.py:94-197 and 256-303:

* FEM body: bulk modulus ``K = E / (3 (1 - 2 min(nu, 0.4999)))``, 2D segment
  penalty ``k = 2 K`` (independent of mesh size), contact stable time
  ``0.9 sqrt(m_node / k)`` with ``m_node = rho * (L_min^2 / 2) / 3``;
* FEM-FEM: the smaller of the two bodies' segment penalties, and the smaller of
  the two stable times;
* spring network: ``k = 100 * spring``, stable time ``0.5 sqrt(mass / k)``;
* spring network against a FEM body: ``100 min(spring, 2 K)`` and the FEM body's
  contact stable time.

The force laws themselves run on the device (flatworld_amd/csrc/fw_kernels.hip,
fw_flex.hip).  A FEM-spring pair exists as in the reference but applies no force
there: ExplicitLoop files it among the unbatched contacts, which no substep runs.
A FEM body against a rigid primitive (flatworld/femcontact.py:173-180) takes the
FEM body's segment penalty ``2 K`` and its contact stable time; a spring network
against one (femcontact.py:243-253) takes ``100 spring`` and ``0.5 sqrt(mass / k)``,
with every node a work item of the same primitive kernel.  The rigid is a
kinematic primitive (flatworld_amd.rigid).  Pairs the reference's factory has no
class for (spring-spring, spring-voxel) give no contact, as there.
"""

from .definitions import ContactType, DomainType


class ContactBase:
    def __init__(self, domain1, domain2, type, is_tied=False):
        self.domain1 = domain1
        self.domain2 = domain2
        self.penalty = 1.0
        self.type = type
        self.tied = is_tied

    @staticmethod
    def _bulk_modulus(mat):
        nu = min(mat.nu, 0.4999)
        return mat.E / (3.0 * (1.0 - 2.0 * nu))

    @staticmethod
    def _segment_penalty(dom, slsfac=1.0):
        mat = dom.prop.mat
        K = ContactBase._bulk_modulus(mat)
        L = dom.mesh.charAverageLength
        if dom.d == 2:
            return 2.0 * slsfac * K
        return 1.5 * slsfac * K * L

    def calStableTime(self, penalty, femdomain):
        mat = femdomain.prop.mat
        L = femdomain.mesh.charLength
        if femdomain.d == 2:
            m_node = mat.rho * (L * L * 0.5) / 3.0
        else:
            m_node = mat.rho * (L * L * L / 6.0) / 4.0
        return 0.9 * (m_node / penalty) ** 0.5

    def update(self, dt):
        pass

    def calculate(self, dt):
        pass


class ContactFlexFlex(ContactBase):
    """FEM vs FEM (flatworld/femcontact.py:164-172): symmetric, the more conservative side's penalty."""

    def __init__(self, domain1, domain2, type, is_tied=False):
        super().__init__(domain1, domain2, type, is_tied)
        self.penalty = min(self._segment_penalty(self.domain1), self._segment_penalty(self.domain2)) * self.penalty
        self.stableTime = self.calStableTime(self.penalty, self.domain1)
        self.stableTime = min(self.stableTime, self.calStableTime(self.penalty, self.domain2))


class _FlexGround(ContactBase):
    """FEM body against a static ground; ``domain1`` is always the FEM body."""

    ground_type = None

    def __init__(self, domain1, domain2, type, is_tied=False):
        super().__init__(domain1, domain2, type, is_tied)
        if self.domain1.type == self.ground_type:
            self.domain1, self.domain2 = self.domain2, self.domain1
        self.penalty = self._segment_penalty(self.domain1) * self.penalty
        self.stableTime = self.calStableTime(self.penalty, self.domain1)


class ContactFlexAnalytical(_FlexGround):
    """FEM vs plane (flatworld/femcontact.py:155-161)."""

    ground_type = DomainType.ANALYTICAL


class ContactFlexHeightField(_FlexGround):
    """FEM vs height field (flatworld/femcontact.py:182-188)."""

    ground_type = DomainType.HEIGHTFIELD


class ContactFlexVoxelMap(_FlexGround):
    """FEM vs voxel grid (flatworld/femcontact.py:191-197)."""

    ground_type = DomainType.VOXELMAP


class _SpringGround(ContactBase):
    """Spring network against a static ground; ``domain1`` is always the springs."""

    ground_type = None
    _SPRING_PENALTY_SCALE = 100.0

    def __init__(self, domain1, domain2, type, is_tied=False):
        super().__init__(domain1, domain2, type, is_tied)
        if self.domain1.type == self.ground_type:
            self.domain1, self.domain2 = self.domain2, self.domain1
        self.penalty = self.domain1.spring * self._SPRING_PENALTY_SCALE
        self.stableTime = 0.5 * (self.domain1.mass / self.penalty) ** 0.5


class ContactSpringAnalytical(_SpringGround):
    """Springs vs plane (flatworld/femcontact.py:256-265)."""

    ground_type = DomainType.ANALYTICAL


class ContactSpringHeightField(_SpringGround):
    """Springs vs height field (flatworld/femcontact.py:269-279)."""

    ground_type = DomainType.HEIGHTFIELD


class ContactSpringRigid(_SpringGround):
    """Springs vs rigid primitive (flatworld/femcontact.py:243-253); every spring node is a
    work item of the rigid-primitive kernel (flatworld/mixedcontact.py:2001-2011)."""

    ground_type = DomainType.RIGID


class ContactSpringFlex(ContactBase):
    """Springs vs FEM body (flatworld/femcontact.py:200-213); ``domain1`` is always the
    springs.  The reference files this pair among its unbatched contacts
    (flatworld/mixedcontact.py:1789-1792), which no substep applies, and its zero-dt
    warm-up ``calculate`` returns at once (femcontact.py:214-217): the pair lowers the
    loop's stable time and turns the BVH gate on, and moves no node."""

    _SPRING_PENALTY_SCALE = 100.0

    def __init__(self, domain1, domain2, type, is_tied=False):
        super().__init__(domain1, domain2, type, is_tied)
        if self.domain1.type == DomainType.FEM:
            self.domain1, self.domain2 = self.domain2, self.domain1
        self.penalty = min(self.domain1.spring, self._segment_penalty(self.domain2)) * self._SPRING_PENALTY_SCALE
        self.stableTime = self.calStableTime(self.penalty, self.domain2)


class ContactFlexRigid(ContactBase):
    """FEM body against a rigid primitive (flatworld/femcontact.py:173-180): domain1 is the
    FEM body, domain2 the rigid; penalty 2 K of the FEM body, its contact stable time."""

    def __init__(self, domain1, domain2, type, is_tied=False):
        super().__init__(domain1, domain2, type, is_tied)
        if self.domain1.type == DomainType.RIGID:
            self.domain1, self.domain2 = self.domain2, self.domain1
        self.penalty = self._segment_penalty(self.domain1) * self.penalty
        self.stableTime = self.calStableTime(self.penalty, self.domain1)


def contactHelper(domain1, domain2, type):
    """Pair factory (flatworld/femcontact.py:283-303); rigid-rigid and rigid-plane pairs
    belong to the rigid-body dynamics and give no contact object here."""
    if type == ContactType.ANALYTICALSPRING:
        return ContactSpringAnalytical(domain1, domain2, type)
    if type == ContactType.SPRINGHEIGHTFIELD:
        return ContactSpringHeightField(domain1, domain2, type)
    if type in (ContactType.SPRINGSPRING, ContactType.SPRINGVOXELMAP):
        return None  # no contact class in the reference either
    if type == ContactType.FLEXSPRING:
        return ContactSpringFlex(domain1, domain2, type)
    if type == ContactType.FLEXANLAYTICAL:
        return ContactFlexAnalytical(domain1, domain2, type)
    if type == ContactType.FEXHEIGHTFIELD:
        return ContactFlexHeightField(domain1, domain2, type)
    if type == ContactType.FLEXVOXELMAP:
        return ContactFlexVoxelMap(domain1, domain2, type)
    if type == ContactType.FLEXFLEX:
        return ContactFlexFlex(domain1, domain2, type)
    if type == ContactType.FLEXRIGID:
        return ContactFlexRigid(domain1, domain2, type)
    if type == ContactType.RIGIDSPRING:
        return ContactSpringRigid(domain1, domain2, type)
    if type in (ContactType.RIGIDANLAYTICAL, ContactType.RIGIDRIGID, ContactType.ANALYTICALANALYTICAL):
        return None
    raise NotImplementedError(f"contact type {type:#b} is not part of the explicit-FEM path")
