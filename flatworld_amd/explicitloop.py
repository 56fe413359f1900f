"""ExplicitLoop drop-in for FEM bodies and spring networks on static grounds
(flatworld/explicitloop.py:71-589).

Setup follows flatworld/explicitloop.py:72-398 for the body-vs-ground subset:

* loop stable time = min(1.0, 1e-3 when a plane exists (RigidManager default,
  flatworld/rigidmanager.py:1780), every FEM / spring domain's stable time),
  raised to ``mass_scaling_dt`` when selective mass scaling is asked for
  (explicitloop.py:164-170), then lowered to every contact pair's stable time;
* contact pairs are formed over domains with ``considerContact`` in the
  reference's O(D^2) order and split into the batched plane / height-field /
  voxel work lists (flatworld/mixedcontact.py:1756-2269).  As in the reference,
  a loop with more than one distinct height field (or voxel grid) leaves those
  contacts unbatched, which means they never run; a warning says so;
* FEM-FEM pairs become one contact group over the loop's FEM bodies
  (flatworld/mixedcontact.py:1857-1957): a spatial hash over their boundary
  edges, built after the warm-up (explicitloop.py:383-385) and rebuilt every
  1/500 s of simulated time, and a contact pass before every substep.  Its
  pairs are gated by the bodies' AABB overlap when the loop runs its BVH (any
  ground contact present) and always active otherwise (explicitloop.py:367-381).

* FEM-rigid pairs (a FEM body against a ``flatworld_amd.rigid.RigidBodyDomain``
  ball, box or capsule) become contact records after the FEM-FEM record and
  before the height-field and voxel records, where the reference launches its
  rigid kernel (flatworld/mixedcontact.py:1697-1714).  The rigid is kinematic:
  its pose is pushed to the device before every substep and advanced at its
  constant velocity after it; the reference's RigidManager dynamics are not part
  of this package.

``advance`` is one fused substep launch: activation gate, ground forces,
element forces, symplectic Euler and AABB, all on the device.
``advanceWithTime`` computes the reference's float64 substep sequence on the
host and enqueues it without any device synchronisation.
"""

import warnings

import numpy as np

from .definitions import ContactType, DomainType
from .femcontact import ContactFlexFlex, ContactSpringFlex, contactHelper
from .femspringmanager import FemSpringManager
from .flexcontact import build_flex_group

_GROUND_TYPES = (DomainType.ANALYTICAL, DomainType.HEIGHTFIELD, DomainType.VOXELMAP, DomainType.RIGID)
_BODY_TYPES = (DomainType.FEM, DomainType.SPRINGMASS)


class ExplicitLoop:
    def __init__(
        self,
        dt,
        domains,
        joints=[],
        ties=None,
        damping=0.0,
        useAdapativeDT=True,
        skip_spatial_hash=False,
        considerRigidRigidContact=True,
        use_pd=0,
        mass_scaling_dt=0.0,
        max_contact_velocity=50.0,
        device=0,
        arena_kwargs=None,
    ):
        if len(joints) > 0:
            raise NotImplementedError("joints act on rigid-body dynamics, which are not part of this package")
        for d in domains:
            if d.type not in _BODY_TYPES + _GROUND_TYPES:
                raise NotImplementedError(f"domain type {d.type:#b} is not part of this package")
            if d.type == DomainType.RIGID and not hasattr(d, "record"):
                raise NotImplementedError("rigid domains are kinematic flatworld_amd.rigid.RigidBodyDomain bodies here")
        self.domains = domains
        self.d = domains[0].d
        self.numDomains = len(domains)
        # ties: index pairs of domains whose contact is tied (flatworld/explicitloop.py:111-113, 400-439);
        # only FEM-FEM pairs act on it, as in the reference's batched kernels
        self.ties = list(ties) if ties is not None else []
        self._tie_pairs = set()
        self._refresh_tie_pairs()
        self._max_domains = self.numDomains + 10
        self.damping = damping
        self.useAdapative = useAdapativeDT
        self._counter_py = 0

        self.stableTime = 1.0
        # a RigidManager exists with any rigid or analytical domain; its default stable time
        # is 1e-3 (flatworld/explicitloop.py:145-160, rigidmanager.py:1780)
        if any(d.type in (DomainType.ANALYTICAL, DomainType.RIGID) for d in domains):
            self.stableTime = min(1.0 / 1000.0, self.stableTime)

        # contact pairs in the reference's order (flatworld/explicitloop.py:237-279)
        self.contacts = []
        self.contactPairIds = []
        contact_domains = [(i, d) for i, d in enumerate(domains) if d.considerContact]
        for idx_a, (i, dom_i) in enumerate(contact_domains):
            for idx_b in range(idx_a):
                j, dom_j = contact_domains[idx_b]
                ctype = dom_i.type | dom_j.type
                if dom_i.category_bits & dom_j.collide_bits == 0:
                    continue
                if dom_i.collide_bits & dom_j.category_bits == 0:
                    continue
                if ctype in (ContactType.RIGIDRIGID, ContactType.RIGIDANLAYTICAL, ContactType.ANALYTICALANALYTICAL):
                    continue
                if hasattr(dom_i, "env_id") and hasattr(dom_j, "env_id") and dom_i.env_id != dom_j.env_id:
                    continue
                if dom_i.type not in _BODY_TYPES and dom_j.type not in _BODY_TYPES:
                    continue  # ground-ground: no contact object in the reference either
                c = contactHelper(dom_i, dom_j, ctype)
                if c is not None:
                    if self._is_tie_pair(i, j):
                        c.tied = True
                    self.contacts.append(c)
                    self.contactPairIds.append((i, j))

        # batched work lists: planes, then one height field, then one voxel grid
        grounds_of = {}
        anal = [(c, p) for c, p in zip(self.contacts, self.contactPairIds) if c.domain2.type == DomainType.ANALYTICAL]
        hf = [(c, p) for c, p in zip(self.contacts, self.contactPairIds) if c.domain2.type == DomainType.HEIGHTFIELD]
        vox = [(c, p) for c, p in zip(self.contacts, self.contactPairIds) if c.domain2.type == DomainType.VOXELMAP]
        rig = [(c, p) for c, p in zip(self.contacts, self.contactPairIds) if c.domain2.type == DomainType.RIGID]
        self._unbatched_contacts = []
        for name, group in (("HeightField", hf), ("VoxelMap", vox)):
            if len({id(c.domain2) for c, _ in group}) > 1:
                warnings.warn(
                    f"{name}: {len({id(c.domain2) for c, _ in group})} distinct grounds in one loop; as in "
                    "flatworld/mixedcontact.py:2140-2261 these contacts stay unbatched and are never applied"
                )
                self._unbatched_contacts += [c for c, _ in group]
                group.clear()
        # FEM-spring pairs: unbatched in the reference and never applied
        # (flatworld/mixedcontact.py:1789-1792, 2263-2268)
        spring_fem = [c for c in self.contacts if isinstance(c, ContactSpringFlex)]
        if spring_fem:
            warnings.warn(
                f"{len(spring_fem)} FEM-spring contact pair(s): as in flatworld/mixedcontact.py:1789-1792 they stay "
                "unbatched and apply no force; they lower the stable time and turn the BVH gate on"
            )
            self._unbatched_contacts += spring_fem
        # rigid records as the reference launches their kernels: primitives, then the
        # mesh ties, then the mesh contacts, each in pair order (mixedcontact.py:1711-1742)
        def rig_order(cp):
            c = cp[0]
            if not getattr(c.domain2, "is_mesh", False):
                return 0
            return 1 if getattr(c, "tied", False) else 2
        rig = sorted(rig, key=rig_order)
        mesh_set = self._mesh_rigids(domains, rig, mass_scaling_dt)
        for c, _ in anal + rig + hf + vox:
            fem_idx = self.domains.index(c.domain1)
            grounds_of.setdefault(fem_idx, []).append((c.domain2, c))

        # FEM-FEM: one group over the loop's FEM bodies (flatworld/mixedcontact.py:1857-1957)
        flex_contacts = [c for c in self.contacts if isinstance(c, ContactFlexFlex)]
        flex = None
        if flex_contacts:
            managed = [d for d in domains if d.type in _BODY_TYPES]
            fem_bodies = [d for d in managed if d.type == DomainType.FEM]
            if managed[: len(fem_bodies)] != fem_bodies:
                raise NotImplementedError(
                    "FEM-FEM contact with a spring-mass domain listed before a FEM domain: the reference's "
                    "spatial hash indexes the FEM bodies by manager position (flatworld/femspringmanager.py:455-466)"
                )
            # the BVH runs when any ground contact exists (flatworld/explicitloop.py:367-381)
            gated = len(anal) + len(rig) + len(hf) + len(vox) + len(self._unbatched_contacts) > 0
            if gated:
                for b, dom in enumerate(fem_bodies):
                    if self.domains.index(dom) != b:
                        raise NotImplementedError(
                            "FEM-FEM contact in a loop with grounds needs the FEM domains listed first: the "
                            "reference's per-domain BVH mask is read at the body's FEM index "
                            "(flatworld/mixedcontact.py:925, explicitloop.py:501-509)"
                        )
            count_nodes = sum(d.nnodes for d in managed)
            flex = build_flex_group(fem_bodies, flex_contacts, gated, count_nodes)

        self.femSpringManager = FemSpringManager(self.d, domains, grounds_of=grounds_of, device=device,
                                                 arena_kwargs=arena_kwargs, flex=flex, mesh_set=mesh_set)
        self.stableTime = min(self.femSpringManager.stableTime, self.stableTime)
        self.femSpringManager.setGlobalAABB(None)
        if mass_scaling_dt > 0.0:
            self.femSpringManager.applyMassScaling(mass_scaling_dt)
            self.stableTime = max(self.stableTime, mass_scaling_dt)
        if flex is not None:
            # initial FEM spatial hash, so that the first substep has one (flatworld/explicitloop.py:383-385)
            self.femSpringManager.populate_fem_spatial_hash(velocity_buffer=0.0)
        self.numContacts = len(self.contacts)
        for c in self.contacts:
            self.stableTime = min(self.stableTime, c.stableTime)
        if mesh_set is not None:
            # initial rigid spatial hash over the elements near the FEM nodes in mesh
            # contact (flatworld/explicitloop.py:386-396)
            mesh_set.initial_build(*mesh_set.fem_contact_aabb(self._mesh_node_sets))
        self.mesh_set = mesh_set

        if useAdapativeDT:
            self.dt = self.stableTime
        else:
            self.dt = dt
        self.default_dt = dt

    # ------------------------------------------------------------ mesh rigids
    def _mesh_rigids(self, domains, rig, mass_scaling_dt):
        """The loop's RigidMeshSet, when a FEM body or spring network touches a mesh
        rigid; its contacts learn the set, their rigid id and the full-query subcycle
        max(2, min(8, rigid-contact stable time / domain stable time))
        (flatworld/explicitloop.py:176-182, 308-327)."""
        from .rigid import RigidMeshSet

        mesh_contacts = [c for c, _ in rig if getattr(c.domain2, "is_mesh", False)]
        if not mesh_contacts:
            return None
        rigid_domains = [d for d in domains if d.type == DomainType.RIGID]
        meshes = [d for d in rigid_domains if getattr(d, "is_mesh", False) and d.considerContact]
        mesh_set = RigidMeshSet(meshes, rigid_domains)
        managed = [d for d in domains if d.type in _BODY_TYPES]
        dom_min = min([1.0] + ([1.0 / 1000.0] if any(d.type in (DomainType.ANALYTICAL, DomainType.RIGID) for d in domains)
                               else []) + [d.stableTime for d in managed])
        if mass_scaling_dt > 0.0:
            dom_min = max(dom_min, mass_scaling_dt)
        rigid_min = min(c.stableTime for c, _ in rig)
        subcycle = int(max(2, min(8, rigid_min / max(dom_min, 1e-12))))
        self._mesh_node_sets = []
        for c in mesh_contacts:
            c.mesh_set = mesh_set
            c.mesh_did = mesh_set.did(c.domain2)
            c.mesh_subcycle = subcycle
            body = c.domain1
            if body.type == DomainType.SPRINGMASS:
                self._mesh_node_sets.append((np.asarray(body.coords, np.float32), np.arange(body.nnodes)))
            else:
                self._mesh_node_sets.append((np.asarray(body.mesh.coords, np.float32), np.asarray(body.mesh.boundaryNodes)))
        return mesh_set

    # ------------------------------------------------------------ ties
    def _parse_tie_entry(self, tie):
        """A pair of domain indices (negative ones count from the end) as (low, high), or
        None for anything else, an index out of range or a domain tied to itself
        (flatworld/explicitloop.py:400-420)."""
        if not isinstance(tie, (tuple, list)) or len(tie) != 2:
            return None
        a, b = tie
        if not isinstance(a, (int, np.integer)) or not isinstance(b, (int, np.integer)):
            return None
        n = len(self.domains)
        ia, ib = int(a), int(b)
        ia += n if ia < 0 else 0
        ib += n if ib < 0 else 0
        if not (0 <= ia < n and 0 <= ib < n) or ia == ib:
            return None
        return (min(ia, ib), max(ia, ib))

    def _refresh_tie_pairs(self):
        self._tie_pairs.clear()
        bad = 0
        for tie in self.ties:
            pair = self._parse_tie_entry(tie)
            if pair is None:
                bad += 1
            else:
                self._tie_pairs.add(pair)
        if bad:
            print(f"Warning: {bad} tie entries could not be resolved and were ignored")

    def _is_tie_pair(self, idx_a, idx_b):
        return (min(idx_a, idx_b), max(idx_a, idx_b)) in self._tie_pairs

    # ------------------------------------------------------------ stepping
    def advance(self):
        """One substep: gate, ground forces, element forces, integrate (flatworld/explicitloop.py:473-555)."""
        self.femSpringManager.advance_substeps(1, self.dt, self.damping, diag=True, dt64=[self.dt])
        self._counter_py += 1

    def substep_sequence(self, frame_dt):
        """The float64 dt sequence advanceWithTime runs (flatworld/explicitloop.py:573-586)."""
        seq = []
        remaining = float(frame_dt)
        eps = 1e-12
        while remaining > eps:
            current_stable = self.stableTime if self.useAdapative else self.default_dt
            sub_dt = max(min(current_stable, remaining), 1e-8)
            seq.append(sub_dt)
            remaining -= sub_dt
        return seq

    def advanceWithTime(self, frame_dt=1.0 / 60.0, verbose=False):
        """Advance by ``frame_dt`` of simulated time (flatworld/explicitloop.py:558-589)."""
        seq = self.substep_sequence(frame_dt)
        if not seq:
            return
        dts = np.asarray(seq, dtype=np.float64).astype(np.float32)
        # runs of equal float32 dt become one multi-substep enqueue
        starts = [0] + [k for k in range(1, len(dts)) if dts[k] != dts[k - 1]]
        ends = starts[1:] + [len(dts)]
        for r, (s, e) in enumerate(zip(starts, ends)):
            self.dt = seq[e - 1]
            self.femSpringManager.advance_substeps(e - s, float(dts[s]), self.damping, diag=(r == len(starts) - 1),
                                                   dt64=seq[s:e])
        self._counter_py += len(seq)
        if verbose:
            print(f"Completed {len(seq)} substeps for {frame_dt} seconds of simulation time.")

    def reset(self):
        """Restart the step counter (FEM state is kept, as in flatworld/explicitloop.py:635-679)."""
        self._counter_py = 0
        self.femSpringManager.reset_gate()
