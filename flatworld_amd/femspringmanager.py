"""FemSpringManager drop-in over the HIP arena (flatworld/femspringmanager.py:495-1033).

The reference packs every FEM and spring-mass domain of a loop into global Warp
arrays and launches force -> springs -> integrate -> AABB -> clear kernels per
substep.  Here every such domain is one scene of a ``flatworld_amd._native.Arena``
and a substep is one fused kernel launch over all of them.  Readbacks keep the
reference's attribute names and global layout: ``coords.numpy()`` returns the
``MAX_NODES x 2`` array with each domain at its node offset, exactly as the Warp
array would hold it.
"""

import numpy as np

from . import _native
from .definitions import DomainType
from .flexcontact import REBUILD_INTERVAL
from .rigid import KinematicRigids
from .scene import GROUND_RIGID, add_flex_record, pack_domain, update_rigid_spec


class DeviceField:
    """Read-only view with the ``.numpy()`` accessor of a Warp array."""

    def __init__(self, manager, field, capacity, width, dtype, per_node=True):
        self._m = manager
        self._field = field
        self._capacity = capacity
        self._width = width
        self._dtype = dtype
        self._per_node = per_node

    @property
    def shape(self):
        return (self._capacity, self._width) if self._width > 1 else (self._capacity,)

    def numpy(self):
        return self._m._gather(self._field, self._capacity, self._width, self._dtype, self._per_node)


class ScalarField:
    """Length-1 counter array read as ``arr.numpy()[0]`` (flatworld/femspringmanager.py:558-563)."""

    def __init__(self, value):
        self._v = int(value)

    def numpy(self):
        return np.array([self._v], dtype=np.int32)


class FemSpringManager:
    """FEM and spring-mass domains of one loop, resident in one HIP arena.

    ``grounds_of`` maps a domain index (position in ``domains``) to its ordered
    list of ``(ground_domain, contact)`` pairs; ExplicitLoop fills it.  ``flex``,
    when the loop has FEM-FEM contacts, is the ``flexcontact.FlexGroupSpec`` over
    the loop's FEM bodies in this manager's order.
    """

    def __init__(self, d, domains, spatial_hash=None, grounds_of=None, device=0, track_disp=True, arena_kwargs=None,
                 flex=None, mesh_set=None):
        if d != 2:
            raise ValueError(f"FemSpringManager supports d=2 only (got {d})")
        self.d = d
        self.aabb = None
        # FEM and spring domains in the order given (flatworld/femspringmanager.py:647-658)
        fem = [(idx, dom) for idx, dom in enumerate(domains) if dom.type in (DomainType.FEM, DomainType.SPRINGMASS)]
        count_nodes = sum(dom.nnodes for _, dom in fem)
        count_elements = sum(dom.nelements for _, dom in fem)
        self.MAX_NODES = max(count_nodes + 1024, 2048)
        self.MAX_ELEMENTS = max(count_elements + 2048, 4096)
        self.MAX_DOMAINS = max(len(domains), 128)
        self.domains = []
        self.domainMeshes = []
        self.stableTime = 1.0
        # the scenes are tiled and packed on the host first, so that the arena takes the
        # HBM they need (plus the FEM-FEM group's block), not a share of the free memory:
        # several loops then live side by side on one GPU
        arena_kwargs = dict(arena_kwargs or {})
        shape = {k: arena_kwargs[k] for k in ("tile_nodes", "tile_slots", "block", "grid_per_cu") if k in arena_kwargs}
        shape = dict(dict(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0), **shape)
        grounds_of = grounds_of or {}
        self._specs = []  # the packed scenes, in manager order
        fem_body = 0
        for idx, dom in fem:
            spec = pack_domain(dom, grounds_of.get(idx, []))
            if flex is not None and dom.type == DomainType.FEM:
                if flex.in_pair(fem_body):
                    add_flex_record(spec, flex.flex_cap(fem_body))
                fem_body += 1
            self._specs.append(spec)
        prepared = [_native.Prepared(spec, shape, track_disp) for spec in self._specs]
        if "capacity_bytes" not in arena_kwargs:
            need = sum(p.footprint for p in prepared)
            if flex is not None:
                need += flex.block_bytes_bound()
            arena_kwargs["capacity_bytes"] = need + (16 << 20)
        self._arena = _native.Arena(device=device, **arena_kwargs)
        self._scene_ids = []
        node_off, elem_off = 0, 0
        self._node_offsets = []
        self._elem_offsets = []
        conn_rows = []
        for local_idx, ((idx, dom), spec, prep) in enumerate(zip(fem, self._specs, prepared)):
            sid = self._arena.add_prepared(prep)
            self._scene_ids.append(sid)
            self._node_offsets.append(node_off)
            self._elem_offsets.append(elem_off)
            # global connectivity rows: node ids offset by the domain, -1 padded to 4
            # (flatworld/femspringmanager.py:744-749, 855-858)
            rows = np.full((dom.nelements, 4), -1, dtype=np.int32)
            rows[:, : spec.conn.shape[1]] = spec.conn + node_off
            conn_rows.append(rows)
            node_off += dom.nnodes
            elem_off += dom.nelements
            self.domains.append(dom)
            self.domainMeshes.append(getattr(dom, "mesh", None))
            dom.attach(self, local_idx)
            self.stableTime = min(dom.stableTime, self.stableTime)
        self._n_nodes = node_off
        self._n_elems = elem_off
        self._spring_rest = {k: np.asarray(dom.restLength, np.float32)
                             for k, (_, dom) in enumerate(fem) if dom.type == DomainType.SPRINGMASS}
        n_spring = len(self._spring_rest)
        self.femDomainIds = ScalarField(0)
        self.numFemDomains = ScalarField(len(fem) - n_spring)
        self.numSpringDomains = ScalarField(n_spring)
        self.totalNodes = ScalarField(node_off)
        self.totalElements = ScalarField(elem_off)
        self.totalFEMElements = ScalarField(sum(dom.nelements for _, dom in fem if dom.type == DomainType.FEM))
        self.totalFEMNodes = ScalarField(sum(dom.nnodes for _, dom in fem if dom.type == DomainType.FEM))
        self.domainNodeOffset = _OffsetField(self._node_offsets, self.MAX_DOMAINS)
        self.domainElemOffset = _OffsetField(self._elem_offsets, self.MAX_DOMAINS)
        self._global_ids = [idx for idx, _ in fem]

        N, E = self.MAX_NODES, self.MAX_ELEMENTS
        self.coords = DeviceField(self, _native.FIELD_X, N, 2, np.float32)
        self.V = DeviceField(self, _native.FIELD_V, N, 2, np.float32)
        self.U = DeviceField(self, _native.FIELD_U, N, 2, np.float32)
        self.Fext = DeviceField(self, _native.FIELD_FORCE, N, 2, np.float32)
        self.mass = DeviceField(self, _native.FIELD_MASS, N, 1, np.float32)
        self.elemStress = DeviceField(self, _native.FIELD_STRESS, E, 3, np.float32, per_node=False)
        self.elemStrain = DeviceField(self, _native.FIELD_STRAIN, E, 3, np.float32, per_node=False)
        self.plasticStrain = DeviceField(self, _native.FIELD_EPS_P, E, 3, np.float32, per_node=False)
        self.eqPlasticStrain = DeviceField(self, _native.FIELD_EQPS, E, 1, np.float32, per_node=False)
        self.elemVolume = DeviceField(self, _native.FIELD_AREA, E, 1, np.float32, per_node=False)
        self.J_inv = DeviceField(self, _native.FIELD_JINV, E, 4, np.float32, per_node=False)
        self.yieldFlag = DeviceField(self, _native.FIELD_YIELD, E, 1, np.uint8, per_node=False)
        self.contactFlag = DeviceField(self, _native.FIELD_CONTACT, N, 1, np.uint8)
        self.A = _ZeroField((N, 2))
        conn = np.zeros((E, 4), dtype=np.int32)
        if conn_rows:
            allrows = np.concatenate(conn_rows)
            conn[: allrows.shape[0]] = allrows
        self.connectivity = _HostField(conn)
        # FEM-FEM contact group and its hash schedule (flatworld/femspringmanager.py:520-524)
        self._flex = flex
        self._flex_gid = None
        self._sh_rebuild_interval = REBUILD_INTERVAL
        self._sh_fem_elapsed = 0.0
        self._ties_done = False
        if flex is not None:
            bodies = [sid for sid, (_, dom) in zip(self._scene_ids, fem) if dom.type == DomainType.FEM]
            self._flex_gid = self._arena.flex_add(flex, bodies)
        # kinematic rigid records: (scene, record index, record) per FEM-rigid pair
        self._rigid_records = [(sid, gi, g) for sid, spec in zip(self._scene_ids, self._specs)
                               for gi, g in enumerate(spec.grounds) if g.kind == GROUND_RIGID]
        self._rigids = []
        for _, _, g in self._rigid_records:
            if all(g.rigid_domain is not r for r in self._rigids):
                self._rigids.append(g.rigid_domain)
        # the host schedule of the kinematic rigids, and the mesh set versions each
        # mesh record last received
        self._mesh_set = mesh_set
        self._kin = KinematicRigids(self._rigids, mesh_set)
        self._mesh_pushed = {}

    def sync_rigids(self, reset_at=-1):
        """Push every rigid domain's current pose to its contact records (stream-ordered);
        a mesh record also takes what changed in its mesh set since the last push and
        the step before which its skip epochs reset."""
        for sid, gi, g in self._rigid_records:
            if g.mesh_set is None:
                update_rigid_spec(g)
                self._arena.set_rigid(sid, gi, g)
                continue
            m = g.mesh_set
            r = g.rigid_domain
            g.rigid_center = r.center.astype(np.float32)
            g.rigid_vel = r.velocity.astype(np.float32)
            g.rigid_omega = np.float32(r.omega)
            g.aabb_lo, g.aabb_hi = m.aabb_of(r)
            g.mesh_reset_at = int(reset_at)
            last = self._mesh_pushed.get((sid, gi), (-1, -1, -1))
            upd = ((_native.MESH_UPDATE_COORDS if last[0] != m.v_coords else 0)
                   | (_native.MESH_UPDATE_EDGES if last[1] != m.v_edges else 0)
                   | (_native.MESH_UPDATE_HASH if last[2] != m.v_hash else 0))
            self._arena.set_rigid(sid, gi, g, mesh_update=upd)
            self._mesh_pushed[(sid, gi)] = (m.v_coords, m.v_edges, m.v_hash)

    def _moving_rigids(self):
        return [r for r in self._rigids if r.omega != 0.0 or np.any(r.velocity != 0.0)]

    # ------------------------------------------------------------ stepping
    def setGlobalAABB(self, aabb=None):
        """Attach the loop's AABB table and run the zero-dt warm-up femStep (flatworld/femspringmanager.py:872-879)."""
        self.aabb = aabb
        self._arena.step(1, 0.0, _native.FW_STEP_NO_CONTACT | _native.FW_STEP_DIAG, damping_override=0.0)

    def substep(self, dt, damping):
        """One femStep without ground contact (flatworld/femspringmanager.py:938-939)."""
        self.femStep(float(dt), float(damping))

    def femStep(self, dt, damping):
        self._arena.step(1, float(np.float32(dt)), _native.FW_STEP_NO_CONTACT | _native.FW_STEP_DIAG,
                         damping_override=float(np.float32(damping)))

    def advance_substeps(self, n, dt, damping, diag=True, dt64=None):
        """n full substeps (ground contact + femStep) with one fused launch each.

        With FEM-FEM contact, ``dt64`` (one float64 per substep, default ``dt``)
        advances the hash schedule: a rebuild due at a substep is enqueued before
        it, as run_batched_contacts does (flatworld/mixedcontact.py:1672-1678).
        Loops with moving kinematic rigids or mesh rigids are stepped one substep at
        a time: poses (and the mesh set's positions, edge boxes, hash and skip reset)
        are pushed before the substep, and the rigids advance after it
        (``rigid.KinematicRigids``)."""
        seq = [float(dt)] * int(n) if dt64 is None else [float(v) for v in dt64]
        if not self._moving_rigids() and self._mesh_set is None:
            self._advance_block(n, dt, damping, diag, dt64)
            self._kin.counter += int(n)
            if seq:
                self._kin.prev_dt = seq[-1]
            return
        for k in range(int(n)):
            self.sync_rigids(self._kin.before_substep())
            self._advance_block(1, dt, damping, diag and k == int(n) - 1, [seq[k]])
            self._kin.after_substep(seq[k])

    def _advance_block(self, n, dt, damping, diag=True, dt64=None):
        flags = _native.FW_STEP_DIAG if diag else 0
        dt32, damp32 = float(np.float32(dt)), float(np.float32(damping))
        n = int(n)
        if self._flex_gid is None:
            self._arena.step(n, dt32, flags, damping_override=damp32)
            return
        seq = [float(dt)] * n if dt64 is None else [float(v) for v in dt64]
        n_tied = self._flex.n_tied
        pending = 0
        for k in range(n):
            # the hash is refreshed while sliding items exist, or until the ties are
            # resolved (flatworld/mixedcontact.py:1668-1689)
            if n_tied < self._flex.n_items or not self._ties_done:
                self._sh_fem_elapsed += seq[k]
                if self._sh_fem_elapsed >= self._sh_rebuild_interval:
                    if pending:
                        self._arena.step(pending, dt32, flags & ~_native.FW_STEP_DIAG, damping_override=damp32)
                        pending = 0
                    self._arena.flex_rebuild(self._flex_gid)
                    self._sh_fem_elapsed = 0.0
            if n_tied > 0 and not self._ties_done:
                if pending:
                    self._arena.step(pending, dt32, flags & ~_native.FW_STEP_DIAG, damping_override=damp32)
                    pending = 0
                self._arena.flex_init_ties(self._flex_gid)
                self._ties_done = True
            pending += 1
        if pending:
            self._arena.step(pending, dt32, flags, damping_override=damp32)

    def populate_fem_spatial_hash(self, velocity_buffer=0.0):
        """Build the FEM spatial hash at the current state (flatworld/femspringmanager.py:1041-1065)."""
        if velocity_buffer != 0.0:
            raise NotImplementedError("the FEM spatial hash is built with velocity_buffer=0 only, as the reference calls it")
        if self._flex_gid is not None:
            self._arena.flex_rebuild(self._flex_gid)

    def maybe_rebuild_fem_spatial_hash(self, dt):
        """flatworld/femspringmanager.py:1035-1039."""
        self._sh_fem_elapsed += dt
        if self._sh_fem_elapsed >= self._sh_rebuild_interval:
            self.populate_fem_spatial_hash(velocity_buffer=0.0)
            self._sh_fem_elapsed = 0.0

    def flex_ties(self):
        """Resolved tie edge (-1: none) and (weight 0, weight 1, normal gap) of every work item."""
        seg = self._arena.flex_read(self._flex_gid, _native.FLEX_TIE_SEG)
        state = self._arena.flex_read(self._flex_gid, _native.FLEX_TIE_STATE)
        return seg, state[:, :3].copy()

    def flex_cache(self):
        """Cached target segment of every FEM-FEM work item (-1: none)."""
        return self._arena.flex_read(self._flex_gid, _native.FLEX_CACHE)

    def applyMassScaling(self, dt_target):
        """Selective mass scaling of every domain at the current positions
        (flatworld/femspringmanager.py:141-201, 899-924); prints the added mass."""
        before = self.mass.numpy().astype(np.float64)
        for sid in self._scene_ids:
            self._arena.mass_scaling(sid, dt_target)
        after = self.mass.numpy().astype(np.float64)
        added = float(np.sum(after - before))
        total = float(np.sum(after))
        pct = added / total * 100.0 if total > 0 else 0.0
        print(f"[MassScaling] dt_target={dt_target:.3e}  added_mass={added:.4e}  "
              f"total_mass={total:.4e}  ({pct:.2f}%)")
        return added

    def reset_gate(self):
        for sid in self._scene_ids:
            self._arena.reset_gate(sid)

    def sync(self):
        self._arena.sync()

    # ------------------------------------------------------------ readback
    def _gather(self, field, capacity, width, dtype, per_node):
        out = np.zeros((capacity, width) if width > 1 else (capacity,), dtype=dtype)
        offs = self._node_offsets if per_node else self._elem_offsets
        for k, (sid, off) in enumerate(zip(self._scene_ids, offs)):
            if field == _native.FIELD_AREA and k in self._spring_rest:
                a = self._spring_rest[k]  # a spring's elemVolume is its rest length (femspringmanager.py:860)
            else:
                a = self._arena.read(sid, field)
            out[off : off + a.shape[0]] = a
        return out

    def domain_coords(self, local_idx):
        return self._arena.read(self._scene_ids[local_idx], _native.FIELD_X)

    def domain_aabb(self, local_idx):
        bb = self._arena.read(self._scene_ids[local_idx], _native.FIELD_AABB)
        return bb[:2].copy(), bb[2:].copy()

    def arena_info(self):
        return self._arena.info()


class _OffsetField:
    def __init__(self, offsets, cap):
        self._a = np.zeros(cap, dtype=np.int32)
        self._a[: len(offsets)] = offsets

    def numpy(self):
        return self._a.copy()


class _HostField:
    """Host-held array with the ``.numpy()`` accessor (constant after construction)."""

    def __init__(self, a):
        self._a = a
        self.shape = a.shape

    def numpy(self):
        return self._a.copy()


class _ZeroField:
    """The reference's acceleration array A, which no kernel ever writes (always zero)."""

    def __init__(self, shape):
        self.shape = shape

    def numpy(self):
        return np.zeros(self.shape, dtype=np.float32)
