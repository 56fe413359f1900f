"""Shared helpers for the GPU-vs-oracle parity tests.

The HIP path is driven through the C ABI (flatworld_amd._native.Arena); the
oracle is the serial CPU restatement in oracle/ (the checker).  Both receive
the same SceneSpec arrays.

``check_fields`` is the bit-exact comparison every parity case uses: positions,
velocities, the per-substep nodal force, the accumulated displacement U, the
contact-active flags and the boundary AABB for every scene; for triangle scenes
also the Cauchy stress, the Green strain and the plastic-yield flag of the last
substep, and for J2 scenes the plastic strain and the equivalent plastic strain.
"""

import sys

import numpy as np

from flatworld_amd import _native
from flatworld_amd.definitions import MaterialType
from oracle.oracle import OracleScene  # noqa: F401  (re-exported for the tests)

# Tolerances stated by the benchmark definition for floating-point parity.
FORCE_RTOL = 1e-5  # per-substep nodal force, relative to the largest nodal force
POS_TOL_EDGE = 1e-4  # positions after 100 substeps, in units of the minimum edge length

MAT_SPRING = 4


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def min_edge(spec):
    c, x = spec.conn, spec.x0
    d = np.concatenate([x[c[:, 0]] - x[c[:, 1]], x[c[:, 1]] - x[c[:, 2]], x[c[:, 2]] - x[c[:, 0]]])
    return float(np.sqrt((d * d).sum(1)).min())


class GpuScene:
    """One scene alone in its own arena, sized to the scene (so that arenas a test
    has not released yet do not hold the GPU's memory)."""

    def __init__(self, spec, arena=None, **arena_kw):
        if arena is None:
            shape = {k: v for k, v in arena_kw.items() if k != "capacity_bytes"}
            prep = _native.Prepared(spec, dict(dict(tile_nodes=512, tile_slots=1024, block=256, grid_per_cu=0), **shape))
            arena_kw.setdefault("capacity_bytes", prep.footprint + (16 << 20))
            arena = _native.Arena(**arena_kw)
            self.sid = arena.add_prepared(prep)
        else:
            self.sid = arena.add(spec)
        self.arena = arena
        self.spec = spec

    def warmup(self):
        self.arena.step(1, 0.0, _native.FW_STEP_NO_CONTACT | _native.FW_STEP_DIAG, damping_override=0.0)

    def run(self, n, dt=None):
        dt = self.spec.dt if dt is None else dt
        self.arena.step(int(n), float(np.float32(dt)), _native.FW_STEP_DIAG)

    def read(self, field):
        return self.arena.read(self.sid, field)


def _same(a, b):
    """Bitwise-equal float32 up to the sign of zero; NaN matches NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


def _assert_same(name, a, b, label):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{label}: {name} shape {a.shape} vs oracle {b.shape}"
    bad = np.flatnonzero(~_same(a, b).reshape(-1))
    if bad.size:
        af, bf = a.reshape(-1), b.reshape(-1)
        log(f"  {label}: {name} differs at {bad.size} of {af.size} entries, first {bad[:6]}: "
            f"gpu {af[bad[:6]]} oracle {bf[bad[:6]]}")
    assert bad.size == 0, f"{label}: {name} differs at {bad.size} entries"


def check_fields(read, ora, label="", diag=True):
    """Bit-exact comparison of one scene through ``read(field)`` against an OracleScene.

    ``diag`` says the last substep launched on the GPU carried FW_STEP_DIAG, so the
    last-substep diagnostics (nodal force, stress, strain, yield and contact flags)
    are current on both sides.  Returns the number of contact-active nodes and
    yielded elements in the last substep (so that a caller can assert that its
    scenario exercised them)."""
    spec = ora.spec
    # a parity check on a blown-up scene would compare NaNs: require a finite state
    assert np.isfinite(ora.x).all() and np.isfinite(ora.v).all(), f"{label}: oracle state is not finite"
    fields = [("x", _native.FIELD_X, ora.x), ("v", _native.FIELD_V, ora.v), ("U", _native.FIELD_U, ora.u),
              ("aabb", _native.FIELD_AABB, ora.aabb)]
    if diag:
        assert np.isfinite(ora.fext).all(), f"{label}: oracle nodal force is not finite"
        fields += [("force", _native.FIELD_FORCE, ora.fext), ("contact", _native.FIELD_CONTACT, ora.contact_flag)]
    if spec.mat_type != MAT_SPRING and diag:
        fields += [("stress", _native.FIELD_STRESS, ora.stress), ("strain", _native.FIELD_STRAIN, ora.strain),
                   ("yield", _native.FIELD_YIELD, ora.yield_flag)]
    if spec.mat_type == MaterialType.MISES:
        fields += [("eps_p", _native.FIELD_EPS_P, ora.eps_p), ("eqps", _native.FIELD_EQPS, ora.eqps)]
    for name, fld, ref in fields:
        _assert_same(name, read(fld), ref, label)
    return dict(contact=int(np.count_nonzero(ora.contact_flag)), yielded=int(np.count_nonzero(ora.yield_flag)))


def compare(gpu, ora, exact, label=""):
    """exact=True: check_fields on every field; exact=False: the tolerance figures."""
    if exact:
        return check_fields(gpu.read, ora, label)
    assert np.isfinite(ora.x).all() and np.isfinite(ora.fext).all(), f"{label}: oracle state is not finite"
    x = gpu.read(_native.FIELD_X)
    f = gpu.read(_native.FIELD_FORCE)
    h = min_edge(gpu.spec)
    dx = float(np.abs(x - ora.x).max())
    fmax = float(np.abs(ora.fext).max())
    df = float(np.abs(f - ora.fext).max())
    return dict(pos_err_edges=dx / h, force_rel=df / max(fmax, 1e-30))
