"""Shared helpers for the GPU-vs-oracle parity tests.

The HIP path is driven through the C ABI (flatworld_amd._native.Arena); the
oracle is the serial CPU restatement in oracle/ (the checker).  Both receive
the same SceneSpec arrays.
"""

import sys

import numpy as np

from flatworld_amd import _native
from flatworld_amd.definitions import MaterialType
from oracle.oracle import OracleScene

# Tolerances stated by the benchmark definition for floating-point parity.
FORCE_RTOL = 1e-5  # per-substep nodal force, relative to the largest nodal force
POS_TOL_EDGE = 1e-4  # positions after 100 substeps, in units of the minimum edge length


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def min_edge(spec):
    c, x = spec.conn, spec.x0
    d = np.concatenate([x[c[:, 0]] - x[c[:, 1]], x[c[:, 1]] - x[c[:, 2]], x[c[:, 2]] - x[c[:, 0]]])
    return float(np.sqrt((d * d).sum(1)).min())


class GpuScene:
    """One scene alone in its own arena."""

    def __init__(self, spec, arena=None, **arena_kw):
        self.arena = arena or _native.Arena(**arena_kw)
        self.sid = self.arena.add(spec)
        self.spec = spec

    def warmup(self):
        self.arena.step(1, 0.0, _native.FW_STEP_NO_CONTACT | _native.FW_STEP_DIAG, damping_override=0.0)

    def run(self, n, dt=None):
        dt = self.spec.dt if dt is None else dt
        self.arena.step(int(n), float(np.float32(dt)), _native.FW_STEP_DIAG)

    def read(self, field):
        return self.arena.read(self.sid, field)


def compare(gpu, ora, exact, label=""):
    """Compare state; exact=True demands bitwise-equal float32 (up to the sign of zero)."""
    x = gpu.read(_native.FIELD_X)
    v = gpu.read(_native.FIELD_V)
    f = gpu.read(_native.FIELD_FORCE)
    res = {}
    # a parity check on a blown-up scene would compare NaNs: require a finite state
    assert np.isfinite(ora.x).all() and np.isfinite(ora.fext).all(), f"{label}: oracle state is not finite"
    if exact:
        for name, a, b in (("x", x, ora.x), ("v", v, ora.v), ("force", f, ora.fext)):
            bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))).reshape(-1))
            assert bad.size == 0, f"{label}: {name} differs at {bad.size} entries, first {bad[:5]}: gpu {a.reshape(-1)[bad[:3]]} oracle {b.reshape(-1)[bad[:3]]}"
        if gpu.spec.mat_type == MaterialType.MISES:
            ep = gpu.read(_native.FIELD_EPS_P)
            eq = gpu.read(_native.FIELD_EQPS)
            for name, a, b in (("eps_p", ep, ora.eps_p), ("eqps", eq, ora.eqps)):
                bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))).reshape(-1))
                if bad.size:
                    af, bf = a.reshape(-1), b.reshape(-1)
                    log(f"  {label}: {name} mismatch count={bad.size} of {af.size}, max|d|={np.abs(af - bf).max():.3e}, "
                        f"max|ref|={np.abs(bf).max():.3e}, first={bad[:6]}, gpu={af[bad[:6]]}, oracle={bf[bad[:6]]}, "
                        f"oracle nonzero={np.count_nonzero(bf)}, gpu nonzero={np.count_nonzero(af)}")
                assert bad.size == 0, f"{label}: {name} differs"
    else:
        h = min_edge(gpu.spec)
        dx = float(np.abs(x - ora.x).max())
        fmax = float(np.abs(ora.fext).max())
        df = float(np.abs(f - ora.fext).max())
        res = dict(pos_err_edges=dx / h, force_rel=df / max(fmax, 1e-30))
    return res
