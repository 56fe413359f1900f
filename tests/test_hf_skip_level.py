"""The height-field skip level: the substep kernel skips the height-field query for a
node above the highest sample plus (|highest| + |lowest|) 2^-22 (below the lowest
sample minus that margin, for a reversed field), rounded outward to float32.  The
oracle always runs the full query (flatworld/mixedcontact.py:166-217); these checks
show, on adversarial sample sets, that it reports no contact for any node at or
beyond that level, so the skip cannot change a result."""

import math

import numpy as np
import pytest

from flatworld_amd.scene import GROUND_HEIGHTFIELD, GroundSpec
from oracle.oracle import ground_force


def skip_level(h, reversed_):
    hi, lo = float(np.max(h)), float(np.min(h))
    m = (abs(hi) + abs(lo)) * 2.0 ** -22 + 1e-37
    if reversed_:
        return np.nextafter(np.float32(lo - m), np.float32(-np.inf))
    return np.nextafter(np.float32(hi + m), np.float32(np.inf))


def _hf(h, lb, ub, reverse):
    lo = np.array([lb, np.min(h) - 1.0], np.float32)
    hi = np.array([ub, np.max(h) + 1.0], np.float32)
    return GroundSpec(GROUND_HEIGHTFIELD, np.float32(1e4), lo, hi, heights=np.ascontiguousarray(h, np.float32),
                      hf_lb_x=np.float32(lb), hf_ub_x=np.float32(ub), reverse=1 if reverse else 0)


CASES = {
    "tiny-then-one": np.array([1e-8, 1.0, 1e-8, 1.0], np.float32),
    "wide-range": np.array([-3.0e6, 2.5e6, -1.0, 7.0e5, 2.5e6], np.float32),
    "flat-zero": np.zeros(8, np.float32),
    "negative": np.array([-0.3, -0.1000001, -0.2, -0.1], np.float32),
    "random": np.random.default_rng(5).normal(0.0, 0.05, 64).astype(np.float32),
    "steep": np.array([0.0, 1.0e4, 0.0, 1.0e4, 3.3e3], np.float32),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("reverse", [False, True])
def test_no_contact_beyond_the_skip_level(name, reverse):
    h = CASES[name]
    lb, ub = -0.5, 1.5
    g = _hf(h, lb, ub, reverse)
    z = skip_level(h, reverse)
    rng = np.random.default_rng(11)
    # x across the field, at the samples, between them and just inside each sample
    xs = np.concatenate([rng.uniform(lb - 0.1, ub + 0.1, 4000),
                         np.linspace(lb, ub, len(h)),
                         np.nextafter(np.linspace(lb, ub, len(h)).astype(np.float32), np.float32(-np.inf))])
    step = np.float32(-np.inf) if reverse else np.float32(np.inf)
    far = np.float32(z - 1.0) if reverse else np.float32(z + 1.0)
    for zz in (z, np.nextafter(z, step), far):
        for x in xs:
            f, hit = ground_force(g, [x, zz], [0.0, -1.0], 1e-4)
            assert not hit, (name, reverse, float(x), float(zz))
    # and the level is tight enough to matter: just inside the sample range the query does touch
    zin = np.float32(np.min(h) + 0.25 * (np.max(h) - np.min(h))) if reverse else np.float32(np.max(h) - 0.25 * (np.max(h) - np.min(h)))
    if np.max(h) > np.min(h):
        hits = [ground_force(g, [x, zin], [0.0, 0.0], 1e-4)[1] for x in np.linspace(lb, ub, 200)]
        assert any(hits)


def test_level_bounds_every_interpolated_height():
    """float32 h0 + (h1 - h0) t over a dense t grid never passes the level."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        scale = 10.0 ** rng.uniform(-8, 6)
        h = (rng.normal(0.0, 1.0, 2) * scale).astype(np.float32)
        if rng.uniform() < 0.3:
            h[0] = np.float32(h[0] * 1e-7)
        t = np.concatenate([np.linspace(0.0, 1.0, 4097), 1.0 - np.logspace(-9, -1, 64)]).astype(np.float32)
        t = t[t < 1.0]
        hh = h[0] + (h[1] - h[0]) * t  # float32 throughout
        assert hh.dtype == np.float32
        assert hh.max() <= skip_level(h, False)
        assert hh.min() >= skip_level(h, True)
