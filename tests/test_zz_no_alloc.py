"""fw_step_noalloc: fw_step without allocation, its FEM-FEM contact scratch in a fixed
per-GPU array.  Its results must equal fw_step's bit for bit, on two calls in a row
whose second has the smaller input (so the array still holds the first call's
contents), when calls alternate with fw_step on one arena, and without contact groups.
"""

import numpy as np
import pytest

import flatworld_amd as fw
from flatworld_amd import _native
from flatworld_amd import workloads as W
from tests.parity import log

pytestmark = pytest.mark.gpu

FIELDS = (_native.FIELD_X, _native.FIELD_V, _native.FIELD_FORCE, _native.FIELD_CONTACT)


def _loops(n, seed):
    return W.contact_batched_config(n_groups=n, seed=seed)


def _arena(groups):
    a = _native.Arena(tile_nodes=256)
    sids, gids = [], []
    for sp, g in groups:
        ids = [a.add(s) for s in sp]
        gids.append(a.flex_add(g, ids))
        sids.append(ids)
    a.step(1, -1.0, 0)  # warm-up substep
    a.flex_rebuild(-1)
    a.step(0)  # tables uploaded, as fw_step_noalloc requires after admission
    return a, sids, gids


def _assert_same(a, sa, ga, b, sb, gb, label):
    for ia, ib in zip(sa, sb):
        for x, y in zip(ia, ib):
            for fld in FIELDS:
                assert np.array_equal(a.read(x, fld), b.read(y, fld)), f"{label}: scene {x} field {fld}"
    for g, h in zip(ga, gb):
        assert np.array_equal(a.flex_read(g, _native.FLEX_CACHE), b.flex_read(h, _native.FLEX_CACHE)), f"{label}: cache"


def test_noalloc_equals_step_two_calls_second_smaller():
    big = _loops(6, seed=11)
    log("[case] fw_step_noalloc vs fw_step: 6 loops, then 2 loops")
    a, sa, ga = _arena(big)
    b, sb, gb = _arena(big)
    # the upper blocks reach the lower ones after a few hundred substeps; the hashes are
    # rebuilt between calls about as often as the reference's schedule does
    for rep in range(4):
        if rep:
            a.flex_rebuild(-1)
            b.flex_rebuild(-1)
        a.step(150, -1.0, _native.FW_STEP_DIAG)
        b.step_noalloc(150, -1.0, _native.FW_STEP_DIAG)
        b.sync()
        _assert_same(a, sa, ga, b, sb, gb, f"6 loops, call {rep}")
    touched = sum(int(np.count_nonzero(a.flex_read(g, _native.FLEX_CACHE) >= 0)) for g in ga)
    assert touched > 0, "the loops should be in FEM-FEM contact"
    # second call, smaller input: a new arena with fewer loops; the shared array still
    # holds the first call's scratch
    small = _loops(2, seed=12)
    c, sc, gc = _arena(small)
    d, sd, gd = _arena(small)
    for rep in range(4):
        if rep:
            c.flex_rebuild(-1)
            d.flex_rebuild(-1)
        c.step(150, -1.0, _native.FW_STEP_DIAG)
        d.step_noalloc(150, -1.0, _native.FW_STEP_DIAG)
        d.sync()
        _assert_same(c, sc, gc, d, sd, gd, f"2 loops, call {rep}")
    assert sum(int(np.count_nonzero(c.flex_read(g, _native.FLEX_CACHE) >= 0)) for g in gc) > 0
    for ar in (a, b, c, d):
        ar.close()


def test_noalloc_alternating_with_step():
    """Calls of both entry points on one arena hand the contribution lists over correctly."""
    groups = _loops(4, seed=13)
    log("[case] fw_step_noalloc alternating with fw_step on one arena")
    a, sa, ga = _arena(groups)
    b, sb, gb = _arena(groups)
    for rep, n in enumerate((150, 145, 150, 150)):
        a.step(n, -1.0, _native.FW_STEP_DIAG)
        if rep % 2 == 0:
            b.step_noalloc(n, -1.0, _native.FW_STEP_DIAG)
        else:
            b.step(n, -1.0, _native.FW_STEP_DIAG)
        a.flex_rebuild(-1)  # a rebuild between calls, about as often as the reference's schedule
        b.flex_rebuild(-1)
        b.sync()
        _assert_same(a, sa, ga, b, sb, gb, f"call {rep}")
    assert sum(int(np.count_nonzero(a.flex_read(g, _native.FLEX_CACHE) >= 0)) for g in ga) > 0
    a.close()
    b.close()


def test_noalloc_without_groups_and_refusals():
    specs = W.batched_config(n_scenes=5, seed=3, side_lo=6, side_hi=30)
    log("[case] fw_step_noalloc without contact groups; refusals")
    a = _native.Arena(tile_nodes=256)
    b = _native.Arena(tile_nodes=256)
    ia = [a.add(s) for s in specs]
    ib = [b.add(s) for s in specs]
    with pytest.raises(ValueError, match="fw_step"):
        b.step_noalloc(1)  # tables not uploaded yet
    b.step(0)
    a.step(61, -1.0, _native.FW_STEP_DIAG)
    b.step_noalloc(61, -1.0, _native.FW_STEP_DIAG)
    b.sync()
    for x, y in zip(ia, ib):
        for fld in FIELDS:
            assert np.array_equal(a.read(x, fld), b.read(y, fld))
    a.close()
    b.close()
