"""Hoare partition on cells whose stoppers are unevenly spread: already
partitioned and reverse-sorted keys (K = 0, every key a stopper), a handful
of misplaced keys far apart, cuts near either end (the swap units then sit
on one side or the other), all-equal and two-valued keys.  Each case is
compared with the oracle's serial loop (partition.cpp:28-60): the whole
permutation, bit for bit.  Needs an MI355X."""
import numpy as np
import pytest
from gpu_helpers import fbits, run_both, assert_same_build

pytestmark = pytest.mark.gpu

F32 = np.float32


def _cell(orb, d, cut):
    c = np.zeros(1, orb.CELL_DTYPE)
    c["id"] = 0
    c["nLeafCells"] = d
    c["prevCutAxis"] = -1
    c["cutAxis"] = 0
    c["lower"] = (-1.0, -1.0, -1.0)
    c["upper"] = (1.0, 1.0, 1.0)
    # margins (c, c) give getCut() = fl(2c) / 2 = c exactly
    c["cutMarginLeft"], c["cutMarginRight"] = cut, cut
    return c


def _check(orb, oracle, keys, cut):
    n = len(keys)
    keys = keys.astype(F32)
    y = np.arange(n, dtype=F32)           # the input index rides along
    z = (-np.arange(n, dtype=F32)) - 0.5
    rc, split, ox, oy, oz = oracle.hoare(keys, y, z, 0, F32(cut))
    ctx = orb.Context(n, 4, generate=False)
    try:
        ctx.set_particles(keys, y, z)
        cells = _cell(orb, 4, F32(cut))
        if rc != 0:
            with pytest.raises(orb.OrbError) as e:
                ctx.partition(cells)
            assert e.value.code == orb.ERR_UB
            return
        ctx.partition(cells)
        gx, gy, gz = ctx.get_particles()
        assert np.array_equal(fbits(gx), fbits(ox))
        assert np.array_equal(gy, oy)
        assert np.array_equal(fbits(gz), fbits(oz))
        rg = ctx.ranges()
        assert tuple(rg[1]) == (0, split) and tuple(rg[2]) == (split, n)
    finally:
        ctx.close()


SIZES = [4097, 50001, 300007]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("q", [0.003, 0.5, 0.997])
def test_sorted_and_reversed(orb, oracle, n, q):
    keys = np.linspace(-0.9, 0.9, n, dtype=F32)
    cut = keys[int(q * (n - 1))]
    _check(orb, oracle, keys, cut)            # K = 0: nothing to swap
    _check(orb, oracle, keys[::-1].copy(), cut)  # every key a stopper


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("swaps", [1, 3, 40])
def test_few_misplaced_far_apart(orb, oracle, n, swaps):
    rng = np.random.default_rng(n + swaps)
    keys = np.sort(rng.random(n).astype(F32) - F32(0.5))
    m = n // 3
    cut = keys[m]
    left = rng.choice(m, swaps, replace=False)
    right = m + 1 + rng.choice(n - m - 1, swaps, replace=False)
    keys[left], keys[right] = keys[right].copy(), keys[left].copy()
    _check(orb, oracle, keys, cut)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("q", [0.01, 0.2, 0.8, 0.99])
def test_uniform_cut_near_the_ends(orb, oracle, n, q):
    rng = np.random.default_rng(7 * n)
    keys = rng.random(n).astype(F32) - F32(0.5)
    cut = np.sort(keys)[int(q * (n - 1))]
    _check(orb, oracle, keys, cut)


@pytest.mark.parametrize("n", SIZES)
def test_equal_and_two_valued_keys(orb, oracle, n):
    _check(orb, oracle, np.full(n, 0.25, F32), F32(0.25))
    rng = np.random.default_rng(n)
    two = np.where(rng.random(n) < 0.3, F32(1.0), F32(0.0))
    for cut in (0.0, 0.5, 1.0):
        _check(orb, oracle, two, F32(cut))


@pytest.mark.parametrize("form", [0, 1])
def test_rebuild_of_built_order(orb, oracle, form):
    """A second build from the order the first one left: every cut finds its
    cell (nearly) partitioned already, so K is tiny at every level."""
    n = (1 << 20) + 37
    x, y, z = oracle.generate(n)
    ctx = orb.Context(n, 256, generate=False)
    ctx.set_particles(x, y, z)
    ctx.build(form, 1)
    bx, by, bz = ctx.get_particles()
    ctx.close()
    g, r = run_both(orb, oracle, bx, by, bz, 256, 1, form=form)
    assert_same_build(g, r)
