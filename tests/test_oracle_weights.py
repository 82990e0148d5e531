"""Pins of the oracle's weighted build (workload weights w(p_i),
documentation/notes.tex:259-271, 326-336), independent of its own code:

* every weight 1 is the unweighted build (itself pinned in test_oracle.py);
* a particle of integer weight k counts as k particles at the same place.
  Under the stable <= partition (o=2) the copies of one particle stay
  together, so the weighted build of (p_i, w_i) and the unweighted build of
  the multiset that repeats p_i w_i times make the same tree: the same cells
  (every field), passes and per-pass left sums, and the weighted final order
  expanded by the weights is the multiset build's final order.  Under o=0
  (Hoare) the same holds for the tree whenever no key equals its cell's cut,
  which the test checks on the result;
* the weighted termination rule against exact rational arithmetic with the
  float32 roundings applied by hand, for sums beyond 2^32."""
from fractions import Fraction

import numpy as np
import pytest

from gpu_helpers import assert_heaps_equal, fbits


def expand(x, y, z, w):
    rep = np.repeat(np.arange(len(x)), w)
    return x[rep], y[rep], z[rep]


def keys_tie_a_cut(r, xyz):
    """True if some cell's particles hold a key equal to its cut (ranges keep
    each cell's own range; later partitions only permute inside it)."""
    for cid in np.nonzero(r.exists)[0]:
        c = r.heap[cid]
        lc = 2 * cid + 1
        if lc >= len(r.exists) or not r.exists[lc]:
            continue
        b, e = r.ranges[cid]
        cut = np.float32((np.float32(c["cutMarginLeft"]) + np.float32(c["cutMarginRight"]))
                         .astype(np.float64) / 2.0)
        if np.any(xyz[int(c["cutAxis"])][b:e] == cut):
            return True
    return False


@pytest.mark.parametrize("n,d,levels_mode,cmp,part", [
    (5000, 64, 1, 0, 0), (20011, 300, 1, 1, 1), (1 << 14, 1 << 10, 0, 0, 0),
    (777, 8, 1, 1, 0)])
def test_unit_weights_are_the_unweighted_build(oracle, n, d, levels_mode, cmp, part):
    x, y, z = oracle.generate(n)
    a = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part,
                     pass_counts=True)
    b = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part,
                     weights=np.ones(n, np.uint32), pass_weights=True)
    assert a.rc == b.rc == 0
    assert list(a.level_passes) == list(b.level_passes)
    assert_heaps_equal(a.heap, b.heap)
    assert np.array_equal(a.ranges, b.ranges)
    for u, v in zip((a.x, a.y, a.z), (b.x, b.y, b.z)):
        assert np.array_equal(fbits(u), fbits(v))
    used = a.pass_counts != 0xFFFFFFFF
    assert np.array_equal(a.pass_counts[used].astype(np.uint64), b.pass_weights[used])
    assert np.all(b.pass_weights[~used] == np.uint64(0xFFFFFFFFFFFFFFFF))


@pytest.mark.parametrize("seed,n,d,wmax", [(1, 3000, 32, 5), (2, 12000, 128, 9),
                                           (3, 500, 16, 40), (4, 40000, 512, 3)])
def test_weight_k_is_k_copies_stable(oracle, seed, n, d, wmax):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-0.5, 0.5, n).astype(np.float32) for _ in range(3))
    w = rng.integers(1, wmax + 1, n).astype(np.uint32)
    rw = oracle.build(x, y, z, d, levels_mode=1, cmp=1, partition=1, weights=w,
                      pass_weights=True)
    ex, ey, ez = expand(x, y, z, w)
    rd = oracle.build(ex, ey, ez, d, levels_mode=1, cmp=1, partition=1, pass_counts=True)
    assert rw.rc == rd.rc == 0
    assert list(rw.level_passes) == list(rd.level_passes)
    assert_heaps_equal(rw.heap, rd.heap)
    used = rd.pass_counts != 0xFFFFFFFF
    assert np.array_equal(rd.pass_counts[used].astype(np.uint64), rw.pass_weights[used])
    # each cell's weight is the multiset cell's size
    cw = np.concatenate([[0], np.cumsum(rw.weights.astype(np.int64))])
    ok = rw.exists.astype(bool)
    assert np.array_equal(cw[rw.ranges[ok, 1]] - cw[rw.ranges[ok, 0]],
                          (rd.ranges[ok, 1] - rd.ranges[ok, 0]).astype(np.int64))
    # the weighted final order, expanded, is the multiset's final order
    fx, fy, fz = expand(rw.x, rw.y, rw.z, rw.weights)
    for u, v in zip((fx, fy, fz), (rd.x, rd.y, rd.z)):
        assert np.array_equal(fbits(u), fbits(v))


@pytest.mark.parametrize("seed,n,d,wmax", [(5, 4000, 64, 6), (6, 9000, 100, 4)])
def test_weight_k_is_k_copies_hoare(oracle, seed, n, d, wmax):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-0.5, 0.5, n).astype(np.float32) for _ in range(3))
    w = rng.integers(1, wmax + 1, n).astype(np.uint32)
    rw = oracle.build(x, y, z, d, levels_mode=1, cmp=0, partition=0, weights=w,
                      pass_weights=True)
    ex, ey, ez = expand(x, y, z, w)
    rd = oracle.build(ex, ey, ez, d, levels_mode=1, cmp=0, partition=0, pass_counts=True)
    assert rw.rc == rd.rc == 0
    assert not keys_tie_a_cut(rw, (rw.x, rw.y, rw.z))
    assert list(rw.level_passes) == list(rd.level_passes)
    assert_heaps_equal(rw.heap, rd.heap)
    used = rd.pass_counts != 0xFFFFFFFF
    assert np.array_equal(rd.pass_counts[used].astype(np.uint64), rw.pass_weights[used])
    # the weights moved with their particles
    key = {(a, b, c): k for a, b, c, k in zip(fbits(x), fbits(y), fbits(z), w)}
    moved = np.array([key[(a, b, c)] for a, b, c in zip(fbits(rw.x), fbits(rw.y), fbits(rw.z))])
    assert np.array_equal(moved, rw.weights)


def f32_round(q):
    """Round a Fraction to the nearest float32, ties to even (normal range)."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    ulp = Fraction(2) ** (e - 23)
    m = q / ulp
    lo = m.numerator // m.denominator
    rem = m - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    return sign * lo * ulp


def test_weighted_termination_rule_exact(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(17)
    cases = [(int(rng.integers(0, 1 << 40)), int(rng.integers(1, 1 << 41)),
              int(rng.integers(1, 1 << 20))) for _ in range(3000)]
    cases += [((1 << 33) + 1, (1 << 34) + 3, 3), (5, 10, 2), (0, 1, 1), (2 ** 50, 2 ** 51 + 1, 7)]
    for wl, wc, nl in cases:
        ratio = f32_round(Fraction(-(-nl // 2), nl))
        prod = f32_round(f32_round(Fraction(wc)) * ratio)
        diff = f32_round(f32_round(Fraction(wl)) - prod)
        want = int(diff)  # truncation toward zero
        assert L.orc_difference_w(wl, wc, nl) == want, (wl, wc, nl)
        if wl < 2 ** 32 and wc < 2 ** 32:
            assert L.orc_difference(wl, wc, nl) == want
