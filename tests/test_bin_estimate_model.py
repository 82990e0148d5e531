"""The first histogram round bins most keys from the arithmetic estimate
(v - lo) * 256 / (hi - lo) alone and consults the 255-entry candidate table
only near a bin edge.  This states that shortcut in numpy float32 (the GPU
code rounds the same way: no contraction, correctly rounded division) and
checks it against the table walk on margins of every kind: unit boxes,
windows a few ulps wide, large magnitudes, subnormal ones, margins near the
overflow guard.  Claim: whenever the shortcut gives bin e, candidate e is
strictly below the key and candidate e + 1 strictly above (so the key ties
no candidate), for the `<` and `<=` histograms alike.

Candidates follow the bisection controller's cut, cell.h:74-76: float sum,
halved in double, narrowed to float."""
import numpy as np
import pytest

F32 = np.float32
KBINS = 256


def cell_cut(lo, hi):
    s = (np.asarray(hi, F32) + np.asarray(lo, F32)).astype(F32)
    return (s.astype(np.float64) / 2.0).astype(F32)


def candidates(lo, hi):
    """cand[t] for t in [1, 255] by descending the depth-8 midpoint tree."""
    t = np.arange(1, KBINS)
    L = np.full(t.shape, lo, F32)
    R = np.full(t.shape, hi, F32)
    node = np.full(t.shape, KBINS // 2)
    step = KBINS // 4
    done = node == t
    for _ in range(8):
        m = cell_cut(L, R)
        left = ~done & (t < node)
        right = ~done & (t > node)
        R = np.where(left, m, R)
        L = np.where(right, m, L)
        node = np.where(left, node - step, np.where(right, node + step, node))
        step //= 2
        done = node == t
    assert done.all()
    c = cell_cut(L, R)
    return np.concatenate([[-np.inf], c, [np.inf]]).astype(F32)


def tolerance(lo, hi, scale):
    m = max(abs(F32(lo)), abs(F32(hi)))
    if not m < F32(2.0 ** 126):
        return F32(1.0)
    with np.errstate(over="ignore"):
        return F32(F32(2.0 ** -13) + F32(F32(m + F32(2.0 ** -120)) * F32(F32(2.0 ** -20) * scale)))


def estimate(v, lo, scale, tol):
    """The bin, or -1: the fraction of the estimate must lie within
    lim = 0.5 - tol of 0.5 (rounded as float32 does)."""
    lim = F32(F32(0.5) - tol)
    with np.errstate(over="ignore", invalid="ignore"):
        ef = ((v - F32(lo)).astype(F32) * scale).astype(F32)
        fr = (ef - np.floor(ef)).astype(F32)
        ok = np.abs((fr - F32(0.5)).astype(F32)) < lim
    return np.where(ok, np.nan_to_num(ef, nan=0.0, posinf=0.0).astype(np.int64), -1)


def check(lo, hi, keys):
    lo, hi = F32(lo), F32(hi)
    width = F32(hi - lo)
    with np.errstate(over="ignore", divide="ignore"):
        scale = F32(F32(KBINS) / width) if width > 0 else F32(0.0)
    tol = tolerance(lo, hi, scale)
    cand = candidates(lo, hi)
    keys = keys[(keys >= lo) & (keys <= hi)].astype(F32)
    e = estimate(keys, lo, scale, tol)
    hit = e >= 0
    eh, vh = e[hit], keys[hit]
    # the claim: strictly between candidates e and e + 1
    assert np.all(cand[eh] < vh), (lo, hi)
    assert np.all(vh < cand[eh + 1]), (lo, hi)
    # and the table's own answer agrees for both comparisons
    assert np.array_equal(np.searchsorted(cand[1:KBINS], vh, side="right"), eh)
    assert np.array_equal(np.searchsorted(cand[1:KBINS], vh, side="left"), eh)
    return hit.mean() if len(keys) else 1.0


def probe_keys(lo, hi, cand, rng, n=20000):
    """Random keys plus every candidate and its float neighbours."""
    lo, hi = F32(lo), F32(hi)
    u = rng.random(n)
    with np.errstate(over="ignore", invalid="ignore"):
        r = (lo.astype(np.float64) + u * (hi.astype(np.float64) - lo.astype(np.float64))).astype(F32)
    c = cand[1:KBINS]
    near = np.concatenate([c, np.nextafter(c, F32(-np.inf)), np.nextafter(c, F32(np.inf)),
                           np.nextafter(np.nextafter(c, F32(np.inf)), F32(np.inf)),
                           [lo, hi, np.nextafter(lo, F32(np.inf)), np.nextafter(hi, F32(-np.inf))]])
    return np.concatenate([r, near.astype(F32)])


MARGINS = [
    (-0.5, 0.5), (0.0, 1.0), (-1.0, 0.0), (-12.3, 87.1), (0.25, 0.2539),
    (1.0, float(np.nextafter(F32(1.0), F32(2.0)))),
    (1.0, float(F32(1.0) + F32(2.0 ** -16))), (3.0e5, 3.0e5 + 1.0),
    (-3.0e30, 2.0e30), (1.0e37, 1.7e38), (-1.7e38, -1.0e38),
    (float(F32(2.0 ** 125)), float(F32(2.0 ** 126) * F32(0.999))),
    (1.0e-40, 5.0e-39), (-3.0e-42, 3.0e-42), (0.0, 1.0e-44),
    (-1.0e-30, 1.0e-30), (1.0e-20, 1.0000001e-20),
]


@pytest.mark.parametrize("lo,hi", MARGINS)
def test_shortcut_matches_table(lo, hi):
    rng = np.random.default_rng(abs(hash((lo, hi))) % (2 ** 32))
    lo, hi = F32(lo), F32(hi)
    cand = candidates(lo, hi)
    check(lo, hi, probe_keys(lo, hi, cand, rng))


def test_random_margins():
    rng = np.random.default_rng(5)
    for _ in range(400):
        e = rng.integers(-140, 127)
        a = F32(rng.uniform(-2, 2) * 2.0 ** e)
        w = F32(abs(a) * 2.0 ** -rng.integers(0, 26) + 2.0 ** -149)
        lo, hi = (a, F32(a + w)) if rng.random() < 0.5 else (F32(a - w), a)
        if not np.isfinite(lo) or not np.isfinite(hi) or not lo < hi:
            continue
        cand = candidates(lo, hi)
        check(lo, hi, probe_keys(lo, hi, cand, rng, 2000))


def test_unit_box_takes_the_shortcut():
    """On a box of ordinary size nearly every key skips the table."""
    rng = np.random.default_rng(1)
    keys = (rng.random(200000) - 0.5).astype(F32)
    assert check(F32(-0.5), F32(0.5), keys) > 0.999


def test_a_dropped_term_fails():
    """Without the candidates' error term the shortcut misbins keys next to
    candidates on a window 100 ulps wide, where the midpoints round: the
    bound is not slack."""
    lo, hi = F32(1.0e6), F32(1.0e6 + 6.25)
    cand = candidates(lo, hi)
    keys = probe_keys(lo, hi, cand, np.random.default_rng(3))
    keys = keys[(keys >= lo) & (keys <= hi)]
    scale = F32(F32(KBINS) / F32(hi - lo))
    e = estimate(keys, lo, scale, F32(2.0 ** -13))
    hit = e >= 0
    wrong = ~((cand[e[hit]] < keys[hit]) & (keys[hit] < cand[e[hit] + 1]))
    assert wrong.any()


def estimate_clamped(v, lo, scale, tol):
    """The first round's form: the estimate clamped to [-0.5, 255.5] (a
    clamped value stands for a key half a bin or more outside the margins),
    trusted only while lim = 0.5 - tol exceeds a quarter bin; -1 where the
    table decides."""
    lim = F32(F32(0.5) - tol)
    limx = lim if lim > F32(0.25) else F32(-1.0)
    with np.errstate(over="ignore", invalid="ignore"):
        ef = ((v - F32(lo)).astype(F32) * scale).astype(F32)
        ef = np.minimum(np.maximum(ef, F32(-0.5)), F32(255.5)).astype(F32)
        fr = (ef - np.floor(ef)).astype(F32)
        ok = np.abs((fr - F32(0.5)).astype(F32)) < limx
    return np.where(ok, np.trunc(ef).astype(np.int64), -1)


def check_clamped(lo, hi, keys):
    lo, hi = F32(lo), F32(hi)
    width = F32(hi - lo)
    with np.errstate(over="ignore", divide="ignore"):
        scale = F32(F32(KBINS) / width) if width > 0 else F32(0.0)
    tol = tolerance(lo, hi, scale)
    cand = candidates(lo, hi)
    keys = keys[np.isfinite(keys)].astype(F32)
    e = estimate_clamped(keys, lo, scale, tol)
    hit = e >= 0
    eh, vh = e[hit], keys[hit]
    assert np.all((eh >= 0) & (eh < KBINS)), (lo, hi)
    assert np.all(cand[eh] < vh), (lo, hi)
    assert np.all(vh < cand[eh + 1]), (lo, hi)
    assert np.array_equal(np.searchsorted(cand[1:KBINS], vh, side="right"), eh)
    assert np.array_equal(np.searchsorted(cand[1:KBINS], vh, side="left"), eh)
    return hit.mean() if len(keys) else 1.0


def outside_keys(lo, hi, rng, n=4000):
    """Keys below lo and above hi, near the margins and far from them."""
    lo, hi = F32(lo), F32(hi)
    w = max(float(hi) - float(lo), 1e-30)
    with np.errstate(over="ignore", invalid="ignore"):
        near = np.concatenate([np.nextafter(lo, F32(-np.inf)) - F32(0.0),
                               np.nextafter(hi, F32(np.inf)) + F32(0.0)], axis=None)
        d = rng.random(n) * w * np.where(rng.random(n) < 0.5, 0.01, 4.0)
        far = np.concatenate([float(lo) - d, float(hi) + d]).astype(F32)
    return np.concatenate([near.astype(F32), far])


@pytest.mark.parametrize("lo,hi", MARGINS)
def test_clamped_shortcut_matches_table(lo, hi):
    rng = np.random.default_rng(abs(hash((lo, hi, 1))) % (2 ** 32))
    lo, hi = F32(lo), F32(hi)
    cand = candidates(lo, hi)
    keys = np.concatenate([probe_keys(lo, hi, cand, rng), outside_keys(lo, hi, rng)])
    check_clamped(lo, hi, keys)


def test_clamped_random_margins():
    rng = np.random.default_rng(9)
    for _ in range(300):
        e = rng.integers(-140, 127)
        a = F32(rng.uniform(-2, 2) * 2.0 ** e)
        w = F32(abs(a) * 2.0 ** -rng.integers(0, 26) + 2.0 ** -149)
        lo, hi = (a, F32(a + w)) if rng.random() < 0.5 else (F32(a - w), a)
        if not np.isfinite(lo) or not np.isfinite(hi) or not lo < hi:
            continue
        cand = candidates(lo, hi)
        keys = np.concatenate([probe_keys(lo, hi, cand, rng, 1000), outside_keys(lo, hi, rng, 500)])
        check_clamped(lo, hi, keys)


def test_clamped_unit_box_takes_the_shortcut():
    rng = np.random.default_rng(2)
    keys = (rng.random(200000) - 0.5).astype(F32)
    assert check_clamped(F32(-0.5), F32(0.5), keys) > 0.999


def test_clamping_without_the_quarter_bin_guard_fails():
    """Trusting clamped estimates on a window where the estimate itself is
    off by more than a quarter bin misbins keys outside the margins: the
    guard is needed."""
    lo, hi = F32(1.0e6), F32(1.0e6 + 0.5)       # 8 ulps: tol well above 1/4
    width = F32(hi - lo)
    scale = F32(F32(KBINS) / width)
    cand = candidates(lo, hi)
    keys = np.concatenate([probe_keys(lo, hi, cand, np.random.default_rng(4)),
                           outside_keys(lo, hi, np.random.default_rng(4))])
    with np.errstate(over="ignore", invalid="ignore"):
        ef = ((keys - lo).astype(F32) * scale).astype(F32)
        ef = np.minimum(np.maximum(ef, F32(-0.5)), F32(255.5)).astype(F32)
        fr = (ef - np.floor(ef)).astype(F32)
    e = np.trunc(ef).astype(np.int64)
    ok = np.abs(fr - F32(0.5)) < F32(0.45)        # no guard: a fixed limit
    wrong = ok & ~((cand[e] < keys) & (keys < cand[np.minimum(e + 1, KBINS)]))
    assert wrong.any()
