"""The chunk-level Hoare split used by the GPU partition, stated in numpy and
checked against the serial loop of partition.cpp:28-60 (the oracle's
hoare()) on random, tie-heavy and sorted keys with small chunk sizes, so
every boundary case of the search shows up.

Model (per cell, chunk size CH): A(k) = L-stoppers (key not < cut) before
chunk k, B(k) = R-stoppers (key not > cut) at or after chunk k.  k* is the
last chunk with A(k*) <= B(k*); every t below max(A(k*), B(k*+1)) is a
swapped pair, none from min(B(k*), A(k*+1)) on, and in between both stoppers
of pair t lie in chunk k*.  Then i = min(L_K, R_{K-1}) (L_0 when K = 0).
The swap side is the larger of [0, i) and [i, n); every pair is owned by
exactly one chunk of that side."""
import numpy as np
import pytest

F32 = np.float32


def split_model(keys, cut, CH):
    n = len(keys)
    Lm = ~(keys < cut)
    Rm = ~(keys > cut)
    nch = (n + CH - 1) // CH
    A = np.concatenate([[0], np.cumsum([Lm[k * CH:(k + 1) * CH].sum() for k in range(nch)])])
    Rb = np.concatenate([[0], np.cumsum([Rm[k * CH:(k + 1) * CH].sum() for k in range(nch)])])
    TL, TR = int(A[-1]), int(Rb[-1])
    if n == 0 or TL == 0:
        return None
    B = TR - Rb
    ks = max(k for k in range(nch) if A[k] <= B[k])
    lo_t, hi_t = max(A[ks], B[ks + 1]), min(B[ks], A[ks + 1])
    assert lo_t <= hi_t
    Lpos = np.nonzero(Lm)[0]
    Rpos = np.nonzero(Rm)[0][::-1]          # R_t: t-th from the right
    in_chunk = lambda p: ks * CH <= p < (ks + 1) * CH
    K = int(lo_t)
    for t in range(int(lo_t), int(hi_t)):
        assert in_chunk(Lpos[t]) and in_chunk(Rpos[t])
        if Lpos[t] < Rpos[t]:
            K = t + 1
        else:
            break
    i = int(Lpos[0]) if K == 0 else int(Rpos[K - 1])
    if 0 < K < TL:
        i = min(i, int(Lpos[K]))
    # ownership: chunks of the larger side own each pair exactly once
    owner_left = 2 * i >= n
    owned = np.zeros(K, np.int64)
    for k in range(nch):
        if owner_left:
            t0, t1 = A[k], min(A[k + 1], K)
        else:
            t0, t1 = TR - Rb[k + 1], min(TR - Rb[k], K)
        if t0 < t1:
            owned[t0:t1] += 1
    assert (owned == 1).all()
    a = keys.copy()
    for t in range(K):
        a[Lpos[t]], a[Rpos[t]] = a[Rpos[t]], a[Lpos[t]]
    a[i], a[n - 1] = a[n - 1], a[i]
    return a, i


@pytest.mark.parametrize("seed", range(4))
def test_split_model_matches_serial_loop(oracle, seed):
    rng = np.random.default_rng(seed)
    checked = 0
    for it in range(250):
        n = int(rng.integers(1, 400))
        kind = it % 4
        if kind == 0:
            keys = rng.integers(0, 6, n).astype(F32)
        elif kind == 1:
            keys = rng.random(n).astype(F32)
        elif kind == 2:
            keys = np.sort(rng.integers(0, 50, n)).astype(F32)
        else:
            keys = rng.random(n).astype(F32)
            keys[rng.random(n) < 0.9] = F32(0.3)
        cut = F32(rng.choice(keys)) if rng.random() < 0.5 else F32(rng.random() * (6 if kind == 0 else 1))
        CH = int(rng.choice([1, 2, 3, 8, 64, 4096]))
        z = np.zeros(n, F32)
        rc, split, ox, _, _ = oracle.hoare(keys, z, z, 0, cut)
        m = split_model(keys, cut, CH)
        if rc != 0:
            assert m is None  # no key >= cut: the reference's loop leaves the cell
            continue
        assert m is not None
        a, i = m
        assert i == split
        assert np.array_equal(a.view(np.uint32), ox.view(np.uint32))
        checked += 1
    assert checked > 150
