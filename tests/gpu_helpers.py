"""Shared helpers for the GPU parity tests."""
import numpy as np

F32 = np.float32


def fbits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def assert_heaps_equal(h_gpu, h_orc, exists=None):
    assert h_gpu.shape == h_orc.shape
    for name in h_orc.dtype.names:
        a, b = h_gpu[name], h_orc[name]
        if a.dtype.kind == "f":
            a, b = fbits(a), fbits(b)
        if exists is not None:
            a, b = a[exists.astype(bool)], b[exists.astype(bool)]
        bad = np.nonzero(a != b)[0] if a.ndim == 1 else np.nonzero((a != b).any(axis=1))[0]
        assert bad.size == 0, f"field {name} differs at {bad[:8]}"


def run_both(orb, oracle, x, y, z, d, levels_mode, cmp=0, partition=0, form=0, trace=True):
    n = len(x)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_semantics(cmp, partition)
    ctx.set_particles(x, y, z)
    if trace:
        ctx.enable_trace(True)
    info = ctx.build(form, levels_mode)
    r = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=partition,
                     pass_counts=trace)
    assert r.rc == 0
    g = {"info": info, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
         "xyz": ctx.get_particles(), "pc": ctx.pass_counts() if trace else None}
    ctx.close()
    return g, r


def assert_same_build(g, r, trace=True):
    assert g["info"]["nlevels"] == r.nlevels
    assert g["info"]["level_passes"] == list(r.level_passes)
    assert g["info"]["total_passes"] == r.total_passes
    assert np.array_equal(g["exists"], r.exists)
    assert_heaps_equal(g["heap"], r.heap, r.exists)
    assert np.array_equal(g["ranges"], r.ranges)
    for a, b in zip(g["xyz"], (r.x, r.y, r.z)):
        assert np.array_equal(fbits(a), fbits(b))
    if trace:
        assert np.array_equal(g["pc"], r.pass_counts)


def bin_edge_keys(n, rng):
    """Keys at the root cell's first-round candidate cuts (-0.5 + j/256, exact
    midpoints of [-0.5, 0.5]) and one float step either side: the arithmetic
    bin estimate of every key sits on or next to a bin edge, so the exact
    table or walk decides every bin (the first level then runs into the
    32-pass cap)."""
    F = np.float32
    c = (F(-0.5) + np.arange(1, 256, dtype=F) / F(256)).astype(F)
    pool = np.concatenate([c, np.nextafter(c, F(-1)), np.nextafter(c, F(1))])
    return rng.choice(pool, n).astype(F)
