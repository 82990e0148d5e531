"""Workload-weighted builds (w(p_i), documentation/notes.tex:259-271, 326-336)
on the GPU against the oracle's weighted build, bit for bit: the cell heap,
the range map, the particles and their weights in final order, the passes
per level and the left weight of every pass.  Plain and batched forms, the
three reference semantics, the LDS subtree path, weight sums beyond 2^32
(the float rounding of the termination rule on 64-bit sums), unit weights
against the unweighted build, and the weighted services."""
import os

import numpy as np
import pytest
from gpu_helpers import assert_heaps_equal, bin_edge_keys, fbits

import inputs

pytestmark = pytest.mark.gpu

SEMANTICS = [(0, 0), (1, 0), (1, 1)]


def host_threads():
    return max(1, len(os.sched_getaffinity(0)))


def run_weighted(orb, oracle, x, y, z, w, d, levels_mode, cmp=0, part=0, form=0, graph=True):
    n = len(x)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_semantics(cmp, part)
    ctx.set_particles(x, y, z)
    ctx.set_weights(w)
    ctx.enable_trace(True)
    ctx.enable_graph(graph)
    info = ctx.build(form, levels_mode)
    g = {"info": info, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
         "xyz": ctx.get_particles(), "w": ctx.weights(), "pw": ctx.pass_weights()}
    ctx.close()
    r = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part, weights=w,
                     pass_weights=True, threads=host_threads())
    return g, r


def assert_same_weighted(g, r):
    assert r.rc == 0
    assert g["info"]["level_passes"] == list(r.level_passes)
    assert np.array_equal(g["exists"], r.exists)
    assert_heaps_equal(g["heap"], r.heap, r.exists)
    assert np.array_equal(g["ranges"], r.ranges)
    for a, b in zip(g["xyz"], (r.x, r.y, r.z)):
        assert np.array_equal(fbits(a), fbits(b))
    assert np.array_equal(g["w"], r.weights)
    assert np.array_equal(g["pw"], r.pass_weights)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("cmp,part", SEMANTICS)
def test_weighted_orbit_20_10(orb, oracle, form, cmp, part):
    n, d = 1 << 20, 1024
    x, y, z = oracle.generate(n)
    w = np.random.default_rng(3).integers(1, 9, n).astype(np.uint32)
    g, r = run_weighted(orb, oracle, x, y, z, w, d, 1, cmp, part, form)
    assert_same_weighted(g, r)


@pytest.mark.parametrize("form", [0, 1])
def test_weighted_reference_levels(orb, oracle, form):
    n, d = (1 << 19) + 3, 1024
    x, y, z = oracle.generate(n)
    w = np.random.default_rng(4).integers(1, 100, n).astype(np.uint32)
    g, r = run_weighted(orb, oracle, x, y, z, w, d, 0, form=form)
    assert_same_weighted(g, r)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("form", [0, 1])
def test_weighted_sums_beyond_2e32(orb, oracle, form):
    """2^24 + 5 particles of weight up to 1000: the root's weight is ~8.4e9,
    so every pass rounds 64-bit sums to float32."""
    n, d = (1 << 24) + 5, 1024
    x, y, z = orb.generate(n, seed=24)
    w = np.random.default_rng(5).integers(1, 1001, n).astype(np.uint32)
    assert int(w.astype(np.uint64).sum()) > 1 << 32
    g, r = run_weighted(orb, oracle, x, y, z, w, d, 1, form=form)
    used = r.pass_weights[r.pass_weights != np.uint64(2 ** 64 - 1)]
    assert (used > np.uint64(1 << 32)).any()
    assert_same_weighted(g, r)


@pytest.mark.parametrize("cmp,part", [(0, 0), (1, 0)])
def test_weighted_subtree_levels(orb, oracle, cmp, part):
    """A deep tree: the last levels of the batched form run in the LDS
    subtree kernel, which sums weights in LDS."""
    n, d = 1 << 18, 1 << 12
    x, y, z = inputs.clustered(n, n_spheres=16, a=0.02, seed=6)
    w = np.random.default_rng(6).integers(1, 50, n).astype(np.uint32)
    g, r = run_weighted(orb, oracle, x, y, z, w, d, 1, cmp, part, form=1)
    assert_same_weighted(g, r)


def test_weighted_subtree_root_on_bin_edges(orb, oracle):
    """A weighted root cell within the LDS subtree kernel's limit (5000 keys)
    whose keys all sit on or next to its first-round candidate cuts: the
    queued keys and their weights fill the wave's queue many times over."""
    rng = np.random.default_rng(32)
    n = 5000
    x = bin_edge_keys(n, rng)
    y = (rng.random(n) - 0.5).astype(np.float32)
    z = (rng.random(n) - 0.5).astype(np.float32)
    w = rng.integers(1, 1000, n).astype(np.uint32)
    g, r = run_weighted(orb, oracle, x, y, z, w, 32, 1, 0, 0, form=1)
    assert_same_weighted(g, r)


@pytest.mark.parametrize("form", [0, 1])
def test_unit_weights_equal_unweighted_build(orb, oracle, form):
    n, d = (1 << 18) + 7, 300
    x, y, z = oracle.generate(n)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    ctx.build(form, 1)
    plain = (ctx.heap(), ctx.ranges(), ctx.get_particles())
    ctx.set_particles(x, y, z)
    ctx.set_weights(np.ones(n, np.uint32))
    ctx.build(form, 1)
    weighted = (ctx.heap(), ctx.ranges(), ctx.get_particles())
    ctx.close()
    assert_heaps_equal(weighted[0], plain[0])
    assert np.array_equal(weighted[1], plain[1])
    for a, b in zip(weighted[2], plain[2]):
        assert np.array_equal(fbits(a), fbits(b))


def test_weighted_graph_replays_and_restore(orb, oracle):
    n, d = (1 << 18) + 9, 500
    x, y, z = oracle.generate(n)
    w = np.random.default_rng(8).integers(1, 20, n).astype(np.uint32)
    r = oracle.build(x, y, z, d, levels_mode=1, weights=w)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    ctx.set_weights(w)
    ctx.save_particles()
    for rep in range(3):                         # capture, then replays
        ctx.restore_particles()
        ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
        assert np.array_equal(ctx.ranges(), r.ranges), rep
        assert np.array_equal(ctx.weights(), r.weights), rep
        ctx.restore_particles()
        ctx.build(orb.FORM_PLAIN, orb.LEVELS_FULL)  # eager work between replays
        assert np.array_equal(ctx.ranges(), r.ranges), rep
    ctx.set_weights(w)                           # weights newer than the snapshot
    with pytest.raises(orb.OrbError) as e:
        ctx.restore_particles()
    assert e.value.code == orb.ERR_STATE
    ctx.close()


def test_weighted_services(orb, oracle):
    n = 200003
    x, y, z = orb.generate(n, seed=13)
    w = np.random.default_rng(13).integers(0, 70000, n).astype(np.uint32)
    ctx = orb.Context(n, 8, generate=False)
    ctx.set_particles(x, y, z)
    with pytest.raises(orb.OrbError) as e:       # weights not set yet
        ctx.count_weight(np.zeros(1, orb.CELL_DTYPE))
    assert e.value.code == orb.ERR_STATE
    ctx.set_weights(w)
    cells = np.zeros(3, orb.CELL_DTYPE)
    for i, (ax, lo, hi, found) in enumerate(((0, -0.5, 0.5, 0), (1, -0.25, 0.1, 1),
                                            (2, -0.5, -0.3, 0))):
        cells[i]["id"] = 0
        cells[i]["cutAxis"] = ax
        cells[i]["cutMarginLeft"], cells[i]["cutMarginRight"] = lo, hi
        cells[i]["foundCut"] = found
        cells[i]["nLeafCells"] = 8
    total = ctx.count_weight(cells)
    assert (total == w.astype(np.uint64).sum()).all()
    sentinel = np.uint64(12345)
    lt = ctx.count_left_weight(cells, orb.CMP_LT, out=np.full(3, sentinel, np.uint64))
    le = ctx.count_left_weight(cells, orb.CMP_LE)
    for i, c in enumerate(cells):
        k = (x, y, z)[int(c["cutAxis"])]
        cut = np.float32((np.float64(np.float32(c["cutMarginLeft"]) + np.float32(c["cutMarginRight"]))) / 2)
        want_le = int(w[k <= cut].astype(np.uint64).sum())
        want_lt = int(w[k < cut].astype(np.uint64).sum())
        assert int(le[i]) == want_le
        assert int(lt[i]) == (int(sentinel) if c["foundCut"] else want_lt)
    ctx.close()
