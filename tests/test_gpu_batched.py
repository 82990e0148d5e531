"""Parity of the HIP batched form (device-driven build: histogram replay,
LDS-resident subtrees) with the CPU oracle, bit for bit.  Needs an MI355X."""
import numpy as np
import pytest
from gpu_helpers import assert_same_build, bin_edge_keys, fbits, run_both

import inputs

pytestmark = pytest.mark.gpu

BATCHED = 1


@pytest.mark.parametrize("levels_mode", [0, 1])
def test_batched_orbit_20_10(orb, oracle, levels_mode):
    x, y, z = oracle.generate(1 << 20)
    g, r = run_both(orb, oracle, x, y, z, 1024, levels_mode, form=BATCHED)
    assert_same_build(g, r)


@pytest.mark.parametrize("cmp,partition", [(1, 0), (1, 1)])
def test_batched_gpu_semantics(orb, oracle, cmp, partition):
    x, y, z = oracle.generate(1 << 18)
    g, r = run_both(orb, oracle, x, y, z, 256, 1, cmp, partition, form=BATCHED)
    assert_same_build(g, r)


def test_batched_tie_heavy(orb, oracle):
    x, y, z = inputs.quantized((1 << 18) + 17, levels=512, seed=8)
    g, r = run_both(orb, oracle, x, y, z, 512, 1, form=BATCHED)
    assert_same_build(g, r)


@pytest.mark.parametrize("n,d", [(5, 2), (1000, 7), (4099, 12), (33333, 100), (70000, 1),
                                 (3000, 64), (2048, 128), (2049, 128)])
def test_batched_small_and_odd(orb, oracle, n, d):
    x, y, z = oracle.generate(n)
    g, r = run_both(orb, oracle, x, y, z, d, 1, form=BATCHED)
    assert_same_build(g, r)


@pytest.mark.parametrize("n,d,levels_mode", [(1 << 16, 1024, 1), (1 << 16, 1024, 0),
                                             (1 << 18, 1 << 14, 1), (100000, 777, 1)])
def test_batched_subtree_levels(orb, oracle, n, d, levels_mode):
    """Deep trees: the last levels run in the LDS-resident subtree kernel."""
    x, y, z = oracle.generate(n)
    g, r = run_both(orb, oracle, x, y, z, d, levels_mode, form=BATCHED)
    assert_same_build(g, r)


def test_batched_clustered_mixed_paths(orb, oracle):
    """Uneven cells: some reach the subtree size early, others stay large."""
    x, y, z = inputs.clustered(1 << 18, n_spheres=16, a=0.01, seed=5)
    g, r = run_both(orb, oracle, x, y, z, 1 << 10, 1, form=BATCHED)
    assert_same_build(g, r)


def test_batched_repeatable_and_equal_to_plain(orb, oracle):
    n, d = 1 << 19, 512
    ctx = orb.Context(n, d, generate=True)
    ctx.save_particles()
    outs = []
    for form in (0, 1, 1):
        ctx.restore_particles()
        info = ctx.build(form, orb.LEVELS_FULL)
        outs.append((info["level_passes"], ctx.heap(), ctx.ranges(), ctx.get_particles()))
    ctx.close()
    for o in outs[1:]:
        assert o[0] == outs[0][0]
        assert np.array_equal(o[2], outs[0][2])
        for a, b in zip(o[3], outs[0][3]):
            assert np.array_equal(fbits(a), fbits(b))
        for name in ("cutMarginLeft", "cutMarginRight"):
            assert np.array_equal(fbits(o[1][name]), fbits(outs[0][1][name]))


def test_batched_graph_replays_match_eager(orb, oracle):
    """The captured build replayed several times, with eager work between the
    replays, gives the eager build's results every time."""
    n, d = (1 << 18) + 5, 300
    x, y, z = oracle.generate(n)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    ctx.save_particles()
    ctx.enable_graph(False)
    ctx.enable_trace(True)
    ref_info = ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
    ref = (ctx.heap(), ctx.ranges(), ctx.get_particles(), ctx.pass_counts())
    ctx.enable_graph(True)
    for rep in range(4):
        ctx.restore_particles()
        info = ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
        assert info["level_passes"] == ref_info["level_passes"], rep
        assert np.array_equal(ctx.ranges(), ref[1]), rep
        for a, b in zip(ctx.get_particles(), ref[2]):
            assert np.array_equal(fbits(a), fbits(b)), rep
        for name in ("cutMarginLeft", "cutMarginRight"):
            assert np.array_equal(fbits(ctx.heap()[name]), fbits(ref[0][name])), rep
        assert np.array_equal(ctx.pass_counts(), ref[3]), rep
        # eager work on the same context between replays
        ctx.restore_particles()
        ctx.build(orb.FORM_PLAIN, orb.LEVELS_REFERENCE)
        ctx.restore_particles()
    ctx.close()
    r = oracle.build(x, y, z, d, levels_mode=1)
    assert np.array_equal(ref[1], r.ranges)


def test_autotune_keeps_results_exact(orb, oracle):
    """Every shape the autotuner chooses between is exact: the tuned build
    equals the oracle bit for bit, and the tuner leaves the particles as it
    found them."""
    n, d = 1 << 18, 1 << 12   # deep enough for the subtree kernel
    x, y, z = oracle.generate(n)
    r = oracle.build(x, y, z, d, levels_mode=1, pass_counts=True)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    ctx.enable_trace(True)
    t = ctx.autotune(orb.LEVELS_FULL, repeats=1)
    assert t["subtree_max_keys"] in (-1, 4400, 2200, 0)
    assert set(t["chunk_blocks"]) <= {0, 16384, 4096, 1024}
    assert set(t["subtree_blocks"]) <= {0, 256, 1024, 4096}
    assert t["tuned_ms"] <= t["default_ms"]
    gx, _, _ = ctx.get_particles()
    assert np.array_equal(fbits(gx), fbits(x))           # input order restored
    assert tuple(ctx.ranges()[0]) == (0, n)
    for _ in range(2):                                    # eager, then graph replay
        ctx.set_particles(x, y, z)
        info = ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
        g = {"info": info, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
             "xyz": ctx.get_particles(), "pc": ctx.pass_counts()}
        assert_same_build(g, r)
    ctx.close()


def test_tune_rows_persist_and_apply(orb, oracle, tmp_path):
    """The autotuner's choices as keyed rows survive the text table and apply
    to a fresh context: level by level the same shapes, and the build stays
    exact; a context of other n takes the rows whose cell-size keys it
    shares."""
    n, d = 1 << 18, 1 << 12
    x, y, z = oracle.generate(n)
    r = oracle.build(x, y, z, d, levels_mode=1)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    t = ctx.autotune(orb.LEVELS_FULL, repeats=1)
    rows = ctx.tune_rows()
    assert [row["depth"] for row in rows] == list(range(1, len(rows) + 1))
    assert all(row["ms"] > 0 for row in rows)
    ctx.close()
    p = tmp_path / "tune.txt"
    orb.write_tune_table(p, rows)
    back = orb.read_tune_table(p)
    assert back == rows
    fresh = orb.Context(n, d, generate=False)
    assert fresh.set_tune_rows(back) == len(rows)
    got = fresh.tuning()
    assert got["subtree_max_keys"] == t["subtree_max_keys"]
    assert got["chunk_blocks"] == t["chunk_blocks"]
    assert got["subtree_blocks"] == t["subtree_blocks"]
    fresh.set_particles(x, y, z)
    info = fresh.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
    g = {"info": info, "heap": fresh.heap(), "ranges": fresh.ranges(), "exists": fresh.exists(),
         "xyz": fresh.get_particles(), "pc": None}
    assert_same_build(g, r, trace=False)
    fresh.close()
    other = orb.Context(2 * n, 2 * d, generate=False)  # same cell sizes, one level deeper
    matched = other.set_tune_rows(back)
    assert matched == len(rows)                         # every key but the new root's
    other.close()


@pytest.mark.parametrize("shape", [
    {"subtree_max_keys": 0, "nlevels": 12, "chunk_blocks": [1024] * 12,
     "subtree_blocks": [1024] * 12},
    {"subtree_max_keys": 2200, "nlevels": 12, "chunk_blocks": [4096] * 12,
     "subtree_blocks": [256] * 12},
])
def test_explicit_shapes_are_exact(orb, oracle, shape):
    n, d = (1 << 18) + 5, 3000
    x, y, z = oracle.generate(n)
    r = oracle.build(x, y, z, d, levels_mode=1)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_tuning(shape)
    got = ctx.tuning()
    assert got["subtree_max_keys"] == shape["subtree_max_keys"]
    assert got["chunk_blocks"] == shape["chunk_blocks"]
    for _ in range(2):
        ctx.set_particles(x, y, z)
        info = ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
        g = {"info": info, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
             "xyz": ctx.get_particles(), "pc": None}
        assert_same_build(g, r, trace=False)
    ctx.set_tuning(None)
    assert ctx.tuning()["subtree_max_keys"] == -1
    ctx.close()


@pytest.mark.parametrize("grid_bits", [13, 16])
def test_batched_capped_and_found_cells(orb, oracle, grid_bits):
    """Keys on a 2^-13 / 2^-16 grid: some cells end at the 32-pass cap (cut at
    a midpoint no round counted), others are found, some at a candidate equal
    to its neighbour; the partition's counts must be exact in every case."""
    rng = np.random.default_rng(12)
    n = (1 << 20) + 3
    x, y, z = ((rng.integers(0, 1 << grid_bits, n) / (1 << grid_bits) - 0.5).astype(np.float32)
               for _ in range(3))
    g, r = run_both(orb, oracle, x, y, z, 2048, 1, form=BATCHED)
    assert_same_build(g, r)


@pytest.mark.parametrize("cmp", [0, 1])
def test_batched_keys_on_bin_edges(orb, oracle, cmp):
    rng = np.random.default_rng(11)
    n = (1 << 17) + 3
    x = bin_edge_keys(n, rng)
    y = (rng.random(n) - 0.5).astype(np.float32)
    z = (rng.random(n) - 0.5).astype(np.float32)
    g, r = run_both(orb, oracle, x, y, z, 64, 1, cmp, 0, form=BATCHED)
    assert_same_build(g, r)


@pytest.mark.parametrize("cmp", [0, 1])
def test_batched_dense_window(orb, oracle, cmp):
    """Almost every key of the root cell lies within 1e-6 of the median while
    a few outliers at +-0.5 set its box: the second round's window holds
    nearly the whole cell, so the gather for rounds 3 and 4 fills its lists to
    their full span and those rounds bin every key of the cell."""
    rng = np.random.default_rng(21)
    n = (1 << 18) + 5
    x = (rng.standard_normal(n) * 1e-7).astype(np.float32)
    x[:6] = np.float32(-0.5)
    x[6:12] = np.float32(0.5)
    y = ((rng.random(n) - 0.5) * 0.9).astype(np.float32)
    z = ((rng.random(n) - 0.5) * 0.9).astype(np.float32)
    g, r = run_both(orb, oracle, x, y, z, 256, 1, cmp, 0, form=BATCHED)
    assert_same_build(g, r)


@pytest.mark.parametrize("cmp", [0, 1])
def test_batched_subtree_root_on_bin_edges(orb, oracle, cmp):
    """A root cell small enough for the LDS subtree kernel (8000 keys, 32
    leaves) whose keys all sit on or next to its first-round candidate cuts:
    every key of the first round takes the exact walk, so each wave's queue of
    such keys fills and is walked in the middle of its part as well as at its
    end."""
    rng = np.random.default_rng(31)
    n = 8000
    x = bin_edge_keys(n, rng)
    y = (rng.random(n) - 0.5).astype(np.float32)
    z = (rng.random(n) - 0.5).astype(np.float32)
    g, r = run_both(orb, oracle, x, y, z, 32, 1, cmp, 0, form=BATCHED)
    assert_same_build(g, r)

