"""Tree export and particle ids: ids that move with the particles through
every partition (bit-exact against the oracle, which moves them the same
way), the stable partition's buffer flips, leaves, membership and tight leaf
boxes (checked against numpy on the oracle's final particles).  Needs an
MI355X."""
import numpy as np
import pytest
from gpu_helpers import assert_same_build, fbits, run_both

pytestmark = pytest.mark.gpu


def build_with_ids(orb, oracle, x, y, z, d, levels_mode, form, cmp=0, partition=0):
    ctx = orb.Context(len(x), d, generate=False)
    ctx.set_semantics(cmp, partition)
    ctx.enable_ids(True)
    ctx.set_particles(x, y, z)
    info = ctx.build(form, levels_mode)
    r = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=partition, ids=True)
    assert r.rc == 0
    return ctx, info, r


def oracle_leaves(r):
    cap = len(r.exists)
    out = []
    for i in range(cap):
        if not r.exists[i]:
            continue
        lc, rc = 2 * i + 1, 2 * i + 2
        if (lc < cap and r.exists[lc]) or (rc < cap and r.exists[rc]):
            continue
        out.append(i)
    out = np.array(out, np.int32)
    return out, r.ranges[out]


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("cmp,partition", [(0, 0), (1, 1)])
@pytest.mark.parametrize("n,d,levels_mode", [(1 << 16, 1024, 1), (100003, 100, 1),
                                             (1 << 16, 512, 0)])
def test_ids_follow_particles(orb, oracle, form, cmp, partition, n, d, levels_mode):
    """Hoare (with the LDS subtree in the batched form) and the stable
    partition (buffer flips on levels that hold every particle, copy-back
    on the others), plain and batched."""
    x, y, z = oracle.generate(n)
    ctx, info, r = build_with_ids(orb, oracle, x, y, z, d, levels_mode, form, cmp, partition)
    gx, gy, gz = ctx.get_particles()
    for a, b in zip((gx, gy, gz), (r.x, r.y, r.z)):
        assert np.array_equal(fbits(a), fbits(b))
    ids = ctx.ids()
    assert np.array_equal(ids, r.ids)
    assert np.array_equal(np.sort(ids), np.arange(n, dtype=np.uint32))   # a permutation
    for a, orig in zip((gx, gy, gz), (x, y, z)):
        assert np.array_equal(fbits(a), fbits(orig[ids]))                 # payload follows
    assert np.array_equal(ctx.ranges(), r.ranges)
    ctx.close()


def test_stable_semantics_builds_without_ids(orb, oracle):
    """The stable partition's buffer flips with ids off, odd and even level
    counts, both forms."""
    for d in (2, 8, 16, 100):
        x, y, z = oracle.generate(1 << 15)
        for form in (0, 1):
            g, r = run_both(orb, oracle, x, y, z, d, 1, cmp=1, partition=1, form=form)
            assert_same_build(g, r)


def test_leaves_membership_bounds(orb, oracle):
    n, d = 200001, 300
    x, y, z = oracle.generate(n)
    for form in (0, 1):
        ctx, info, r = build_with_ids(orb, oracle, x, y, z, d, 1, form)
        lids, lrg = ctx.leaves()
        oids, org = oracle_leaves(r)
        assert np.array_equal(lids, oids) and np.array_equal(lrg, org)
        assert len(lids) == d
        # membership by position: the leaf whose range holds the position
        mem = ctx.membership()
        want = np.full(n, -1, np.int32)
        for i, (b, e) in zip(oids, org):
            want[b:e] = i
        assert np.array_equal(mem, want) and (want >= 0).all()
        # by input order through the oracle's ids
        mem_in = ctx.membership(by_input_order=True)
        want_in = np.empty(n, np.int32)
        want_in[r.ids] = want
        assert np.array_equal(mem_in, want_in)
        # tight boxes: min/max of the final particles in each leaf range
        lo, hi = ctx.leaf_bounds()
        P = np.stack([r.x, r.y, r.z], axis=1)
        for j, (b, e) in enumerate(org):
            if e > b:
                assert np.array_equal(lo[j], P[b:e].min(axis=0))
                assert np.array_equal(hi[j], P[b:e].max(axis=0))
            else:
                assert np.isposinf(lo[j]).all() and np.isneginf(hi[j]).all()
        # every box lies inside its cell's box (the cut planes bound it)
        cells = r.heap[oids]
        assert (lo >= cells["lower"]).all() and (hi <= cells["upper"]).all()
        ctx.close()


def test_ids_graph_replays_and_restore(orb, oracle):
    n, d = 1 << 17, 2048
    x, y, z = oracle.generate(n)
    r = oracle.build(x, y, z, d, levels_mode=1, ids=True)
    ctx = orb.Context(n, d, generate=False)
    ctx.enable_ids(True)
    ctx.set_particles(x, y, z)
    ctx.save_particles()
    for k in range(3):
        ctx.restore_particles()
        ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
        assert np.array_equal(ctx.ids(), r.ids), k
        assert np.array_equal(fbits(ctx.get_particles()[0]), fbits(r.x)), k
    ctx.close()


def test_tree_export_errors(orb, oracle):
    x, y, z = oracle.generate(1000)
    ctx = orb.Context(1000, 4, generate=False)
    ctx.set_particles(x, y, z)
    with pytest.raises(orb.OrbError) as e:
        ctx.leaves()                     # no build yet
    assert e.value.code == orb.ERR_STATE
    ctx.build(orb.FORM_PLAIN, orb.LEVELS_FULL)
    with pytest.raises(orb.OrbError) as e:
        ctx.ids()                        # ids were never enabled
    assert e.value.code == orb.ERR_STATE
    with pytest.raises(orb.OrbError) as e:
        ctx.membership(by_input_order=True)
    assert e.value.code == orb.ERR_STATE
    assert len(ctx.leaves()[0]) == 4
    ctx.close()


def test_tipsy_snapshot_build(orb, oracle, tmp_path):
    """A clustered snapshot read from a Tipsy file builds like the same
    positions handed over directly."""
    import inputs
    x, y, z = inputs.clustered(150000, n_spheres=6, a=0.02, seed=9)
    P = np.stack([x, y, z], 1)
    path = tmp_path / "c.std"
    inputs.write_tipsy(path, gas=P[:20000], dark=P[20000:140000], star=P[140000:])
    ctx = orb.Context.from_tipsy(path, 500)
    assert ctx.n == 150000
    info = ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
    r = oracle.build(x, y, z, 500, levels_mode=1)
    g = {"info": info, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
         "xyz": ctx.get_particles(), "pc": None}
    assert_same_build(g, r, trace=False)
    ctx.close()
