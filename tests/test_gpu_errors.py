"""Error behaviour on the GPU, against the CPU oracle.

* Undefined inputs (an empty cell, or a cell with no key at or above its
  cut: the reference's Hoare loop of partition.cpp:28-60 then runs outside
  the cell): both build forms name the same cell as the oracle, the first in
  the reference's order, even when many cells of a level are undefined.
* A NaN coordinate is refused (ORB_ERR_ARG) by orb_set_particles and, per
  request, by the server; infinities are ordinary keys.
* A captured batched build is recaptured when a buffer it uses is freed and
  allocated again (particle ids turned off and on, level tables grown by a
  service call with more cells than any level)."""
import numpy as np
import pytest
from gpu_helpers import assert_same_build, fbits, run_both

import inputs

pytestmark = pytest.mark.gpu

SEMANTICS = [(0, 0), (1, 0), (1, 1)]


def undefined_cases(seed, count):
    rng = np.random.default_rng(seed)
    for k in range(count):
        dist = ("uniform", "quantized")[k % 2]
        if dist == "uniform":
            n = int(rng.integers(20, 3000))           # fewer particles than leaves: empty cells
            d = int(2 ** rng.uniform(6, 12))
        else:
            n = int(rng.integers(2000, 40000))        # few distinct keys: cuts pass every key
            d = int(2 ** rng.uniform(4, 11))
        cmp, part = SEMANTICS[int(rng.integers(0, 3))]
        yield n, d, dist, int(rng.integers(0, 1 << 30)), cmp, part, int(rng.integers(0, 2))


@pytest.mark.parametrize("seed", [3, 4])
def test_undefined_cell_named_as_the_oracle_names_it(orb, oracle, seed):
    undefined = exact = 0
    for n, d, dist, s, cmp, part, levels_mode in undefined_cases(seed, 12):
        if dist == "uniform":
            x, y, z = orb.generate(n, seed=s)
        else:
            x, y, z = inputs.quantized(n, levels=int(s % 6) + 2, seed=s)
        r = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part)
        for form in (0, 1):
            case = (n, d, dist, s, cmp, part, levels_mode, form)
            ctx = orb.Context(n, d, generate=False)
            ctx.set_semantics(cmp, part)
            ctx.set_particles(x, y, z)
            try:
                ctx.build(form, levels_mode)
                rc, cell = 0, -1
            except orb.OrbError as e:
                rc, cell = e.code, e.error_cell
            if r.rc == 0 and rc == 0:
                assert np.array_equal(ctx.ranges(), r.ranges), case
                for a, b in zip(ctx.get_particles(), (r.x, r.y, r.z)):
                    assert np.array_equal(fbits(a), fbits(b)), case
                exact += 1
            ctx.close()
            assert rc == r.rc, case
            assert cell == r.error_cell, (case, cell, r.error_cell)
            undefined += rc != 0
    assert undefined >= 6 and exact >= 2, (undefined, exact)


def test_many_empty_cells_on_one_level(orb, oracle):
    """200 particles, 1024 leaves: a level of 256 cells holds many empty ones;
    the reference's order names the one with the lowest heap id."""
    x, y, z = orb.generate(200, seed=9)
    r = oracle.build(x, y, z, 1024, levels_mode=1)
    assert r.rc == orb.ERR_UB
    for form in (0, 1):
        ctx = orb.Context(200, 1024, generate=False)
        ctx.set_particles(x, y, z)
        with pytest.raises(orb.OrbError) as e:
            ctx.build(form, 1)
        ctx.close()
        assert e.value.code == orb.ERR_UB and e.value.error_cell == r.error_cell, form


def test_nan_is_refused_and_infinities_build(orb, oracle):
    n, d = 50000, 64
    x, y, z = orb.generate(n, seed=5)
    ctx = orb.Context(n, d, generate=False)
    bad = y.copy()
    bad[1234] = np.nan
    with pytest.raises(orb.OrbError) as e:
        ctx.set_particles(x, bad, z)
    assert e.value.code == orb.ERR_ARG
    ctx.close()
    x = x.copy()
    x[[3, 77, 4000]] = [np.inf, -np.inf, np.inf]
    z = z.copy()
    z[[10, 20]] = [-np.inf, np.inf]
    for form in (0, 1):
        for cmp, part in SEMANTICS:
            g, r = run_both(orb, oracle, x, y, z, d, 1, cmp=cmp, partition=part, form=form)
            assert_same_build(g, r)


def test_server_refuses_a_nan_request_alone(orb, oracle):
    srv = orb.Server(max_active=4, levels_mode=1)
    good = [orb.generate(30000 + k, seed=k + 1) for k in range(2)]
    bad = list(orb.generate(20000, seed=77))
    bad[2] = bad[2].copy()
    bad[2][555] = np.nan
    ids = [srv.submit(*good[0], 100), srv.submit(*bad, 100), srv.submit(*good[1], 37)]
    while srv.step():
        pass
    res = [srv.fetch(i, n, d) for i, n, d in zip(ids, (30000, 20000, 30001), (100, 100, 37))]
    srv.close()
    assert res[1]["error"] == orb.ERR_ARG
    for k, (ri, d) in enumerate(((res[0], 100), (res[2], 37))):
        assert ri["error"] == 0
        r = oracle.build(*good[k], d, levels_mode=1)
        assert np.array_equal(ri["ranges"], r.ranges)
        for a, b in zip((ri["x"], ri["y"], ri["z"]), (r.x, r.y, r.z)):
            assert np.array_equal(fbits(a), fbits(b))


def test_graph_recaptured_when_buffers_move(orb, oracle):
    n, d = (1 << 18) + 11, 500
    x, y, z = orb.generate(n, seed=21)
    r = oracle.build(x, y, z, d, levels_mode=1, ids=True)
    ctx = orb.Context(n, d, generate=False)
    ctx.enable_graph(True)

    def check():
        ctx.set_particles(x, y, z)
        ctx.build(orb.FORM_BATCHED, orb.LEVELS_FULL)
        assert np.array_equal(ctx.ranges(), r.ranges)
        assert np.array_equal(ctx.ids(), r.ids)
        for a, b in zip(ctx.get_particles(), (r.x, r.y, r.z)):
            assert np.array_equal(fbits(a), fbits(b))

    ctx.enable_ids(True)
    check()                      # captured with these id buffers
    ctx.enable_ids(False)        # ids freed ...
    ctx.enable_ids(True)         # ... and allocated again
    check()
    # a service call with more cells than any level grows the level tables
    cells = np.zeros(3 * ctx.cap, orb.CELL_DTYPE)
    cells["id"] = np.arange(3 * ctx.cap) % ctx.cap
    ctx.count(cells)
    check()
    ctx.close()
