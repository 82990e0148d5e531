"""Parity of the HIP plain form (service entry points + host-driven build)
with the CPU oracle: bit-exact counts, cut values, cell heap, ranges and the
final particle order.  Needs an MI355X."""
import numpy as np
import pytest
from gpu_helpers import assert_same_build, fbits, run_both

pytestmark = pytest.mark.gpu

F32 = np.float32


def test_generator_matches_oracle(orb, oracle):
    n = (1 << 16) + 3
    ctx = orb.Context(n, 16, generate=True)
    gx, gy, gz = ctx.get_particles()
    ox, oy, oz = oracle.generate(n)
    for a, b in zip((gx, gy, gz), (ox, oy, oz)):
        assert np.array_equal(fbits(a), fbits(b))
    ctx.close()


def _root_cells(orb, d, margins=(-0.5, 0.5), axis=0):
    c = np.zeros(1, orb.CELL_DTYPE)
    c["id"] = 0
    c["nLeafCells"] = d
    c["prevCutAxis"] = -1
    c["cutAxis"] = axis
    c["lower"] = (-0.5, -0.5, -0.5)
    c["upper"] = (0.5, 0.5, 0.5)
    c["cutMarginLeft"], c["cutMarginRight"] = margins
    return c


def test_count_services_against_numpy(orb, oracle):
    n = 100003  # not a multiple of 4: misaligned chunk heads and tails
    x, y, z = oracle.generate(n)
    ctx = orb.Context(n, 8, generate=False)
    ctx.set_particles(x, y, z)
    for axis, col in enumerate((x, y, z)):
        for lo, hi in [(-0.5, 0.5), (-0.5, 0.0), (-0.123, 0.377), (0.2, 0.2000001)]:
            cells = _root_cells(orb, 8, (lo, hi), axis)
            cut = F32(F32(F32(lo) + F32(hi)) * F32(0.5))
            assert ctx.count(cells)[0] == n
            assert ctx.count_left(cells)[0] == np.count_nonzero(col < cut)
            assert ctx.count_left_gpu(cells)[0] == np.count_nonzero(col <= cut)
            ctx.make_axis(cells)
            assert ctx.count_left_gpu_axis(cells)[0] == np.count_nonzero(col <= cut)
            rc, out = ctx.run_service(orb.PST_COUNTLEFT, cells, 1)
            assert rc == 4 and out[0] == np.count_nonzero(col < cut)
    # found cells are skipped by ServiceCountLeft: the output slot is untouched
    cells = _root_cells(orb, 8)
    cells["foundCut"] = 1
    out = np.full(1, 12345, np.uint32)
    assert ctx.count_left(cells, out)[0] == 12345
    ctx.close()


def test_partition_service_tie_heavy(orb, oracle):
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 7, 100, 4095, 4096, 4097, 50001):
        keys = (rng.integers(0, 5, n) / 4.0 - 0.5).astype(np.float32)
        y = np.arange(n, dtype=np.float32)
        z = rng.standard_normal(n).astype(np.float32)
        ctx = orb.Context(n, 4, generate=False)
        ctx.set_particles(keys, y, z)
        for m in [(-0.5, 0.5), (-0.5, 0.0), (-0.25, 0.25)]:
            cut = F32(F32(F32(m[0]) + F32(m[1])) * F32(0.5))
            ctx.set_particles(keys, y, z)
            cells = _root_cells(orb, 4, m, 0)
            rc, split, ox, oy, oz = oracle.hoare(keys, y, z, 0, cut)
            if rc != 0:
                with pytest.raises(orb.OrbError) as e:
                    ctx.partition(cells)
                assert e.value.code == orb.ERR_UB
                continue
            ctx.partition(cells)
            gx, gy, gz = ctx.get_particles()
            assert np.array_equal(fbits(gx), fbits(ox))
            assert np.array_equal(gy, oy)
            assert np.array_equal(fbits(gz), fbits(oz))
            rg = ctx.ranges()
            assert tuple(rg[1]) == (0, split) and tuple(rg[2]) == (split, n)
        ctx.close()


def test_partition_service_many_cells(orb, oracle):
    """One level of 64 cells of uneven sizes, each with its own axis and cut."""
    x, y, z = oracle.generate(1 << 16)
    # full tree for d=128: level 7 splits cells 63..126
    r = oracle.build(x, y, z, 128, levels_mode=1)
    ids = np.arange(63, 127)
    cells = r.heap[ids].copy()   # level-7 cells with their final margins
    n = len(x)
    ctx = orb.Context(n, 128, generate=False)
    ctx.set_particles(x, y, z)
    # the reference schedule for d=128 stops after level 6: the device then
    # holds the particles and ranges the level-7 partition starts from
    info = ctx.build(orb.FORM_PLAIN, orb.LEVELS_REFERENCE)
    assert info["nlevels"] == 6
    ctx.partition(cells)
    gx, gy, gz = ctx.get_particles()
    assert np.array_equal(fbits(gx), fbits(r.x))
    assert np.array_equal(fbits(gy), fbits(r.y))
    assert np.array_equal(fbits(gz), fbits(r.z))
    ctx.close()


@pytest.mark.parametrize("levels_mode", [0, 1])
def test_plain_build_orbit_20_10(orb, oracle, levels_mode):
    """`orbit 20 10 0`: the reference's CPU path, bit for bit."""
    x, y, z = oracle.generate(1 << 20)
    g, r = run_both(orb, oracle, x, y, z, 1024, levels_mode)
    assert_same_build(g, r)


@pytest.mark.parametrize("cmp,partition", [(1, 0), (1, 1)])
def test_plain_build_gpu_semantics(orb, oracle, cmp, partition):
    """o=1 (<= counts, Hoare) and o=2 (<= counts, stable <= partition)."""
    x, y, z = oracle.generate(1 << 18)
    g, r = run_both(orb, oracle, x, y, z, 256, 1, cmp, partition)
    assert_same_build(g, r)


def test_plain_build_tie_heavy(orb, oracle):
    rng = np.random.default_rng(8)
    n = (1 << 18) + 17
    x, y, z = ((rng.integers(0, 1 << 9, n) / (1 << 9) - 0.5).astype(np.float32) for _ in range(3))
    g, r = run_both(orb, oracle, x, y, z, 512, 1)
    assert_same_build(g, r)


@pytest.mark.parametrize("n,d", [(5, 2), (1000, 7), (4099, 12), (33333, 100), (70000, 1)])
def test_plain_build_small_and_odd(orb, oracle, n, d):
    x, y, z = oracle.generate(n)
    g, r = run_both(orb, oracle, x, y, z, d, 1)
    assert_same_build(g, r)


def test_undefined_input_is_reported(orb, oracle):
    """A cut above every key: the reference's Hoare loop runs past the cell."""
    n = 1000
    x, y, z = oracle.generate(n)
    ctx = orb.Context(n, 4, generate=False)
    ctx.set_particles(x, y, z)
    cells = _root_cells(orb, 4, (0.9, 1.0), 0)
    with pytest.raises(orb.OrbError) as e:
        ctx.partition(cells)
    assert e.value.code == orb.ERR_UB
    # the context stays usable
    ctx.set_particles(x, y, z)
    info = ctx.build(orb.FORM_PLAIN, orb.LEVELS_REFERENCE)
    assert info["nlevels"] == 1
    ctx.close()


def test_measure_stream_leaves_particles_alone(orb, oracle):
    """The bandwidth reference reads the particle arrays and copies them to
    scratch: plausible rates, and the context's particles unchanged."""
    n = (1 << 22) + 64
    ctx = orb.Context(n, 16, generate=True)
    before = ctx.get_particles()
    read_gbs, copy_gbs = ctx.measure_stream(3)
    assert 100.0 < read_gbs < 8000.0
    assert 100.0 < copy_gbs < 8000.0
    after = ctx.get_particles()
    for a, b in zip(before, after):
        assert np.array_equal(fbits(a), fbits(b))
    ctx.close()
