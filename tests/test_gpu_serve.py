"""Serving with continuous batching: every build's result is the same alone
and in a mix of concurrent builds of different sizes, bit for bit (zero
tolerance).  Needs an MI355X."""
import numpy as np
import pytest
from gpu_helpers import assert_heaps_equal, fbits

import inputs

pytestmark = pytest.mark.gpu


def make_requests(seed, count, lo=10, hi=17):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        n = int(round(2 ** rng.uniform(lo, hi)))
        d = min(int(round(2 ** rng.uniform(2, 9))), n // 128)   # cells keep >= ~100 keys
        if k % 2 == 0:
            x, y, z = (rng.uniform(-0.5, 0.5, n).astype(np.float32) for _ in range(3))
        else:
            x, y, z = inputs.clustered(n, n_spheres=4, a=0.02, seed=seed * 100 + k)
        reqs.append((x, y, z, d))
    return reqs


def alone(orb, x, y, z, d, levels_mode):
    ctx = orb.Context(len(x), d, generate=False)
    ctx.set_particles(x, y, z)
    info = ctx.build(orb.FORM_BATCHED, levels_mode)
    out = {"x": None, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
           "info": info}
    out["x"], out["y"], out["z"] = ctx.get_particles()
    ctx.close()
    return out


def assert_same(a, b):
    assert np.array_equal(a["exists"], b["exists"])
    assert_heaps_equal(a["heap"], b["heap"], b["exists"])
    assert np.array_equal(a["ranges"], b["ranges"])
    for k in ("x", "y", "z"):
        assert np.array_equal(fbits(a[k]), fbits(b[k]))


@pytest.mark.parametrize("levels_mode", [1])
def test_serve_mix_equals_alone(orb, levels_mode):
    reqs = make_requests(seed=11, count=12)
    srv = orb.Server(particle_capacity=1 << 20, cell_capacity=1 << 14, max_active=5,
                     levels_mode=levels_mode)
    ids = []
    # staggered arrivals: a few now, more while the first ones run
    for k, (x, y, z, d) in enumerate(reqs):
        ids.append(srv.submit(x, y, z, d))
        if k % 3 == 2:
            srv.step()
    for _ in range(10000):
        if srv.step() == 0:
            break
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["max_batch"] >= 2          # builds shared level steps
    for rid, (x, y, z, d) in zip(ids, reqs):
        assert srv.poll(rid)
        got = srv.fetch(rid, len(x), d)
        ref = alone(orb, x, y, z, d, levels_mode)
        assert got["error"] == 0
        assert got["level_passes"] == ref["info"]["level_passes"]
        assert_same(got, ref)
    srv.close()


def test_serve_admission_waits_for_memory(orb, oracle):
    """A pool that holds one build at a time still serves a queue of them."""
    x, y, z = oracle.generate(30000)
    srv = orb.Server(particle_capacity=32768, cell_capacity=1024, max_active=4,
                     levels_mode=orb.LEVELS_FULL)
    ids = [srv.submit(x, y, z, 100) for _ in range(3)]
    srv.step()
    st = srv.stats()
    assert st["active"] + st["finishing"] <= 1 and st["pending"] >= 2
    for _ in range(10000):
        if srv.step() == 0:
            break
    r = oracle.build(x, y, z, 100, levels_mode=1)
    for rid in ids:
        got = srv.fetch(rid, len(x), 100)
        assert np.array_equal(got["ranges"], r.ranges)
        assert np.array_equal(fbits(got["x"]), fbits(r.x))
    with pytest.raises(orb.OrbError):
        srv.submit(*oracle.generate(40000), 100)   # larger than the pool: never admissible
    srv.close()


def test_serve_device_requests(orb):
    """Device-resident requests: reordered in place, same as host requests."""
    reqs = make_requests(seed=5, count=6, lo=12, hi=16)
    srv = orb.Server(particle_capacity=1 << 19, cell_capacity=1 << 13, max_active=6,
                     levels_mode=orb.LEVELS_FULL)
    dev, ids = [], []
    for x, y, z, d in reqs:
        t = [orb.DeviceArray(a) for a in (x, y, z)]
        dev.append(t)
        ids.append(srv.submit_device(t[0].address, t[1].address, t[2].address, len(x), d))
    orb.hip().hipDeviceSynchronize()
    host = [srv.submit(x, y, z, d) for x, y, z, d in reqs]
    for _ in range(10000):
        if srv.step() == 0:
            break
    for rid, hid, t, (x, y, z, d) in zip(ids, host, dev, reqs):
        a = srv.fetch(rid, len(x), d)
        b = srv.fetch(hid, len(x), d)
        for k, key in enumerate("xyz"):
            assert np.array_equal(fbits(t[k].numpy()), fbits(b[key]))
            assert np.array_equal(fbits(a[key]), fbits(b[key]))
        assert np.array_equal(a["ranges"], b["ranges"])
        assert_heaps_equal(a["heap"], b["heap"], b["exists"])
        for arr in t:
            arr.free()
    srv.close()
