"""Parity at the benchmark workloads' full sizes, bit for bit against the CPU
oracle (run on every host thread the process may use: OpenMP over cells and
over a large cell's keys gives the serial result).

  C2  2^27 reference-generator particles -> 1024 leaves, full tree: plain and
      batched forms.  The top levels hold cells of more than 2^24 keys, where
      the termination rule's float arithmetic (src/orbit.cpp:204-208) rounds
      the counts; the test checks that such rounded counts occur.
  C3  2^26 particles, 90% from 64 Plummer spheres (a = 0.01) + 10% uniform
      -> 2^14 leaves, batched form (uneven cells, LDS subtrees).
  C4  2^28 particles -> 2^20 leaves, batched form; one level ends at the
      32-pass cap (src/orbit.cpp:149,219,227): a midpoint no pass counted
      becomes the cut.
  C5  requests of the serving workload (N in [2^16, 2^24], d in [2^4, 2^12],
      uniform from a request seed or Plummer) built together by one server,
      each fetched result against the oracle.

Each test needs a few GB of host memory and up to a minute of oracle time."""
import os

import numpy as np
import pytest
from gpu_helpers import assert_same_build, fbits

import inputs

pytestmark = pytest.mark.gpu

FULL = 1


def host_threads():
    return max(1, len(os.sched_getaffinity(0)))


def build_gpu(orb, ctx, form, trace=True):
    ctx.restore_particles()
    info = ctx.build(form, FULL)
    return {"info": info, "heap": ctx.heap(), "ranges": ctx.ranges(), "exists": ctx.exists(),
            "xyz": ctx.get_particles(), "pc": ctx.pass_counts() if trace else None}


def rounded_counts(pass_counts):
    """Per-pass left counts above 2^24 that float32 does not hold exactly."""
    pc = pass_counts[pass_counts != 0xFFFFFFFF].astype(np.uint64)
    big = pc[pc > (1 << 24)]
    return int((big.astype(np.float32).astype(np.uint64) != big).sum())


@pytest.mark.timeout(900)
def test_c2_uniform_2e27_both_forms(orb, oracle):
    n, d = 1 << 27, 1024
    x, y, z = orb.generate(n)                      # the reference stream (seed 0)
    print(f"C2: n={n} d={d}", flush=True)
    r = oracle.build(x, y, z, d, levels_mode=FULL, pass_counts=True, threads=host_threads())
    assert r.rc == 0
    assert rounded_counts(r.pass_counts) > 0      # the H4 regime is exercised
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    del x, y, z
    ctx.save_particles()
    ctx.enable_trace(True)
    for form in (orb.FORM_PLAIN, orb.FORM_BATCHED):
        print(f"C2: form {form}", flush=True)
        g = build_gpu(orb, ctx, form)
        assert_same_build(g, r)
        del g
    ctx.close()


@pytest.mark.timeout(900)
def test_c3_plummer_2e26_batched(orb, oracle):
    n, d = 1 << 26, 1 << 14
    x, y, z = inputs.clustered(n, n_spheres=64, a=0.01, background=0.1, seed=2022)
    print(f"C3: n={n} d={d}", flush=True)
    r = oracle.build(x, y, z, d, levels_mode=FULL, pass_counts=True, threads=host_threads())
    assert r.rc == 0
    sizes = (r.ranges[:, 1] - r.ranges[:, 0])[r.exists.astype(bool)]
    assert sizes.max() > 1 << 24
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    del x, y, z
    ctx.save_particles()
    ctx.enable_trace(True)
    g = build_gpu(orb, ctx, orb.FORM_BATCHED)
    assert_same_build(g, r)
    ctx.close()


@pytest.mark.timeout(900)
def test_c4_uniform_2e28_batched_with_capped_level(orb, oracle):
    n, d = 1 << 28, 1 << 20
    x, y, z = orb.generate(n)
    print(f"C4: n={n} d={d}", flush=True)
    r = oracle.build(x, y, z, d, levels_mode=FULL, pass_counts=True, threads=host_threads())
    assert r.rc == 0
    assert 32 in list(r.level_passes)            # a level ends at the pass cap
    capped = np.nonzero((r.pass_counts[:, 31] != 0xFFFFFFFF) & ~r.heap["foundCut"].astype(bool)
                        & r.exists.astype(bool))[0]
    assert capped.size > 0                         # cells cut at a never-counted midpoint
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    del x, y, z
    ctx.save_particles()
    ctx.enable_trace(True)
    g = build_gpu(orb, ctx, orb.FORM_BATCHED)
    assert_same_build(g, r)
    ctx.close()


def c5_requests(seed, count):
    """(n, d, kind, input seed): the fixed corners of the serving workload and
    draws log-uniform in N and d, half uniform, half Plummer."""
    reqs = [(1 << 24, 1 << 12, "uniform", 101), (1 << 16, 16, "plummer", 102),
            ((1 << 24) - 3, 4000, "plummer", 103), (65537, 1 << 12, "uniform", 104)]
    rng = np.random.default_rng(seed)
    for k in range(count):
        n = int(round(2.0 ** rng.uniform(16, 24)))
        dd = int(round(2.0 ** rng.uniform(4, 12)))
        reqs.append((n, dd, ("uniform", "plummer")[k % 2], int(rng.integers(1, 1 << 40))))
    return reqs


def c5_input(orb, n, kind, s):
    if kind == "uniform":
        return orb.generate(n, seed=s)
    return inputs.clustered(n, n_spheres=64, a=0.01, background=0.1, seed=s)


@pytest.mark.timeout(900)
def test_c5_served_requests_match_oracle(orb, oracle):
    reqs = c5_requests(55, 6)
    srv = orb.Server(max_active=16, levels_mode=FULL)
    ids = []
    for n, d, kind, s in reqs:
        x, y, z = c5_input(orb, n, kind, s)
        ids.append(srv.submit(x, y, z, d))
    done = {}
    while srv.step():
        pass
    while True:
        rid = srv.next_done()
        if rid is None:
            break
        done[rid] = True
    assert set(done) == set(ids)
    assert srv.stats()["max_batch"] > 1             # requests shared level steps
    for rid, (n, d, kind, s) in zip(ids, reqs):
        print(f"C5: n={n} d={d} {kind} seed={s}", flush=True)
        res = srv.fetch(rid, n, d)
        assert res["error"] == 0
        x, y, z = c5_input(orb, n, kind, s)
        r = oracle.build(x, y, z, d, levels_mode=FULL, threads=host_threads())
        assert r.rc == 0
        g = {"info": {"nlevels": res["nlevels"], "level_passes": res["level_passes"],
                      "total_passes": res["total_passes"]},
             "heap": res["heap"], "ranges": res["ranges"], "exists": res["exists"],
             "xyz": (res["x"], res["y"], res["z"]), "pc": None}
        assert_same_build(g, r, trace=False)
    srv.close()


def test_generator_streams_on_device_equal_host(orb):
    """orb_init(generate=true) fills the device with orb_generate's seed-0
    stream."""
    n = (1 << 20) + 3
    ctx = orb.Context(n, 64, generate=True)
    gx, gy, gz = ctx.get_particles()
    ctx.close()
    hx, hy, hz = orb.generate(n)
    for a, b in ((gx, hx), (gy, hy), (gz, hz)):
        assert np.array_equal(fbits(a), fbits(b))
