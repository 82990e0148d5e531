"""Randomised parity sweep: plain and batched builds of random sizes, leaf
counts, distributions (uniform, clustered, quantized with heavy ties) and
semantics (o=0, o=1, o=2), against the CPU oracle, bit for bit.

Where the reference is undefined (an empty cell or a cell with no key at or
above its cut, partition.cpp:28-60 reading past the cell), both sides must
say so and name the same cell: the first in the reference's order."""
import numpy as np
import pytest
from gpu_helpers import assert_same_build, bin_edge_keys, fbits, run_both

import inputs

pytestmark = pytest.mark.gpu

SEMANTICS = [(0, 0), (1, 0), (1, 1)]  # (comparison, partition): o=0, o=1, o=2


def random_cases(seed, count):
    rng = np.random.default_rng(seed)
    for k in range(count):
        n = int(2 ** rng.uniform(10, 20))
        d = int(2 ** rng.uniform(1, 12))
        dist = ("uniform", "clustered", "quantized")[k % 3]
        cmp, part = SEMANTICS[int(rng.integers(0, 3))]
        levels_mode = int(rng.integers(0, 2))
        form = int(rng.integers(0, 2))
        s = int(rng.integers(0, 1 << 30))
        if dist == "clustered":
            args = dict(n_spheres=int(rng.integers(2, 20)), a=float(rng.uniform(0.005, 0.1)), seed=s)
        elif dist == "quantized":
            args = dict(levels=int(rng.integers(4, 600)), seed=s)
        else:
            args = {}
        yield n, d, dist, args, cmp, part, levels_mode, form


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_random_configs_match_oracle(orb, oracle, seed):
    exact = 0
    for n, d, dist, args, cmp, part, levels_mode, form in random_cases(seed, 8):
        if dist == "uniform":
            x, y, z = oracle.generate(n)
        elif dist == "clustered":
            x, y, z = inputs.clustered(n, **args)
        else:
            x, y, z = inputs.quantized(n, **args)
        case = (n, d, dist, cmp, part, levels_mode, form)
        try:
            g, r = run_both(orb, oracle, x, y, z, d, levels_mode, cmp=cmp, partition=part,
                            form=form)
        except orb.OrbError as e:
            r = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part)
            assert r.rc == e.code, case
            assert r.error_cell == e.error_cell, (case, r.error_cell, e.error_cell)
            continue
        assert_same_build(g, r)
        exact += 1
    assert exact >= 4


def random_payload_cases(seed, count):
    """Batched builds with a payload: workload weights or particle ids, over
    uniform, clustered, quantized and bin-edge keys (every key on or next to a
    first-round candidate cut of the root), sizes from 2^8 to 2^20."""
    rng = np.random.default_rng(seed)
    for k in range(count):
        n = int(2 ** rng.uniform(8, 20))
        d = int(2 ** rng.uniform(1, 13))
        dist = ("uniform", "clustered", "quantized", "binedge")[k % 4]
        cmp, part = SEMANTICS[int(rng.integers(0, 3))]
        levels_mode = int(rng.integers(0, 2))
        payload = ("weights", "ids")[(k // 4) % 2]
        yield n, d, dist, int(rng.integers(0, 1 << 30)), cmp, part, levels_mode, payload


@pytest.mark.parametrize("seed", [21, 22])
def test_random_payload_configs_match_oracle(orb, oracle, seed):
    from test_gpu_tree import build_with_ids
    from test_gpu_weights import assert_same_weighted, run_weighted
    exact = 0
    for n, d, dist, s, cmp, part, levels_mode, payload in random_payload_cases(seed, 16):
        r2 = np.random.default_rng(s)
        if dist == "uniform":
            x, y, z = oracle.generate(n)
        elif dist == "clustered":
            x, y, z = inputs.clustered(n, n_spheres=int(r2.integers(2, 20)),
                                       a=float(r2.uniform(0.005, 0.1)), seed=s)
        elif dist == "quantized":
            x, y, z = inputs.quantized(n, levels=int(r2.integers(4, 600)), seed=s)
        else:
            x = bin_edge_keys(n, r2)
            y = (r2.random(n) - 0.5).astype(np.float32)
            z = (r2.random(n) - 0.5).astype(np.float32)
        case = (n, d, dist, cmp, part, levels_mode, payload)
        w = r2.integers(1, 1000, n).astype(np.uint32)
        if payload == "weights":
            ref = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part,
                               weights=w)
        else:
            ref = oracle.build(x, y, z, d, levels_mode=levels_mode, cmp=cmp, partition=part)
        if ref.rc != 0:
            # undefined input: the GPU build must refuse it with the same cell
            ctx = orb.Context(n, d, generate=False)
            ctx.set_semantics(cmp, part)
            ctx.set_particles(x, y, z)
            if payload == "weights":
                ctx.set_weights(w)
            else:
                ctx.enable_ids(True)
            with pytest.raises(orb.OrbError) as ei:
                ctx.build(1, levels_mode)
            ctx.close()
            assert ei.value.code == ref.rc, case
            assert ei.value.error_cell == ref.error_cell, case
            continue
        if payload == "weights":
            g, r = run_weighted(orb, oracle, x, y, z, w, d, levels_mode, cmp, part, form=1)
            assert_same_weighted(g, r)
        else:
            ctx, info, r = build_with_ids(orb, oracle, x, y, z, d, levels_mode, 1, cmp, part)
            assert np.array_equal(ctx.ids(), r.ids), case
            for a, b in zip(ctx.get_particles(), (r.x, r.y, r.z)):
                assert np.array_equal(fbits(a), fbits(b)), case
            assert np.array_equal(ctx.ranges(), r.ranges), case
            ctx.close()
        exact += 1
    assert exact >= 8
