"""Randomised parity sweep: plain and batched builds of random sizes, leaf
counts, distributions (uniform, clustered, quantized with heavy ties) and
semantics (o=0, o=1, o=2), against the CPU oracle, bit for bit.

Where the reference is undefined (an empty cell or a cell with no key at or
above its cut, partition.cpp:28-60 reading past the cell), both sides must
say so and name the same cell: the first in the reference's order."""
import numpy as np
import pytest
from gpu_helpers import assert_same_build, run_both

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
