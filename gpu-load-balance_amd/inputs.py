"""Synthetic particle sets for the ORB workloads.

uniform()   -- the reference generator (xorshf96, seeds of
               src/services/init.cu:11) is implemented in the HIP library
               (orb_init with generate=true); this module covers the others.
clustered() -- 90% of the particles drawn from Plummer spheres (scale radius
               a, centres uniform in [-0.5, 0.5)^3, equal split) plus 10%
               uniform background; draws outside [-0.5, 0.5)^3 are redrawn
               (the library's orb_generate_plummer, host code).
quantized() -- coordinates on a coarse grid: many equal keys (ties at cuts).
write_tipsy() -- a Tipsy snapshot (standard big-endian or native layout).
"""
import struct

import numpy as np


def clustered(n, n_spheres=64, a=0.01, background=0.1, seed=0):
    """Plummer spheres plus a uniform background, float32 in [-0.5, 0.5)^3
    (the library's host generator, orb_generate_plummer)."""
    import orb_amd
    return orb_amd.generate_plummer(n, seed=seed, n_spheres=n_spheres, a=a, background=background)


def quantized(n, levels=512, seed=0):
    rng = np.random.default_rng(seed)
    return tuple((rng.integers(0, levels, n) / levels - 0.5).astype(np.float32) for _ in range(3))


def write_tipsy(path, gas=None, dark=None, star=None, big_endian=True, time=0.0, seed=0):
    """Writes positions ([m][3] float32 arrays per particle kind) as a Tipsy
    file; the other fields get seeded random values."""
    rng = np.random.default_rng(seed)
    kinds = [(gas, 12), (dark, 9), (star, 11)]
    counts = [0 if a is None else len(a) for a, _ in kinds]
    e = ">" if big_endian else "<"
    with open(path, "wb") as f:
        f.write(struct.pack(e + "diiiiii", time, sum(counts), 3, counts[0], counts[1],
                            counts[2], 0))
        for (a, nf), m in zip(kinds, counts):
            if m == 0:
                continue
            rec = rng.standard_normal((m, nf)).astype(np.float32)
            rec[:, 1:4] = np.asarray(a, np.float32)
            f.write(rec.astype(e + "f4").tobytes())
