"""Synthetic particle sets for the ORB workloads.

uniform()   -- the reference generator (xorshf96, seeds of
               src/services/init.cu:11) is implemented in the HIP library
               (orb_init with generate=true); this module covers the others.
clustered() -- 90% of the particles drawn from Plummer spheres (scale radius
               a, centres uniform in [-0.5, 0.5)^3, equal split) plus 10%
               uniform background; draws outside [-0.5, 0.5)^3 are redrawn.
quantized() -- coordinates on a coarse grid: many equal keys (ties at cuts).
write_tipsy() -- a Tipsy snapshot (standard big-endian or native layout).
"""
import struct

import numpy as np


def clustered(n, n_spheres=64, a=0.01, background=0.1, seed=0):
    rng = np.random.default_rng(seed)
    n_bg = int(round(n * background))
    n_cl = n - n_bg
    centres = rng.uniform(-0.5, 0.5, (n_spheres, 3))
    per = np.full(n_spheres, n_cl // n_spheres)
    per[: n_cl - per.sum()] += 1
    out = np.empty((n, 3), np.float32)
    k = 0
    for s in range(n_spheres):
        need = int(per[s])
        while need > 0:
            m = max(int(need * 1.2) + 16, 64)
            u = rng.uniform(1e-12, 1.0, m)
            r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)          # Plummer radius
            d = rng.normal(size=(m, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)       # isotropic direction
            p = centres[s] + r[:, None] * d
            ok = np.all((p >= -0.5) & (p < 0.5), axis=1)
            p = p[ok][:need].astype(np.float32)
            ok32 = np.all((p >= -0.5) & (p < 0.5), axis=1)     # after rounding to f32
            p = p[ok32]
            out[k:k + len(p)] = p
            k += len(p)
            need -= len(p)
    out[k:] = rng.uniform(-0.5, 0.5, (n - k, 3)).astype(np.float32)
    out[k:] = np.minimum(out[k:], np.float32(np.nextafter(np.float32(0.5), np.float32(0))))
    perm = rng.permutation(n)
    out = out[perm]
    return (np.ascontiguousarray(out[:, 0]), np.ascontiguousarray(out[:, 1]),
            np.ascontiguousarray(out[:, 2]))


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
