"""Writes tests/golden/rng_vectors.json from the reference's own generator.

Runs in the build container only, where oracle/_ref/libref_rng.so is built by
`make -C oracle` from the xorshf96 definition of src/services/init.cu:11-25.
The fixture holds the reference's outputs (float32 values as uint32 bit
patterns): the first 1024 particles drawn as init.cu:47-53 draws them
(x_i, y_i, z_i in turn), and the SHA-256 of the first 2^20 particles' x, y, z
arrays (little-endian float32, x then y then z).
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def reference_particles(n):
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_rng.so"))
    ref.ref_generate.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    x, y, z = (np.empty(n, np.float32) for _ in range(3))
    ref.ref_generate(n, x.ctypes.data, y.ctypes.data, z.ctypes.data)
    return x, y, z


def digest(x, y, z):
    h = hashlib.sha256()
    for a in (x, y, z):
        h.update(np.ascontiguousarray(a, "<f4").tobytes())
    return h.hexdigest()


def main():
    n = 1 << 20
    x, y, z = reference_particles(n)
    out = {"source": "src/services/init.cu:11-25 (xorshf96), draws of init.cu:47-53",
           "first_1024": {k: a[:1024].view(np.uint32).tolist() for k, a in
                          (("x", x), ("y", y), ("z", z))},
           "n_digest": n, "sha256_xyz": digest(x, y, z)}
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
