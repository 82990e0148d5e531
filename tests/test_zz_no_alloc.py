"""orb_count_left_static (the count-left service over caller device memory,
no allocation) against the context services it mirrors, orb_count_left
(< cut, found cells skipped) and orb_count_left_gpu (<= cut): two calls in a
row on the same GPU, the second with a smaller input, both comparisons."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENTINEL = 0xDEADBEEF


class DeviceBytes:
    """A numpy array copied into hipMalloc'ed memory (the caller's buffers)."""

    def __init__(self, orb, a):
        self.orb = orb
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert orb.hip().hipMalloc(C.byref(self.ptr), max(self.a.nbytes, 4)) == 0
        if self.a.nbytes:
            assert orb.hip().hipMemcpy(self.ptr, self.a.ctypes.data_as(C.c_void_p),
                                       self.a.nbytes, 1) == 0

    def numpy(self):
        out = np.empty_like(self.a)
        assert self.orb.hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                        out.nbytes, 2) == 0
        return out

    def free(self):
        self.orb.hip().hipFree(self.ptr)


def level_cells(heap, exists, depth, rng):
    """The cells of one depth of a built tree, with fresh margins inside their
    boxes (so the cut is not the one the build found) and random found flags."""
    ids = [i for i in range((1 << depth) - 1, (1 << (depth + 1)) - 1) if exists[i]]
    cells = heap[ids].copy()
    for c in cells:
        ax = int(c["cutAxis"])
        lo, hi = float(c["lower"][ax]), float(c["upper"][ax])
        a, b = sorted(rng.uniform(lo, hi, 2))
        c["cutMarginLeft"], c["cutMarginRight"] = np.float32(a), np.float32(b)
        c["foundCut"] = int(rng.integers(0, 2))
    return cells


def one_case(orb, n, d, depth, seed):
    rng = np.random.default_rng(seed)
    x, y, z = orb.generate(n, seed=seed)
    ctx = orb.Context(n, d, generate=False)
    ctx.set_particles(x, y, z)
    ctx.build(orb.FORM_PLAIN, orb.LEVELS_FULL)
    heap, exists, ranges = ctx.heap(), ctx.exists(), ctx.ranges()
    cells = level_cells(heap, exists, depth, rng)
    # a repeated cell (overlapping ranges) and a cell id the tree never
    # created (its range is empty)
    empty = cells[:1].copy()
    empty["id"] = int(np.nonzero(exists == 0)[0][0])
    # (concatenate drops the struct padding: back to the Cell layout)
    cells = orb.cells_array(np.concatenate([cells, cells[:1], empty]))
    assert cells.dtype == orb.CELL_DTYPE and cells.itemsize == 52
    want_lt = np.full(len(cells), SENTINEL, np.uint32)
    ctx.count_left(cells, out=want_lt)
    want_le = ctx.count_left_gpu(cells)
    px, py, pz = ctx.get_particles()
    ctx.close()
    bufs = [DeviceBytes(orb, a) for a in (px, py, pz, ranges, cells)]
    got = {}
    for cmp in (orb.CMP_LT, orb.CMP_LE):
        out = DeviceBytes(orb, np.full(len(cells), SENTINEL, np.uint32))
        rc = orb.lib().orb_count_left_static(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n,
                                             bufs[3].ptr, len(ranges), bufs[4].ptr, len(cells),
                                             cmp, out.ptr, None)
        assert rc == 4 * len(cells)
        assert orb.hip().hipDeviceSynchronize() == 0
        got[cmp] = out.numpy()
        out.free()
    for b in bufs:
        b.free()
    assert np.array_equal(got[orb.CMP_LT], want_lt)
    assert np.array_equal(got[orb.CMP_LE], want_le)
    assert (want_lt == SENTINEL).any() and (want_lt != SENTINEL).any()


def test_static_count_left_matches_services_twice(orb):
    one_case(orb, (1 << 20) + 37, 60, 4, seed=5)    # 18 cells, 2^20 particles
    one_case(orb, 50000, 12, 2, seed=6)             # then a smaller input: 6 cells


def test_static_count_left_rejects_sizes(orb):
    L = orb.lib()
    cells = np.zeros(8097, orb.CELL_DTYPE)
    out = np.zeros(8097, np.uint32)
    # sizes are checked before any device memory is touched
    assert L.orb_count_left_static(None, None, None, 10, None, 1, cells.ctypes.data, 8097,
                                   orb.CMP_LT, out.ctypes.data, None) == orb.ERR_ARG
    assert L.orb_count_left_static(None, None, None, (1 << 30) + 1, None, 1, cells.ctypes.data,
                                   1, orb.CMP_LT, out.ctypes.data, None) == orb.ERR_ARG
    assert L.orb_count_left_static(None, None, None, 10, None, 1, cells.ctypes.data, 1, 7,
                                   out.ctypes.data, None) == orb.ERR_ARG
    assert L.orb_count_left_static(None, None, None, 10, None, 1, None, 0, orb.CMP_LT, None,
                                   None) == 0
