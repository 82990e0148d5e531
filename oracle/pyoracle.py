"""ctypes binding of the CPU oracle (liborb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry and by the
benchmark's CPU-baseline leg.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborb_oracle.so")
REF_CELL_PATH = os.path.join(HERE, "_ref", "libref_cell.so")

MAX_PASSES = 32
LEVELS_REFERENCE, LEVELS_FULL = 0, 1
CMP_LT, CMP_LE = 0, 1
PART_HOARE, PART_STABLE_LE = 0, 1
OK, ERR_ARG, ERR_UB = 0, -1, -4


class Cell(C.Structure):
    """Same layout as `struct Cell` (src/cell.h:9-17), 52 bytes."""
    _fields_ = [
        ("id", C.c_int32),
        ("nLeafCells", C.c_int32),
        ("prevCutAxis", C.c_int32),
        ("cutAxis", C.c_int32),
        ("foundCut", C.c_bool),
        ("cutMarginLeft", C.c_float),
        ("cutMarginRight", C.c_float),
        ("lower", C.c_float * 3),
        ("upper", C.c_float * 3),
    ]

    def fields(self):
        return (self.id, self.nLeafCells, self.prevCutAxis, self.cutAxis, bool(self.foundCut),
                _bits(self.cutMarginLeft), _bits(self.cutMarginRight),
                tuple(_bits(v) for v in self.lower), tuple(_bits(v) for v in self.upper))


assert C.sizeof(Cell) == 52


def _bits(f):
    return int(np.float32(f).view(np.uint32))


# numpy structured dtype with the Cell layout (for heap arrays)
CELL_DTYPE = np.dtype({
    "names": ["id", "nLeafCells", "prevCutAxis", "cutAxis", "foundCut",
              "cutMarginLeft", "cutMarginRight", "lower", "upper"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "u1", "<f4", "<f4", ("<f4", 3), ("<f4", 3)],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 40],
    "itemsize": 52,
})


class Opts(C.Structure):
    _fields_ = [("levels_mode", C.c_int32), ("cmp", C.c_int32),
                ("partition", C.c_int32), ("max_passes", C.c_int32), ("threads", C.c_int32)]


class Trace(C.Structure):
    _fields_ = [
        ("heap", C.c_void_p), ("ranges", C.c_void_p), ("exists", C.c_void_p),
        ("pass_counts", C.c_void_p), ("level_passes", C.c_void_p),
        ("heap_cap", C.c_int64), ("nlevels", C.c_int32), ("total_passes", C.c_int32),
        ("error_cell", C.c_int32), ("pad", C.c_int32), ("ids", C.c_void_p),
        ("weights", C.c_void_p), ("pass_weights", C.c_void_p),
        ("count_ms", C.c_double), ("partition_ms", C.c_double),
    ]


class RNG(C.Structure):
    _fields_ = [("x", C.c_uint64), ("y", C.c_uint64), ("z", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        L.orc_generate.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_rng_reference_seed.argtypes = [C.POINTER(RNG)]
        L.orc_rng_step.argtypes = [C.POINTER(RNG)]
        L.orc_rng_value.argtypes = [C.c_uint64]
        L.orc_rng_value.restype = C.c_float
        L.orc_get_cut.argtypes = [C.POINTER(Cell)]
        L.orc_get_cut.restype = C.c_float
        L.orc_split.argtypes = [C.POINTER(Cell), C.POINTER(Cell), C.POINTER(Cell)]
        L.orc_root.argtypes = [C.c_int32]
        L.orc_root.restype = Cell
        L.orc_cell_make.argtypes = [C.c_int32, C.c_int32, fp, fp]
        L.orc_cell_make.restype = Cell
        L.orc_set_cut_axis.argtypes = [C.POINTER(Cell)]
        L.orc_n_levels.argtypes = [C.c_int32]
        L.orc_n_cells_on_last_level.argtypes = [C.c_int32]
        L.orc_left_child_id.argtypes = [C.c_int32]
        L.orc_right_child_id.argtypes = [C.c_int32]
        L.orc_difference.argtypes = [C.c_uint32, C.c_uint32, C.c_int32]
        L.orc_difference.restype = C.c_int32
        L.orc_difference_w.argtypes = [C.c_uint64, C.c_uint64, C.c_int32]
        L.orc_difference_w.restype = C.c_int64
        L.orc_hoare.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                                C.c_int, C.c_float, C.POINTER(C.c_int64)]
        L.orc_stable_le.argtypes = L.orc_hoare.argtypes
        L.orc_hoare_closed_form.argtypes = [C.c_void_p, C.c_int64, C.c_float,
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p]
        L.orc_heap_capacity.argtypes = [C.c_int32, C.c_int32]
        L.orc_heap_capacity.restype = C.c_int64
        L.orc_build.argtypes = [C.c_int64, C.c_int32, C.POINTER(Opts), C.c_void_p, C.c_void_p,
                                C.c_void_p, C.POINTER(Trace)]
        L.orc_level_cells.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int64]
        L.orc_level_cells.restype = C.c_int64
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def generate(n):
    """Reference particles (init.cu:47-53): returns float32 arrays x, y, z."""
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    z = np.empty(n, np.float32)
    lib().orc_generate(n, ptr(x), ptr(y), ptr(z))
    return x, y, z


class BuildResult:
    pass


def build(x, y, z, d, levels_mode=LEVELS_REFERENCE, cmp=CMP_LT, partition=PART_HOARE,
          max_passes=MAX_PASSES, pass_counts=False, ids=False, threads=1, weights=None,
          pass_weights=False):
    """Runs the oracle build in place on copies of x, y, z (and of the
    weights, when given: uint32 per particle, moved with the particles)."""
    L = lib()
    n = len(x)
    x = np.ascontiguousarray(x, np.float32).copy()
    y = np.ascontiguousarray(y, np.float32).copy()
    z = np.ascontiguousarray(z, np.float32).copy()
    cap = L.orc_heap_capacity(d, levels_mode)
    heap = np.zeros(cap, CELL_DTYPE)
    ranges = np.zeros((cap, 2), np.uint32)
    exists = np.zeros(cap, np.uint8)
    pc = np.zeros((cap, MAX_PASSES), np.uint32) if pass_counts else None
    lp = np.zeros(64, np.int32)
    pid = np.zeros(max(n, 1), np.uint32) if ids else None
    wts = None
    if weights is not None:
        wts = np.ascontiguousarray(weights, np.uint32).copy()
        assert len(wts) == n
    pw = np.zeros((cap, MAX_PASSES), np.uint64) if pass_weights else None
    tr = Trace(heap=heap.ctypes.data, ranges=ranges.ctypes.data, exists=exists.ctypes.data,
               pass_counts=(pc.ctypes.data if pc is not None else None),
               level_passes=lp.ctypes.data, heap_cap=cap,
               ids=(pid.ctypes.data if pid is not None else None),
               weights=(wts.ctypes.data if wts is not None and n > 0 else None),
               pass_weights=(pw.ctypes.data if pw is not None else None))
    opts = Opts(levels_mode, cmp, partition, max_passes, threads)
    rc = L.orc_build(n, d, C.byref(opts), ptr(x), ptr(y), ptr(z), C.byref(tr))
    r = BuildResult()
    r.rc = rc
    r.x, r.y, r.z = x, y, z
    r.heap, r.ranges, r.exists = heap, ranges, exists
    r.pass_counts = pc
    r.ids = pid[:n] if pid is not None else None
    r.weights = wts
    r.pass_weights = pw
    r.count_ms = tr.count_ms
    r.partition_ms = tr.partition_ms
    r.level_passes = lp[:max(tr.nlevels, 0)].copy()
    r.nlevels = tr.nlevels
    r.total_passes = tr.total_passes
    r.error_cell = tr.error_cell
    return r


def hoare(key_x, key_y, key_z, axis, cut, begin=0, end=None):
    x, y, z = (np.ascontiguousarray(a, np.float32).copy() for a in (key_x, key_y, key_z))
    end = len(x) if end is None else end
    s = C.c_int64(0)
    rc = lib().orc_hoare(ptr(x), ptr(y), ptr(z), begin, end, axis, np.float32(cut), C.byref(s))
    return rc, s.value, x, y, z


def hoare_closed_form(key, cut):
    key = np.ascontiguousarray(key, np.float32)
    perm = np.zeros(len(key), np.int64)
    K = C.c_int64(0)
    s = C.c_int64(0)
    rc = lib().orc_hoare_closed_form(ptr(key), len(key), np.float32(cut), C.byref(K), C.byref(s),
                                     ptr(perm))
    return rc, K.value, s.value, perm


def stable_le(key_x, key_y, key_z, axis, cut, begin=0, end=None):
    x, y, z = (np.ascontiguousarray(a, np.float32).copy() for a in (key_x, key_y, key_z))
    end = len(x) if end is None else end
    s = C.c_int64(0)
    rc = lib().orc_stable_le(ptr(x), ptr(y), ptr(z), begin, end, axis, np.float32(cut), C.byref(s))
    return rc, s.value, x, y, z
