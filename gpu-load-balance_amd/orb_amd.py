"""ctypes binding of liborb_hip.so (the C-ABI declared in include/orb.h).

This is the Python face of the MI355X ORB build used by the tests and the
benchmark.  It loads the in-tree HIP library and raises if it is missing or
no GPU is usable: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborb_hip.so")

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_UB, ERR_STATE, ERR_NODEV, ERR_BUSY = 0, -1, -2, -3, -4, -5, -6, -7
LEVELS_REFERENCE, LEVELS_FULL = 0, 1
CMP_LT, CMP_LE = 0, 1
PART_HOARE, PART_STABLE_LE = 0, 1
FORM_PLAIN, FORM_BATCHED = 0, 1
MAX_PASSES = 32

# enum pst_service (src/services/pst.h:66-81)
(PST_SRV_STOP, PST_SETADD, PST_INIT, PST_INITGPU, PST_COPYPARTICLES, PST_COPYCELLS,
 PST_COUNTLEFTGPU, PST_COUNTLEFTAXISGPU, PST_COUNTLEFT, PST_PARTITION, PST_PARTITIONGPU,
 PST_COUNT, PST_FINALIZE, PST_MAKEAXIS) = range(14)


class Cell(C.Structure):
    """struct Cell (src/cell.h:9-17), 52 bytes."""
    _fields_ = [
        ("id", C.c_int32), ("nLeafCells", C.c_int32), ("prevCutAxis", C.c_int32),
        ("cutAxis", C.c_int32), ("foundCut", C.c_bool), ("cutMarginLeft", C.c_float),
        ("cutMarginRight", C.c_float), ("lower", C.c_float * 3), ("upper", C.c_float * 3),
    ]


assert C.sizeof(Cell) == 52

CELL_DTYPE = np.dtype({
    "names": ["id", "nLeafCells", "prevCutAxis", "cutAxis", "foundCut",
              "cutMarginLeft", "cutMarginRight", "lower", "upper"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "u1", "<f4", "<f4", ("<f4", 3), ("<f4", 3)],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 40],
    "itemsize": 52,
})


class MetaParams(C.Structure):
    """struct META_PARAMS (src/constants.h:13-17)."""
    _fields_ = [("GPU_COUNT", C.c_bool), ("GPU_PARTITION", C.c_bool), ("FAST_MEDIAN", C.c_bool)]


class BuildInfo(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("total_passes", C.c_int32),
                ("level_passes", C.c_int32 * 64), ("error_cell", C.c_int32), ("pad", C.c_int32),
                ("ms_total", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("count_ms", C.c_double), ("partition_ms", C.c_double), ("other_ms", C.c_double),
                ("count_bytes", C.c_int64), ("count_launches", C.c_int32), ("pad", C.c_int32),
                ("swap_ms", C.c_double), ("swap_launches", C.c_int32), ("pad2", C.c_int32),
                ("swap_pairs", C.c_int64), ("swap_bytes", C.c_int64)]


class ServerConfig(C.Structure):
    _fields_ = [("particle_capacity", C.c_int64), ("cell_capacity", C.c_int64),
                ("max_active", C.c_int32), ("levels_mode", C.c_int32), ("cmp", C.c_int32),
                ("partition", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in (
        "submitted", "admitted", "completed", "steps", "cells_stepped", "max_batch", "pending",
        "active", "finishing", "particle_capacity", "particles_free", "cell_capacity",
        "cells_free")]


class RequestInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("d", C.c_int32), ("error", C.c_int32),
                ("error_cell", C.c_int32), ("nlevels", C.c_int32), ("total_passes", C.c_int32),
                ("level_passes", C.c_int32 * 64), ("pad", C.c_int32), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("done_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("subtree_max_keys", C.c_int32), ("nlevels", C.c_int32),
                ("chunk_blocks", C.c_int32 * 64), ("subtree_blocks", C.c_int32 * 64),
                ("default_ms", C.c_double), ("tuned_ms", C.c_double)]


class TuneRow(C.Structure):
    _fields_ = [("depth", C.c_int32), ("log2_min_keys", C.c_int32), ("log2_max_keys", C.c_int32),
                ("chunk_blocks", C.c_int32), ("subtree_blocks", C.c_int32),
                ("subtree_max_keys", C.c_int32), ("ms", C.c_double)]


TUNE_FIELDS = ("depth", "log2_min_keys", "log2_max_keys", "chunk_blocks", "subtree_blocks",
               "subtree_max_keys", "ms")


def _rows_in(rows):
    arr = (TuneRow * max(1, len(rows)))()
    for i, r in enumerate(rows):
        for k in TUNE_FIELDS:
            setattr(arr[i], k, float(r[k]) if k == "ms" else int(r[k]))
    return arr


def _rows_out(arr, n):
    return [{k: getattr(arr[i], k) for k in TUNE_FIELDS} for i in range(n)]


class TipsyHeader(C.Structure):
    _fields_ = [("time", C.c_double), ("nbodies", C.c_int32), ("ndim", C.c_int32),
                ("nsph", C.c_int32), ("ndark", C.c_int32), ("nstar", C.c_int32),
                ("big_endian", C.c_int32)]


class OrbError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = lib().orb_error_string(code).decode() if _lib else str(code)
        super().__init__(f"{what}: {msg} ({code})")


_lib = None

# exported symbols and their prototypes (restype, argtypes)
_vp, _i32, _i64, _b = C.c_void_p, C.c_int32, C.c_int64, C.c_bool
PROTOTYPES = {
    "orb_init": (C.c_int, [C.POINTER(_vp), _i64, _i32, _b, MetaParams]),
    "orb_finalize": (C.c_int, [_vp]),
    "orb_set_semantics": (C.c_int, [_vp, _i32, _i32]),
    "orb_set_particles": (C.c_int, [_vp, _vp, _vp, _vp, _i64]),
    "orb_get_particles": (C.c_int, [_vp, _vp, _vp, _vp, _i64]),
    "orb_reset_ranges": (C.c_int, [_vp]),
    "orb_save_particles": (C.c_int, [_vp]),
    "orb_restore_particles": (C.c_int, [_vp]),
    "orb_measure_stream": (C.c_int, [_vp, _i32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "orb_copy_particles": (C.c_int, [_vp]),
    "orb_copy_cells": (C.c_int, [_vp, _vp, _i32]),
    "orb_count": (C.c_int, [_vp, _vp, _i32, _vp]),
    "orb_make_axis": (C.c_int, [_vp, _vp, _i32]),
    "orb_count_left": (C.c_int, [_vp, _vp, _i32, _vp]),
    "orb_count_left_gpu_axis": (C.c_int, [_vp, _vp, _i32, _vp]),
    "orb_count_left_gpu": (C.c_int, [_vp, _vp, _i32, _vp]),
    "orb_partition": (C.c_int, [_vp, _vp, _i32]),
    "orb_partition_gpu": (C.c_int, [_vp, _vp, _i32]),
    "orb_run_service": (C.c_int, [_vp, _i32, _vp, _i32, _vp, _i32]),
    "orb_count_left_static": (C.c_int, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i32, _i32, _vp,
                                        _vp]),
    "orb_heap_capacity": (C.c_int64, [_i32]),
    "orb_build": (C.c_int, [_vp, _i32, _i32, C.POINTER(BuildInfo)]),
    "orb_get_heap": (C.c_int, [_vp, _vp, _i64]),
    "orb_get_ranges": (C.c_int, [_vp, _vp, _i64]),
    "orb_get_exists": (C.c_int, [_vp, _vp, _i64]),
    "orb_autotune": (C.c_int, [_vp, _i32, _i32, C.POINTER(Tuning)]),
    "orb_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "orb_get_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "orb_level_shape": (C.c_int, [C.c_int64, _i32, _i32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "orb_tune_table_write": (C.c_int, [C.c_char_p, C.POINTER(TuneRow), _i32]),
    "orb_tune_table_read": (C.c_int, [C.c_char_p, C.POINTER(TuneRow), _i32]),
    "orb_get_tune_rows": (C.c_int, [_vp, C.POINTER(TuneRow), _i32]),
    "orb_set_tune_rows": (C.c_int, [_vp, C.POINTER(TuneRow), _i32]),
    "orb_enable_ids": (C.c_int, [_vp, _b]),
    "orb_get_ids": (C.c_int, [_vp, _vp, _i64]),
    "orb_get_leaves": (C.c_int, [_vp, _vp, _vp, _i64, C.POINTER(C.c_int64)]),
    "orb_get_membership": (C.c_int, [_vp, _vp, _i64, _b]),
    "orb_get_leaf_bounds": (C.c_int, [_vp, _vp, _vp, _i64]),
    "orb_enable_trace": (C.c_int, [_vp, _b]),
    "orb_set_weights": (C.c_int, [_vp, _vp, _i64]),
    "orb_get_weights": (C.c_int, [_vp, _vp, _i64]),
    "orb_count_weight": (C.c_int, [_vp, _vp, _i32, _vp]),
    "orb_count_left_weight": (C.c_int, [_vp, _vp, _i32, _i32, _vp]),
    "orb_get_pass_weights": (C.c_int, [_vp, _vp, _i64]),
    "orb_get_pass_counts": (C.c_int, [_vp, _vp, _i64]),
    "orb_get_timing": (C.c_int, [_vp, C.POINTER(Timing)]),
    "orb_enable_timing": (C.c_int, [_vp, _b]),
    "orb_enable_graph": (C.c_int, [_vp, _b]),
    "orb_get_level_timing": (C.c_int, [_vp, _vp, _i32]),
    "orb_server_create": (C.c_int, [C.POINTER(_vp), C.POINTER(ServerConfig)]),
    "orb_server_destroy": (C.c_int, [_vp]),
    "orb_server_submit": (C.c_int, [_vp, _vp, _vp, _vp, _i64, _i32, C.POINTER(C.c_uint64)]),
    "orb_server_submit_device": (C.c_int, [_vp, _vp, _vp, _vp, _i64, _i32,
                                           C.POINTER(C.c_uint64)]),
    "orb_server_step": (C.c_int, [_vp]),
    "orb_server_poll": (C.c_int, [_vp, C.c_uint64]),
    "orb_server_next_done": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "orb_server_fetch": (C.c_int, [_vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                   C.POINTER(RequestInfo)]),
    "orb_server_get_stats": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "orb_read_tipsy": (C.c_int, [C.c_char_p, C.POINTER(TipsyHeader), _vp, _vp, _vp, _i64]),
    "orb_init_tipsy": (C.c_int, [C.POINTER(_vp), C.c_char_p, _i32, MetaParams]),
    "orb_generate": (C.c_int, [C.c_uint64, _i64, _vp, _vp, _vp]),
    "orb_generate_plummer": (C.c_int, [C.c_uint64, _i64, _i32, C.c_float, C.c_float, _vp, _vp,
                                       _vp]),
    "orb_error_string": (C.c_char_p, [C.c_int]),
    "orb_last_hip_error": (C.c_int, []),
    "orb_device_count": (C.c_int, []),
}


def lib():
    """Loads liborb_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH} (run make -C gpu-load-balance_amd)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, what):
    if rc < 0:
        raise OrbError(rc, what)
    return rc


def cells_array(cells):
    """Converts a numpy CELL_DTYPE array (or list of Cell) to a contiguous array."""
    if isinstance(cells, np.ndarray):
        return np.ascontiguousarray(cells, dtype=CELL_DTYPE)
    a = np.zeros(len(cells), CELL_DTYPE)
    for i, c in enumerate(cells):
        a[i] = (c.id, c.nLeafCells, c.prevCutAxis, c.cutAxis, c.foundCut, c.cutMarginLeft,
                c.cutMarginRight, tuple(c.lower), tuple(c.upper))
    return a


class Context:
    """One device-resident particle set (ServiceInit) and its ORB state."""

    def __init__(self, n, d, generate=True, params=(False, False, False)):
        L = lib()
        if L.orb_device_count() < 1:
            raise OrbError(ERR_NODEV, "orb_init")
        h = C.c_void_p()
        _check(L.orb_init(C.byref(h), int(n), int(d), bool(generate), MetaParams(*params)), "orb_init")
        self.h = h
        self.n = int(n)
        self.d = int(d)
        self.cap = int(L.orb_heap_capacity(int(d)))

    @classmethod
    def from_tipsy(cls, path, d, params=(False, False, False)):
        """The particles of a Tipsy snapshot (ServiceInit's generate=false)."""
        L = lib()
        if L.orb_device_count() < 1:
            raise OrbError(ERR_NODEV, "orb_init_tipsy")
        h = C.c_void_p()
        _check(L.orb_init_tipsy(C.byref(h), str(path).encode(), int(d), MetaParams(*params)),
               "orb_init_tipsy")
        self = cls.__new__(cls)
        self.h = h
        self.n = int(read_tipsy(path)[0]["nbodies"])
        self.d = int(d)
        self.cap = int(L.orb_heap_capacity(int(d)))
        return self

    def close(self):
        if self.h:
            lib().orb_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- particles / semantics
    def set_semantics(self, cmp=-1, partition=-1):
        _check(lib().orb_set_semantics(self.h, cmp, partition), "orb_set_semantics")

    def set_particles(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        _check(lib().orb_set_particles(self.h, _p(x), _p(y), _p(z), self.n), "orb_set_particles")

    def get_particles(self):
        x = np.empty(self.n, np.float32)
        y = np.empty(self.n, np.float32)
        z = np.empty(self.n, np.float32)
        _check(lib().orb_get_particles(self.h, _p(x), _p(y), _p(z), self.n), "orb_get_particles")
        return x, y, z

    def save_particles(self):
        _check(lib().orb_save_particles(self.h), "orb_save_particles")

    def restore_particles(self):
        _check(lib().orb_restore_particles(self.h), "orb_restore_particles")

    def reset_ranges(self):
        _check(lib().orb_reset_ranges(self.h), "orb_reset_ranges")

    # ---- services (Cell[] in, uint32[] out)
    def _cells_call(self, fn, cells, with_out, out=None):
        a = cells_array(cells)
        if with_out:
            if out is None:
                out = np.zeros(len(a), np.uint32)
            rc = fn(self.h, _p(a), len(a), _p(out))
            _check(rc, fn.__name__)
            return out
        return _check(fn(self.h, _p(a), len(a)), fn.__name__)

    def copy_particles(self):
        return _check(lib().orb_copy_particles(self.h), "orb_copy_particles")

    def copy_cells(self, cells):
        return self._cells_call(lib().orb_copy_cells, cells, False)

    def count(self, cells):
        return self._cells_call(lib().orb_count, cells, True)

    def make_axis(self, cells):
        return self._cells_call(lib().orb_make_axis, cells, False)

    def count_left(self, cells, out=None):
        return self._cells_call(lib().orb_count_left, cells, True, out)

    def count_left_gpu_axis(self, cells):
        return self._cells_call(lib().orb_count_left_gpu_axis, cells, True)

    def count_left_gpu(self, cells):
        return self._cells_call(lib().orb_count_left_gpu, cells, True)

    def partition(self, cells):
        return self._cells_call(lib().orb_partition, cells, False)

    def partition_gpu(self, cells):
        return self._cells_call(lib().orb_partition_gpu, cells, False)

    def run_service(self, sid, cells=None, n_out=0):
        a = cells_array(cells if cells is not None else [])
        out = np.zeros(max(n_out, 1), np.uint32)
        rc = lib().orb_run_service(self.h, sid, _p(a), a.nbytes, _p(out), 4 * n_out)
        _check(rc, "orb_run_service")
        return rc, out[:n_out]

    # ---- whole build
    def build(self, form=FORM_PLAIN, levels_mode=LEVELS_REFERENCE):
        info = BuildInfo()
        rc = lib().orb_build(self.h, form, levels_mode, C.byref(info))
        if rc < 0:
            err = OrbError(rc, "orb_build")
            err.error_cell = info.error_cell
            raise err
        return {
            "nlevels": info.nlevels,
            "total_passes": info.total_passes,
            "level_passes": list(info.level_passes[:info.nlevels]),
            "ms_total": info.ms_total,
        }

    def heap(self):
        h = np.zeros(self.cap, CELL_DTYPE)
        _check(lib().orb_get_heap(self.h, _p(h), self.cap), "orb_get_heap")
        return h

    def ranges(self):
        r = np.zeros((self.cap, 2), np.uint32)
        _check(lib().orb_get_ranges(self.h, _p(r), self.cap), "orb_get_ranges")
        return r

    def exists(self):
        e = np.zeros(self.cap, np.uint8)
        _check(lib().orb_get_exists(self.h, _p(e), self.cap), "orb_get_exists")
        return e

    # ---- autotuner (batched form)
    @staticmethod
    def _tuning_dict(t):
        n = t.nlevels
        return {"subtree_max_keys": t.subtree_max_keys, "nlevels": n,
                "chunk_blocks": list(t.chunk_blocks[:n]),
                "subtree_blocks": list(t.subtree_blocks[:n]),
                "default_ms": t.default_ms, "tuned_ms": t.tuned_ms}

    def autotune(self, levels_mode=LEVELS_FULL, repeats=2):
        t = Tuning()
        _check(lib().orb_autotune(self.h, int(levels_mode), int(repeats), C.byref(t)),
               "orb_autotune")
        return self._tuning_dict(t)

    def tuning(self):
        t = Tuning()
        _check(lib().orb_get_tuning(self.h, C.byref(t)), "orb_get_tuning")
        return self._tuning_dict(t)

    def set_tuning(self, d=None):
        if d is None:
            _check(lib().orb_set_tuning(self.h, None), "orb_set_tuning")
            return
        t = Tuning()
        t.subtree_max_keys = int(d.get("subtree_max_keys", -1))
        t.nlevels = int(d.get("nlevels", 0))
        for i, v in enumerate(d.get("chunk_blocks", [])[:64]):
            t.chunk_blocks[i] = int(v)
        for i, v in enumerate(d.get("subtree_blocks", [])[:64]):
            t.subtree_blocks[i] = int(v)
        _check(lib().orb_set_tuning(self.h, C.byref(t)), "orb_set_tuning")

    def tune_rows(self):
        """The current choices as rows keyed by level shape (orb_get_tune_rows)."""
        arr = (TuneRow * 64)()
        n = _check(lib().orb_get_tune_rows(self.h, arr, 64), "orb_get_tune_rows")
        return _rows_out(arr, n)

    def set_tune_rows(self, rows):
        """Applies keyed rows; returns how many levels found a row."""
        arr = _rows_in(rows)
        return _check(lib().orb_set_tune_rows(self.h, arr, len(rows)), "orb_set_tune_rows")

    # ---- tree export
    def enable_ids(self, on=True):
        _check(lib().orb_enable_ids(self.h, bool(on)), "orb_enable_ids")

    def ids(self):
        out = np.zeros(self.n, np.uint32)
        _check(lib().orb_get_ids(self.h, _p(out), self.n), "orb_get_ids")
        return out

    def leaves(self):
        """(heap ids, [count][2] ranges) of the last build's leaves."""
        cnt = C.c_int64()
        _check(lib().orb_get_leaves(self.h, None, None, 0, C.byref(cnt)), "orb_get_leaves")
        ids = np.zeros(cnt.value, np.int32)
        rg = np.zeros((cnt.value, 2), np.uint32)
        if cnt.value:
            _check(lib().orb_get_leaves(self.h, _p(ids), _p(rg), cnt.value, C.byref(cnt)),
                   "orb_get_leaves")
        return ids, rg

    def membership(self, by_input_order=False):
        out = np.zeros(self.n, np.int32)
        _check(lib().orb_get_membership(self.h, _p(out), self.n, bool(by_input_order)),
               "orb_get_membership")
        return out

    def leaf_bounds(self):
        ids, _ = self.leaves()
        lo = np.zeros((len(ids), 3), np.float32)
        hi = np.zeros((len(ids), 3), np.float32)
        if len(ids):
            _check(lib().orb_get_leaf_bounds(self.h, _p(lo), _p(hi), len(ids)),
                   "orb_get_leaf_bounds")
        return lo, hi

    # ---- workload weights
    def set_weights(self, w):
        if w is None:
            _check(lib().orb_set_weights(self.h, None, self.n), "orb_set_weights")
            return
        w = np.ascontiguousarray(w, np.uint32)
        assert len(w) == self.n
        _check(lib().orb_set_weights(self.h, _p(w), self.n), "orb_set_weights")

    def weights(self):
        out = np.zeros(self.n, np.uint32)
        _check(lib().orb_get_weights(self.h, _p(out), self.n), "orb_get_weights")
        return out

    def count_weight(self, cells):
        a = cells_array(cells)
        out = np.zeros(len(a), np.uint64)
        _check(lib().orb_count_weight(self.h, _p(a), len(a), _p(out)), "orb_count_weight")
        return out

    def count_left_weight(self, cells, cmp, out=None):
        a = cells_array(cells)
        if out is None:
            out = np.zeros(len(a), np.uint64)
        _check(lib().orb_count_left_weight(self.h, _p(a), len(a), int(cmp), _p(out)),
               "orb_count_left_weight")
        return out

    def pass_weights(self):
        pw = np.zeros((self.cap, MAX_PASSES), np.uint64)
        _check(lib().orb_get_pass_weights(self.h, _p(pw), self.cap), "orb_get_pass_weights")
        return pw

    def enable_trace(self, on=True):
        _check(lib().orb_enable_trace(self.h, bool(on)), "orb_enable_trace")

    def pass_counts(self):
        pc = np.zeros((self.cap, MAX_PASSES), np.uint32)
        _check(lib().orb_get_pass_counts(self.h, _p(pc), self.cap), "orb_get_pass_counts")
        return pc

    def level_timing(self):
        ms = np.zeros(64, np.float64)
        k = _check(lib().orb_get_level_timing(self.h, _p(ms), 64), "orb_get_level_timing")
        return ms[:k].copy()

    def measure_stream(self, reps=10):
        """(read GB/s, copy GB/s) of 16-byte streams over the particle arrays."""
        r, w = C.c_double(0.0), C.c_double(0.0)
        _check(lib().orb_measure_stream(self.h, int(reps), C.byref(r), C.byref(w)),
               "orb_measure_stream")
        return r.value, w.value

    def enable_graph(self, on=True):
        _check(lib().orb_enable_graph(self.h, bool(on)), "orb_enable_graph")

    def enable_timing(self, on=True):
        _check(lib().orb_enable_timing(self.h, bool(on)), "orb_enable_timing")

    def timing(self):
        t = Timing()
        _check(lib().orb_get_timing(self.h, C.byref(t)), "orb_get_timing")
        return {"count_ms": t.count_ms, "partition_ms": t.partition_ms, "other_ms": t.other_ms,
                "count_bytes": t.count_bytes, "count_launches": t.count_launches,
                "swap_ms": t.swap_ms, "swap_launches": t.swap_launches,
                "swap_pairs": t.swap_pairs, "swap_bytes": t.swap_bytes}


_hip = None


def hip():
    """The HIP runtime liborb_hip.so is linked against (one runtime per
    process: a second one, such as a framework's bundled copy, would not see
    the GPU once this one owns it)."""
    global _hip
    if _hip is None:
        lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(_vp), C.c_size_t]
        h.hipFree.argtypes = [_vp]
        h.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A float32 array in device memory (hipMalloc), for device requests."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, np.float32)
        self.n = host.size
        self.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(self.ptr), max(self.n, 1) * 4) != 0:
            raise OrbError(ERR_NOMEM, "hipMalloc")
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, np.float32)
        if hip().hipMemcpy(self.ptr, _p(host), self.n * 4, 1) != 0:
            raise OrbError(ERR_HIP, "hipMemcpy")

    def numpy(self):
        out = np.empty(self.n, np.float32)
        if hip().hipMemcpy(_p(out), self.ptr, self.n * 4, 2) != 0:
            raise OrbError(ERR_HIP, "hipMemcpy")
        return out

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching ORB server (include/orb_serve.h)."""

    def __init__(self, particle_capacity=0, cell_capacity=0, max_active=16,
                 levels_mode=LEVELS_FULL, cmp=CMP_LT, partition=PART_HOARE):
        L = lib()
        if L.orb_device_count() < 1:
            raise OrbError(ERR_NODEV, "orb_server_create")
        cfg = ServerConfig(int(particle_capacity), int(cell_capacity), int(max_active),
                           int(levels_mode), int(cmp), int(partition))
        h = C.c_void_p()
        _check(L.orb_server_create(C.byref(h), C.byref(cfg)), "orb_server_create")
        self.h = h

    def close(self):
        if self.h:
            lib().orb_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, x, y, z, d):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        rid = C.c_uint64()
        _check(lib().orb_server_submit(self.h, _p(x), _p(y), _p(z), len(x), int(d),
                                       C.byref(rid)), "orb_server_submit")
        return rid.value

    def submit_device(self, xp, yp, zp, n, d):
        """Particles already in device memory (integer addresses, e.g. a
        torch tensor's data_ptr()); reordered in place when the build ends."""
        rid = C.c_uint64()
        _check(lib().orb_server_submit_device(self.h, C.c_void_p(xp), C.c_void_p(yp),
                                              C.c_void_p(zp), int(n), int(d), C.byref(rid)),
               "orb_server_submit_device")
        return rid.value

    def step(self):
        return _check(lib().orb_server_step(self.h), "orb_server_step")

    def poll(self, rid):
        return _check(lib().orb_server_poll(self.h, rid), "orb_server_poll") == 1

    def next_done(self):
        """Id of a request finished since the last call, or None."""
        rid = C.c_uint64()
        rc = _check(lib().orb_server_next_done(self.h, C.byref(rid)), "orb_server_next_done")
        return rid.value if rc == 1 else None

    def fetch(self, rid, n, d, particles=True):
        cap = int(lib().orb_heap_capacity(int(d)))
        x = np.empty(n, np.float32) if particles else None
        y = np.empty(n, np.float32) if particles else None
        z = np.empty(n, np.float32) if particles else None
        ranges = np.zeros((cap, 2), np.uint32)
        heap = np.zeros(cap, CELL_DTYPE)
        exists = np.zeros(cap, np.uint8)
        info = RequestInfo()
        rc = lib().orb_server_fetch(self.h, rid, _p(x), _p(y), _p(z), _p(ranges), _p(heap),
                                    _p(exists), cap, C.byref(info))
        if rc < 0 and rc != info.error:  # a failed call, not the build's own status
            raise OrbError(rc, "orb_server_fetch")
        return {"x": x, "y": y, "z": z, "ranges": ranges, "heap": heap, "exists": exists,
                "error": info.error, "error_cell": info.error_cell, "nlevels": info.nlevels,
                "level_passes": list(info.level_passes[:info.nlevels]),
                "total_passes": info.total_passes, "submit_ms": info.submit_ms,
                "admit_ms": info.admit_ms, "done_ms": info.done_ms}

    def stats(self):
        st = ServerStats()
        _check(lib().orb_server_get_stats(self.h, C.byref(st)), "orb_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}


def level_shape(n, d, level):
    """(cells, fewest keys, most keys) of a level of a full build (orb_level_shape)."""
    lo, hi = C.c_int64(), C.c_int64()
    cells = _check(lib().orb_level_shape(int(n), int(d), int(level), C.byref(lo), C.byref(hi)),
                   "orb_level_shape")
    return cells, lo.value, hi.value


def write_tune_table(path, rows):
    arr = _rows_in(rows)
    return _check(lib().orb_tune_table_write(os.fsencode(path), arr, len(rows)),
                  "orb_tune_table_write")


def read_tune_table(path, cap=256):
    arr = (TuneRow * cap)()
    n = _check(lib().orb_tune_table_read(os.fsencode(path), arr, cap), "orb_tune_table_read")
    return _rows_out(arr, n)


def generate(n, seed=0):
    """xorshf96 particles (host code, no GPU): seed 0 is the reference's
    stream (src/services/init.cu:11-25, 47-53)."""
    x, y, z = (np.empty(int(n), np.float32) for _ in range(3))
    _check(lib().orb_generate(int(seed) & (2**64 - 1), int(n), _p(x), _p(y), _p(z)),
           "orb_generate")
    return x, y, z


def generate_plummer(n, seed=0, n_spheres=64, a=0.01, background=0.1):
    """Plummer spheres plus a uniform background (host code, no GPU)."""
    x, y, z = (np.empty(int(n), np.float32) for _ in range(3))
    _check(lib().orb_generate_plummer(int(seed) & (2**64 - 1), int(n), int(n_spheres), float(a),
                                      float(background), _p(x), _p(y), _p(z)),
           "orb_generate_plummer")
    return x, y, z


def read_tipsy(path):
    """(header dict, x, y, z) of a Tipsy snapshot (host code, no GPU)."""
    h = TipsyHeader()
    p = str(path).encode()
    _check(lib().orb_read_tipsy(p, C.byref(h), None, None, None, 0), "orb_read_tipsy")
    n = h.nbodies
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    z = np.empty(n, np.float32)
    _check(lib().orb_read_tipsy(p, C.byref(h), _p(x), _p(y), _p(z), n), "orb_read_tipsy")
    hd = {k: getattr(h, k) for k, _ in TipsyHeader._fields_}
    return hd, x, y, z
