"""The C-ABI library loads and exports every entry point include/*.h declares.

No compute calls: this runs without a GPU.
"""
import ctypes as C
import glob
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b(orb_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("orb_init", "orb_finalize", "orb_count_left", "orb_partition", "orb_build",
                 "orb_run_service", "orb_count_left_gpu", "orb_partition_gpu"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import orb_amd
    lib = orb_amd.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    import orb_amd
    assert declared_symbols() <= set(orb_amd.PROTOTYPES)


def test_cell_layout_matches_reference():
    import orb_amd
    import pyoracle
    assert C.sizeof(orb_amd.Cell) == 52
    assert orb_amd.CELL_DTYPE.itemsize == 52
    for (n1, t1), (n2, t2) in zip(orb_amd.Cell._fields_, pyoracle.Cell._fields_):
        assert n1 == n2
        assert getattr(orb_amd.Cell, n1).offset == getattr(pyoracle.Cell, n2).offset


def test_no_gpu_is_reported_not_faked():
    import orb_amd
    lib = orb_amd.lib()
    if lib.orb_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    rc = lib.orb_init(C.byref(h), 1024, 16, True, orb_amd.MetaParams(False, False, False))
    assert rc == orb_amd.ERR_NODEV and not h.value
    with pytest.raises(orb_amd.OrbError):
        orb_amd.Context(1024, 16)


def test_error_strings():
    import orb_amd
    lib = orb_amd.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        assert lib.orb_error_string(code)
