"""Writes tests/golden/cell_vectors.json from the reference's own src/cell.h.

Runs in the build container only, where oracle/_ref/libref_cell.so is built
from the reference tree by `make -C oracle`.  The fixture holds inputs and the
reference's outputs (float32 values as uint32 bit patterns) for:
  - Cell::getCut on random and edge margins        (src/cell.h:74-76)
  - Cell::cut + setCutAxis + setCutMargin (split)   (src/cell.h:78-126, orbit.cpp:235-243)
  - getNLevels / getNCellsOnLastLevel / getTotalNumberOfCells / child ids
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402  (Cell layout only)

Cell = pyoracle.Cell


def bits(v):
    return int(np.float32(v).view(np.uint32))


def cell_to_json(c):
    return {"id": c.id, "nLeafCells": c.nLeafCells, "prevCutAxis": c.prevCutAxis,
            "cutAxis": c.cutAxis, "foundCut": bool(c.foundCut),
            "cutMarginLeft": bits(c.cutMarginLeft), "cutMarginRight": bits(c.cutMarginRight),
            "lower": [bits(v) for v in c.lower], "upper": [bits(v) for v in c.upper]}


def main():
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_cell.so"))
    fp = C.POINTER(C.c_float)
    ref.ref_make.argtypes = [C.c_int, C.c_int, fp, fp, C.POINTER(Cell)]
    ref.ref_get_cut.argtypes = [C.POINTER(Cell)]
    ref.ref_get_cut.restype = C.c_float
    ref.ref_split.argtypes = [C.POINTER(Cell), C.POINTER(Cell), C.POINTER(Cell)]
    for f in ("ref_n_levels", "ref_n_cells_on_last_level", "ref_total_cells",
              "ref_left_child_id", "ref_right_child_id"):
        getattr(ref, f).argtypes = [C.c_int]
    assert ref.ref_cell_size() == 52

    rng = np.random.default_rng(20220621)
    out = {"source": "reference src/cell.h compiled by oracle/Makefile", "get_cut": [],
           "split": [], "levels": []}

    # getCut: random margins, halving edge cases, subnormals, opposite signs
    margins = [(-0.5, 0.5), (0.0, 0.0), (-1e-45, 1e-45), (1e-45, 3e-45), (0.25, 0.25000003),
               (-0.3, 0.7), (3.4e38, 3.4e38), (1e-40, 2e-40)]
    for _ in range(200):
        a, b = np.sort(rng.uniform(-0.5, 0.5, 2).astype(np.float32))
        margins.append((float(a), float(b)))
    for _ in range(100):
        a = np.float32(rng.uniform(-0.5, 0.5))
        margins.append((float(a), float(np.nextafter(a, np.float32(1)))))
    for lo, hi in margins:
        c = Cell()
        c.cutMarginLeft = np.float32(lo)
        c.cutMarginRight = np.float32(hi)
        out["get_cut"].append({"left": bits(lo), "right": bits(hi),
                               "cut": bits(ref.ref_get_cut(C.byref(c)))})

    # split: random boxes, axes, margins and leaf counts (incl. odd d)
    for k in range(300):
        lo = rng.uniform(-0.5, 0.0, 3).astype(np.float32)
        hi = (lo + rng.uniform(1e-6, 0.5, 3)).astype(np.float32)
        if k % 7 == 0:  # equal extents: tie-break on the first max
            hi = (lo + np.float32(0.25)).astype(np.float32)
        p = Cell()
        ref.ref_make(int(rng.integers(0, 5000)), int(rng.integers(2, 4097)),
                     lo.ctypes.data_as(fp), hi.ctypes.data_as(fp), C.byref(p))
        p.cutAxis = int(rng.integers(0, 3))
        ax = p.cutAxis
        a, b = np.sort(rng.uniform(lo[ax], hi[ax], 2).astype(np.float32))
        p.cutMarginLeft, p.cutMarginRight = a, b
        p.foundCut = bool(rng.integers(0, 2))
        l, r = Cell(), Cell()
        rc = ref.ref_split(C.byref(p), C.byref(l), C.byref(r))
        if rc != 0:
            continue
        out["split"].append({"parent": cell_to_json(p), "left": cell_to_json(l),
                             "right": cell_to_json(r)})

    for d in list(range(1, 70)) + [127, 128, 129, 1000, 1023, 1024, 1025, 4096, 16384, 1 << 20]:
        out["levels"].append({"d": d, "n_levels": ref.ref_n_levels(d),
                              "n_cells_on_last_level": ref.ref_n_cells_on_last_level(d),
                              "total_cells": ref.ref_total_cells(d),
                              "left_child": ref.ref_left_child_id(d),
                              "right_child": ref.ref_right_child_id(d)})

    with open(os.path.join(HERE, "cell_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {len(out['get_cut'])} getCut, {len(out['split'])} split, "
          f"{len(out['levels'])} level vectors")


if __name__ == "__main__":
    main()
