"""The autotuner's persisted table (csrc/orb_tune_table.cpp), host side:
the level key (a level's cell-size histogram under the count rule) pinned
against the oracle's own builds, and the text table's round trip and
rejection of malformed files.  No GPU needed."""
import numpy as np
import pytest

import orb_amd as orb


def actual_level_sizes(r, level):
    """Keys per split cell of a level in an oracle build: the cells of depth
    level-1 with at least two leaves (the ones the level splits)."""
    a, e = (1 << (level - 1)) - 1, (1 << level) - 2
    ids = [i for i in range(a, min(e, len(r.exists) - 1) + 1)
           if r.exists[i] and r.heap[i]["nLeafCells"] >= 2]
    return np.array([int(r.ranges[i][1]) - int(r.ranges[i][0]) for i in ids], np.int64)


@pytest.mark.parametrize("n,d", [(200003, 300), (1 << 17, 1024), (65537, 7), (100000, 1000)])
def test_level_shape_matches_oracle_builds(oracle, n, d):
    """Count and sizes of every level's cells: the shape's cell count is
    exact, and the fewest / most keys agree with the oracle's build within
    the termination rule's slack (|left - target| < 3 per split)."""
    x, y, z = oracle.generate(n)
    r = oracle.build(x, y, z, d, levels_mode=1)
    level = 1
    while True:
        cells, lo, hi = orb.level_shape(n, d, level)
        sizes = actual_level_sizes(r, level)
        assert cells == len(sizes), level
        if cells == 0:
            break
        slack = 3 * level
        assert abs(int(sizes.min()) - lo) <= slack, (level, sizes.min(), lo)
        assert abs(int(sizes.max()) - hi) <= slack, (level, sizes.max(), hi)
        level += 1
    assert level - 1 == int(np.ceil(np.log2(d))), level


def test_level_shape_uneven_leaves():
    """d = 300: level 9 splits only the 44 cells of two leaves (300 - 256),
    each holding 2/300 of the keys."""
    cells, lo, hi = orb.level_shape(1000003, 300, 9)
    assert cells == 44
    assert lo == hi == 1000003 * 2 // 300
    cells, lo, hi = orb.level_shape(1000003, 300, 8)
    assert cells == 128 and lo < hi       # 2- and 3-leaf cells


def test_level_shape_rejects_bad_arguments():
    for args in ((-1, 8, 1), (100, 0, 1), (100, 8, 0)):
        with pytest.raises(orb.OrbError) as e:
            orb.level_shape(*args)
        assert e.value.code == orb.ERR_ARG


ROWS = [
    {"depth": 1, "log2_min_keys": 26, "log2_max_keys": 26, "chunk_blocks": 0,
     "subtree_blocks": 0, "subtree_max_keys": -1, "ms": 2.0625},
    {"depth": 4, "log2_min_keys": 23, "log2_max_keys": 24, "chunk_blocks": 16384,
     "subtree_blocks": 1024, "subtree_max_keys": 4400, "ms": 0.1 + 0.2},
]


def test_table_round_trip(tmp_path):
    p = tmp_path / "tune.txt"
    assert orb.write_tune_table(p, ROWS) == 2
    text = p.read_text()
    assert text.startswith("#")
    p.write_text(text + "\n   # a comment line\n\n")
    back = orb.read_tune_table(p)
    assert back == ROWS                   # ms bit-exact through %.17g


@pytest.mark.parametrize("line", [
    "1 26 26 0 0 -1",                     # a field missing
    "1 26 26 0 0 -1 1.0 7",               # one too many
    "0 26 26 0 0 -1 1.0",                 # depth 0
    "1 27 26 0 0 -1 1.0",                 # max below min
    "1 26 26 -5 0 -1 1.0",                # negative grid cap
    "1 26 26 0 0 -1 nan",                 # not a time
    "1 26 26 0 0 -1 1.0x",                # trailing junk
])
def test_table_rejects_malformed_rows(tmp_path, line):
    p = tmp_path / "bad.txt"
    p.write_text("# header\n1 26 26 0 0 -1 1.0\n" + line + "\n")
    with pytest.raises(orb.OrbError) as e:
        orb.read_tune_table(p)
    assert e.value.code == orb.ERR_ARG


def test_table_read_limits(tmp_path):
    p = tmp_path / "t.txt"
    orb.write_tune_table(p, ROWS)
    with pytest.raises(orb.OrbError):
        orb.read_tune_table(p, cap=1)      # more rows than the array holds
    with pytest.raises(orb.OrbError):
        orb.read_tune_table(tmp_path / "missing.txt")
