"""Tipsy snapshot reader (host code in the HIP library; no GPU needed):
positions of gas, dark and star records in file order, both byte orders,
and the malformed-file errors."""
import numpy as np
import pytest

import inputs


def pos(rng, m):
    return (rng.random((m, 3)) - 0.5).astype(np.float32)


@pytest.mark.parametrize("big_endian", [True, False])
def test_read_tipsy_positions(orb_host, tmp_path, big_endian):
    orb = orb_host
    rng = np.random.default_rng(3)
    g, d, s = pos(rng, 17), pos(rng, 1000), pos(rng, 5)
    path = tmp_path / "snap.std"
    inputs.write_tipsy(path, g, d, s, big_endian=big_endian, time=0.25)
    h, x, y, z = orb.read_tipsy(path)
    assert (h["nbodies"], h["nsph"], h["ndark"], h["nstar"]) == (1022, 17, 1000, 5)
    assert h["ndim"] == 3 and h["time"] == 0.25 and h["big_endian"] == int(big_endian)
    want = np.concatenate([g, d, s])
    assert np.array_equal(np.stack([x, y, z], 1).view(np.uint32), want.view(np.uint32))


def test_read_tipsy_dark_only(orb_host, tmp_path):
    orb = orb_host
    rng = np.random.default_rng(4)
    d = pos(rng, 4096)
    path = tmp_path / "dark.std"
    inputs.write_tipsy(path, dark=d)
    h, x, y, z = orb.read_tipsy(path)
    assert h["nbodies"] == 4096 and h["nsph"] == 0 and h["nstar"] == 0
    assert np.array_equal(np.stack([x, y, z], 1), d)


def test_read_tipsy_errors(orb_host, tmp_path):
    orb = orb_host
    rng = np.random.default_rng(5)
    path = tmp_path / "t.std"
    inputs.write_tipsy(path, dark=pos(rng, 100))
    raw = path.read_bytes()
    (tmp_path / "trunc.std").write_bytes(raw[:-4])
    (tmp_path / "bad.std").write_bytes(raw[:12] + b"\x00\x00\x00\x02" + raw[16:])  # ndim 2
    (tmp_path / "short.std").write_bytes(raw[:20])
    for name in ("trunc.std", "bad.std", "short.std", "missing.std"):
        with pytest.raises(orb.OrbError) as e:
            orb.read_tipsy(tmp_path / name)
        assert e.value.code == orb.ERR_ARG, name
