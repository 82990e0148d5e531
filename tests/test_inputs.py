"""The library's synthetic-input generators (host code, no GPU).

orb_generate(seed 0) is the reference stream, checked against the oracle's
restatement of src/services/init.cu:11-25, 47-53; other seeds are
deterministic, distinct streams of the same recurrence.  orb_generate_plummer
is checked against the workload's definition: split, box, profile."""
import numpy as np
import pytest


def test_seed_zero_is_the_reference_stream(orb_host, oracle):
    for n in (1, 7, 1000, (1 << 16) + 5):
        a = orb_host.generate(n)
        b = oracle.generate(n)
        for u, v in zip(a, b):
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32)), n


def test_seeded_streams_deterministic_and_distinct(orb_host):
    a = orb_host.generate(50000, seed=12345)
    b = orb_host.generate(50000, seed=12345)
    c = orb_host.generate(50000, seed=12346)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert not np.array_equal(a[0], c[0])
    for u in a:
        assert u.min() >= -0.5 and u.max() <= 0.5
        assert abs(float(u.mean())) < 0.01 and abs(float(u.var()) - 1 / 12) < 0.002
    # the prefix of a longer stream is the shorter stream (i-major draws)
    d = orb_host.generate(1000, seed=12345)
    for u, v in zip(a, d):
        assert np.array_equal(u[:1000], v)


def test_plummer_box_split_and_determinism(orb_host):
    n = 300001
    x, y, z = orb_host.generate_plummer(n, seed=7, n_spheres=64, a=0.01, background=0.1)
    x2, _, z2 = orb_host.generate_plummer(n, seed=7, n_spheres=64, a=0.01, background=0.1)
    assert np.array_equal(x, x2) and np.array_equal(z, z2)
    for u in (x, y, z):
        assert u.dtype == np.float32
        assert u.min() >= -0.5 and u.max() < 0.5
    x3, _, _ = orb_host.generate_plummer(n, seed=8)
    assert not np.array_equal(x, x3)


def test_plummer_profile_single_sphere(orb_host):
    """One sphere of a = 0.02, no background: the fraction of particles inside
    the half-mass radius a / sqrt(2^(2/3) - 1) is one half (the box clips
    the far tail only)."""
    n = 200000
    a = 0.02
    x, y, z = orb_host.generate_plummer(n, seed=3, n_spheres=1, a=a, background=0.0)
    p = np.stack([x, y, z], 1).astype(np.float64)
    c = np.median(p, axis=0)
    r = np.linalg.norm(p - c, axis=1)
    r_half = a / np.sqrt(2 ** (2 / 3) - 1)
    assert abs(float((r < r_half).mean()) - 0.5) < 0.02
    # isotropy: the direction's mean is near zero
    dirs = (p - c) / r[:, None]
    assert np.abs(dirs.mean(axis=0)).max() < 0.02


def test_plummer_background_only_is_uniform(orb_host):
    x, y, z = orb_host.generate_plummer(100000, seed=4, n_spheres=8, a=0.01, background=1.0)
    for u in (x, y, z):
        assert abs(float(u.mean())) < 0.01 and abs(float(u.var()) - 1 / 12) < 0.003


def test_plummer_sphere_shares_are_equal(orb_host):
    """Four far-apart tight spheres: each holds a quarter of the clustered
    particles (the first spheres take the remainder)."""
    n = 40003
    x, y, z = orb_host.generate_plummer(n, seed=11, n_spheres=4, a=1e-4, background=0.0)
    p = np.stack([x, y, z], 1)
    # group by nearest of the four densest points
    from collections import Counter
    keys = Counter(map(tuple, np.round(p, 2)))
    centres = np.array([k for k, _ in keys.most_common(4)])
    owner = np.argmin(((p[:, None, :] - centres[None]) ** 2).sum(-1), axis=1)
    counts = sorted(np.bincount(owner, minlength=4).tolist())
    assert counts == [10000, 10001, 10001, 10001]


def test_generator_argument_errors(orb_host):
    with pytest.raises(orb_host.OrbError):
        orb_host.generate_plummer(10, n_spheres=0)
    with pytest.raises(orb_host.OrbError):
        orb_host.generate_plummer(10, a=-1.0)
    with pytest.raises(orb_host.OrbError):
        orb_host.generate_plummer(10, background=1.5)
