"""The reference's master loop (src/orbit.cpp:74-275) bound to the C-ABI, as a
maintainer of the reference would write it: tests/binding/orbit_binding.cpp,
compiled against the reference's own src/cell.h and include/orb.h (the
static_asserts there are the Cell drop-in claim), calls orb_run_service with
the reference's service ids and Cell[] payloads, keeps the reference's own
termination rule and Cell::cut split, and compares the tree, range map and
particles it ends with against orb_build (plain and batched) on the same
particles.  `orbit 20 10 <o>` for o = 0, 1, 2.

The program is built by `make -C oracle` where the reference tree exists (it
includes the reference's header) into oracle/_ref; without it the test is
skipped."""
import os
import subprocess

import pytest

from conftest import ORACLE

pytestmark = pytest.mark.gpu

BIN = os.path.join(ORACLE, "_ref", "orbit_binding")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/orbit_binding not built")
@pytest.mark.parametrize("o", [0, 1, 2])
def test_reference_master_loop_on_the_abi(orb, o):
    r = subprocess.run([BIN, "20", "10", str(o)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "binding ok" in r.stdout
