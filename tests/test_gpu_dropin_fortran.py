"""The reference driver's call sequence in Fortran (oracle/dropin_driver.f90,
built by `make -C oracle dropin` against the reference's own modules and
graph routines) runs against libwtomo3d_hip.so: ASMBLE through the shim
wtomo3d_amd/fortran/asmble.f90 and STRUMPACK_SERIAL_FINTER jobs 1/2/3/-1."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin_driver")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in driver not built (needs the reference sources)")
def test_fortran_driver_against_library():
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    out = r.stdout
    assert r.returncode == 0, out + r.stderr
    for job in ("job 1 ierr", "asmble ierr", "job 2 ierr", "job 3 ierr", "job -1 ierr"):
        line = [l for l in out.splitlines() if job in l][0]
        assert int(line.split()[-1]) == 0, out
    assert "DROPIN OK" in out, out
