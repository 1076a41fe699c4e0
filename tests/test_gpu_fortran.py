"""GPU: the Fortran host (ISO_C_BINDING drop-in modules CalPotPlus/RhoOptN,
profess_amd/fortran) on the reference's test/optIon input."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "profess_amd", "fortran", "ofd_fortran_test")


def test_fortran_host_optIon():
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built (amdflang)")
    r = subprocess.run([EXE, os.path.join(ROOT, "tests", "golden", "al_HC.lda.recpot")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
