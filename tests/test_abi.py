"""CPU checks of the drop-in boundary: the library loads and exports every
entry point include/ofd.h declares (no compute call: there is no GPU here),
and the host-side helpers behave."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "ofd.h")) as fh:
        txt = fh.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ofd_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_reference_surface():
    names = _declared()
    for n in ["ofd_calpot", "ofd_rho_opt", "ofd_create", "ofd_set_ion_terms", "ofd_ion_forces",
              "ofd_build_ion_potential", "ofd_set_wgc_table"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from profess_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == _declared()


def test_request_bytes_is_monotone_and_host_only():
    from profess_amd import _lib, request_bytes
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    b32 = request_bytes((32, 32, 32))
    b128 = request_bytes((128, 128, 128))
    bw = request_bytes((128, 128, 128), kinetic=5)
    assert 0 < b32 < b128 < bw
    # about 25 real grids at 128^3
    assert 0.2e9 < b128 < 0.6e9


def test_server_thread_starts_waits_and_stops_without_pools():
    """The progress server's thread and its wait/stop protocol are host-only
    (no pool, no device call): a wait with nothing to drive times out with 0."""
    import ctypes as C
    import time
    from profess_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    h = C.c_void_p()
    assert L.ofd_server_create(C.byref(h)) == 0 and h.value
    t = time.perf_counter()
    assert L.ofd_server_wait(h, 0.05) == 0
    assert time.perf_counter() - t >= 0.04
    assert L.ofd_server_last_error(h) == b""
    assert L.ofd_server_add_pool(h, None) < 0
    L.ofd_server_destroy(h)


def test_input_generators():
    from profess_amd import inputs
    lat, frac = inputs.fcc_supercell(7.65, 2)
    assert frac.shape == (32, 3) and np.allclose(np.diag(lat), 15.3)
    lat, frac = inputs.hcp_supercell(6.06, 9.84, (2, 2, 1))
    assert frac.shape == (8, 3)
    c = inputs.al_fcc_case((1, 1, 1), (16, 16, 16))
    vol = abs(np.linalg.det(c["lattice"]))
    rho0 = c["num_electrons"] / vol
    assert c["rho"].shape == (16, 16, 16)
    assert np.all(c["rho"] >= 0.7 * rho0 - 1e-15) and np.all(c["rho"] <= 1.3 * rho0 + 1e-15)
    mg, v = inputs.synthetic_recpot(3, 1.2)
    assert v.size == 15001 and abs(mg * inputs.BOHR - 30.0) < 1e-12


def test_ctypes_structs_match_the_header(tmp_path):
    """Every struct _lib.py passes through ctypes has the size and field
    offsets the C compiler gives the same struct in include/ofd.h."""
    import ctypes as C
    import shutil
    import subprocess
    from profess_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = [v for k, v in vars(_lib).items()
               if k.startswith("ofd_") and isinstance(v, type) and issubclass(v, C.Structure)]
    assert len(structs) >= 10
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "ofd.h"', "int main(void) {"]
    for s in structs:
        n = s.__name__
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (n, n))
        for f, _ in s._fields_:
            # a C field named as a Python keyword carries a trailing _ in ctypes
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (n, f, n, f.rstrip("_")))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for s in structs:
        n = s.__name__
        assert int(got[n]) == C.sizeof(s), n
        for f, _ in s._fields_:
            assert int(got["%s.%s" % (n, f)]) == getattr(s, f).offset, (n, f)


def test_fortran_bind_c_types_match_the_header(tmp_path):
    """The Fortran binding's BIND(C) types (profess_amd/fortran/ofd_profess.f90)
    have the size and the field offsets of the header's structs they stand
    for (offsets matter: a missing trailing int can hide in the padding)."""
    import shutil
    import subprocess
    from profess_amd import _lib
    fc = "/opt/rocm/bin/amdflang"
    cc = shutil.which("gcc") or shutil.which("cc")
    if not os.path.exists(fc) or cc is None:
        pytest.skip("no Fortran or C compiler")
    pairs = {"ofd_grid_t": "ofd_grid", "ofd_functional_t": "ofd_functional",
             "ofd_opt_params_t": "ofd_opt_params", "ofd_opt_stats_t": "ofd_opt_stats"}
    fields = {c: [f.rstrip("_") for f, _ in getattr(_lib, c)._fields_] for c in pairs.values()}
    f90 = ["PROGRAM layout", "  USE ISO_C_BINDING", "  USE ofd_c", "  INTEGER(C_INTPTR_T) :: b"]
    f90 += ["  TYPE(%s), TARGET :: v%d" % (t, k) for k, t in enumerate(pairs)]
    for k, (t, c) in enumerate(pairs.items()):
        f90.append('  PRINT *, "%s", C_SIZEOF(v%d)' % (c, k))
        f90.append("  b = TRANSFER(C_LOC(v%d), b)" % k)
        for f in fields[c]:
            f90.append('  PRINT *, "%s.%s", TRANSFER(C_LOC(v%d%%%s), b) - b' % (c, f, k, f))
    f90 += ["END PROGRAM layout"]
    (tmp_path / "layout.f90").write_text("\n".join(f90) + "\n")
    src = os.path.join(ROOT, "profess_amd", "fortran", "ofd_profess.f90")
    subprocess.run([fc, "-c", "-module-dir", str(tmp_path), src, "-o", str(tmp_path / "ofd_profess.o")], check=True)
    subprocess.run([fc, "-I", str(tmp_path), str(tmp_path / "layout.f90"), "-o", str(tmp_path / "layout_f")],
                   check=True, cwd=str(tmp_path))
    out = subprocess.run([str(tmp_path / "layout_f")], capture_output=True, text=True, check=True).stdout.split()
    fort = {out[i]: int(out[i + 1]) for i in range(0, len(out), 2)}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "ofd.h"', "int main(void) {"]
    for c in pairs.values():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (c, c))
        lines += ['  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (c, f, c, f) for f in fields[c]]
    lines += ["  return 0;", "}"]
    (tmp_path / "layout.c").write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(tmp_path / "layout.c"), "-o",
                    str(tmp_path / "layout_c")], check=True)
    cout = dict(l.split() for l in subprocess.run([str(tmp_path / "layout_c")], capture_output=True, text=True,
                                                   check=True).stdout.splitlines())
    assert len(fort) == len(cout) >= 30
    for k, v in cout.items():
        assert fort[k] == int(v), (k, fort[k], v)
