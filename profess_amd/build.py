"""Builds profess_amd/libprofess_amd.so (gfx950) with hipcc, in tree.

Usage: python -m profess_amd.build    (or profess_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libprofess_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["ofd_kernels.hip", "ofd_fast.hip", "ofd_fast_d3.hip", "ofd_fast_d4.hip", "ofd_fast_f3.hip",
           "ofd_fast_f4.hip", "ofd_fast_d8.hip", "ofd_fast_f8.hip",
           "ofd_fast_d5.hip", "ofd_fast_f5.hip", "ofd_fast_d7.hip", "ofd_fast_f7.hip", "ofd_tables.hip", "ofd_plan.cpp",
           "ofd_api.cpp", "ofd_io.cpp"]
COMMON = ["-std=c++17", "-O3", "-fPIC", "-pthread", "-Wno-unused-result", "-Wno-unused-value"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _stale(src, out):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(os.path.dirname(HERE), "include", "ofd.h"))
    return any(os.path.getmtime(p) > t for p in deps if os.path.exists(p))


def _compile(src, force=True):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if not force and not _stale(src, out):
        return out
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return out


def build(verbose=False, force=True):
    """Compile every source (force=True, the default) or only stale ones."""
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def build_fortran(verbose=False):
    """ISO_C_BINDING drop-in modules + the Fortran host test program, linked
    against libprofess_amd.so (amdflang)."""
    if not os.path.exists(FC):
        return None
    out = os.path.join(FORTRAN, "ofd_fortran_test")
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmd = [FC, "-O2", "-module-dir", mod, "-I" + mod, os.path.join(FORTRAN, "ofd_profess.f90"),
           os.path.join(FORTRAN, "ofd_fortran_test.f90"), "-o", out, "-L" + HERE, "-lprofess_amd",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + r.stderr[-4000:])
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True, force="--incremental" not in sys.argv)
    build_fortran(verbose=True)
    sys.exit(0)
