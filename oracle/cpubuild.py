"""Builds oracle/cpu/libofd_cpu.so, the CPU restatement of one WT
evaluation (oracle/cpu/ofd_cpu.cpp) that bench.py times as its CPU baseline.
Test infrastructure only. Built with g++ and OpenMP for x86-64-v3 (AVX2 and
FMA), which the build container and the GPU hosts both run; complex products
without the C99 infinity/NaN recovery (-fcx-limited-range), as Fortran's are."""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu", "ofd_cpu.cpp")
OUT = os.path.join(HERE, "cpu")
LIB = os.path.join(OUT, "libofd_cpu.so")


def build(force=False):
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    os.makedirs(OUT, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fcx-limited-range", "-fopenmp", "-shared", "-fPIC", "-std=c++17", SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.ofd_cpu_create.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.POINTER(C.c_void_p)]
        L.ofd_cpu_destroy.argtypes = [C.c_void_p]
        L.ofd_cpu_destroy.restype = None
        L.ofd_cpu_calpot.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


class CpuEval:
    """One grid and functional: lattice columns in bohr (as ofd_grid.cell),
    rho0 the average density the WT kernel is built for."""

    def __init__(self, lattice, grid, rho0, lam=1.0, mu=1.0, alpha=-1.0, beta=-1.0):
        import numpy as np
        self.shape = (grid[2], grid[1], grid[0])
        cell = np.ascontiguousarray(np.asarray(lattice, dtype=np.float64).T.ravel())
        h = C.c_void_p()
        rc = lib().ofd_cpu_create(int(grid[0]), int(grid[1]), int(grid[2]),
                                  cell.ctypes.data_as(C.POINTER(C.c_double)), lam, mu, alpha, beta, rho0,
                                  C.byref(h))
        if rc != 0:
            raise RuntimeError("ofd_cpu_create failed (%d)" % rc)
        self._h = h

    def calpot(self, rho, vion, e_ionion=0.0, opt_sqrt=True, threads=1):
        import numpy as np
        r = np.ascontiguousarray(rho, dtype=np.float64)
        v = np.ascontiguousarray(vion, dtype=np.float64)
        pot = np.empty_like(r)
        et = np.zeros(9)
        dp = C.POINTER(C.c_double)
        rc = lib().ofd_cpu_calpot(self._h, r.ctypes.data_as(dp), v.ctypes.data_as(dp), float(e_ionion),
                                  1 if opt_sqrt else 0, int(threads), pot.ctypes.data_as(dp), et.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError("ofd_cpu_calpot failed (%d)" % rc)
        return pot, et

    def close(self):
        if self._h is not None and self._h.value:
            lib().ofd_cpu_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
