"""Loader for oracle/_ref/libprofess_ref.so (test infrastructure only).

The library holds the reference's own self-contained Fortran files compiled
from their sources by oracle/ref/Makefile (MathSplines.f90, MathDCsrch.f,
MathDCstep.f, MathRKsuite.f, MathLNsrch.f90, Constants.f90, and the serial
OutputFiles.f90 / MPI_Functions.f90 that MathLNsrch uses) plus
oracle/ref/ref_driver.f90.
``available()`` is False when it was not built (no /root/reference here).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libprofess_ref.so")
REF_SOURCE = "/root/reference/Source"
_lib = None


def build(force=False):
    """Compile the reference's Fortran files into oracle/_ref (needs the tree)."""
    if not os.path.isdir(REF_SOURCE):
        return False
    if os.path.exists(LIB_PATH) and not force:
        return True
    subprocess.run(["make", "-C", os.path.join(_HERE, "ref"), "-s"], check=True)
    return os.path.exists(LIB_PATH)


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        i = ctypes.POINTER(ctypes.c_int)
        _lib.ref_spline_set.argtypes = [ctypes.c_int, d, d, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double, d]
        _lib.ref_spline_val.argtypes = [ctypes.c_int, d, d, d, ctypes.c_int, d, d]
        _lib.ref_dcsrch.argtypes = [d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, i, ctypes.c_double,
                                    ctypes.c_double, i, d]
        _lib.ref_wgc_eta_table.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, d, d, d, d]
        _lib.ref_wgc_eta_table.restype = ctypes.c_int
        _lib.ref_lnsrch.argtypes = [d, ctypes.c_double, ctypes.c_double, ctypes.c_char_p, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, d]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def spline_set(t, y, ibcbeg, ybcbeg, ibcend, ybcend):
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ypp = np.zeros_like(t)
    lib().ref_spline_set(t.size, _p(t), _p(y), ibcbeg, ybcbeg, ibcend, ybcend, _p(ypp))
    return ypp


def spline_val(t, y, ypp, tvals):
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ypp = np.ascontiguousarray(ypp, dtype=np.float64)
    tv = np.ascontiguousarray(tvals, dtype=np.float64).ravel()
    out = np.zeros_like(tv)
    lib().ref_spline_val(t.size, _p(t), _p(y), _p(ypp), tv.size, _p(tv), _p(out))
    return out


class RefDcsrch:
    """Same interface as profess_oracle.Dcsrch, backed by the Fortran dcsrch."""

    def __init__(self, ftol, gtol, xtol, stpmin, stpmax):
        self.args = (ftol, gtol, xtol)
        self.stpmin, self.stpmax = stpmin, stpmax
        self.code = ctypes.c_int(0)
        self.isave = (ctypes.c_int * 2)()
        self.dsave = (ctypes.c_double * 13)()

    def step(self, stp, f, g):
        s = ctypes.c_double(stp)
        lib().ref_dcsrch(ctypes.byref(s), f, g, self.args[0], self.args[1], self.args[2],
                         ctypes.byref(self.code), self.stpmin, self.stpmax, self.isave, self.dsave)
        task = {1: "FG", 2: "CONVERGENCE", 3: "WARNING", 4: "ERROR"}[self.code.value]
        return s.value, task


class RefLnSrch:
    """Same interface as profess_oracle.LnSrch, backed by the Fortran LnSrch
    (its module state lives in the library: one instance at a time)."""

    def __init__(self):
        self.points = np.full((2, 3), -1.0)

    def step(self, alp, f, g, task, c1, c2, stpmin, stpmax):
        a = ctypes.c_double(alp)
        buf = ctypes.create_string_buffer(task.ljust(70).encode()[:70], 70)
        pts = np.zeros(6)
        lib().ref_lnsrch(ctypes.byref(a), f, g, buf, c1, c2, 1e-12, stpmin, stpmax, _p(pts))
        self.points = pts.reshape(3, 2).T.copy()    # Fortran points(2,3), column-major
        return a.value, buf.raw.decode().rstrip()


def wgc_eta_table(alpha, beta, gamma):
    """1D WGC kernel table (eta, w, w', w'') integrated by the reference's RKSUITE."""
    nmax = 20000
    eta = np.zeros(nmax)
    w = np.zeros(nmax)
    w1 = np.zeros(nmax)
    w2 = np.zeros(nmax)
    n = lib().ref_wgc_eta_table(alpha, beta, gamma, nmax, _p(eta), _p(w), _p(w1), _p(w2))
    if n <= 0:
        raise RuntimeError("RKSUITE failed (code %d)" % n)
    return eta[:n].copy(), w[:n].copy(), w1[:n].copy(), w2[:n].copy()
