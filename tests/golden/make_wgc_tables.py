"""Writes the WGC kernel-table fixtures the GPU tests read.

The 1D table (eta, w, w', w'') of the second-order WGC kernel is what
makeKernel2 (KEDF_KernelODE.f90:254-478) integrates with RKSUITE at TOL 5e-5.
The reference's own RKSUITE, compiled from its sources into oracle/_ref by
oracle/ref/Makefile, produces it here; the GPU tests and the CPU tests that
run without that library read the committed result instead, so nothing built
from the reference has to be present where the tests run.

  python tests/golden/make_wgc_tables.py      (needs oracle/_ref built)

Table written (alpha, beta, gamma):
  wgc_table_default.npz  (5+sqrt5)/6, (5-sqrt5)/6, 2.7  (SetupKEDF.f90:285-314).
The alpha and beta written out in test/optDen/test.inpt:6-7 are these values
to 32 digits, so the same table serves that sample.
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

TABLES = {
    "default": ((5.0 + math.sqrt(5.0)) / 6.0, (5.0 - math.sqrt(5.0)) / 6.0, 2.7),
}


def path_of(name):
    return os.path.join(HERE, "wgc_table_%s.npz" % name)


def main():
    sys.path.insert(0, ROOT)
    from oracle import refbuild as R
    if not R.available() and not R.build():
        raise SystemExit("oracle/_ref/libprofess_ref.so is not built and /root/reference is absent")
    for name, (a, b, g) in TABLES.items():
        eta, w, w1, w2 = R.wgc_eta_table(a, b, g)
        np.savez_compressed(path_of(name), params=np.array([a, b, g]), eta=eta, w=w, w1=w1, w2=w2)
        print("wrote %s: %d knots" % (path_of(name), eta.size))


if __name__ == "__main__":
    main()
