"""The committed WGC kernel-table fixtures (tests/golden/wgc_table_*.npz,
written by tests/golden/make_wgc_tables.py from the reference's RKSUITE
integration of makeKernel2, KEDF_KernelODE.f90:254-478)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("default",)


def load(name):
    """(eta, w, w', w'') of fixture `name`."""
    with np.load(os.path.join(GOLDEN, "wgc_table_%s.npz" % name), allow_pickle=False) as z:
        return z["eta"].copy(), z["w"].copy(), z["w1"].copy(), z["w2"].copy()


def params(name):
    with np.load(os.path.join(GOLDEN, "wgc_table_%s.npz" % name), allow_pickle=False) as z:
        return tuple(float(x) for x in z["params"])


def for_params(alpha, beta, gamma):
    """The fixture integrated for (alpha, beta, gamma); KeyError when none is."""
    for name in NAMES:
        a, b, g = params(name)
        if abs(a - alpha) <= 1e-15 and abs(b - beta) <= 1e-15 and abs(g - gamma) <= 1e-15:
            return load(name)
    raise KeyError("no WGC table fixture for alpha=%r beta=%r gamma=%r" % (alpha, beta, gamma))
