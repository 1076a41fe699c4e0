"""Synthetic inputs of the benchmark workloads and the .recpot reader.

fcc / hcp supercells with seeded Gaussian displacements, the low-|G| cosine
density rho = rho0 (1 + 0.3 f), the analytic local pseudopotential
V_loc(q) = -4 pi Z/q^2 exp(-q^2 rc^2/4) written in .recpot units, and a parser
for .recpot text that follows ReadPseudo's reading rules (ReadIonFile.f90:830-870).
"""
from __future__ import annotations

import math

import numpy as np

BOHR = 0.529177208607388                     # Constants.f90:50
_QE, _C, _MU0 = 1.602176462e-19, 299792458.0, 4.0 * math.pi * 1.0e-7
HARTREE_TO_EV = _QE * 1.0e10 / (4.0 * math.pi * (1.0 / (_MU0 * _C ** 2)) * BOHR)


def read_recpot(text):
    """Returns (maxG in 1/A, values in eV A^3) as the file holds them."""
    lines = text.splitlines()
    i = 0
    while "END COMMENT" not in lines[i]:
        i += 1
    i += 2
    maxg = float(lines[i].split()[0])
    vals = []
    for ln in lines[i + 1:]:
        t = ln.split()
        if not t:
            continue
        if t[0] == "1000":
            break
        vals.extend(float(x) for x in t)
    return maxg, np.array(vals)


def synthetic_recpot(Z, rc, qmax_bohr=30.0, dq=0.002):
    """Analytic V_loc table in .recpot units; q = 0 holds lim (V + 4 pi Z/q^2) = pi Z rc^2."""
    n = int(round(qmax_bohr / dq)) + 1
    q = np.linspace(0.0, qmax_bohr, n)
    v = np.empty(n)
    v[0] = math.pi * Z * rc * rc
    v[1:] = -4.0 * math.pi * Z / q[1:] ** 2 * np.exp(-q[1:] ** 2 * rc * rc / 4.0)
    return qmax_bohr / BOHR, v * (BOHR ** 3 * HARTREE_TO_EV)


def fcc_supercell(a, reps):
    """Cubic fcc supercell: lattice (columns, bohr), fractional positions."""
    r = np.array(reps, dtype=int) if np.ndim(reps) else np.array([reps] * 3)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(r[0]) for j in range(r[1]) for k in range(r[2])], dtype=float)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) / r
    lat = np.diag(a * r.astype(float))
    return lat, frac


def hcp_supercell(a, c, reps):
    """hcp supercell of the primitive (non-orthogonal) cell: a1 = a(1,0,0),
    a2 = a(-1/2, sqrt3/2, 0), a3 = c(0,0,1); 2 atoms per primitive cell."""
    r = np.array(reps, dtype=int) if np.ndim(reps) else np.array([reps] * 3)
    prim = np.array([[a, 0, 0], [-a / 2, a * math.sqrt(3) / 2, 0], [0, 0, c]]).T
    basis = np.array([[1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]])
    cells = np.array([[i, j, k] for i in range(r[0]) for j in range(r[1]) for k in range(r[2])], dtype=float)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) / r
    lat = prim * r[None, :]
    return lat, frac


def displace(lat, frac, sigma, seed):
    """Seeded Gaussian Cartesian displacements (bohr), wrapped into [0,1)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(0.0, sigma, size=frac.shape)
    f = frac + np.linalg.solve(lat, d.T).T
    return np.mod(f, 1.0)


def cosine_density(shape, num_electrons, vol, seed, nmodes=64, amp=0.3, mmax=3):
    """rho = rho0 (1 + amp f), f = sum_k a_k cos(2 pi m_k.s + phi_k) / sum|a_k|,
    m_k in Z^3 with 1 <= |m|_inf <= mmax; |f| <= 1, clipped at 1e-12."""
    n3, n2, n1 = shape
    rng = np.random.default_rng(seed)
    rho0 = num_electrons / vol
    s1 = np.arange(n1) / n1
    s2 = np.arange(n2) / n2
    s3 = np.arange(n3) / n3
    f = np.zeros(shape)
    ak = rng.uniform(0.0, 1.0, nmodes)
    phik = rng.uniform(0.0, 2 * math.pi, nmodes)
    ms = []
    while len(ms) < nmodes:
        m = rng.integers(-mmax, mmax + 1, 3)
        if 1 <= np.max(np.abs(m)) <= mmax:
            ms.append(m)
    for a_, p_, m in zip(ak, phik, ms):
        # separable phase: exp(i 2pi (m1 s1 + m2 s2 + m3 s3))
        e1 = np.exp(2j * math.pi * m[0] * s1)
        e2 = np.exp(2j * math.pi * m[1] * s2)
        e3 = np.exp(2j * math.pi * m[2] * s3)
        f += a_ * np.real(np.exp(1j * p_) * e3[:, None, None] * e2[None, :, None] * e1[None, None, :])
    f /= np.sum(np.abs(ak))
    rho = rho0 * (1.0 + amp * f)
    return np.maximum(rho, 1e-12)


# ---- benchmark workloads
AL_A = 7.65          # bohr, fcc Al cubic lattice constant
AL_Z, AL_RC = 3, 1.2
MG_A, MG_C = 6.06, 9.84   # bohr, hcp Mg
MG_Z, MG_RC = 2, 1.4


def al_fcc_case(reps, grid, sigma=0.1, seed_ion=1, seed_rho=2):
    lat, frac = fcc_supercell(AL_A, reps)
    frac = displace(lat, frac, sigma, seed_ion)
    ne = AL_Z * frac.shape[0]
    vol = abs(np.linalg.det(lat))
    shape = (grid[2], grid[1], grid[0])
    rho = cosine_density(shape, ne, vol, seed_rho)
    return dict(lattice=lat, frac=frac, num_electrons=ne, grid=tuple(grid), rho=rho,
                recpot=synthetic_recpot(AL_Z, AL_RC), Z=AL_Z)


def mg_hcp_case(reps, grid, sigma=0.1, seed_ion=3, seed_rho=4):
    lat, frac = hcp_supercell(MG_A, MG_C, reps)
    frac = displace(lat, frac, sigma, seed_ion)
    ne = MG_Z * frac.shape[0]
    vol = abs(np.linalg.det(lat))
    shape = (grid[2], grid[1], grid[0])
    rho = cosine_density(shape, ne, vol, seed_rho)
    return dict(lattice=lat, frac=frac, num_electrons=ne, grid=tuple(grid), rho=rho,
                recpot=synthetic_recpot(MG_Z, MG_RC), Z=MG_Z)


# ---------------------------------------------------------------- files
def _errbuf():
    import ctypes as C
    return C.create_string_buffer(512)


def read_geometry(path):
    """A .ion file through the library's reader (ofd_read_geometry,
    ReadGeometry ReadIonFile.f90:80-705). Returns a dict: lattice (columns,
    bohr), frac (ions grouped by species), type_start, file_index, species,
    pseudo_files."""
    import ctypes as C
    from ._lib import OFDError, lib, ofd_geometry
    g = ofd_geometry()
    e = _errbuf()
    if lib().ofd_read_geometry(str(path).encode(), C.byref(g), e, 512) != 0:
        raise OFDError(e.value.decode())
    try:
        n, ns = g.nion, g.nspecies
        out = {"lattice": np.array(g.cell[:]).reshape(3, 3).T.copy(),
               "frac": np.array(g.frac[:3 * n]).reshape(n, 3),
               "type_start": [g.type_start[k] for k in range(ns + 1)],
               "file_index": [g.file_index[k] for k in range(n)],
               "species": [g.species[k].s.decode() for k in range(ns)],
               "pseudo_files": [g.pseudo_file[k].s.decode() for k in range(ns)]}
    finally:
        lib().ofd_free_geometry(C.byref(g))
    return out


def read_recpot_file(path):
    """(maxG in 1/A, values in eV A^3) of a .recpot file through the
    library's reader (ofd_read_recpot, ReadPseudo ReadIonFile.f90:708-930)."""
    import ctypes as C
    from ._lib import OFDError, dptr, lib
    mg = C.c_double()
    n = C.c_int()
    e = _errbuf()
    if lib().ofd_read_recpot(str(path).encode(), C.byref(mg), None, 0, C.byref(n), e, 512) != 0:
        raise OFDError(e.value.decode())
    vals = np.zeros(n.value)
    if lib().ofd_read_recpot(str(path).encode(), C.byref(mg), dptr(vals), vals.size, C.byref(n), e, 512) != 0:
        raise OFDError(e.value.decode())
    return mg.value, vals


def write_density(path, rho):
    """rho (n3, n2, n1) in the reference's density file format (Print3DArray)."""
    from ._lib import OFDError, dptr, lib
    r = np.ascontiguousarray(rho, dtype=np.float64)
    n3, n2, n1 = r.shape
    e = _errbuf()
    if lib().ofd_write_density(str(path).encode(), n1, n2, n3, 1, dptr(r), e, 512) != 0:
        raise OFDError(e.value.decode())


def read_density(path, grid, lattice=None, num_electrons=0.0, repeat=(1, 1, 1)):
    """A density file on grid (n1, n2, n3) (ReadDensity): copied when the
    file's grid (times repeat) matches, interpolated otherwise; scaled to
    num_electrons when lattice (columns, bohr) and num_electrons > 0 are given.
    Returns (n3, n2, n1)."""
    from ._lib import OFDError, dptr, lib
    n1, n2, n3 = grid
    out = np.zeros((n3, n2, n1))
    cell = None if lattice is None else np.ascontiguousarray(np.asarray(lattice, dtype=np.float64).T).ravel()
    e = _errbuf()
    if lib().ofd_read_density(str(path).encode(), n1, n2, n3, dptr(cell), float(num_electrons), int(repeat[0]),
                              int(repeat[1]), int(repeat[2]), dptr(out), e, 512) != 0:
        raise OFDError(e.value.decode())
    return out


def recpot_charge(maxg_invA, values_ev_a3):
    """The valence charge ReadPseudo derives from a .recpot table
    (ReadIonFile.f90:895-900): NINT(-(V(2) - V(1)) (maxG / (n-1))^2 / 4 pi)
    with V in Ha bohr^3 and maxG in 1/bohr."""
    v = np.asarray(values_ev_a3, dtype=np.float64) / (BOHR ** 3 * HARTREE_TO_EV)
    mg = maxg_invA * BOHR
    return int(round(-(v[1] - v[0]) * (mg / (v.size - 1)) ** 2 / (4.0 * math.pi)))


def system_from_files(ion_path, grid, pseudo_dir=None, kinetic=4, precision=0, spline_order=10, ewald=True):
    """An OFDSystem of a PROFESS .ion file on `grid` (n1, n2, n3): geometry and
    species from the file, each species' .recpot from its SPECIES_POT entry
    (relative to pseudo_dir, default the .ion file's folder), electron count
    from the tables' charges, ionic potential built on the device and, with
    ewald, the ion-ion energy from the Ewald sum. Returns (system, geometry)."""
    import os
    from .calpot import EwaldEnergy, OFDSystem
    g = read_geometry(ion_path)
    base = pseudo_dir if pseudo_dir is not None else os.path.dirname(os.path.abspath(str(ion_path)))
    species, ne = [], 0.0
    for k, fname in enumerate(g["pseudo_files"]):
        mg, vals = read_recpot_file(os.path.join(base, fname))
        lo, hi = g["type_start"][k], g["type_start"][k + 1]
        species.append((mg, vals, g["frac"][lo:hi]))
        ne += recpot_charge(mg, vals) * (hi - lo)
    s = OFDSystem(g["lattice"], tuple(grid), ne, kinetic=kinetic, precision=precision)
    s.build_ion_potential(species, spline_order=spline_order)
    if ewald:
        EwaldEnergy(s)
    return s, g
