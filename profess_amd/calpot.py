"""Python mirror of the reference's call surface for the energy/potential path.

Names follow PROFESS: ``CalculatePotentialPlus`` (CalPotential.f90:41),
``ProjNormRhoMinimization`` (RhoOptN.f90:60), ``IonElectronForcesSpline``
(IonElectronSpline.f90:219). Arrays are numpy float64 of shape (n3, n2, n1),
C order, which is the reference's Fortran rho(n1,n2,n3) memory layout.
Everything runs through libprofess_amd.so on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (OFDError, dptr, iptr, lib, ofd_functional, ofd_grid, ofd_opt_params, ofd_opt_stats)

ETABLE_NAMES = ["total", "kinetic", "ion-electron", "hartree", "xc", "ion-ion", "tf", "vw", "nonlocal"]
CONV = {"ENE": 1, "POT": 2, "BOTH": 3}


class OFDSystem:
    """One cell + grid + functional: the module state CalculatePotentialPlus reads.

    lattice: 3x3, columns are the lattice vectors in bohr (cell%cellReal).
    kinetic: 4 = Wang-Teter, 5 = WGC. Negative parameters take the reference's
    defaults (SetupKEDF.f90:217-322). precision: 0 FP64, 1 FP32 FFT passes.
    """

    def __init__(self, lattice, grid, num_electrons, kinetic=4, lam=-100.0, mu=-100.0, alpha=-100.0,
                 beta=-100.0, gamma=-1.0, rho0=-1.0, rhoS=-1.0, precision=0):
        L = lib()
        self.n1, self.n2, self.n3 = (int(x) for x in grid)
        self.lattice = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
        g = ofd_grid(self.n1, self.n2, self.n3, (C.c_double * 9)(*self.lattice.T.ravel()))
        f = ofd_functional(kinetic, lam, mu, alpha, beta, gamma, rho0, rhoS, precision)
        h = C.c_void_p()
        rc = L.ofd_create(C.byref(g), C.byref(f), float(num_electrons), C.byref(h))
        self._h = h
        self.kinetic = kinetic
        self.num_electrons = float(num_electrons)
        if rc != 0:
            msg = L.ofd_last_error(h).decode() if h.value else "ofd_create failed"
            self.close()
            raise OFDError(msg)
        v = C.c_double(0.0)
        L.ofd_get_volume(h, C.byref(v))
        self.vol = v.value   # the library's cell%vol: uniform densities agree bit for bit
        self.dV = self.vol / (self.n1 * self.n2 * self.n3)
        self.nspin = 1

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ofd_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def detach(self):
        """Hand the native context over (to a pool); this object no longer owns it."""
        h = self._h
        self._h = None
        return h

    def _check(self, rc):
        if rc != 0:
            raise OFDError(lib().ofd_last_error(self._h).decode())

    @property
    def shape(self):
        return (self.n3, self.n2, self.n1)

    @property
    def rho_shape(self):
        """Shape of density and potential arrays: (2, n3, n2, n1) with spin."""
        return self.shape if self.nspin == 1 else (2,) + self.shape

    def set_spin(self, nspin=2, lsda=1, fixed_mag=False, nmag=-1, tolm=-1.0):
        """numSpin (input SPIN), the LSDA form (1 Perdew-Zunger, 2 Perdew-Wang
        92), FIXMAG, and the magnetization loop's nmag and tolm (ofd_set_spin)."""
        from ._lib import ofd_spin
        sp = ofd_spin(int(nspin), int(lsda), int(bool(fixed_mag)), int(nmag), float(tolm))
        self._check(lib().ofd_set_spin(self._h, C.byref(sp)))
        self.nspin = int(nspin)

    def spin_split(self, rho, magmom):
        """Initial channels from a total density and MAGMOM (System.f90:326-329)."""
        r = np.ascontiguousarray(rho, dtype=np.float64)
        assert r.shape == self.shape
        out = np.empty((2,) + self.shape)
        self._check(lib().ofd_spin_split(self._h, dptr(r), float(magmom), dptr(out)))
        return out

    def magnetization(self):
        """After a spin solve: (magnetization, N_up, N_down, magnetization-loop solves)."""
        out = np.zeros(3)
        it = C.c_int(0)
        self._check(lib().ofd_get_magnetization(self._h, dptr(out), C.byref(it)))
        return out[0], out[1], out[2], it.value

    @property
    def params(self):
        out = np.zeros(7)
        self._check(lib().ofd_get_params(self._h, dptr(out)))
        return dict(zip(["alpha", "beta", "gamma", "rho0", "rhoS", "lambda", "mu"], out.tolist()))

    # -- per-geometry tables
    def set_ion_terms(self, vion, e_ionion=0.0):
        v = np.ascontiguousarray(vion, dtype=np.float64)
        assert v.shape == self.shape
        self._check(lib().ofd_set_ion_terms(self._h, dptr(v), float(e_ionion)))

    def build_ion_potential(self, species, spline_order=10, e_ionion=0.0):
        """species: list of (maxg_invA, values_eVA3, frac (nion, 3)) in .recpot file units."""
        starts = [0]
        fr, vals, npts, offs, maxg = [], [], [], [], []
        for mg, v, f in species:
            f = np.asarray(f, dtype=np.float64).reshape(-1, 3)
            fr.append(f)
            starts.append(starts[-1] + f.shape[0])
            offs.append(sum(len(x) for x in vals))
            vals.append(np.asarray(v, dtype=np.float64))
            npts.append(len(v))
            maxg.append(float(mg))
        frac = np.ascontiguousarray(np.concatenate(fr), dtype=np.float64)
        ts = np.array(starts, dtype=np.int32)
        self._check(lib().ofd_build_ion_potential(
            self._h, len(species), iptr(ts), dptr(frac), dptr(np.array(maxg)),
            iptr(np.array(npts, dtype=np.int32)), iptr(np.array(offs, dtype=np.int32)),
            dptr(np.ascontiguousarray(np.concatenate(vals))), int(spline_order), float(e_ionion)))
        self.nion = frac.shape[0]

    def ion_potential(self):
        out = np.empty(self.shape)
        self._check(lib().ofd_get_ion_potential(self._h, dptr(out)))
        return out

    def autotune(self, reps=5):
        """Pick the evaluation passes' tile shapes for this grid by timing
        (ofd_autotune); later systems of the same shape start from them."""
        from ._lib import ofd_tune_result
        r = ofd_tune_result()
        self._check(lib().ofd_autotune(self._h, int(reps), C.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    def set_wgc_table(self, eta, w, w1, w2):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (eta, w, w1, w2)]
        self._check(lib().ofd_set_wgc_table(self._h, a[0].size, *[dptr(x) for x in a]))

    def build_wgc_table(self):
        n = C.c_int(0)
        self._check(lib().ofd_build_wgc_table(self._h, C.byref(n)))
        return n.value


def wgc_kernel_table(alpha, beta, gamma):
    """The library's 1D WGC kernel table (eta, w, w', w'') on makeKernel2's knots
    (KEDF_KernelODE.f90:254-478), integrated on the host (ofd_wgc_kernel_table)."""
    L = lib()
    n = L.ofd_wgc_kernel_table(alpha, beta, gamma, 0, None, None, None, None)
    if n <= 0:
        raise OFDError("WGC table integration failed (%d)" % n)
    out = [np.zeros(n) for _ in range(4)]
    got = L.ofd_wgc_kernel_table(alpha, beta, gamma, n, *[dptr(x) for x in out])
    if got != n:
        raise OFDError("WGC table integration failed (%d)" % got)
    return tuple(out)


def CalculatePotentialPlus(system: OFDSystem, rho, optSqrt=True, eTable=True):
    """CalculatePotentialPlus(rho, optSqrt, potential[, eTable]) (CalPotential.f90:41).
    Returns (potential, eTable(9) in Hartree or None)."""
    r = np.ascontiguousarray(rho, dtype=np.float64)
    if r.shape != system.rho_shape:
        raise ValueError("rho has shape %r, expected %r" % (r.shape, system.rho_shape))
    pot = np.empty_like(r)
    et = np.zeros(9) if eTable else None
    system._check(lib().ofd_calpot(system._h, dptr(r), 1 if optSqrt else 0, dptr(pot), dptr(et)))
    return pot, et


def CalculatePotentialPlusDevice(system: OFDSystem, rho_ptr, pot_ptr, optSqrt=True, eTable=True):
    """Device-resident form: rho_ptr/pot_ptr are device addresses (e.g. tensor.data_ptr())."""
    et = np.zeros(9) if eTable else None
    system._check(lib().ofd_calpot_device(system._h, C.c_void_p(rho_ptr), 1 if optSqrt else 0,
                                          C.c_void_p(pot_ptr), dptr(et)))
    return et


def opt_params(max_iter=-1, pot_tol=-1.0, tole=-1.0, conv_check="ENE", max_line=-1, mixed_switch=0.0,
               fused_dots=False, cg_alg="HZ"):
    return ofd_opt_params(int(max_iter), float(pot_tol), float(tole), CONV[conv_check.upper()], int(max_line),
                          float(mixed_switch), int(bool(fused_dots)), {"PR": 1, "HZ": 2}[cg_alg.upper()[:2]])


def ProjNormRhoMinimization(system: OFDSystem, rhoR, dir_type=1, **kw):
    """ProjNormRhoMinimization(energy, dir_type) (RhoOptN.f90:60) with
    dir_type 1 (NTN), 2 (NCG; cg_alg="HZ" or "PR" as DENC 2 or 1) or 3 (NBF).
    Returns (rho, energy(9), rhoOutcome, stats dict)."""
    r = np.array(rhoR, dtype=np.float64, order="C", copy=True)
    if r.shape != system.rho_shape:
        raise ValueError("rhoR has shape %r, expected %r" % (r.shape, system.rho_shape))
    et = np.zeros(9)
    outc = C.c_int(-1)
    st = ofd_opt_stats()
    p = opt_params(**kw)
    system._check(lib().ofd_rho_opt(system._h, dptr(r), int(dir_type), C.byref(p), dptr(et), C.byref(outc),
                                    C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return r, et, outc.value, stats


def IonElectronForcesSpline(system: OFDSystem, rho):
    """Ion-electron forces (Ha/bohr), ions in the order given to build_ion_potential."""
    r = np.ascontiguousarray(rho, dtype=np.float64)
    if r.shape != system.rho_shape:
        raise ValueError("rho has shape %r, expected %r" % (r.shape, system.rho_shape))
    out = np.zeros((system.nion, 3))
    system._check(lib().ofd_ion_forces(system._h, dptr(r), dptr(out)))
    return out


def EwaldEnergy(system: OFDSystem, forces=False):
    """Ion-ion Ewald energy (Hartree) of the system's ions with the particle-mesh
    reciprocal sum (Ewald.f90:465-596), and with forces=True also the Ewald
    forces (Ha/bohr, EwaldForces Ewald.f90:1196-1310). The energy becomes the
    system's ion-ion energy (eTable(6)) for later evaluations. Returns
    (energy, forces or None, info) with info the Ewald parameters."""
    e = C.c_double()
    out = np.zeros((system.nion, 3)) if forces else None
    info = np.zeros(5)
    system._check(lib().ofd_ewald(system._h, C.byref(e), dptr(out), dptr(info)))
    return e.value, out, dict(zip(("eta", "real_cutoff", "recip_cutoff", "real_error", "recip_error"), info))


def ConjugateGradient(system: OFDSystem, rhoR, frac, cg_type=2, force_cutoff=-1.0, max_steps=-1, nn_dist=-1.0,
                      max_geometries=1000, rho_dir_type=1, **rho_kw):
    """Ion relaxation (IonOptCG2::ConjugateGradient, IonOptCG2.f90:60-520)
    from fractional positions `frac` (ions in the order and species grouping
    given to build_ion_potential) and the starting density rhoR, with the
    density optimizer of rho_dir_type (1 NTN, 2 NCG, 3 NBF) at every geometry
    (rho_kw as for ProjNormRhoMinimization). Returns (rho, frac, energy9, forces, energies of
    every geometry, stats)."""
    from ._lib import ofd_ion_params, ofd_ion_stats
    r = np.ascontiguousarray(rhoR, dtype=np.float64).copy()
    fr = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3).copy()
    ip = ofd_ion_params(int(cg_type), float(force_cutoff), int(max_steps), float(nn_dist), int(rho_dir_type))
    p = opt_params(**rho_kw)
    et = np.zeros(9)
    forces = np.zeros_like(fr)
    en = np.zeros(max_geometries)
    st = ofd_ion_stats()
    system._check(lib().ofd_ion_relax(system._h, dptr(r), dptr(fr), C.byref(p), C.byref(ip), dptr(et), dptr(forces),
                                      dptr(en), int(max_geometries), C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return r, fr, et, forces, en[:st.geometries].copy(), stats


def request_bytes(grid, kinetic=4, precision=0):
    return int(lib().ofd_request_bytes(int(grid[0]), int(grid[1]), int(grid[2]), int(kinetic), int(precision)))


def CalculatePotentialPlusMany(systems, rho_ptrs, pot_ptrs, optSqrt=True, eTable=True):
    """One batched evaluation of several systems of one grid shape/functional
    (device-resident rho/potential addresses). Returns eTables (nb x 9) or None."""
    nb = len(systems)
    hs = (C.c_void_p * nb)(*[s._h.value for s in systems])
    ra = (C.c_void_p * nb)(*rho_ptrs)
    pa_ = (C.c_void_p * nb)(*pot_ptrs)
    et = np.zeros((nb, 9)) if eTable else None
    rc = lib().ofd_calpot_many(hs, nb, ra, pa_, 1 if optSqrt else 0, dptr(et) if eTable else None)
    if rc != 0:
        raise OFDError(lib().ofd_last_error(systems[0]._h).decode())
    return et
