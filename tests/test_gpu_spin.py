"""GPU parity of the spin-polarized path (numSpin = 2) through the C ABI
against the oracle: the evaluation (both LSDA forms, alpha = beta and
alpha != beta, both optSqrt forms, densities across rs = 1) at the FP64 bar
(terms 1e-10 relative, potentials 1e-10 of max|V|), the fixed-magnetization
NTN solve and the magnetization loop against the oracle's runs, and the
forces of a spin density."""
import numpy as np
import pytest

from oracle import profess_oracle as O
from tests.test_gpu_parity import POT_RTOL, _al_psp_text, _check_terms, _say, _tri_lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import profess_amd
    profess_amd.lib()
    return profess_amd


def _spin_case(grid, alpha=-100.0, beta=-100.0, seed=0, nion=3, dense=False):
    n1, n2, n3 = grid
    cell = O.Cell(_tri_lattice(), n1, n2, n3)
    rng = np.random.default_rng(seed)
    frac = rng.random((nion, 3))
    psp = O.make_psp(*O.read_recpot_text(_al_psp_text()))
    vion = O.ion_potential(cell, [(psp, frac)])
    ne = 3.0 * nion
    f = O.Functional(kinetic=4, alpha=alpha, beta=beta).resolve(ne, cell.vol)
    if dense:
        # totals in [0.15, 0.35] (rs = 1 at 0.2387), polarization up to 0.6
        tot = 0.15 + 0.2 * rng.random(cell.shape)
    else:
        tot = ne / cell.vol * (1.0 + 0.3 * (rng.random(cell.shape) - 0.5))
    z = 0.6 * (rng.random(cell.shape) - 0.5) * 2
    rho2 = np.stack([tot * (1 + z) / 2, tot * (1 - z) / 2])
    return cell, vion, ne, f, rho2


SPIN_CASES = [((16, 16, 16), -100.0, -100.0, 1, False), ((24, 20, 18), -100.0, -100.0, 2, False),
              ((32, 30, 28), 0.9, 0.7, 1, False), ((20, 18, 16), 0.9, 0.7, 2, False),
              ((64, 32, 64), -100.0, -100.0, 2, False), ((16, 16, 16), -100.0, -100.0, 1, True),
              ((18, 16, 20), -100.0, -100.0, 2, True)]


@pytest.mark.parametrize("grid,alpha,beta,lsda,dense", SPIN_CASES)
def test_spin_calpot_matches_oracle(pa, grid, alpha, beta, lsda, dense):
    _say("spin calpot grid %s alpha %g beta %g lsda %d dense %s" % (grid, alpha, beta, lsda, dense))
    cell, vion, ne, f, rho2 = _spin_case(grid, alpha, beta, dense=dense)
    sysm = O.System(cell, f, vion, e_ionion=-1.25)
    s = pa.OFDSystem(cell.lattice, grid, ne, kinetic=4, alpha=alpha, beta=beta)
    s.set_ion_terms(vion, -1.25)
    s.set_spin(2, lsda=lsda)
    for opt_sqrt in (True, False):
        pot, et = pa.CalculatePotentialPlus(s, rho2, optSqrt=opt_sqrt, eTable=True)
        pr, er = O.calculate_potential_plus_spin(rho2, sysm, opt_sqrt, True, lsda=lsda)
        _check_terms(et, er)
        for k in range(2):
            assert np.max(np.abs(pot[k] - pr[k])) <= POT_RTOL * np.max(np.abs(pr[k]))
    s.close()


def test_spin_back_to_one_channel(pa):
    """A context switched to numSpin = 2 and back evaluates as before."""
    cell, vion, ne, f, rho2 = _spin_case((16, 16, 16))
    rho = rho2[0] + rho2[1]
    sysm = O.System(cell, f, vion)
    s = pa.OFDSystem(cell.lattice, (16, 16, 16), ne, kinetic=4)
    s.set_ion_terms(vion, 0.0)
    s.set_spin(2)
    pa.CalculatePotentialPlus(s, rho2)
    s.set_spin(1)
    pot, et = pa.CalculatePotentialPlus(s, rho)
    pr, er = O.calculate_potential_plus(rho, sysm, True, True)
    _check_terms(et, er)
    assert np.max(np.abs(pot - pr)) <= POT_RTOL * np.max(np.abs(pr))
    s.close()


def _al4(pa, n, lsda, fixed, nmag=-1):
    from profess_amd import inputs
    c = inputs.al_fcc_case((1, 1, 1), (n, n, n), sigma=0.0)
    ne = c["num_electrons"]
    s = pa.OFDSystem(c["lattice"], c["grid"], ne, kinetic=4)
    s.build_ion_potential([(c["recpot"][0], c["recpot"][1], c["frac"])])
    s.set_spin(2, lsda=lsda, fixed_mag=fixed, nmag=nmag)
    cell = O.Cell(c["lattice"], *c["grid"])
    f = O.Functional(kinetic=4).resolve(ne, cell.vol)
    sysm = O.System(cell, f, s.ion_potential())
    return s, sysm, ne


@pytest.mark.parametrize("lsda", [1, 2])
def test_spin_ntn_fixed_magnetization_matches_oracle(pa, lsda):
    """fcc Al4 on 16^3 at a fixed magnetization of 0.4: the device solve takes
    the oracle's iterations and ends within 1e-6 eV/atom of its energy, with
    the same channel densities and electron numbers."""
    s, sysm, ne = _al4(pa, 16, lsda, True)
    _say("spin NTN fixed m = 0.4, lsda %d" % lsda)
    rho0 = np.full(s.shape, ne / s.vol)
    r2 = s.spin_split(rho0, 0.4)
    np.testing.assert_allclose(r2, O.spin_split(rho0, 0.4, s.vol), rtol=1e-15)
    rho, et, oc, st = pa.ProjNormRhoMinimization(s, r2, conv_check="POT", pot_tol=1e-6)
    full, pot = O.make_spin_evaluators(sysm, lsda)
    res = O.rho_minimization_spin(r2, full, pot, s.dV, O.OptParams(conv_check="POT", pot_tol=1e-6))
    _say("  iterations gpu %d oracle %d, E gpu %.12f oracle %.12f" % (st["iterations"], res.iters, et[0],
                                                                      res.energy[0]))
    assert oc == 0 and res.outcome == 0
    assert st["iterations"] == res.iters
    assert abs(et[0] - res.energy[0]) * O.HARTREE_TO_EV / 4 < 1e-6
    # a solve stopped at a potential norm of 1e-6 fixes the density only to
    # about 1e-5 of its maximum (the energy, quadratic in that error, far better)
    assert np.max(np.abs(rho - res.rho)) < 1e-4 * np.max(res.rho)
    m, nup, ndn, it = s.magnetization()
    assert it == 1 and m == pytest.approx(0.4, abs=1e-12)
    assert np.sum(rho[0]) * s.dV == pytest.approx((ne + 0.4) / 2, rel=1e-10)
    s.close()


def test_spin_magnetization_loop_matches_oracle(pa):
    """The magnetization loop from m = 0 on fcc Al4 at 8^3: the same number of
    solves as the oracle, final magnetization and energy."""
    s, sysm, ne = _al4(pa, 8, 1, False)
    _say("spin magnetization loop 8^3")
    rho0 = np.full(s.shape, ne / s.vol)
    r2 = np.stack([rho0 / 2, rho0 / 2])
    prm = dict(conv_check="POT", pot_tol=1e-6)
    rho, et, oc, st = pa.ProjNormRhoMinimization(s, r2, **prm)
    full, pot = O.make_spin_evaluators(sysm, 1)
    res = O.rho_minimization_spin_mag(r2, full, pot, s.dV, O.OptParams(**prm))
    m, nup, ndn, it = s.magnetization()
    _say("  solves gpu %d oracle %d, m gpu %.10f oracle %.10f" % (it, res.mag_iters, m, res.magnetization))
    assert it == res.mag_iters
    assert m == pytest.approx(res.magnetization, abs=1e-6)
    assert abs(et[0] - res.energy[0]) * O.HARTREE_TO_EV / 4 < 1e-6
    s.close()


def test_spin_forces_use_the_total_density(pa):
    """ofd_ion_forces on a spin density equals the forces on rho_up + rho_down
    (CalForce.f90:180)."""
    from profess_amd import inputs
    c = inputs.al_fcc_case((1, 1, 1), (16, 16, 16), sigma=0.1)
    ne = c["num_electrons"]
    s1 = pa.OFDSystem(c["lattice"], c["grid"], ne, kinetic=4)
    s2 = pa.OFDSystem(c["lattice"], c["grid"], ne, kinetic=4)
    for s in (s1, s2):
        s.build_ion_potential([(c["recpot"][0], c["recpot"][1], c["frac"])])
    s2.set_spin(2)
    rng = np.random.default_rng(3)
    rho2 = ne / s1.vol / 2 * (1 + 0.2 * rng.random((2,) + s1.shape))
    f1 = pa.IonElectronForcesSpline(s1, rho2[0] + rho2[1])
    f2 = pa.IonElectronForcesSpline(s2, rho2)
    np.testing.assert_allclose(f2, f1, rtol=1e-12, atol=1e-14)
    s1.close()
    s2.close()


@pytest.mark.parametrize("grid,alpha,beta,lsda", [((64, 64, 64), -100.0, -100.0, 1), ((48, 40, 60), 0.9, 0.7, 2),
                                                  ((120, 64, 48), 0.9, 0.7, 1), ((32, 96, 32), -100.0, -100.0, 2)])
def test_fast_spin_matches_generic_path(pa, monkeypatch, grid, alpha, beta, lsda):
    """The spin kinds on the register FFT passes and on the generic LDS passes
    agree to rounding in FP64 (and the FP32 passes within FP32 rounding); the
    FP64 result also meets the oracle at the FP64 bar."""
    _say("spin fast/generic grid %s alpha %g beta %g lsda %d" % (grid, alpha, beta, lsda))
    cell, vion, ne, f, rho2 = _spin_case(grid, alpha, beta, seed=4)
    res = {}
    for prec in (0, 1):
        for flag in ("0", "1"):
            monkeypatch.setenv("OFD_GENERIC", flag)
            s = pa.OFDSystem(cell.lattice, grid, ne, kinetic=4, alpha=alpha, beta=beta, precision=prec)
            s.set_ion_terms(vion, -1.25)
            s.set_spin(2, lsda=lsda)
            res[prec, flag] = pa.CalculatePotentialPlus(s, rho2, eTable=True)
            s.close()
    (p0, e0), (p1, e1) = res[0, "0"], res[0, "1"]
    for k in range(2):
        assert np.max(np.abs(p0[k] - p1[k])) <= 1e-13 * np.max(np.abs(p1[k]))
    _check_terms(e0, e1, rtol=1e-13)
    for flag in ("0", "1"):
        p32, e32 = res[1, flag]
        for k in range(2):
            assert np.max(np.abs(p32[k] - p1[k])) <= 1e-5 * np.max(np.abs(p1[k]))
        _check_terms(e32, e1, rtol=1e-5, floor=1e-8)
    sysm = O.System(cell, f, vion, e_ionion=-1.25)
    pr, er = O.calculate_potential_plus_spin(rho2, sysm, True, True, lsda=lsda)
    _check_terms(e0, er)
    for k in range(2):
        assert np.max(np.abs(p0[k] - pr[k])) <= POT_RTOL * np.max(np.abs(pr[k]))


def test_fast_spin_solve_matches_generic_path(pa, monkeypatch):
    """fcc Al4 on 32^3 at a fixed magnetization of 0.4, solved with the spin
    evaluations (trial points, signed and Newton forms) on the register FFT
    passes and on the generic passes. The two differ in the last bits of each
    evaluation, and the stop at a potential norm of 1e-6 can fall one
    iteration apart: the iteration counts agree within one, the energies
    within 1e-8 eV/atom and the densities within 1e-4 of max rho (a potential
    norm of 1e-6 fixes the density to about 1e-5 of its maximum)."""
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("OFD_GENERIC", flag)
        s, sysm, ne = _al4(pa, 32, 2, True)
        _say("spin NTN fixed m = 0.4 on 32^3, OFD_GENERIC=%s" % flag)
        rho0 = np.full(s.shape, ne / s.vol)
        out[flag] = pa.ProjNormRhoMinimization(s, s.spin_split(rho0, 0.4), conv_check="POT", pot_tol=1e-6)
        s.close()
    (rf, ef, of, sf), (rg, eg, og, sg) = out["0"], out["1"]
    _say("  iterations fast %d generic %d, E %.14f %.14f" % (sf["iterations"], sg["iterations"], ef[0], eg[0]))
    assert of == 0 and og == 0
    assert abs(sf["iterations"] - sg["iterations"]) <= 1
    assert abs(ef[0] - eg[0]) * O.HARTREE_TO_EV / 4 < 1e-8
    assert np.max(np.abs(rf - rg)) <= 1e-4 * np.max(rg)
