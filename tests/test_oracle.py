"""Pins the CPU oracle (oracle/profess_oracle.py) before it is trusted as the
checker: closed forms, brute force, finite differences, the reference's own
Fortran (oracle/_ref) and the reference's recorded sample outputs."""
import math
import os

import numpy as np
import pytest

from oracle import profess_oracle as O
from oracle import refbuild as R
from tests import wgc_tables as W
from tests import sample_cases as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _al_psp():
    with open(os.path.join(GOLDEN, "al_HC.lda.recpot")) as fh:
        return O.make_psp(*O.read_recpot_text(fh.read()))


def _tri_cell(n=(6, 5, 4)):
    lat = np.array([[7.3, 0.2, 0.1], [0.3, 7.5, 0.0], [-0.2, 0.4, 7.6]]).T
    return O.Cell(lat, *n)


def test_fft_matches_brute_force_dft():
    rng = np.random.default_rng(1)
    for shape in [(4, 5, 6), (3, 6, 10), (7, 2, 8)]:
        a = rng.standard_normal(shape)
        n3, n2, n1 = shape
        c = O.fft_forward(a)
        i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
        for k3, k2, k1 in [(0, 0, 0), (1, 1, 1), (n3 - 1, 0, n1 // 2), (2 % n3, 1, 2)]:
            ph = np.exp(-2j * np.pi * (k1 * i1 / n1 + k2 * i2 / n2 + k3 * i3 / n3))
            ref = np.sum(a * ph) / a.size
            assert abs(c[k3, k2, k1] - ref) < 1e-13
        back = O.fft_backward(c, shape)
        np.testing.assert_allclose(back, a, rtol=0, atol=1e-12)


def test_g_tables_triclinic():
    cell = _tri_cell((6, 5, 4))
    qv, qt, mask = O.g_tables(cell)
    b = cell.recip
    # recip . lattice = 2 pi I
    np.testing.assert_allclose(b.T @ cell.lattice, 2 * np.pi * np.eye(3), atol=1e-12)
    # index (k1=2, i2=3 -> m2=-2 (5//2=2, 3>2), i3=2 -> m3=2 (4//2=2 stays positive))
    g = 2 * b[:, 0] - 2 * b[:, 1] + 2 * b[:, 2]
    np.testing.assert_allclose(qv[2, 3, 2], g, atol=1e-13)
    assert abs(qt[2, 3, 2] - np.linalg.norm(g)) < 1e-13
    # qMask: every G paired with -G counted once; G = 0 excluded
    assert not mask[0, 0, 0]


def test_lind_g_branches():
    lam, mu = 1.0, 1.0
    eta = np.array([0.3, 0.9, 1.5, 3.0])
    closed = 1.0 / (0.5 + (1 - eta ** 2) / (4 * eta) * np.log(np.abs((1 + eta) / (1 - eta)))) \
        - 3 * mu * eta ** 2 - lam
    np.testing.assert_allclose(O.lind_g(eta, lam, mu), closed, rtol=1e-14)
    # the Taylor branch continues the closed form at the 3.65 switch point
    lo = O.lind_g(np.array([3.65]), lam, mu)[0]
    e = 3.65
    cf = 1.0 / (0.5 + (1 - e * e) / (4 * e) * math.log(abs((1 + e) / (1 - e)))) - 3 * mu * e * e - lam
    hi = O.lind_g(np.array([3.65 + 1e-12]), lam, mu)[0]
    assert abs(lo - cf) < 1e-12 and abs(hi - cf) < 1e-9
    assert O.lind_g(np.array([0.0]), 0.3, 1.0)[0] == pytest.approx(0.7)
    assert O.lind_g(np.array([1.0]), 1.0, 1.0)[0] == pytest.approx(-2.0)


def _cos_field(cell, m, base, amp):
    n3, n2, n1 = cell.shape
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    arg = 2 * np.pi * (m[0] * i1 / n1 + m[1] * i2 / n2 + m[2] * i3 / n3)
    g = cell.recip @ np.array(m, dtype=float)
    return base + amp * np.cos(arg), np.cos(arg), g @ g


def test_hartree_and_vw_single_cosine_mode():
    cell = _tri_cell((8, 6, 10))
    _, qt, _ = O.g_tables(cell)
    rho, cosr, g2 = _cos_field(cell, (1, 2, -1), 0.03, 0.004)
    vh, eh = O.hartree(rho, qt)
    np.testing.assert_allclose(vh, 4 * np.pi * 0.004 / g2 * cosr, atol=1e-13)
    assert eh * cell.dV == pytest.approx(np.pi * 0.004 ** 2 * cell.vol / g2, rel=1e-12)
    t, evw = O.vw_sqrt(rho, qt, 0.7)
    np.testing.assert_allclose(t, 0.7 * 0.004 * g2 * cosr, atol=1e-13)
    # E = 1/2 mu sum(phi t) dV = mu b^2 G^2 Omega / 4 for phi = a + b cos
    assert evw * cell.dV == pytest.approx(0.7 * 0.004 ** 2 * g2 * cell.vol / 4, rel=1e-12)


def test_uniform_density_closed_forms():
    cell = _tri_cell((6, 6, 6))
    rho0 = 0.027
    rho = np.full(cell.shape, rho0)
    f = O.Functional(kinetic=4).resolve(rho0 * cell.vol, cell.vol)
    sysm = O.System(cell, f, np.zeros(cell.shape))
    pot, et = O.calculate_potential_plus(rho, sysm)
    assert et[6] == pytest.approx(O.CTF * rho0 ** (5 / 3) * cell.vol, rel=1e-13)
    assert abs(et[7]) < 1e-14 and abs(et[8]) < 1e-14 and abs(et[3]) < 1e-14
    rs = 0.62035049089940009 / rho0 ** (1 / 3)
    exc = -0.73855876638202234 * rho0 ** (1 / 3) + (-0.1423) / (1 + 1.0529 * math.sqrt(rs) + 0.3334 * rs)
    assert et[4] == pytest.approx(exc * rho0 * cell.vol, rel=1e-13)


def _pz81_eps_c(rs):
    """Perdew-Zunger correlation per electron (Perdew & Zunger, PRB 23, 5048
    (1981), Appendix C): A ln rs + B + C rs ln rs + D rs for rs < 1,
    gamma / (1 + beta1 sqrt(rs) + beta2 rs) for rs >= 1 (unpolarized)."""
    A, B, C, D = 0.0311, -0.048, 0.0020, -0.0116
    g, b1, b2 = -0.1423, 1.0529, 0.3334
    hi = A * np.log(rs) + B + C * rs * np.log(rs) + D * rs
    lo = g / (1 + b1 * np.sqrt(rs) + b2 * rs)
    dhi = A / rs + C * (np.log(rs) + 1) + D
    dlo = -g * (b1 / (2 * np.sqrt(rs)) + b2) / (1 + b1 * np.sqrt(rs) + b2 * rs) ** 2
    return np.where(rs < 1, hi, lo), np.where(rs < 1, dhi, dlo)


def test_lda_both_branches_from_definitions():
    """LDA (XC_LDA.f90:51-310) at uniform densities on both sides of rs = 1,
    against the definitions: Dirac exchange eps_x = -(3/4)(3/pi)^(1/3) rho^(1/3),
    rs = (3/(4 pi rho))^(1/3), PZ81 correlation, and the potential
    v = d(rho eps)/d rho = 4/3 eps_x + eps_c - (rs/3) d eps_c/d rs. This pins
    the high-density branch (rs < 1, XC_LDA.f90:141-142 and :283-286) that
    no sample run reaches, including its derived potential constants."""
    cell = _tri_cell((4, 4, 4))
    for rho0 in (0.5, 0.3, 0.027, 2.0):
        rho = np.full(cell.shape, rho0)
        rs = (3.0 / (4.0 * math.pi * rho0)) ** (1.0 / 3.0)
        ex = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * rho0 ** (1.0 / 3.0)
        ec, dec = _pz81_eps_c(np.array(rs))
        assert O.lda_energy(rho) == pytest.approx((ex + ec) * rho0 * rho.size, rel=1e-13)
        v = 4.0 / 3.0 * ex + ec - rs / 3.0 * dec
        np.testing.assert_allclose(O.lda_potential(rho), v, rtol=1e-13)
    assert (3.0 / (4.0 * math.pi * 0.5)) ** (1.0 / 3.0) < 1.0


def test_lda_potential_is_energy_derivative_across_rs_1():
    """Finite-difference dE_xc/d rho against the LDA potential on densities
    in [0.2, 0.6], on both sides of rs = 1 (rho = 0.2387)."""
    rng = np.random.default_rng(9)
    rho = 0.2 + 0.4 * rng.random(400)
    rho = rho[np.abs(rho - 3.0 / (4.0 * math.pi)) > 1e-3]
    assert np.any(rho < 0.2387) and np.any(rho > 0.2387)
    h = 1e-6
    fd = np.array([(O.lda_energy(np.array([r + h])) - O.lda_energy(np.array([r - h]))) / (2 * h) for r in rho])
    np.testing.assert_allclose(O.lda_potential(rho), fd, rtol=1e-8)


def test_potential_is_energy_derivative():
    cell = _tri_cell((8, 6, 6))
    psp = _al_psp()
    frac = np.array([[0.1, 0.2, 0.3], [0.6, 0.5, 0.8]])
    vion = O.ion_potential(cell, [(psp, frac)])
    rng = np.random.default_rng(3)
    for kin in (4, 5):
        f = O.Functional(kinetic=kin).resolve(6.0, cell.vol)
        sysm = O.System(cell, f, vion)
        if kin == 5:
            # smooth synthetic 1D table standing in for the ODE output (shape only)
            eta = O.wgc_eta_knots()
            w = -1.0 * eta ** 2 / (1 + eta ** 2)
            sysm.wgc_kernels = O.wgc_kernels_from_table(sysm.qtable, eta, w, np.gradient(w, eta),
                                                        np.gradient(np.gradient(w, eta), eta), f)
        rho = 6.0 / cell.vol * (1 + 0.3 * rng.random(cell.shape))
        d = rng.standard_normal(cell.shape)
        pot, _ = O.calculate_potential_plus(rho, sysm, opt_sqrt=False)
        h = 1e-6
        ep = O.calculate_potential_plus(rho + h * d, sysm, False)[1][0]
        em = O.calculate_potential_plus(rho - h * d, sysm, False)[1][0]
        fd = (ep - em) / (2 * h)
        assert fd == pytest.approx(np.sum(pot * d) * cell.dV, rel=1e-7)
        # dE/dphi form: optSqrt multiplies by 2 sqrt(rho) before vW is added
        p2, _ = O.calculate_potential_plus(rho, sysm, opt_sqrt=True)
        np.testing.assert_allclose(p2, 2 * np.sqrt(rho) * pot, rtol=1e-11, atol=1e-13)


def test_ion_potential_spline_converges_to_exact_structure_factor():
    cell = _tri_cell((12, 12, 12))
    psp = _al_psp()
    frac = np.array([[0.1, 0.0, 0.0], [0.5, 0.55, 0.99]])
    exact = O.ion_energy_exact_recip(cell, [(psp, frac)])
    errs = []
    for order in (6, 10, 14):
        v = O.fft_forward(O.ion_potential(cell, [(psp, frac)], order=order))
        errs.append(np.max(np.abs(v - exact)) / np.max(np.abs(exact)))
    assert errs[1] < errs[0] and errs[2] < errs[1] and errs[1] < 1e-3


def test_ion_forces_are_minus_energy_gradient():
    cell = _tri_cell((10, 10, 10))
    psp = _al_psp()
    frac = np.array([[0.13, 0.21, 0.37], [0.61, 0.52, 0.83]])
    rng = np.random.default_rng(5)
    rho = 6.0 / cell.vol * (1 + 0.2 * rng.random(cell.shape))
    forces = O.ion_forces(cell, [(psp, frac)], rho)
    h = 1e-5
    for ion in range(2):
        for c in range(3):
            dr = np.zeros(3)
            dr[c] = h
            ds = np.linalg.solve(cell.lattice, dr)
            fp = frac.copy()
            fp[ion] += ds
            fm = frac.copy()
            fm[ion] -= ds
            ep = np.sum(rho * O.ion_potential(cell, [(psp, fp)])) * cell.dV
            em = np.sum(rho * O.ion_potential(cell, [(psp, fm)])) * cell.dV
            assert -(ep - em) / (2 * h) == pytest.approx(forces[ion, c], rel=1e-6, abs=1e-9)


def test_recpot_charge_and_units():
    psp = _al_psp()
    assert psp.charge == 3 and psp.numPoints == 15003
    assert psp.maxG == pytest.approx(56.6993428892377764 * O.BOHR)
    # synthetic V_loc round trip through the same ReadPseudo semantics
    maxg, vals = O.synthetic_recpot_values(3, 1.2)
    sp = O.make_psp(maxg, vals)
    assert sp.charge == 3
    q = np.array([0.5, 1.3, 4.0])
    np.testing.assert_allclose(O.psp_lookup(sp, q), -4 * np.pi * 3 / q ** 2 * np.exp(-q ** 2 * 1.44 / 4),
                               rtol=1e-7)


needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (no reference tree)")


@needs_ref
def test_splines_bit_exact_against_reference_fortran():
    rng = np.random.default_rng(7)
    t = np.cumsum(rng.random(40) + 0.1)
    y = np.sin(t)
    for ib, ie in [(0, 0), (1, 1), (2, 1)]:
        a = O.spline_cubic_set(t, y, ib, 0.3, ie, -0.2)
        b = R.spline_set(t, y, ib, 0.3, ie, -0.2)
        assert np.array_equal(a, b)
        tv = np.concatenate([rng.random(50) * (t[-1] + 2) - 1, t[:5]])
        va = np.array([O.spline_cubic_val(t, y, a, x) for x in tv])
        vb = R.spline_val(t, y, b, tv)
        assert np.array_equal(va, vb)
        np.testing.assert_allclose(O.spline_cubic_val_vec(t, y, a, tv), vb, rtol=1e-13, atol=1e-13)


@needs_ref
def test_dcsrch_step_for_step_against_reference_fortran():
    # Moré–Thuente test functions (phi(a) = -a/(a^2+b), and a cubic-like one)
    fns = [lambda a: (-a / (a * a + 2.0), -(2.0 - a * a) / (a * a + 2.0) ** 2),
           lambda a: ((a + 0.004) ** 5 - 2 * (a + 0.004) ** 4, 5 * (a + 0.004) ** 4 - 8 * (a + 0.004) ** 3)]
    for fn in fns:
        for stp0 in (1e-3, 0.1, 1.0, 10.0):
            ours = O.Dcsrch(1e-3, 0.1, 1e-10, 0.0, 1e10)
            theirs = R.RefDcsrch(1e-3, 0.1, 1e-10, 0.0, 1e10)
            f, g = fn(0.0)
            s1 = s2 = stp0
            for _ in range(40):
                s1, t1 = ours.step(s1, f, g)
                s2, t2 = theirs.step(s2, f, g)
                assert s1 == s2 and t1[:2] == t2[:2]
                if not t1.startswith("FG"):
                    break
                f, g = fn(s1)


def _option_system():
    psp = _al_psp()
    cell = O.Cell(S.OPTION_LATTICE_A.T / O.BOHR, 16, 16, 16)
    vion = O.ion_potential(cell, [(psp, S.OPTION_FRAC)])
    f = O.Functional(kinetic=4).resolve(12.0, cell.vol)
    return O.System(cell, f, vion, S.OPTION_EII_EV / O.HARTREE_TO_EV)


def test_reference_sample_optIon_wt_triclinic():
    """test/optIon sample: step-0 energy (uniform density) and the converged NTN
    total energy on a triclinic 16^3 cell (reference outputs agree to about
    8 significant figures across machines, per the reference's manual FAQ)."""
    sysm = _option_system()
    rho = np.full(sysm.cell.shape, 12.0 / sysm.cell.vol)
    full, pot = O.make_evaluators(sysm)
    _, et = full(rho)
    assert et[0] * O.HARTREE_TO_EV == pytest.approx(S.OPTION_STEP0_EV, rel=1e-12)
    res = O.rho_minimization(rho, full, pot, sysm.cell.dV)
    assert res.outcome == 0 and res.iters == S.OPTION_ITERS
    assert abs(res.energy[0] * O.HARTREE_TO_EV - S.OPTION_FINAL_EV) / 4 < 1e-6   # eV/atom
    for h, ref in zip(res.history[1:], S.OPTION_ITER_HA):
        assert h["energy"] == pytest.approx(ref, rel=2e-6)


def test_reference_sample_optDen_wgc():
    """test/optDen sample (WGC, second order): same Newton inner-iteration counts
    per step and the same energies within the accuracy of the 1D kernel table
    (RKSUITE at TOL 5e-5, KEDF_KernelODE.f90:376)."""
    psp = _al_psp()
    cell = O.Cell(S.OPTDEN_LATTICE_A.T / O.BOHR, 16, 16, 16)
    vion = O.ion_potential(cell, [(psp, S.OPTDEN_FRAC)])
    f = O.Functional(kinetic=5, alpha=S.OPTDEN_ALPHA, beta=S.OPTDEN_BETA).resolve(12.0, cell.vol)
    sysm = O.System(cell, f, vion, S.OPTDEN_EII_EV / O.HARTREE_TO_EV)
    eta, w, w1, w2 = W.for_params(f.alpha, f.beta, f.gamma)
    sysm.wgc_kernels = O.wgc_kernels_from_table(sysm.qtable, eta, w, w1, w2, f)
    rho = np.full(cell.shape, 12.0 / cell.vol)
    full, pot = O.make_evaluators(sysm)
    res = O.rho_minimization(rho, full, pot, cell.dV)
    assert [h["newton_iter"] for h in res.history[1:]] == S.OPTDEN_NEWTON_ITERS
    for h, ref in zip(res.history, S.OPTDEN_STEP_EV):
        assert abs(h["energy"] * O.HARTREE_TO_EV - ref) / 4 < 5e-6


# ---------------------------------------------------------------- NCG directions
def test_cg_direction_hz_sufficient_descent_and_restarts():
    """Hager-Zhang directions satisfy d.g <= -7/8 |g|^2 for any previous
    direction with d_old.y > 0 (Hager & Zhang, SIAM J. Optim. 16 (2005),
    Theorem 1.1); the reference's InnerProd carries dV, which cancels at
    dV = 1. A flipped sign on the 2|y|^2 term breaks the bound in most draws.
    Steps 1, 101, 201 restart with -g (RhoDirCG.f90:67-70)."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        g, g_old, d_old = (rng.standard_normal(50) for _ in range(3))
        if np.dot(d_old, g - g_old) <= 0:
            d_old = -d_old
        d, g_saved = O.cg_direction(5, g, g_old, d_old, 1.0)
        assert np.dot(d, g) <= -0.875 * np.dot(g, g) + 1e-12
        assert np.array_equal(g_saved, g)
    for it in (1, 101, 201):
        d, _ = O.cg_direction(it, g, g_old, d_old, 0.3)
        assert np.array_equal(d, -g)
    d, _ = O.cg_direction(102, g, g_old, d_old, 0.3)
    assert not np.allclose(d, -g)


@pytest.mark.parametrize("alg", ["HZ", "PR"])
@pytest.mark.parametrize("dV", [1.0, 0.37])
def test_cg_direction_terminates_on_quadratic(alg, dV):
    """With exact line searches on an SPD quadratic both formulas reduce to
    linear CG (the HZ correction term is d_old.g = 0; HS = PR there), which
    reaches the minimiser in n steps; dV cancels in both ratios."""
    rng = np.random.default_rng(3)
    n = 8
    a = rng.standard_normal((n, n))
    a = a @ a.T + n * np.eye(n)
    b = rng.standard_normal(n)
    x, g_old, d_old = np.zeros(n), np.zeros(n), np.zeros(n)
    for it in range(1, n + 1):
        g = a @ x - b
        d, g_old = O.cg_direction(it, g, g_old, d_old, dV, alg)
        x = x - np.dot(g, d) / np.dot(d, a @ d) * d
        d_old = d
    assert np.linalg.norm(a @ x - b) < 1e-12 * np.linalg.norm(b)


def _dense_bfgs_direction(pairs, g):
    """-H g with H built explicitly by the BFGS inverse update
    H <- (I - r s y^T) H (I - r y s^T) + r s s^T over the stored pairs, oldest
    first, from H0 = (s.y / y.y) I of the newest pair (Nocedal & Wright,
    Numerical Optimization, eq. 7.19)."""
    s, y = pairs[-1]
    n = g.size
    h = np.dot(s, y) / np.dot(y, y) * np.eye(n)
    for s, y in pairs:
        r = 1.0 / np.dot(y, s)
        v = np.eye(n) - r * np.outer(y, s)
        h = v.T @ h @ v + r * np.outer(s, s)
    return -h @ g


def test_bfgs_two_loop_matches_dense_update():
    """BFGSDirection's two-loop recursion over the MBFGS = 5 ring, called once
    per step as the optimizer does (1/y.s is stored for the newest pair only,
    RhoDirBFGS.f90:169-171), equals -H g of the dense inverse-BFGS matrix over
    the newest five pairs, through the ring's wrap-around; dV cancels. Reversing
    either loop, or taking H0 from the oldest pair, fails this by 20-80 %. The
    first call returns -g; with g = y of the newest pair the result is -s
    (secant condition)."""
    rng = np.random.default_rng(2)
    n = 9
    mem = O.BfgsMemory((n,))
    g0 = rng.standard_normal(n)
    assert np.array_equal(mem.direction(g0, 0.37), -g0)
    pairs = []
    for k in range(12):
        s = rng.standard_normal(n)
        y = s + 0.5 * rng.standard_normal(n)
        if np.dot(s, y) <= 0:
            y = -y
        mem.s[mem.pt], mem.y[mem.pt] = s, y
        mem.pt = (mem.pt + 1) % mem.m
        pairs.append((s, y))
        g = rng.standard_normal(n)
        d = mem.direction(g, 0.37)
        ref = _dense_bfgs_direction(pairs[-5:], g)
        assert np.max(np.abs(d - ref)) < 1e-13 * np.max(np.abs(ref))
    mem.iter -= 1   # the same step again, now on g = y_newest
    np.testing.assert_allclose(mem.direction(pairs[-1][1], 0.37), -pairs[-1][0], rtol=0, atol=1e-14)


def test_ncg_reaches_ntn_minimum_optIon():
    """ProjNormRhoMinimization with dir_type 2 (NCG, HZ and PR) and 3 (NBF) on the
    test/optIon input converges to the minimum NTN reaches, which the
    reference's recorded output pins (same energy to 1e-7 eV/atom), and no
    accepted step raises the energy. PR's line search stalls (Dcsrch
    warnings) below a potential norm of about 1e-5 here, so it runs to 1e-5."""
    sysm = _option_system()
    rho = np.full(sysm.cell.shape, 12.0 / sysm.cell.vol)
    full, pot = O.make_evaluators(sysm)
    ntn = O.rho_minimization(rho, full, pot, sysm.cell.dV, O.OptParams(conv_check="POT", pot_tol=1e-6))
    assert abs(ntn.energy[0] * O.HARTREE_TO_EV - S.OPTION_FINAL_EV) / 4 < 1e-6
    for alg, tol in (("HZ", 1e-6), ("PR", 1e-5)):
        prm = O.OptParams(conv_check="POT", pot_tol=tol)
        res = O.rho_minimization(rho, full, pot, sysm.cell.dV, prm, dir_type=2, cg_alg=alg)
        assert res.outcome == 0 and res.iters > ntn.iters
        assert abs(res.energy[0] - ntn.energy[0]) * O.HARTREE_TO_EV / 4 < 1e-7
        es = [h["energy"] for h in res.history]
        assert all(b < a + 1e-10 for a, b in zip(es, es[1:]))   # Hartree; rounding near the minimum
    # dir_type 3 (NBF, RhoDirBFGS.f90), the same way
    res = O.rho_minimization(rho, full, pot, sysm.cell.dV, O.OptParams(conv_check="POT", pot_tol=1e-6), dir_type=3)
    assert res.outcome == 0 and res.iters > ntn.iters
    assert abs(res.energy[0] - ntn.energy[0]) * O.HARTREE_TO_EV / 4 < 1e-7


# ---------------------------------------------------------------- Ewald
def _ewald_cell(lat_a, n=16):
    return O.Cell(np.asarray(lat_a).T / O.BOHR, n, n, n)


def test_ewald_reference_samples():
    """The oracle's particle-mesh Ewald energy against the reference's recorded
    ion-ion energies (test/optIon first and final geometry, test/optDen) and
    the printed Ewald parameters."""
    fin = np.array([[0.026242363836079, 0.012513411405787, 0.997499799127206],
                    [0.526213499121448, 0.512582285849498, 0.997494511543261],
                    [0.026254939350756, 0.512439721648613, 0.497479346361002],
                    [0.526294902673969, 0.012517237611835, 0.497501509902651]])
    cases = [(S.OPTION_LATTICE_A, S.OPTION_FRAC, S.OPTION_EII_EV, (0.5862, 8.582, 6.234)),
             (S.OPTION_LATTICE_A, fin, S.OPTION_EII_FINAL_EV, (0.5862, 8.582, 6.234)),
             (S.OPTDEN_LATTICE_A, S.OPTDEN_FRAC, S.OPTDEN_EII_EV, (0.5914, 8.498, 6.283))]
    for lat, frac, ref_ev, printed in cases:
        cell = _ewald_cell(lat)
        e, comps, p = O.ewald_energy(cell, np.full(4, 3.0), frac)
        assert abs(e * O.HARTREE_TO_EV - ref_ev) <= 1e-11 * abs(ref_ev)
        # G12.4 / F formats of EwaldSetup's report (Ewald.f90:451-455)
        assert round(p.eta, 4) == printed[0]
        assert round(p.real_cutoff, 3) == printed[1]
        assert round(p.recip_cutoff, 3) == printed[2]


def test_ewald_madelung_and_eta_independence():
    """fcc lattice of point charges Z in a neutralizing background: energy per
    ion = -0.8958736 Z^2 / r_ws (the fcc Madelung constant referred to the
    Wigner-Seitz radius); the exact-sum energy does not depend on eta; the
    particle-mesh sum agrees with the exact one."""
    cell = O.Cell(np.eye(3) * 7.65, 24, 24, 24)
    fcc = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=float)
    q = np.full(4, 3.0)
    rws = (3.0 * cell.vol / (4.0 * np.pi * 4)) ** (1.0 / 3.0)
    e, _, _ = O.ewald_energy(cell, q, fcc, spline=False)
    assert abs(e / 4 * rws / 9.0 - (-0.8958736)) < 2e-7
    rng = np.random.default_rng(3)
    fr = rng.random((5, 3))
    qq = np.array([3.0, 3.0, 2.0, 2.0, 3.0])
    tri = O.Cell(np.array([[7.3, 0.2, 0.1], [0.3, 7.5, 0.0], [-0.2, 0.4, 7.6]]).T, 24, 24, 24)
    es = []
    for eta in (0.9, 0.7):
        p = O.EwaldParams(eta, 6.5 / eta, 13.0 * eta, 0.0, 0.0)
        es.append(O.ewald_energy(tri, qq, fr, spline=False, params=p)[0])
    assert abs(es[0] - es[1]) < 1e-10 * abs(es[0])
    ep, _, _ = O.ewald_energy(tri, qq, fr, spline=True)
    ex, _, _ = O.ewald_energy(tri, qq, fr, spline=False)
    assert abs(ep - ex) < 1e-7 * abs(ex)


def test_ewald_forces_are_minus_energy_gradient():
    cell = _ewald_cell(S.OPTION_LATTICE_A)
    q = np.full(4, 3.0)
    fr = S.OPTION_FRAC.copy()
    p = O.ewald_setup(cell, q)
    f = O.ewald_forces(cell, q, fr, params=p)
    h = 1e-5
    for i, k in [(0, 0), (1, 1), (3, 2), (2, 0)]:
        dfr = np.linalg.solve(cell.lattice, np.eye(3)[k]) * h
        fp, fm = fr.copy(), fr.copy()
        fp[i] += dfr
        fm[i] -= dfr
        de = (O.ewald_energy(cell, q, fp, params=p)[0] - O.ewald_energy(cell, q, fm, params=p)[0]) / (2 * h)
        assert abs(-de - f[i, k]) < 1e-8


# ---------------------------------------------------------------- ion relaxation
@needs_ref
def test_lnsrch_step_for_step_against_reference_fortran():
    """The oracle's LnSrch against the reference's MathLNsrch.f90 on line
    functions that take every branch kind (bracketing from the left and the
    right, order flips, interpolation failures): same alpha, same task, same
    bracket points after every call."""
    fns = [lambda a: ((a - 0.7) ** 2 - 1.0, 2 * (a - 0.7)),
           lambda a: (np.cos(3 * a) + 0.1 * a, -3 * np.sin(3 * a) + 0.1),
           lambda a: (-a / (a * a + 2.0), -(2.0 - a * a) / (a * a + 2.0) ** 2),
           lambda a: (a ** 4 - 3 * a ** 2 + a, 4 * a ** 3 - 6 * a + 1),
           lambda a: (np.exp(-a) * 0.3 - 0.01 * a, -0.3 * np.exp(-a) - 0.01)]
    c1, c2 = O._f32(1e-4), O._f32(1e-1)
    for fn in fns:
        for a0, smax in ((1.0, 5.0), (0.05, 3.0), (2.5, 10.0), (0.3, 0.5)):
            ours, theirs = O.LnSrch(), R.RefLnSrch()
            f, g = fn(0.0)
            if g > 0:
                continue
            t1 = t2 = "START"
            a1 = a2 = a0
            for _ in range(25):
                a1, t1 = ours.step(a1, f, g, t1, c1, c2, 0.0, smax)
                a2, t2 = theirs.step(a2, f, g, t2, c1, c2, 0.0, smax)
                assert a1 == a2 and t1.strip() == t2.strip(), (t1, t2, a1, a2)
                assert np.array_equal(ours.points, theirs.points)
                if not (t1[:2] == "FG" or t1[:4] == "WARN"):
                    break
                f, g = fn(a1)


def test_nn_min_distance():
    cell = O.Cell(np.eye(3) * 7.65, 16, 16, 16)
    fcc = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=float)
    assert abs(O.nn_min_distance(cell, fcc) - 7.65 / np.sqrt(2.0)) < 1e-12
    one = O.Cell(np.diag([5.0, 6.0, 7.0]), 8, 8, 8)
    assert abs(O.nn_min_distance(one, [[0.1, 0.2, 0.3], [0.9, 0.2, 0.3]]) - 1.0) < 1e-12


def _option_relax_evaluator(state):
    psp = _al_psp()
    cell = O.Cell(S.OPTION_LATTICE_A.T / O.BOHR, 16, 16, 16)
    f = O.Functional(kinetic=4).resolve(12.0, cell.vol)
    q = np.full(4, 3.0)

    def evaluate(frac):
        vion = O.ion_potential(cell, [(psp, frac)])
        eii, _, _ = O.ewald_energy(cell, q, frac)
        full, pot = O.make_evaluators(O.System(cell, f, vion, eii))
        res = O.rho_minimization(state["rho"], full, pot, cell.dV)   # warm start (trashPreviousRho = F)
        state["rho"] = res.rho
        state["iters"].append(res.iters)
        return res.energy[0], O.ion_forces(cell, [(psp, frac)], res.rho) + O.ewald_forces(cell, q, frac)
    return cell, evaluate


def test_reference_sample_optIon_relaxation():
    """test/optIon CG2 relaxation (displaced.inpt: MINI ion, meth ion cg2,
    method ntn): the same 13 geometries with the same density step counts
    but one, the final energy within 1e-6 eV/atom of displaced.out and the
    final geometry within 1e-3 bohr of displaced.final.geom. The intermediate
    energies agree to 5e-5 eV/atom: forces are first order in the density's
    residual error (the solve stops at a potential norm of 1e-5), so the
    first forces already differ from the reference's by about 1e-5 Ha/bohr
    (its recorded dE/dalpha -5.43963e-3 against -5.44135e-3 here), and the
    line-search trial points move with them."""
    cell = O.Cell(S.OPTION_LATTICE_A.T / O.BOHR, 16, 16, 16)
    state = {"rho": np.full(cell.shape, 12.0 / cell.vol), "iters": []}
    cell, evaluate = _option_relax_evaluator(state)
    r = O.ion_relax_cg2(cell, S.OPTION_FRAC, evaluate)
    ev = np.array(r.energies) * O.HARTREE_TO_EV
    assert len(ev) == len(S.OPTION_RELAX_EV)
    assert np.max(np.abs(ev - np.array(S.OPTION_RELAX_EV))) / 4 < 5e-5
    assert abs(ev[-1] - S.OPTION_RELAX_EV[-1]) / 4 < 1e-6
    assert sum(a != b for a, b in zip(state["iters"], S.OPTION_RELAX_NTN_STEPS)) <= 1
    d = np.mod(r.frac - S.OPTION_FINAL_FRAC + 0.5, 1.0) - 0.5
    assert np.max(np.abs(d @ cell.lattice.T)) < 1e-3


@pytest.mark.parametrize("cg_type", [1, 2, 3])
def test_ion_cg_variants_terminate_on_quadratic(cg_type):
    """The ion relaxation's Polak-Ribiere, Hager-Zhang and Dai-Yuan updates
    (IonOptCG2.f90:286-308) on E = 1/2 sum_i r_i^T K r_i: LnSrch's secant
    step is exact on a quadratic, and the Hessian blockdiag(K, K, K) has three
    distinct eigenvalues, so conjugate directions reach the minimum in three
    steps (steepest descent does not)."""
    cell = O.Cell(np.diag([10.0, 11.0, 12.0]), 8, 8, 8)
    k = np.array([[1.0, 0.2, 0.0], [0.2, 0.5, 0.1], [0.0, 0.1, 0.3]])
    c = np.array([[0.2, 0.3, 0.4], [0.6, 0.7, 0.1], [0.45, 0.15, 0.8]])

    def evaluate(frac):
        r = (np.mod(frac - c + 0.5, 1.0) - 0.5) @ cell.lattice.T
        return 0.5 * float(np.sum((r @ k) * r)), -(r @ k)
    start = c + np.array([[0.03, -0.02, 0.01], [-0.01, 0.02, 0.03], [0.02, 0.01, -0.03]])
    res = O.ion_relax_cg2(cell, start, evaluate, cg_type=cg_type, nn_dist=0.0)
    assert res.steps <= 3 and res.max_force < 1e-12
    d = np.mod(res.frac - c + 0.5, 1.0) - 0.5
    assert np.max(np.abs(d @ cell.lattice.T)) < 1e-12


def test_reference_sample_optCell_first_geometry():
    """test/optCell's first geometry (the optIon lattice with the ions at the
    ideal positions, test/optCell/displaced.ion): its ion-ion energy
    (displaced.out:108) to 1e-11 relative, the uniform-density step 0 (:93) to
    1e-12, the same inner Newton counts on the first two NTN steps (2 and 8,
    :94-95), the converged total energy (:112) within 1e-6 eV/atom. The
    reference stops after 7 steps, the oracle after 6: from step 3 on the
    inner CG exits one iteration apart and the ENE test (three energies within
    1e-6 Ha) is met one step earlier; the energy terms at the two stopping
    points differ at first order in the density residual (2e-4 eV/atom)."""
    psp = _al_psp()
    g = os.path.join(GOLDEN, "optCell_displaced.ion")
    with open(g) as fh:
        assert "0.000000000000000   0.500000000000000   0.500000000000000" in fh.read()
    cell = O.Cell(S.OPTION_LATTICE_A.T / O.BOHR, 16, 16, 16)
    eii, _, _ = O.ewald_energy(cell, np.full(4, 3.0), S.OPTCELL_FRAC)
    assert eii * O.HARTREE_TO_EV == pytest.approx(S.OPTCELL_EII_EV, rel=1e-11)
    vion = O.ion_potential(cell, [(psp, S.OPTCELL_FRAC)])
    f = O.Functional(kinetic=4).resolve(12.0, cell.vol)
    sysm = O.System(cell, f, vion, eii)
    rho = np.full(cell.shape, 12.0 / cell.vol)
    full, pot = O.make_evaluators(sysm)
    assert full(rho)[1][0] * O.HARTREE_TO_EV == pytest.approx(S.OPTCELL_STEP0_EV, rel=1e-12)
    res = O.rho_minimization(rho, full, pot, cell.dV)
    assert res.outcome == 0
    assert [h["newton_iter"] for h in res.history[1:3]] == [2, 8]
    assert abs(res.energy[0] * O.HARTREE_TO_EV - S.OPTCELL_ITER_EV[-1]) / 4 < 1e-6
    e = res.energy * O.HARTREE_TO_EV
    for k, i in {"TF": 6, "vW": 7, "WT": 8, "XC": 4, "H": 3, "IE": 2}.items():
        assert abs(e[i] - S.OPTCELL_TERMS_EV[k]) / 4 < 2e-4, k
