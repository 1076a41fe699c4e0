"""GPU parity of the ion-ion Ewald sum (ofd_ewald through the C ABI) against
the oracle and the reference's recorded ion-ion energies. Bar: energy within
1e-10 relative, forces within 1e-10 of max|F| (absolute floor 1e-12 Ha/bohr)."""
import os
import sys

import numpy as np
import pytest

from oracle import profess_oracle as O
from tests import sample_cases as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pa():
    import profess_amd
    profess_amd.lib()
    return profess_amd


def _al():
    with open(os.path.join(GOLDEN, "al_HC.lda.recpot")) as fh:
        return O.read_recpot_text(fh.read())


@pytest.mark.parametrize("case", ["optIon", "optDen"])
def test_ewald_reference_samples(pa, case):
    _say("ewald " + case)
    lat_a, frac, ref_ev = ((S.OPTION_LATTICE_A, S.OPTION_FRAC, S.OPTION_EII_EV) if case == "optIon"
                           else (S.OPTDEN_LATTICE_A, S.OPTDEN_FRAC, S.OPTDEN_EII_EV))
    lat = np.asarray(lat_a).T / O.BOHR
    mg, v = _al()
    s = pa.OFDSystem(lat, (16, 16, 16), 12.0, kinetic=4)
    s.build_ion_potential([(mg, v, frac)])
    e, f, info = pa.EwaldEnergy(s, forces=True)
    cell = O.Cell(lat, 16, 16, 16)
    q = np.full(4, 3.0)
    e_ref, _, p = O.ewald_energy(cell, q, frac)
    f_ref = O.ewald_forces(cell, q, frac)
    assert info["eta"] == p.eta and info["real_cutoff"] == p.real_cutoff
    assert abs(info["recip_cutoff"] - p.recip_cutoff) <= 1e-14 * p.recip_cutoff
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
    assert abs(e * O.HARTREE_TO_EV - ref_ev) <= 1e-10 * abs(ref_ev)
    # the optDen ions sit on a perfect fcc lattice: forces vanish by symmetry
    # and both sides are rounding noise (absolute floor 1e-12 Ha/bohr)
    assert np.max(np.abs(f - f_ref)) <= 1e-10 * np.max(np.abs(f_ref)) + 1e-12
    # the ion-ion energy is eTable(6) of the next evaluation
    rho = np.full(s.shape, 12.0 / s.vol)
    _, et = pa.CalculatePotentialPlus(s, rho)
    assert et[5] == e


def test_ewald_two_species_triclinic(pa):
    """Al (Z = 3) and Mg (Z = 2) ions at random positions in a triclinic cell
    on a ragged grid: energy and forces against the oracle."""
    _say("ewald two species")
    grid = (24, 20, 18)
    lat = np.array([[9.3, 0.4, 0.1], [0.3, 8.5, 0.2], [-0.2, 0.4, 8.1]]).T
    rng = np.random.default_rng(21)
    fa, fm = rng.random((7, 3)), rng.random((5, 3))
    mg_al, v_al = _al()
    mg_mg, v_mg = O.synthetic_recpot_values(2, 1.4)
    s = pa.OFDSystem(lat, grid, 31.0, kinetic=4)
    s.build_ion_potential([(mg_al, v_al, fa), (mg_mg, v_mg, fm)])
    e, f, info = pa.EwaldEnergy(s, forces=True)
    cell = O.Cell(lat, *grid)
    q = np.array([3.0] * 7 + [2.0] * 5)
    frac = np.vstack([fa, fm])
    e_ref, _, p = O.ewald_energy(cell, q, frac)
    f_ref = O.ewald_forces(cell, q, frac)
    assert info["eta"] == p.eta
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
    assert np.max(np.abs(f - f_ref)) <= 1e-10 * np.max(np.abs(f_ref))


def test_system_from_reference_files(pa):
    """The reference's own input files (test/optIon/displaced.ion and its
    al_HC.lda.recpot) through the library's readers, the device ionic
    potential and the Ewald sum: the uniform-density energy of
    displaced.out:92, ion-ion term included."""
    _say("system from files")
    from profess_amd import inputs
    s, g = inputs.system_from_files(os.path.join(GOLDEN, "displaced.ion"), (16, 16, 16))
    rho = np.full(s.shape, 12.0 / s.vol)
    _, et = pa.CalculatePotentialPlus(s, rho)
    assert abs(et[0] * O.HARTREE_TO_EV - S.OPTION_STEP0_EV) <= 1e-10 * abs(S.OPTION_STEP0_EV)
    assert abs(et[5] * O.HARTREE_TO_EV - S.OPTION_EII_EV) <= 1e-10 * abs(S.OPTION_EII_EV)


def test_optCell_first_geometry_from_files(pa):
    """test/optCell/displaced.ion (ideal positions) through the readers, the
    device ionic potential, the device Ewald sum and the NTN solve: its
    recorded ion-ion energy, uniform-density energy and converged total energy
    (test/optCell/sampleoutput/displaced.out:93,108,112), and the oracle's
    iteration count on the same device-built ionic potential."""
    _say("optCell first geometry")
    from profess_amd import inputs
    s, g = inputs.system_from_files(os.path.join(GOLDEN, "optCell_displaced.ion"), (16, 16, 16))
    assert np.array_equal(g["frac"], S.OPTCELL_FRAC)
    rho0 = np.full(s.shape, 12.0 / s.vol)
    _, et = pa.CalculatePotentialPlus(s, rho0)
    assert abs(et[5] * O.HARTREE_TO_EV - S.OPTCELL_EII_EV) <= 1e-10 * abs(S.OPTCELL_EII_EV)
    assert abs(et[0] * O.HARTREE_TO_EV - S.OPTCELL_STEP0_EV) <= 1e-10 * abs(S.OPTCELL_STEP0_EV)
    rho, et, outcome, st = pa.ProjNormRhoMinimization(s, rho0)
    assert outcome == 0
    assert abs(et[0] * O.HARTREE_TO_EV - S.OPTCELL_ITER_EV[-1]) / 4 < 1e-6
    cell = O.Cell(S.OPTION_LATTICE_A.T / O.BOHR, 16, 16, 16)
    f = O.Functional(kinetic=4).resolve(12.0, cell.vol)
    full, pot = O.make_evaluators(O.System(cell, f, s.ion_potential(), et[5]))
    res = O.rho_minimization(rho0, full, pot, cell.dV)
    assert st["iterations"] == res.iters
    assert abs(res.energy[0] - et[0]) * O.HARTREE_TO_EV / 4 < 1e-9
