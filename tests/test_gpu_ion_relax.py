"""GPU: the CG2 ion relaxation (ofd_ion_relax through the C ABI) on the
reference's test/optIon input, against the oracle's relaxation and the
reference's recorded trajectory (displaced.out, displaced.final.geom)."""
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


def _recpot_text():
    with open(os.path.join(GOLDEN, "al_HC.lda.recpot")) as fh:
        return fh.read()


def _oracle_relax(dir_type=1, params=O.OptParams(), cg_type=2):
    psp = O.make_psp(*O.read_recpot_text(_recpot_text()))
    cell = O.Cell(S.OPTION_LATTICE_A.T / O.BOHR, 16, 16, 16)
    f = O.Functional(kinetic=4).resolve(12.0, cell.vol)
    q = np.full(4, 3.0)
    state = {"rho": np.full(cell.shape, 12.0 / cell.vol)}

    def evaluate(frac):
        vion = O.ion_potential(cell, [(psp, frac)])
        eii, _, _ = O.ewald_energy(cell, q, frac)
        full, pot = O.make_evaluators(O.System(cell, f, vion, eii))
        res = O.rho_minimization(state["rho"], full, pot, cell.dV, params, dir_type=dir_type)
        state["rho"] = res.rho
        return res.energy[0], O.ion_forces(cell, [(psp, frac)], res.rho) + O.ewald_forces(cell, q, frac)
    return cell, O.ion_relax_cg2(cell, S.OPTION_FRAC, evaluate, cg_type=cg_type)


def test_ion_relaxation_optIon(pa):
    _say("ion relaxation optIon")
    mg, v = O.read_recpot_text(_recpot_text())
    lat = S.OPTION_LATTICE_A.T / O.BOHR
    s = pa.OFDSystem(lat, (16, 16, 16), 12.0, kinetic=4)
    s.build_ion_potential([(mg, v, S.OPTION_FRAC)])
    rho0 = np.full(s.shape, 12.0 / s.vol)
    rho, frac, et, forces, energies, st = pa.ConjugateGradient(s, rho0, S.OPTION_FRAC)
    _say("geometries %d, steps %d, max force %.3e, %.2f s" % (st["geometries"], st["steps"], st["max_force"],
                                                              st["seconds"]))
    cell, ref = _oracle_relax()
    ev = energies * O.HARTREE_TO_EV
    ev_or = np.array(ref.energies) * O.HARTREE_TO_EV
    # against the oracle's relaxation: same geometries, energies within 1e-6 eV/atom
    assert len(ev) == len(ev_or)
    assert np.max(np.abs(ev - ev_or)) / 4 < 1e-6
    d = np.mod(frac - ref.frac + 0.5, 1.0) - 0.5
    assert np.max(np.abs(d @ lat.T)) < 1e-5
    # against the reference's run (see tests/test_oracle.py for the tolerances)
    assert len(ev) == len(S.OPTION_RELAX_EV)
    assert abs(ev[-1] - S.OPTION_RELAX_EV[-1]) / 4 < 1e-6
    assert np.max(np.abs(ev - np.array(S.OPTION_RELAX_EV))) / 4 < 5e-5
    d = np.mod(frac - S.OPTION_FINAL_FRAC + 0.5, 1.0) - 0.5
    assert np.max(np.abs(d @ lat.T)) < 1e-3
    assert st["max_force"] < float(np.float32(5e-5))
    assert et[0] == energies[-1]


@pytest.mark.parametrize("dir_type", [2, 3])
def test_ion_relaxation_optIon_ncg_nbf(pa, dir_type):
    """The same relaxation with NCG (2) or NBF (3) density solves to a
    potential norm of 1e-6: the oracle's energies at every geometry, and the
    reference's recorded final energy (1e-6 eV/atom) and geometry (1e-3 bohr)."""
    _say("ion relaxation optIon, density dir_type %d" % dir_type)
    mg, v = O.read_recpot_text(_recpot_text())
    lat = S.OPTION_LATTICE_A.T / O.BOHR
    s = pa.OFDSystem(lat, (16, 16, 16), 12.0, kinetic=4)
    s.build_ion_potential([(mg, v, S.OPTION_FRAC)])
    rho0 = np.full(s.shape, 12.0 / s.vol)
    rho, frac, et, forces, energies, st = pa.ConjugateGradient(s, rho0, S.OPTION_FRAC, rho_dir_type=dir_type,
                                                               conv_check="POT", pot_tol=1e-6)
    _, ref = _oracle_relax(dir_type, O.OptParams(conv_check="POT", pot_tol=1e-6))
    ev = energies * O.HARTREE_TO_EV
    ev_or = np.array(ref.energies) * O.HARTREE_TO_EV
    assert len(ev) == len(ev_or) == len(S.OPTION_RELAX_EV)
    assert np.max(np.abs(ev - ev_or)) / 4 < 1e-6
    assert abs(ev[-1] - S.OPTION_RELAX_EV[-1]) / 4 < 1e-6
    d = np.mod(frac - S.OPTION_FINAL_FRAC + 0.5, 1.0) - 0.5
    assert np.max(np.abs(d @ lat.T)) < 1e-3


@pytest.mark.parametrize("cg_type", [1, 3])
def test_ion_relaxation_optIon_pr_dy(pa, cg_type):
    """Polak-Ribiere (1) and Dai-Yuan (3) ion directions on test/optIon: the
    oracle's relaxation with the same update, geometry for geometry, and the
    minimum the reference's Hager-Zhang run records (final energy within
    1e-6 eV/atom)."""
    _say("ion relaxation optIon, cg_type %d" % cg_type)
    mg, v = O.read_recpot_text(_recpot_text())
    lat = S.OPTION_LATTICE_A.T / O.BOHR
    s = pa.OFDSystem(lat, (16, 16, 16), 12.0, kinetic=4)
    s.build_ion_potential([(mg, v, S.OPTION_FRAC)])
    rho0 = np.full(s.shape, 12.0 / s.vol)
    rho, frac, et, forces, energies, st = pa.ConjugateGradient(s, rho0, S.OPTION_FRAC, cg_type=cg_type)
    _, ref = _oracle_relax(cg_type=cg_type)
    ev = energies * O.HARTREE_TO_EV
    ev_or = np.array(ref.energies) * O.HARTREE_TO_EV
    assert len(ev) == len(ev_or)
    assert np.max(np.abs(ev - ev_or)) / 4 < 1e-6
    d = np.mod(frac - ref.frac + 0.5, 1.0) - 0.5
    assert np.max(np.abs(d @ lat.T)) < 1e-5
    assert abs(ev[-1] - S.OPTION_RELAX_EV[-1]) / 4 < 1e-6
    assert st["max_force"] < float(np.float32(5e-5))
