"""The CPU restatement timed as bench.py's CPU baseline (oracle/cpu/ofd_cpu.cpp,
the reference's four-forward, four-inverse transform dataflow on its own
mixed-radix FFT, OpenMP) against the numpy oracle: every eTable term and the
potential at 1e-11, on cubic, mixed-radix and triclinic grids, alpha = beta
and alpha != beta, both optSqrt forms; and the same result on one thread and
on several."""
import numpy as np
import pytest

from oracle import cpubuild
from oracle import profess_oracle as O
from tests.test_oracle import _al_psp, _tri_cell


@pytest.fixture(scope="module")
def cpu():
    cpubuild.build()
    return cpubuild


@pytest.mark.parametrize("grid,alpha,beta", [((16, 16, 16), -1.0, -1.0), ((24, 20, 18), -1.0, -1.0),
                                             ((32, 30, 10), 0.9, 0.7), ((12, 25, 16), -1.0, -1.0)])
def test_cpu_port_matches_oracle(cpu, grid, alpha, beta):
    cell = _tri_cell(grid)
    rng = np.random.default_rng(12)
    frac = rng.random((3, 3))
    vion = O.ion_potential(cell, [(_al_psp(), frac)])
    ne = 9.0
    f = O.Functional(kinetic=4, alpha=alpha if alpha > 0 else -100.0,
                     beta=beta if beta > 0 else -100.0).resolve(ne, cell.vol)
    sysm = O.System(cell, f, vion, e_ionion=-0.75)
    rho = ne / cell.vol * (1.0 + 0.3 * (rng.random(cell.shape) - 0.5))
    ce = cpu.CpuEval(cell.lattice, grid, f.rho0, alpha=f.alpha, beta=f.beta)
    for opt_sqrt in (True, False):
        pot, et = ce.calpot(rho, vion, -0.75, opt_sqrt=opt_sqrt, threads=1)
        pr, er = O.calculate_potential_plus(rho, sysm, opt_sqrt, True)
        for k in range(9):
            assert abs(et[k] - er[k]) <= 1e-11 * abs(er[k]) + 1e-13, (k, et[k], er[k])
        assert np.max(np.abs(pot - pr)) <= 1e-11 * np.max(np.abs(pr))
    p4, e4 = ce.calpot(rho, vion, -0.75, threads=4)
    p1, e1 = ce.calpot(rho, vion, -0.75, threads=1)
    assert np.max(np.abs(p4 - p1)) <= 1e-13 * np.max(np.abs(p1))
    np.testing.assert_allclose(e4, e1, rtol=1e-13, atol=1e-15)
    ce.close()


def test_cpu_port_rejects_unsupported_lengths(cpu):
    with pytest.raises(RuntimeError):
        cpu.CpuEval(np.eye(3) * 8.0, (14, 16, 16), 0.03)   # n1/2 = 7
