"""Inputs of the reference's own example runs (test/optIon, test/optDen) and the
numbers their recorded outputs hold, as data. Lattices in Angstrom (rows are
lattice vectors as written in the .ion files), fractional positions, 16^3 grids."""
import numpy as np

OPTION_LATTICE_A = np.array([[3.869104452400000, 0.100000000000000, 0.0],
                             [0.000000000000000, 3.969104452400000, 0.0],
                             [0.000000000000000, 0.000000000000000, 4.0]])
OPTION_FRAC = np.array([[0.1, 0.0, 0.0], [0.5, 0.55, 0.99], [0.0, 0.5, 0.5], [0.505, 0.0, 0.5]])
# test/optIon/sampleoutput/displaced.out:92-99 (eV) and :119 (ion-ion)
OPTION_STEP0_EV = -2.199483598873e+02
OPTION_FINAL_EV = -2.311425327622e+02
OPTION_EII_EV = -2.986846473280e+02
# displaced.out:934 (ion-ion at the final geometry of displaced.final.geom)
OPTION_EII_FINAL_EV = -3.011568598115e+02
# the CG2 relaxation of displaced.inpt: TOTAL ENERGY of every geometry evaluated
# (displaced.out:123,192,...,938, eV), the NTN steps of each density solve
# (:103,172,...,918) and the final positions (displaced.final.geom)
OPTION_RELAX_EV = [-2.311425327622e+02, -2.312770513588e+02, -2.316147260670e+02, -2.316471183146e+02,
                   -2.316860136007e+02, -2.316996685913e+02, -2.317083720964e+02, -2.317111562797e+02,
                   -2.317116104934e+02, -2.317118205898e+02, -2.317118261145e+02, -2.317118465883e+02,
                   -2.317118498940e+02]
OPTION_RELAX_NTN_STEPS = [7, 5, 6, 5, 5, 5, 5, 5, 3, 4, 3, 3, 3]
OPTION_FINAL_FRAC = np.array([[0.026242363836079, 0.012513411405787, 0.997499799127206],
                              [0.526213499121448, 0.512582285849498, 0.997494511543261],
                              [0.026254939350756, 0.512439721648613, 0.497479346361002],
                              [0.526294902673969, 0.012517237611835, 0.497501509902651]])
OPTION_ITERS = 7
# standard.out:62-68 (Ha), per NTN iteration
OPTION_ITER_HA = [-8.4856061397, -8.4936397786, -8.4942793002, -8.4943223357,
                  -8.4943318670, -8.4943322391, -8.4943322847]

OPTDEN_LATTICE_A = np.eye(3) * 3.9691
OPTDEN_FRAC = np.array([[0, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]], dtype=float)
OPTDEN_ALPHA = 1.2060113295832982827348622781219   # test/optDen/test.inpt:6
OPTDEN_BETA = 0.46065533708336838393180438854479   # test/optDen/test.inpt:7
# test/optDen/sample.output:91-97 (eV), :123-135 (final terms)
OPTDEN_STEP_EV = [-2.226243159654e+02, -2.312025858301e+02, -2.317471544164e+02,
                  -2.317593951594e+02, -2.317606991483e+02, -2.317607243753e+02,
                  -2.317607263207e+02]
OPTDEN_NEWTON_ITERS = [2, 8, 6, 8, 4, 7]
OPTDEN_EII_EV = -2.994046165635e+02

# test/optCell: the same lattice with the ions at the ideal positions
# (test/optCell/displaced.ion), first geometry of its run
# (test/optCell/sampleoutput/displaced.out:93-113), eV
OPTCELL_FRAC = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5]])
OPTCELL_STEP0_EV = -2.224205733115e+02
OPTCELL_ITER_EV = [-2.311374927723e+02, -2.316989773208e+02, -2.317102291421e+02, -2.317117591695e+02,
                   -2.317118278762e+02, -2.317118334728e+02, -2.317118335995e+02]
OPTCELL_EII_EV = -3.011568607522e+02
# TF, vW, WT, LDA, Hartree, ion-electron at the converged density (:104-110)
OPTCELL_TERMS_EV = {"TF": 8.990205775509e+01, "vW": 7.344980972680e+00, "WT": -2.785239973122e+00,
                    "XC": -8.927974807017e+01, "H": 4.503089276774e-01, "IE": 6.381266754054e+01}
