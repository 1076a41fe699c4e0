"""CPU oracle: a numpy restatement of PROFESS's energy/potential evaluation,
its per-geometry tables and the truncated-Newton (NTN) density optimizer.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker; the product never calls it.

Layout conventions (identical to the reference's Fortran arrays):
  real grid   numpy shape (n3, n2, n1), C order  == Fortran rho(n1,n2,n3), x fastest
  half grid   numpy shape (n3, n2, n1//2+1)      == Fortran (k1G,k2G,k3G) r2c layout
  forward FFT  e^{-iG.r} * 1/N   (Fourier_v1.f90:361,388-393)
  backward FFT e^{+iG.r}, unnormalized (Fourier_v1.f90:436)
All energies are in Hartree; eTable indices follow CalPotential.f90:610-644
(0-based here: [0]=total [1]=KE [2]=ion-electron [3]=Hartree [4]=XC
 [5]=ion-ion [6]=TF [7]=vW [8]=nonlocal WT/WGC).

Pinning (what checks each function; see tests/test_oracle_*.py):
  fft_forward/backward   brute-force DFT on small ragged grids, round trip
  g_tables               hand-computed G vectors on a triclinic cell
  lind_g                 closed-form Lindhard response, Taylor branch continuity
  tf/lda/hartree/vw/wt   closed forms (uniform density, single cosine mode),
                         finite-difference potential == dE/drho
  ion potential + forces finite-difference forces; exact structure factor limit;
                         reference sample outputs (E_ie at uniform density)
  ntn optimizer          reference sample trajectories (test/optIon, test/optDen)
  dcsrch                 Fortran reference build (oracle/_ref) step-for-step
  spline_cubic_*         Fortran reference build (oracle/_ref) bit-for-bit
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------
# Constants (Constants.f90:44-79)
# --------------------------------------------------------------------------
PI = 3.141592653589793238462643
BOHR = 0.529177208607388                       # Angstrom per bohr
_QELEC = 1.602176462e-19
_LIGHT = 299792458.0
_MU0 = 4.0 * PI * 1.0e-7
_EPS0 = 1.0 / (_MU0 * _LIGHT ** 2)
HARTREE_TO_EV = _QELEC * 1.0e10 / (4.0 * PI * _EPS0 * BOHR)   # 27.21138342791108
CTF = 2.87123400018819                          # KEDF_TF.f90:42 (literal)
MACH_PREC = float(np.nextafter(1.0, 2.0) - 1.0)  # System.f90:416-438
# single-precision default literals of RhoOptimizers.f90:55,57
POT_TOL_DEFAULT = float(np.float32(1.0e-5))
TOLE_DEFAULT = float(np.float32(1.0e-6))


# --------------------------------------------------------------------------
# Cell and G tables
# --------------------------------------------------------------------------
@dataclass
class Cell:
    """Cell geometry. ``lattice`` columns are the lattice vectors in bohr
    (ReadIonFile.f90:272 reads LATTICE_CART column-major; CellInfo.f90:281)."""
    lattice: np.ndarray
    n1: int
    n2: int
    n3: int

    def __post_init__(self):
        self.lattice = np.asarray(self.lattice, dtype=np.float64).reshape(3, 3)
        # cellRecip = TRANSPOSE(Inverse(cellReal)) * 2 pi      (CellInfo.f90:281)
        self.recip = np.linalg.inv(self.lattice).T * 2.0 * PI
        self.vol = abs(np.linalg.det(self.lattice))
        self.dV = self.vol / float(self.n1 * self.n2 * self.n3)   # CellInfo.f90:288

    @property
    def shape(self):
        return (self.n3, self.n2, self.n1)

    @property
    def nr(self):
        return self.n1 * self.n2 * self.n3


def g_index_vectors(n1, n2, n3):
    """Integer G-index vectors m (PlaneWave.f90:220-236): m1 = i1-1 never wraps;
    m2, m3 wrap to negative when > n/2 (integer division), +n/2 stays positive."""
    k1 = n1 // 2 + 1
    m1 = np.arange(k1, dtype=np.float64)
    m2 = np.arange(n2, dtype=np.float64)
    m2[m2 > n2 // 2] -= n2
    m3 = np.arange(n3, dtype=np.float64)
    m3[m3 > n3 // 2] -= n3
    return m1, m2, m3


def g_tables(cell: Cell):
    """FillQTable (PlaneWave.f90:140-302): qVectors = cellRecip . m, qTable = |q|,
    qMask = serial half-space rule (:278-291). Shapes (n3, n2, k1[,3])."""
    m1, m2, m3 = g_index_vectors(cell.n1, cell.n2, cell.n3)
    M3, M2, M1 = np.meshgrid(m3, m2, m1, indexing="ij")
    b = cell.recip
    qv = np.empty(M1.shape + (3,))
    for c in range(3):
        qv[..., c] = b[c, 0] * M1 + b[c, 1] * M2 + b[c, 2] * M3
    qt = np.sqrt(qv[..., 0] ** 2 + qv[..., 1] ** 2 + qv[..., 2] ** 2)
    return qv, qt, q_mask(cell.n1, cell.n2, cell.n3)


def q_mask(n1, n2, n3, offset=None):
    """Serial qMask rule of PlaneWave.f90:278-291 (1-based indices restated).
    offset = 1 when n1 is odd, 0 when even (Fourier_v1.f90: k1G = n1/2+1)."""
    k1G, k2G, k3G = n1 // 2 + 1, n2, n3
    if offset is None:
        offset = n1 % 2
    mask = np.zeros((n3, n2, k1G), dtype=bool)
    for i3 in range(1, k3G + 1):
        for i2 in range(1, k2G + 1):
            for i1 in range(1, k1G + 1):
                v = ((offset == 1 and i1 >= 2)
                     or (offset == 0 and 2 <= i1 < k1G)
                     or (i1 == 1 and 2 <= i2 <= (k2G - 1) // 2 + 1)
                     or (i1 == 1 and i2 == 1 and 2 <= i3 <= (k3G - 1) // 2 + 1)
                     or (i1 == 1 and k2G % 2 == 0 and i2 == k2G // 2 + 1
                         and 2 <= i3 <= (k3G - 1) // 2 + 1)
                     or (offset == 0 and i1 == k1G and 2 <= i2 <= (k2G - 1) // 2 + 1)
                     or (offset == 0 and i1 == k1G and i2 == 1
                         and 2 <= i3 <= (k3G - 1) // 2 + 1)
                     or (offset == 0 and i1 == k1G and k2G % 2 == 0
                         and i2 == k2G // 2 + 1 and 2 <= i3 <= (k3G - 1) // 2 + 1))
                mask[i3 - 1, i2 - 1, i1 - 1] = v
    return mask


# --------------------------------------------------------------------------
# FFT conventions (Fourier_v1.f90:333-465). numpy's pocketfft stands in for
# FFTW (not vendored by the reference): same transform, same normalization.
# --------------------------------------------------------------------------
# None: numpy.fft, one thread. An integer k: scipy.fft (the same pocketfft)
# with k worker threads; used only to time the oracle on several cores.
FFT_WORKERS = None


def fft_forward(a):
    """R2C, e^{-iG.r}, scaled by 1/N (Fourier_v1.f90:361,388-393)."""
    n = a.size
    if FFT_WORKERS:
        import scipy.fft
        return scipy.fft.rfftn(a, workers=FFT_WORKERS) / float(n)
    return np.fft.rfftn(a) / float(n)


def fft_backward(c, shape):
    """C2R, e^{+iG.r}, unnormalized (Fourier_v1.f90:436). The imaginary parts
    of the x = 0 and x = n1/2 columns are ignored, as in any Hermitian C2R."""
    n = shape[0] * shape[1] * shape[2]
    if FFT_WORKERS:
        import scipy.fft
        return scipy.fft.irfftn(c, s=shape, axes=(0, 1, 2), workers=FFT_WORKERS) * float(n)
    return np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) * float(n)


# --------------------------------------------------------------------------
# Energy terms
# --------------------------------------------------------------------------
def lind_g(eta, lam, mu):
    """LindG, MathFunctions.f90:440-528 (all five branches, same literals)."""
    eta = np.asarray(eta, dtype=np.float64)
    out = np.empty_like(eta)
    neg = eta < 0.0
    small = (~neg) & (eta < 1e-10)
    sing = (~neg) & (~small) & (np.abs(eta - 1.0) < 1e-10)
    big = (~neg) & (~small) & (~sing) & (eta > 3.65)
    mid = ~(neg | small | sing | big)
    out[neg] = 0.0
    e = eta[small]
    out[small] = 1.0 - lam + e ** 2 * (1.0 / 3.0 - 3.0 * mu)
    e = eta[sing]
    out[sing] = 2.0 - lam - 3.0 * mu + 20.0 * (e - 1.0)
    e = eta[big]
    e2 = e ** 2
    i2 = 1.0 / e2
    coeffs = [-0.13714285714285712, -6.39999999999999875e-2, -3.77825602968460128e-2,
              -2.51824061652633074e-2, -1.80879839616166146e-2, -1.36715733124818332e-2,
              -1.07236045520990083e-2, -8.65192783339199453e-3, -7.1372762502456763e-3,
              -5.9945117538835746e-3, -5.10997527675418131e-3, -4.41060829979912465e-3,
              -3.84763737842981233e-3, -3.38745061493813488e-3, -3.00624946457977689e-3]
    acc = np.full_like(e, coeffs[-1])
    for c in reversed(coeffs[:-1]):
        acc = c + i2 * acc
    out[big] = 3.0 * (1.0 - mu) * e2 - lam - 0.6 + i2 * acc
    e = eta[mid]
    out[mid] = (1.0 / (0.5 + 0.25 * (1.0 - e ** 2) * np.log((1.0 + e) / np.abs(1.0 - e)) / e)
                - 3.0 * mu * e ** 2 - lam)
    return out


@dataclass
class Functional:
    """KEDF/XC parameters with SetupKEDF.f90:217-322 defaults.
    kinetic 4 = WT, 5 = WGC (CellInfo kinetic codes)."""
    kinetic: int = 4
    lam: float = 1.0
    mu: float = 1.0
    alpha: float = -100.0
    beta: float = -100.0
    gamma: float = -1.0
    rho0: float = -1.0
    rhoS: float = -1.0
    wgct: int = 2

    def resolve(self, num_ele, vol):
        """SetupKEDF.f90:217-322 defaults and KEDFRefresh (:613-683)."""
        if self.kinetic == 4:
            if self.alpha < 0.0:
                self.alpha = 5.0 / 6.0
            if self.beta < 0.0:
                self.beta = self.alpha
            if self.rho0 <= 0.0:
                self.rho0 = num_ele / vol
        elif self.kinetic == 5:
            if self.alpha < 0.0:
                self.alpha = (5.0 + math.sqrt(5.0)) / 6.0
            if self.beta < 0.0:
                self.beta = (5.0 - math.sqrt(5.0)) / 6.0
            if self.gamma < 0.0:
                self.gamma = 2.7
            if self.alpha < self.beta:
                self.alpha, self.beta = self.beta, self.alpha
            if self.rho0 <= 0.0:
                self.rho0 = num_ele / vol
            if self.rhoS <= 0.0:
                self.rhoS = num_ele / vol      # mrhos = 1 (KEDF_WGCD.f90:64)
        else:
            raise ValueError("only kinetic 4 (WT) and 5 (WGC) are on this path")
        return self


def wt_kernel(qtable, f: Functional):
    """FillWT (KEDF_WTkernel.f90:81-136, fill loop :198-204)."""
    ft = 5.0 / 3.0
    coef = 5.0 / (9.0 * f.alpha * f.beta * f.rho0 ** (f.alpha + f.beta - ft))
    tkf = 2.0 * (3.0 * f.rho0 * PI ** 2) ** (1.0 / 3.0)
    return lind_g(qtable / tkf, f.lam, f.mu) * coef


def tf_potential(rho, lam):
    """TFPotential (KEDF_TF.f90:232): (5/3 cTF lambda) rho^(2/3)."""
    return (5.0 / 3.0 * CTF * lam) * rho ** (2.0 / 3.0)


def tf_energy(rho, lam):
    """TFEnergy (KEDF_TF.f90:148) without dV: lambda cTF sum rho^(5/3)."""
    return lam * CTF * np.sum(rho ** (5.0 / 3.0))


_LDA_A, _LDA_B, _LDA_C, _LDA_D = 3.11e-2, -4.8e-2, 2e-3, -1.16e-2
_LDA_G, _LDA_B1, _LDA_B2 = -1.423e-1, 1.0529, 3.334e-1


def lda_energy(rho):
    """LDAEnergy numSpin=1 (XC_LDA.f90:51-196) without dV."""
    cx = -0.73855876638202234
    cc = 0.62035049089940009
    exch = cx * np.sum(rho ** (4.0 / 3.0))
    rs = cc * rho ** (-1.0 / 3.0)
    lo = rs < 1
    ec = np.empty_like(rho)
    r = rs[lo]
    ec[lo] = np.log(r) * (_LDA_A + _LDA_C * r) + _LDA_B + _LDA_D * r
    r = rs[~lo]
    ec[~lo] = _LDA_G / (1 + _LDA_B1 * np.sqrt(r) + _LDA_B2 * r)
    return exch + np.sum(rho * ec)


def lda_potential(rho):
    """LDAPot numSpin=1 (XC_LDA.f90:199-310)."""
    cx = -0.98474502184269641
    cc = 0.62035049089940009
    cc1 = -5.83666666666666666e-2
    cc2 = 1.33333333333333333e-3
    cc3 = -8.4e-3
    cc4 = -0.174798948333333333
    cc5 = -6.325709333333333333e-2
    r3 = rho ** (1.0 / 3.0)
    rs = cc / r3
    out = np.empty_like(rho)
    lo = rs < 1
    r = rs[lo]
    out[lo] = cx * r3[lo] + np.log(r) * (_LDA_A + cc2 * r) + cc1 + cc3 * r
    r = rs[~lo]
    sq = np.sqrt(r)
    out[~lo] = cx * r3[~lo] + (_LDA_G + cc4 * sq + cc5 * r) / (1.0 + _LDA_B1 * sq + _LDA_B2 * r) ** 2
    return out


def hartree(rho, qtable):
    """JPotentialPlus (Hartree.f90:93-169): 4 pi rho(G)/G^2, G=0 removed;
    energy = sum(V_H rho)/2 (no dV)."""
    c = fft_forward(rho)
    q2 = qtable ** 2
    q2 = q2.copy()
    q2[0, 0, 0] = 1.0                       # the qTable(1,1,1)=1 trick (:133-136)
    c = (4.0 * PI) * c / q2
    c[0, 0, 0] = 0.0
    v = fft_backward(c, rho.shape)
    return v, np.sum(v * rho) / 2.0


def vw_sqrt(phi, qtable, mu):
    """VWPotentialSqrtPlus (KEDF_VW.f90:282-362): dE_vW/dphi and energy (no dV)."""
    c = fft_forward(phi)
    c = c * qtable * qtable
    t = fft_backward(c, phi.shape)
    e = 0.5 * np.sum(phi * t) * mu
    return t * mu, e


def wt_potplus(rho, kernel, f: Functional):
    """WTPotPlus (KEDF_WT.f90:70-170), vacCutoff off. Energy without dV."""
    shp = rho.shape
    p1 = fft_backward(fft_forward(rho ** f.beta) * kernel, shp)
    e = CTF * np.sum(rho ** f.alpha * p1)
    p2 = fft_backward(fft_forward(rho ** f.alpha) * kernel, shp)
    pot = CTF * (f.alpha * rho ** (f.alpha - 1.0) * p1 + f.beta * rho ** (f.beta - 1.0) * p2)
    return pot, e


def wgc_potplus(rho, kernels, f: Functional):
    """WGCPotentialPlus, second-order branch firstOrderWGC = -1 (KEDF_WGC.f90:208-286),
    vacCutoff off, WGCT = 2. kernels: (4, n3, n2, k1) = keKernel(:,:,:,1:4)."""
    shp = rho.shape
    k1, k2, k3, k4 = kernels
    a, b, rs = f.alpha, f.beta, f.rhoS
    F, B = fft_forward, fft_backward
    rhox = rho ** a
    frx = F(rhox)
    theta = rho - rs
    ftrx = F(theta * rhox)
    rhox = theta * theta * rhox / 2.0
    trafo0 = F(rhox)
    work = B(k1 * frx + k2 * ftrx + k3 * trafo0, shp)
    work2 = B(k2 * frx + k4 * ftrx, shp)
    pot = B(k3 * frx, shp)
    rhox = rho ** b
    e = CTF * np.sum(rhox * (work + theta * work2 + theta * theta / 2.0 * pot))
    pot = rho ** (b - 1.0) * (b * work + (rho + b * theta) * work2
                              + theta * (rho + b * theta / 2.0) * pot)
    frx = F(rhox)
    ftrx = F(theta * rhox)
    rhox = theta * theta * rhox / 2.0
    trafo0 = F(rhox)
    work = B(k1 * frx + k2 * ftrx + k3 * trafo0, shp)
    work2 = B(k2 * frx + k4 * ftrx, shp)
    rhox = B(k3 * frx, shp)
    pot = pot + rho ** (a - 1.0) * (a * work + (rho + a * theta) * work2
                                    + theta * (rho + a * theta / 2.0) * rhox)
    return pot * CTF, e


# --------------------------------------------------------------------------
# The aggregator: CalculatePotentialPlus + CalculateEnergy
# --------------------------------------------------------------------------
@dataclass
class System:
    """Everything CalculatePotentialPlus reads from module globals
    (CalPotential.f90:41-116; SURVEY boundary list)."""
    cell: Cell
    func: Functional
    vion: np.ndarray                 # ionPotReal (real grid)
    e_ionion: float = 0.0            # Ewald ionIonEnergy (host constant)
    wgc_kernels: np.ndarray | None = None
    qtable: np.ndarray = field(init=False)
    kernel: np.ndarray | None = field(init=False, default=None)

    def __post_init__(self):
        _, self.qtable, _ = g_tables(self.cell)
        if self.func.kinetic == 4:
            self.kernel = wt_kernel(self.qtable, self.func)


def calculate_potential_plus(rho, sysm: System, opt_sqrt=True, want_energy=True):
    """CalculatePotentialPlus (CalPotential.f90:41-607) for numSpin = 1,
    XC = LDA, kinetic 4 (WT, :393-413) or 5 (WGC, :418-435).
    Returns (potential, eTable(9) in Hartree or None)."""
    f = sysm.func
    et = np.zeros(9)
    # (1) ionic potential (:174-206) and IonElectronEnergyReal (IonElectron.f90:155)
    pot = sysm.vion.copy()
    et[2] = np.sum(rho * pot)
    # (2) Hartree (:208-231)
    vh, et[3] = hartree(rho, sysm.qtable)
    pot = pot + vh
    # (3) XC (:233-300)
    pot = pot + lda_potential(rho)
    et[4] = lda_energy(rho)
    # (4) KEDF (:393-435): TF, nonlocal, chain rule, vW
    pot = pot + tf_potential(rho, f.lam)
    et[6] = tf_energy(rho, f.lam)
    if f.kinetic == 4:
        pnl, et[8] = wt_potplus(rho, sysm.kernel, f)
    elif f.kinetic == 5:
        pnl, et[8] = wgc_potplus(rho, sysm.wgc_kernels, f)
    else:
        raise ValueError("kinetic must be 4 or 5")
    pot = pot + pnl
    phi = np.sqrt(rho)
    if opt_sqrt:
        pot = 2.0 * phi * pot
    tvw, et[7] = vw_sqrt(phi, sysm.qtable, f.mu)
    if not opt_sqrt:
        tvw = tvw / (2.0 * phi)
    pot = pot + tvw
    if not want_energy:
        return pot, None
    # CalculateEnergy (:610-644)
    et[1] = et[6] + et[7] + et[8]
    et = et * sysm.cell.dV
    et[5] = sysm.e_ionion
    et[0] = np.sum(et[1:6])
    return pot, et


# --------------------------------------------------------------------------
# Spin-polarized evaluation, numSpin = 2. Arrays of both channels have numpy
# shape (2, n3, n2, n1) == Fortran rho(n1,n2,n3,2): spin up, then spin down.
# --------------------------------------------------------------------------
LSDA_AX = -0.7385588              # ax of LSDAPotPZ/PW92 (XC_LDA.f90:330, :434)
LSDA_FZZ = 1.709921               # f''(0)
LSDA_GAMMA = 0.5198421            # 2^(4/3) - 2
# Perdew-Zunger 1981 parameters in Rydberg (XC_LDA.f90:453-467): unpolarized
# (c1..c4 for rs < 1; g, b1, b2 for rs >= 1), fully polarized (as..ds; gs, b1s, b2s)
_PZR = dict(c1=0.0622, c2=0.0960, c3=0.0040, c4=0.0232, as_=0.0311, bs=-0.0538, cs=0.0014, ds=-0.0096,
            g=-0.2846, b1=1.0529, b2=0.3334, gs=-0.1686, b1s=1.3981, b2s=0.2611)
# Perdew-Wang 1992 G(rs) parameter rows (a, a1, b1, b2, b3, b4) as the reference
# passes them to spn_gcor (XC_LDA.f90:363-368): -ec unpolarized, -ec polarized,
# alpha_c (spin stiffness, as -alfm)
PW92_U = (0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294)
PW92_P = (0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517)
PW92_A = (0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671)


def spn_gcor(a, a1, b1, b2, b3, b4, p, rs):
    """spn_gcor (XC_LDA.f90:550-592): G(rs) = -2a(1 + a1 rs) ln(1 + 1/(2a(b1 rs^1/2
    + b2 rs + b3 rs^3/2 + b4 rs^(p+1)))) and dG/drs."""
    p1 = p + 1.0
    q0 = -2.0 * a * (1.0 + a1 * rs)
    rs12 = np.sqrt(rs)
    rs32 = rs12 ** 3
    rsp = rs ** p
    q1 = 2.0 * a * (b1 * rs12 + b2 * rs + b3 * rs32 + b4 * rs * rsp)
    q2 = np.log(1.0 + 1.0 / q1)
    gg = q0 * q2
    q3 = a * (b1 / rs12 + 2.0 * b2 + 3.0 * b3 * rs12 + 2.0 * b4 * p1 * rsp)
    ggrs = -2.0 * a * a1 * q2 - q0 * q3 / (q1 ** 2 + q1)
    return gg, ggrs


def _spin_common(rho2):
    """Exchange per channel (on 2 rho_s), rs, zeta and the spin interpolation f,
    f' shared by LSDAPotPZ and LSDAPotPW92 (XC_LDA.f90:361-376, :493-503)."""
    third, ac2 = 1.0 / 3.0, 4.0 / 3.0
    ex = LSDA_AX * (2.0 * rho2) ** third
    vx = ex * ac2
    d = rho2[0] + rho2[1]
    rs = (0.75 / PI / d) ** third
    zet = (rho2[0] - rho2[1]) / d
    f = ((1.0 + zet) ** ac2 + (1.0 - zet) ** ac2 - 2.0) / LSDA_GAMMA
    fz = ac2 * ((1.0 + zet) ** third - (1.0 - zet) ** third) / LSDA_GAMMA
    return ex, vx, d, rs, zet, f, fz


def lsda_pz(rho2):
    """LSDAPotPZ (XC_LDA.f90:417-547), the reference's default LSDA (lsda = 1):
    Perdew-Zunger correlation in Rydberg, halved. Returns (potential (2, ...),
    energy without dV)."""
    ex, vx, d, rs, zet, f, fz = _spin_common(rho2)
    P = _PZR
    eu = np.empty_like(rs)
    vc1 = np.empty_like(rs)
    es = np.empty_like(rs)
    vs = np.empty_like(rs)
    lo = rs < 1.0
    r = rs[lo]
    lrs = np.log(r)
    eu[lo] = P["c1"] * lrs - P["c2"] + (P["c3"] * lrs - P["c4"]) * r
    vc1[lo] = eu[lo] - (P["c1"] + (P["c3"] * lrs + P["c3"] - P["c4"]) * r) / 3.0
    es[lo] = P["as_"] * lrs + P["bs"] + (P["cs"] * lrs + P["ds"]) * r
    vs[lo] = es[lo] - (P["as_"] + (P["cs"] * lrs + P["cs"] + P["ds"]) * r) / 3.0
    r = rs[~lo]
    srs = np.sqrt(r)
    eu[~lo] = P["g"] / (1.0 + P["b1"] * srs + P["b2"] * r)
    vc1[~lo] = eu[~lo] * eu[~lo] * (1.0 + 7.0 * P["b1"] * srs / 6.0 + 4.0 / 3.0 * P["b2"] * r) / P["g"]
    es[~lo] = P["gs"] / (1.0 + P["b1s"] * srs + P["b2s"] * r)
    vs[~lo] = es[~lo] * es[~lo] * (1.0 + 7.0 * P["b1s"] * srs / 6.0 + 4.0 / 3.0 * P["b2s"] * r) / P["gs"]
    vc2 = vc1 + f * (vs - vc1) - (es - eu) * (1.0 + zet) * fz
    vc1 = vc1 + f * (vs - vc1) + (es - eu) * (1.0 - zet) * fz
    ec = eu + f * (es - eu)
    pot = np.stack([0.5 * (2.0 * vx[0] + vc1), 0.5 * (2.0 * vx[1] + vc2)])
    e = 0.5 * np.sum(2.0 * (rho2[0] * ex[0] + rho2[1] * ex[1]) + ec * d)
    return pot, e


def lsda_pw92(rho2):
    """LSDAPotPW92 (XC_LDA.f90:313-414), lsda = 2: Perdew-Wang 1992 correlation.
    Returns (potential (2, ...), energy without dV)."""
    ex, vx, d, rs, zet, f, fz = _spin_common(rho2)
    eu, eurs = spn_gcor(*PW92_U, 1.0, rs)
    ep, eprs = spn_gcor(*PW92_P, 1.0, rs)
    alfm, alfrsm = spn_gcor(*PW92_A, 1.0, rs)
    z4 = zet ** 4
    ec = eu * (1.0 - f * z4) + ep * f * z4 - alfm * f * (1.0 - z4) / LSDA_FZZ
    ecrs = eurs * (1.0 - f * z4) + eprs * f * z4 - alfrsm * f * (1.0 - z4) / LSDA_FZZ
    eczet = 4.0 * (zet ** 3) * f * (ep - eu + alfm / LSDA_FZZ) + fz * (z4 * ep - z4 * eu - (1.0 - z4) * alfm / LSDA_FZZ)
    comm = ec - rs * ecrs / 3.0 - zet * eczet
    pot = np.stack([vx[0] + comm + eczet, vx[1] + comm - eczet])
    e = np.sum(rho2[0] * ex[0] + rho2[1] * ex[1] + ec * d)
    return pot, e


def lda_energy_spin(rho2):
    """LDAEnergy for numSpin = 2 (XC_LDA.f90:152-185): the PW92 energy with the
    exchange constant cX of the unpolarized branch. Without dV."""
    cx = -0.73855876638202234
    third = 1.0 / 3.0
    exa = cx * (2.0 * rho2[0]) ** third
    exb = cx * (2.0 * rho2[1]) ** third
    d = rho2[0] + rho2[1]
    rs = (0.75 / (PI * d)) ** third
    eu, _ = spn_gcor(*PW92_U, 1.0, rs)
    ep, _ = spn_gcor(*PW92_P, 1.0, rs)
    alfm, _ = spn_gcor(*PW92_A, 1.0, rs)
    zet = (rho2[0] - rho2[1]) / d
    f = ((1.0 + zet) ** (4.0 / 3.0) + (1.0 - zet) ** (4.0 / 3.0) - 2.0) / LSDA_GAMMA
    z4 = zet ** 4
    ec = eu * (1.0 - f * z4) + ep * f * z4 - alfm * f * (1.0 - z4) / LSDA_FZZ
    return np.sum(rho2[0] * exa + rho2[1] * exb + ec * d)


def calculate_potential_plus_spin(rho2, sysm: System, opt_sqrt=True, want_energy=True, lsda=1):
    """CalculatePotentialPlus for numSpin = 2 (CalPotential.f90:120-160):
    ionic and Hartree terms on rho_up + rho_down, copied to both channels; LSDA
    exchange-correlation (:262-282, lsda 1 = PZ, 2 = PW92); the kinetic terms
    evaluated on 2 rho_s and halved (CalTF, KEDF_TF.f90:60-78; WT/WGC,
    CalPotential.f90:400-407, :426-431); the chain rule to sqrt(rho_s); spin vW
    on sqrt(2 rho_s) divided by sqrt(2) (CalVW, KEDF_VW.f90:100-140).
    Returns (potential (2, ...), eTable(9) in Hartree or None)."""
    f = sysm.func
    et = np.zeros(9)
    rho = rho2[0] + rho2[1]
    common = sysm.vion.copy()
    et[2] = np.sum(rho * common)
    vh, et[3] = hartree(rho, sysm.qtable)
    common = common + vh
    pot = np.stack([common, common])
    vxc, et[4] = (lsda_pz if lsda == 1 else lsda_pw92)(rho2)
    pot = pot + vxc
    for s in range(2):
        x = 2.0 * rho2[s]
        pot[s] = pot[s] + tf_potential(x, f.lam)
        et[6] += 0.5 * tf_energy(x, f.lam)
    for s in range(2):
        x = 2.0 * rho2[s]
        if f.kinetic == 4:
            pnl, e = wt_potplus(x, sysm.kernel, f)
        elif f.kinetic == 5:
            pnl, e = wgc_potplus(x, sysm.wgc_kernels, f)
        else:
            raise ValueError("kinetic must be 4 or 5")
        pot[s] = pot[s] + pnl
        et[8] += 0.5 * e
    if opt_sqrt:
        pot = 2.0 * np.sqrt(rho2) * pot
    for s in range(2):
        tvw, e = vw_sqrt(np.sqrt(2.0 * rho2[s]), sysm.qtable, f.mu)
        tvw = tvw / math.sqrt(2.0)
        if not opt_sqrt:
            tvw = tvw / (2.0 * np.sqrt(rho2[s]))
        pot[s] = pot[s] + tvw
        et[7] += e * 0.5
    if not want_energy:
        return pot, None
    et[1] = et[6] + et[7] + et[8]
    et = et * sysm.cell.dV
    et[5] = sysm.e_ionion
    et[0] = np.sum(et[1:6])
    return pot, et


def spin_split(rho, magmom, vol):
    """Initial spin densities from a total density and the input MAGMOM
    (System.f90:326-329): rho_s = (rho - (-1)^s magmom / vol) / 2, s = 1, 2."""
    return np.stack([0.5 * (rho - (-1) ** s * magmom / vol) for s in (1, 2)])


# --------------------------------------------------------------------------
# Cubic splines (MathSplines.f90:178-560)
# --------------------------------------------------------------------------
def spline_cubic_set(t, y, ibcbeg, ybcbeg, ibcend, ybcend):
    """spline_cubic_set (MathSplines.f90:178-330) incl. s3_fs tridiagonal solve."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = t.size
    ypp = np.zeros(n)
    diag = np.zeros(n)
    sub = np.zeros(n)
    sup = np.zeros(n)
    if ibcbeg == 0:
        ypp[0], diag[0], sup[0] = 0.0, 1.0, -1.0
    elif ibcbeg == 1:
        ypp[0] = (y[1] - y[0]) / (t[1] - t[0]) - ybcbeg
        diag[0] = (t[1] - t[0]) / 3.0
        sup[0] = (t[1] - t[0]) / 6.0
    else:
        ypp[0], diag[0], sup[0] = ybcbeg, 1.0, 0.0
    ypp[1:n - 1] = (y[2:] - y[1:n - 1]) / (t[2:] - t[1:n - 1]) - (y[1:n - 1] - y[:n - 2]) / (t[1:n - 1] - t[:n - 2])
    sub[1:n - 1] = (t[1:n - 1] - t[:n - 2]) / 6.0
    diag[1:n - 1] = (t[2:] - t[:n - 2]) / 3.0
    sup[1:n - 1] = (t[2:] - t[1:n - 1]) / 6.0
    if ibcend == 0:
        ypp[n - 1], sub[n - 1], diag[n - 1] = 0.0, -1.0, 1.0
    elif ibcend == 1:
        ypp[n - 1] = ybcend - (y[n - 1] - y[n - 2]) / (t[n - 1] - t[n - 2])
        sub[n - 1] = (t[n - 1] - t[n - 2]) / 6.0
        diag[n - 1] = (t[n - 1] - t[n - 2]) / 3.0
    else:
        ypp[n - 1], sub[n - 1], diag[n - 1] = ybcend, 0.0, 1.0
    if n == 2 and ibcbeg == 0 and ibcend == 0:
        return np.zeros(2)
    # s3_fs (:602-660): forward elimination, back substitution (sequential)
    a2 = diag.tolist()
    b = ypp.tolist()
    a1 = sub.tolist()
    a3 = sup.tolist()
    for i in range(1, n - 1):
        xm = a1[i] / a2[i - 1]
        a2[i] = a2[i] - xm * a3[i - 1]
        b[i] = b[i] - xm * b[i - 1]
    xm = a1[n - 1] / a2[n - 2]
    a2[n - 1] = a2[n - 1] - xm * a3[n - 2]
    x = [0.0] * n
    x[n - 1] = (b[n - 1] - xm * b[n - 2]) / a2[n - 1]
    for i in range(n - 2, -1, -1):
        x[i] = (b[i] - a3[i] * x[i + 1]) / a2[i]
    return np.array(x)


def rvec_bracket(t, tval):
    """rvec_bracket (MathSplines.f90:498-545), 0-based (left, right)."""
    n = len(t)
    i1, i3 = 1, n
    i2 = (i3 + i1) // 2
    while i2 != i1:
        x2 = t[i2 - 1]
        if tval == x2:
            return i2 - 1, i2
        elif tval < x2:
            i3 = i2
        elif tval > x2:
            i1 = i2
        i2 = (i3 + i1) // 2
    if tval == t[i3 - 1]:
        return i3 - 1, i3
    if tval == t[i1 - 1]:
        return i1 - 1, i1
    return i1 - 1, i3 - 1


def spline_cubic_val(t, y, ypp, tval):
    """spline_cubic_val (MathSplines.f90:395-470): returns yval only."""
    left, right = rvec_bracket(t, tval)
    dt = tval - t[left]
    h = t[right] - t[left]
    tmp1 = (y[right] - y[left]) / h
    tmp2 = -(ypp[right] / 6.0 + ypp[left] / 3.0) * h
    tmp3 = ypp[right] - ypp[left]
    return y[left] + dt * (tmp1 + tmp2 + dt * (0.5 * ypp[left] + dt * (tmp3 / (6.0 * h))))


def spline_cubic_val_vec(t, y, ypp, tvals):
    """Vectorized spline_cubic_val with the same bracket as rvec_bracket:
    left = last knot with t[left] <= tval (clamped to [0, n-2])."""
    t = np.asarray(t)
    tv = np.asarray(tvals, dtype=np.float64)
    left = np.searchsorted(t, tv, side="right") - 1
    left = np.clip(left, 0, len(t) - 2)
    right = left + 1
    dt = tv - t[left]
    h = t[right] - t[left]
    tmp1 = (y[right] - y[left]) / h
    tmp2 = -(ypp[right] / 6.0 + ypp[left] / 3.0) * h
    tmp3 = ypp[right] - ypp[left]
    return y[left] + dt * (tmp1 + tmp2 + dt * (0.5 * ypp[left] + dt * (tmp3 / (6.0 * h))))


# --------------------------------------------------------------------------
# Pseudopotential (.recpot semantics of ReadPseudo, ReadIonFile.f90:708-1051;
# lookup of LocalPseudoPot.f90:40-204)
# --------------------------------------------------------------------------
@dataclass
class Psp:
    maxG: float            # bohr^-1
    potValues: np.ndarray  # Ha bohr^3
    charge: int
    vqS: np.ndarray
    potDD: np.ndarray

    @property
    def numPoints(self):
        return self.potValues.size


def read_recpot_text(text):
    """Parse a .recpot file as ReadPseudo does (ReadIonFile.f90:830-870):
    skip to END COMMENT, one line, maxG (A^-1), then 3 values per line until
    the '1000' terminator. Returns (maxG_invA, values_eVA3)."""
    lines = text.splitlines()
    i = 0
    while "END COMMENT" not in lines[i]:
        i += 1
    i += 2
    maxg = float(lines[i].split()[0])
    vals = []
    for ln in lines[i + 1:]:
        toks = ln.split()
        if not toks:
            continue
        if toks[0] == "1000":
            break
        vals.extend(float(x) for x in toks)
    return maxg, np.array(vals)


def make_psp(maxg_inv_angstrom, values_ev_a3):
    """ReadPseudo units and spline set-up (ReadIonFile.f90:848-976)."""
    maxG = maxg_inv_angstrom * BOHR
    pv = np.asarray(values_ev_a3, dtype=np.float64) / (BOHR ** 3 * HARTREE_TO_EV)
    n = pv.size
    charge = int(round(-(pv[1] - pv[0]) * (maxG / float(n - 1)) ** 2 / (4.0 * PI)))
    t = np.arange(1, n + 1, dtype=np.float64)
    vqs = np.empty(n)
    vqs[0] = pv[0]
    vqs[1:] = pv[1:] + 4.0 * PI * charge / ((t[1:] - 1.0) / float(n - 1) * maxG) ** 2.0
    pdd = spline_cubic_set(t, vqs, 1, 0.0, 1, 0.0)
    return Psp(maxG, pv, charge, vqs, pdd)


def synthetic_recpot_values(Z, rc, maxg_inv_bohr=30.0, dq=0.002):
    """Synthetic V_loc(q) = -4 pi Z/q^2 exp(-q^2 rc^2/4) as recpot values
    (eV A^3 over a uniform q grid, q=0 entry = pi Z rc^2)."""
    npts = int(round(maxg_inv_bohr / dq)) + 1
    q = np.linspace(0.0, maxg_inv_bohr, npts)
    v = np.empty(npts)
    v[0] = PI * Z * rc * rc
    v[1:] = -4.0 * PI * Z / q[1:] ** 2 * np.exp(-q[1:] ** 2 * rc * rc / 4.0)
    return maxg_inv_bohr / BOHR, v * (BOHR ** 3 * HARTREE_TO_EV)


def psp_lookup(psp: Psp, q):
    """PspRecipCubicSpline (LocalPseudoPot.f90:136-198), vectorized."""
    q = np.asarray(q, dtype=np.float64)
    out = np.zeros_like(q)
    spacing = psp.maxG / float(psp.numPoints - 1)
    zero = q == 0.0
    inside = (q > 0.0) & (q < psp.maxG)
    out[zero] = psp.potValues[0]
    qq = q[inside]
    pos = qq / spacing + 1.0
    left = np.floor(pos).astype(np.int64)
    dt = pos - left
    yl = psp.vqS[left - 1]
    yr = psp.vqS[left]
    pl = psp.potDD[left - 1]
    pr = psp.potDD[left]
    yval = yl + dt * (yr - yl - (pr / 6.0 + pl / 3.0) + dt * (0.5 * pl + dt * ((pr - pl) / 6.0)))
    out[inside] = yval - 4.0 * PI * float(psp.charge) / qq ** 2.0
    return out


# --------------------------------------------------------------------------
# Cardinal B-spline ion potential and forces (CBSpline.f90, IonElectronSpline.f90)
# --------------------------------------------------------------------------
def cardinal_bspline(order, x):
    """GetCardinalBSpline (CBSpline.f90:42-90): array index 0..order."""
    s = [0.0] * (order + 1)
    s[1] = x
    s[2] = 1.0 - x
    for n in range(3, order + 1):
        for k in range(0, n):
            s[n - k] = ((x + float(n - k - 1)) * s[n - k] + (float(k + 1) - x) * s[n - k - 1]) / float(n - 1)
    return s


def fill_b(order, n1, n2, n3):
    """FillB (CBSpline.f90:93-196) called as FillB(b1,b2,b3,order,m1G,m3G,0)
    (SetupKEDF.f90:186): b1 has k1G entries over n1, b2 n2 entries, b3 n3."""
    M = cardinal_bspline(order, 1.0)

    def one(size, tot):
        b = np.empty(size, dtype=np.complex128)
        for ix in range(size):
            m = float(ix)
            den = 0j
            for k in range(order - 1):
                den += M[k + 1] * np.exp(2.0 * PI * 1j * m * float(k) / float(tot))
            b[ix] = np.exp(2.0 * PI * 1j * float(order - 1) * m / float(tot)) / den
        return b
    return one(n1 // 2 + 1, n1), one(n2, n2), one(n3, n3)


def fill_qion_table(frac, shape, order):
    """FillQIonTable (IonElectronSpline.f90:537-638) for one species."""
    n3, n2, n1 = shape
    q = np.zeros(shape)
    for s in np.asarray(frac, dtype=np.float64).reshape(-1, 3):
        g = [s[0] * float(n1), s[1] * float(n2), s[2] * float(n3)]
        fl = [math.floor(v) for v in g]
        M = [cardinal_bspline(order, g[d] - fl[d]) for d in range(3)]
        i = np.arange(order)
        l1 = np.mod(-fl[0] + i, n1)
        l2 = np.mod(-fl[1] + i, n2)
        l3 = np.mod(-fl[2] + i, n3)
        w = (np.array(M[2][1:order + 1])[:, None, None] * np.array(M[1][1:order + 1])[None, :, None]
             * np.array(M[0][1:order + 1])[None, None, :])
        np.add.at(q, (l3[:, None, None], l2[None, :, None], l1[None, None, :]), w)
    return q


def ion_potential(cell: Cell, species, order=10):
    """RefreshIonTerms spline path: IonElectronPotRecipSpline
    (IonElectronSpline.f90:83-216) then C2R (RefreshIons.f90:221).
    species: list of (Psp, frac (nion,3)). Returns ionPotReal."""
    qv, qt, _ = g_tables(cell)
    acc = np.zeros(qt.shape, dtype=np.complex128)
    for psp, frac in species:
        sf = fft_forward(fill_qion_table(frac, cell.shape, order))
        acc = acc + psp_lookup(psp, qt) * sf
    b1, b2, b3 = fill_b(order, cell.n1, cell.n2, cell.n3)
    tmp = acc * b1[None, None, :] * b2[None, :, None] * b3[:, None, None]
    rec = np.conj(tmp) / cell.vol * float(cell.nr)
    return fft_backward(rec, cell.shape)


def ion_forces(cell: Cell, species, rho, order=10):
    """IonElectronForcesSpline (IonElectronSpline.f90:219-355) with
    CalculateSplineForces (:642-753). rhoRecip = FFT(rho) (CalForce.f90:180).
    Returns forces (nion_total, 3) in Ha/bohr, species order."""
    _, qt, _ = g_tables(cell)
    rr = fft_forward(rho)
    b1, b2, b3 = fill_b(order, cell.n1, cell.n2, cell.n3)
    out = []
    for psp, frac in species:
        prec = psp_lookup(psp, qt)
        prec[0, 0, 0] = 0.0
        tmp = rr * prec * b1[None, None, :] * b2[None, :, None] * b3[:, None, None]
        factor = fft_backward(tmp, cell.shape)
        out.extend(spline_force_gather(cell, frac, factor, order))
    return np.array(out)


def spline_force_gather(cell: Cell, frac, factor, order=10):
    """CalculateSplineForces (IonElectronSpline.f90:642-753): per ion,
    -cellInv . sum over the order^3 stencil of the B-spline derivative
    weights times factor (a real grid). Returns (nion, 3)."""
    cell_inv = np.linalg.inv(cell.lattice).T
    n3, n2, n1 = cell.shape
    out = []
    for s in np.asarray(frac, dtype=np.float64).reshape(-1, 3):
        g = [s[0] * n1, s[1] * n2, s[2] * n3]
        fl = [math.floor(v) for v in g]
        M = [np.array(cardinal_bspline(order, g[d] - fl[d])) for d in range(3)]
        D = [np.array(cardinal_bspline(order - 1, g[d] - fl[d]) + [0.0]) for d in range(3)]
        i = np.arange(order)
        l1 = np.mod(fl[0] - i, n1)
        l2 = np.mod(fl[1] - i, n2)
        l3 = np.mod(fl[2] - i, n3)
        fac = factor[l3[:, None, None], l2[None, :, None], l1[None, None, :]]
        m1, m2, m3 = M[0][1:order + 1], M[1][1:order + 1], M[2][1:order + 1]
        d1 = D[0][1:order + 1] - D[0][0:order]
        d2 = D[1][1:order + 1] - D[1][0:order]
        d3 = D[2][1:order + 1] - D[2][0:order]
        q1 = (m2[None, :, None] * m3[:, None, None] * float(n1) * d1[None, None, :])
        q2 = (m1[None, None, :] * m3[:, None, None] * float(n2) * d2[None, :, None])
        q3 = (m1[None, None, :] * m2[None, :, None] * float(n3) * d3[:, None, None])
        qd = np.array([np.sum(q1 * fac), np.sum(q2 * fac), np.sum(q3 * fac)])
        out.append(-(cell_inv @ qd))
    return out


def ion_energy_exact_recip(cell: Cell, species):
    """Exact structure-factor ion potential in G space (IonElectron.f90:162-292
    restated): V(G) = sum_types V_loc(|G|) sum_ions e^{-iG.R} / vol.
    Used only to pin the B-spline approximation as order grows."""
    qv, qt, _ = g_tables(cell)
    acc = np.zeros(qt.shape, dtype=np.complex128)
    for psp, frac in species:
        frac = np.asarray(frac, dtype=np.float64).reshape(-1, 3)
        m1, m2, m3 = g_index_vectors(cell.n1, cell.n2, cell.n3)
        sf = np.zeros(qt.shape, dtype=np.complex128)
        for s in frac:
            ph = (np.exp(-2j * PI * m1 * s[0])[None, None, :] * np.exp(-2j * PI * m2 * s[1])[None, :, None]
                  * np.exp(-2j * PI * m3 * s[2])[:, None, None])
            sf += ph
        acc += psp_lookup(psp, qt) * sf
    return acc / cell.vol


# --------------------------------------------------------------------------
# Ion-ion Ewald sum (Ewald.f90)
# --------------------------------------------------------------------------
EWALD_ERROR_TOL = 1.0e-10 / HARTREE_TO_EV   # errorTolerance, Ewald.f90:112
EWALD_ETA_STEP = 0.01 * BOHR                # etaIncrement, Ewald.f90:121
EWALD_RCUT_STEP = 0.01 * BOHR               # realCutoffIncrement, Ewald.f90:128


def _anint(x):
    """Fortran ANINT: round half away from zero."""
    return np.where(x >= 0.0, np.floor(x + 0.5), -np.floor(-x + 0.5))


@dataclass
class EwaldParams:
    eta: float
    real_cutoff: float
    recip_cutoff: float
    real_error: float
    recip_error: float


def ewald_setup(cell: Cell, charges):
    """EwaldSetup (Ewald.f90:208-460) with the spline (particle-mesh) grid equal
    to the density grid: recipCutoff from the grid (:409-411), eta lowered from
    1.2 until the reciprocal error bound is met (:419-427), realCutoff raised
    from 6 until the real-space bound is met (:431-439)."""
    q = np.asarray(charges, dtype=np.float64)
    nion = q.size
    avg_q2 = float(np.sum(q * q)) / nion
    a = cell.lattice
    norms = [math.sqrt(float(a[0, j] ** 2 + a[1, j] ** 2 + a[2, j] ** 2)) for j in range(3)]
    kc = min((cell.n1 - 1) / norms[0], (cell.n2 - 1) / norms[1], (cell.n3 - 1) / norms[2]) * PI
    sqrt_pi = math.sqrt(PI)
    eta = 1.2
    while True:
        recip_err = float(nion) ** 2 * avg_q2 / sqrt_pi * eta * math.erfc(kc / (2.0 * eta))
        if recip_err < EWALD_ERROR_TOL:
            break
        eta -= EWALD_ETA_STEP
    rc = 6.0
    while True:
        real_err = PI * float(nion) ** 2 * avg_q2 / (cell.vol * eta ** 2) * math.erfc(rc * eta)
        if real_err < EWALD_ERROR_TOL:
            break
        rc += EWALD_RCUT_STEP
    return EwaldParams(eta, rc, kc, real_err, recip_err)


def _ewald_images(cell: Cell, fr, rc):
    """The shell loop of EwaldRealSpace (Ewald.f90:681-760): image offsets
    (i_a, i_b, i_c) of one wrapped fractional separation fr, shell nr = 1, 2,
    ... until a shell adds no image within the cutoff. Yields (r (3,) frac,
    dist^2) of every image with 1e-40 < dist^2 <= rc^2."""
    rmet = cell.lattice.T @ cell.lattice   # Metric3d (MathFunctions.f90:745-749)
    rc2 = rc * rc
    nr = 1
    while True:
        newr = False
        rng = np.arange(-nr, nr + 1)
        A, B, Cc = np.meshgrid(rng, rng, rng, indexing="ij")
        A, B, Cc = A.ravel(), B.ravel(), Cc.ravel()
        if nr > 1:
            on = (np.abs(A) == nr) | (np.abs(B) == nr) | (np.abs(Cc) == nr)
            A, B, Cc = A[on], B[on], Cc[on]
        r = np.stack([A + fr[0], B + fr[1], Cc + fr[2]], axis=1)
        d2 = (rmet[0, 0] * r[:, 0] ** 2 + rmet[1, 1] * r[:, 1] ** 2 + rmet[2, 2] * r[:, 2] ** 2
              + 2.0 * (rmet[1, 0] * r[:, 1] * r[:, 0] + rmet[2, 1] * r[:, 2] * r[:, 1] + rmet[2, 0] * r[:, 2] * r[:, 0]))
        keep = (d2 <= rc2) & (d2 > 1e-40)
        if np.any(keep):
            newr = True
            yield r[keep], d2[keep]
        if not newr:
            return
        nr += 1


def ewald_real_space(cell: Cell, charges, frac, eta, rc):
    """EwaldRealSpace (Ewald.f90:599-775): 1/2 sum_ij q_i q_j erfc(eta d)/d."""
    q = np.asarray(charges, dtype=np.float64)
    fr_all = np.asarray(frac, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for i in range(q.size):
        for j in range(q.size):
            fr = fr_all[i] - fr_all[j]
            fr = fr - _anint(fr)
            for _, d2 in _ewald_images(cell, fr, rc):
                d = np.sqrt(d2)
                e += q[i] * q[j] * float(np.sum(np.array([math.erfc(eta * x) for x in d]) / d))
    return e / 2.0


def ewald_recip_spline(cell: Cell, charges, frac, eta, kc, order=10):
    """EwaldRecipSpaceSpline (Ewald.f90:993-1147): charge-weighted B-spline
    table, R2C (1/N), b-spline moduli, |.|^2, half-space sum (x = 0 plane
    once, every other plane twice) of exp(-G^2/4eta^2)/G^2 for 0 < G^2 <=
    kc^2, times 2 pi / vol * N^2."""
    q = np.asarray(charges, dtype=np.float64)
    fr = np.asarray(frac, dtype=np.float64).reshape(-1, 3)
    table = np.zeros(cell.shape)
    for z in np.unique(q):
        table += fill_qion_table(fr[q == z], cell.shape, order) * z
    rec = fft_forward(table)
    b1, b2, b3 = fill_b(order, cell.n1, cell.n2, cell.n3)
    cc = rec * b1[None, None, :] * b2[None, :, None] * b3[:, None, None]
    s2 = (cc * np.conj(cc)).real
    _, qt, _ = g_tables(cell)
    m2 = qt * qt
    ok = (m2 <= kc * kc) & (m2 > 0.0)
    w = np.full(qt.shape, 2.0)
    w[:, :, 0] = 1.0
    term = np.zeros(qt.shape)
    term[ok] = w[ok] * np.exp(-m2[ok] / eta ** 2 / 4.0) / m2[ok] * s2[ok]
    return float(np.sum(term)) * 2.0 * PI / cell.vol * float(cell.nr) ** 2


def ewald_recip_exact(cell: Cell, charges, frac, eta, kc):
    """EwaldRecipSpace (Ewald.f90:815-990): the exact structure factor over
    the half space of lattice vectors with |G| <= kc, times 4 pi / vol."""
    q = np.asarray(charges, dtype=np.float64)
    fr = np.asarray(frac, dtype=np.float64).reshape(-1, 3)
    a = cell.lattice
    nmax = [int(math.ceil(kc * math.sqrt(float(a[0, j] ** 2 + a[1, j] ** 2 + a[2, j] ** 2)) / PI / 2)) for j in range(3)]
    e = 0.0
    for ia in range(0, nmax[0] + 1):
        for ib in range(0 if ia == 0 else -nmax[1], nmax[1] + 1):
            for ic in range(1 if (ia == 0 and ib == 0) else -nmax[2], nmax[2] + 1):
                m = ia * cell.recip[:, 0] + ib * cell.recip[:, 1] + ic * cell.recip[:, 2]
                if math.sqrt(float(m @ m)) > kc:
                    continue
                ph = 2.0 * PI * (fr @ np.array([ia, ib, ic], dtype=np.float64))
                ss = float(np.sum(q * np.sin(ph)))
                sc = float(np.sum(q * np.cos(ph)))
                m2 = float(m @ m)
                e += math.exp(-(m2 / eta ** 2 / 4.0)) / m2 * (ss * ss + sc * sc)
    return e * 4.0 * PI / cell.vol


def ewald_energy(cell: Cell, charges, frac, order=10, spline=True, params: EwaldParams = None):
    """EwaldEnergy (Ewald.f90:465-596): real space + self term + reciprocal
    (particle-mesh spline by default) + average-potential term, in Hartree.
    Returns (total, [real, self, recip, average], params)."""
    p = params if params is not None else ewald_setup(cell, charges)
    q = np.asarray(charges, dtype=np.float64)
    real = ewald_real_space(cell, q, frac, p.eta, p.real_cutoff)
    self_term = -p.eta / math.sqrt(PI) * float(np.sum(q * q))              # :778-812
    recip = (ewald_recip_spline(cell, q, frac, p.eta, p.recip_cutoff, order) if spline
             else ewald_recip_exact(cell, q, frac, p.eta, p.recip_cutoff))
    avg = -0.5 * PI / p.eta ** 2 / cell.vol * float(np.sum(q)) ** 2        # :1150-1190
    comps = [real, self_term, recip, avg]
    return float(sum(comps)), comps, p


def ewald_forces(cell: Cell, charges, frac, order=10, params: EwaldParams = None):
    """EwaldForces (Ewald.f90:1196-1310) with EwaldRealSpaceForces (:1313-1390)
    and EwaldRecipSpaceForcesSpline (:1649-1851). Returns (nion, 3) Ha/bohr."""
    p = params if params is not None else ewald_setup(cell, charges)
    q = np.asarray(charges, dtype=np.float64)
    fr_all = np.asarray(frac, dtype=np.float64).reshape(-1, 3)
    eta = p.eta
    sqrt_pi = math.sqrt(PI)
    real = np.zeros((q.size, 3))
    for i in range(q.size):
        acc = np.zeros(3)
        for j in range(q.size):
            fr = fr_all[i] - fr_all[j]
            fr = fr - _anint(fr)
            for r, d2 in _ewald_images(cell, fr, p.real_cutoff):
                dvec = r @ cell.lattice.T            # cellReal . r
                ed = eta * np.sqrt(d2)
                f = np.array([math.erfc(x) for x in ed]) / ed ** 3 + 2.0 / sqrt_pi * np.exp(-(ed ** 2)) / ed ** 2
                acc += q[j] * np.sum(dvec * f[:, None], axis=0)
        real[i] = acc * q[i] * eta ** 3
    # reciprocal part: conj(S) |b|^2 exp(-G^2/4eta^2)/G^2, C2R, spline gather
    table = np.zeros(cell.shape)
    for z in np.unique(q):
        table += fill_qion_table(fr_all[q == z], cell.shape, order) * z
    rec = fft_forward(table)
    b1, b2, b3 = fill_b(order, cell.n1, cell.n2, cell.n3)
    bb = ((b1 * np.conj(b1))[None, None, :] * (b2 * np.conj(b2))[None, :, None] * (b3 * np.conj(b3))[:, None, None])
    cc = np.conj(rec) * bb
    _, qt, _ = g_tables(cell)
    m2 = qt * qt
    ok = (m2 <= p.recip_cutoff ** 2) & (m2 > 0.0)
    cc = np.where(ok, cc * np.exp(-np.where(ok, m2, 1.0) / eta ** 2 / 4.0) / np.where(ok, m2, 1.0), 0.0)
    grid = fft_backward(cc, cell.shape)
    recip = np.zeros((q.size, 3))
    for z in np.unique(q):
        idx = np.nonzero(q == z)[0]
        recip[idx] = np.array(spline_force_gather(cell, fr_all[idx], grid * z, order))
    recip *= 4.0 * PI / cell.vol * float(cell.nr)
    return real + recip


# --------------------------------------------------------------------------
# WGC kernels from the 1D eta tables (KEDF_WGCkernel.f90:109-277)
# --------------------------------------------------------------------------
def wgc_kernels_from_table(qtable, eta, w, w1p, w2pp, f: Functional):
    """FillWGC_ReciprocalSpace: int_w1 = eta w', int_w2 = eta^2 w'' (:172-173);
    splines with ibc 0 (:177-179); K2 = -w1/(6 rhoS), K3 = (w2+(7-g)w1)/(36 rhoS^2),
    K4 = (w2+(1+g)w1)/(36 rhoS^2) (:231-245). WGCT = 2 keeps all four (:257)."""
    eta = np.asarray(eta)
    iw1 = eta * np.asarray(w1p)
    iw2 = eta ** 2 * np.asarray(w2pp)
    w = np.asarray(w)
    pp0 = spline_cubic_set(eta, w, 0, 0.0, 0, 0.0)
    pp1 = spline_cubic_set(eta, iw1, 0, 0.0, 0, 0.0)
    pp2 = spline_cubic_set(eta, iw2, 0, 0.0, 0, 0.0)
    tkfs = 2.0 * (3.0 * PI ** 2 * f.rhoS) ** (1.0 / 3.0)
    e = (qtable / tkfs).ravel()
    k1 = spline_cubic_val_vec(eta, w, pp0, e).reshape(qtable.shape)
    s1 = spline_cubic_val_vec(eta, iw1, pp1, e).reshape(qtable.shape)
    s2 = spline_cubic_val_vec(eta, iw2, pp2, e).reshape(qtable.shape)
    g = f.gamma
    rs = f.rhoS
    k4 = (s2 + (1 + g) * s1) / (36.0 * rs ** 2)
    k3 = (s2 + (7 - g) * s1) / (36.0 * rs ** 2)
    k2 = -s1 / (6.0 * rs)
    return np.stack([k1, k2, k3, k4])


def wgc_eta_knots():
    """Knot grid of makeKernel2 (KEDF_KernelODE.f90:316-349)."""
    t_start = [1e-3, 3.0 + 1e-3]
    t_end = [3.0, 100.0]
    t_step = [1e-3, 1e-2]
    npts = [int(math.floor((t_end[i] - t_start[i]) / t_step[i])) - 1 for i in range(2)]
    eta = np.empty(npts[0] + npts[1])
    for i in range(npts[0]):
        eta[i] = t_start[0] + float(i) * t_step[0]
    for i in range(npts[0], eta.size):
        eta[i] = t_start[1] + float(i - npts[0]) * t_step[1]
    return eta


# --------------------------------------------------------------------------
# Moré–Thuente line search (MINPACK-2 dcsrch/dcstep as in MathDCsrch.f,
# MathDCstep.f; algorithm of Moré & Thuente, ACM TOMS 20 (1994) 286-307)
# --------------------------------------------------------------------------
def _dcstep(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """Safeguarded step (MathDCstep.f). Returns updated
    (stx, fx, dx, sty, fy, dy, stp, brackt)."""
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        r = p / q
        stpc = stx + r * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt((theta / s) ** 2 - (dx / s) * (dp / s))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        r = p / q
        stpc = stp + r * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        s = max(abs(theta), abs(dx), abs(dp))
        gamma = s * math.sqrt(max(0.0, (theta / s) ** 2 - (dx / s) * (dp / s)))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0.0 and gamma != 0.0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            if stp > stx:
                stpf = min(stp + 0.66 * (sty - stp), stpf)
            else:
                stpf = max(stp + 0.66 * (sty - stp), stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = min(stpmax, stpf)
            stpf = max(stpmin, stpf)
    else:
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            s = max(abs(theta), abs(dy), abs(dp))
            gamma = s * math.sqrt((theta / s) ** 2 - (dy / s) * (dp / s))
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            r = p / q
            stpc = stp + r * (sty - stp)
            stpf = stpc
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    return stx, fx, dx, sty, fy, dy, stpf, brackt


class Dcsrch:
    """Reverse-communication line search (MathDCsrch.f). ``step(stp, f, g)``
    returns (stp, task) exactly as the Fortran does with its isave/dsave."""

    def __init__(self, ftol, gtol, xtol, stpmin, stpmax):
        self.ftol, self.gtol, self.xtol = ftol, gtol, xtol
        self.stpmin, self.stpmax = stpmin, stpmax
        self.task = "START"

    def step(self, stp, f, g):
        if self.task.startswith("START"):
            task = ""
            if stp < self.stpmin:
                task = "ERROR: STP .LT. STPMIN"
            if stp > self.stpmax:
                task = "ERROR: STP .GT. STPMAX"
            if g >= 0.0:
                task = "ERROR: INITIAL G .GE. ZERO"
            if self.ftol < 0.0:
                task = "ERROR: FTOL .LT. ZERO"
            if self.gtol < 0.0:
                task = "ERROR: GTOL .LT. ZERO"
            if self.xtol < 0.0:
                task = "ERROR: XTOL .LT. ZERO"
            if self.stpmin < 0.0:
                task = "ERROR: STPMIN .LT. ZERO"
            if self.stpmax < self.stpmin:
                task = "ERROR: STPMAX .LT. STPMIN"
            if task.startswith("ERROR"):
                self.task = task
                return stp, task
            self.brackt = False
            self.stage = 1
            self.finit = f
            self.ginit = g
            self.gtest = self.ftol * self.ginit
            self.width = self.stpmax - self.stpmin
            self.width1 = self.width / 0.5
            self.stx, self.fx, self.gx = 0.0, self.finit, self.ginit
            self.sty, self.fy, self.gy = 0.0, self.finit, self.ginit
            self.stmin = 0.0
            self.stmax = stp + 4.0 * stp
            self.task = "FG"
            return stp, self.task
        ftest = self.finit + stp * self.gtest
        if self.stage == 1 and f <= ftest and g >= 0.0:
            self.stage = 2
        task = self.task
        if self.brackt and (stp <= self.stmin or stp >= self.stmax):
            task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS"
        if self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax:
            task = "WARNING: XTOL TEST SATISFIED"
        if stp == self.stpmax and f <= ftest and g <= self.gtest:
            task = "WARNING: STP = STPMAX"
        if stp == self.stpmin and (f > ftest or g >= self.gtest):
            task = "WARNING: STP = STPMIN"
        if f <= ftest and abs(g) <= self.gtol * (-self.ginit):
            task = "CONVERGENCE"
        if task.startswith("WARN") or task.startswith("CONV"):
            self.task = task
            return stp, task
        if self.stage == 1 and f <= self.fx and f > ftest:
            fm = f - stp * self.gtest
            fxm = self.fx - self.stx * self.gtest
            fym = self.fy - self.sty * self.gtest
            gm = g - self.gtest
            gxm = self.gx - self.gtest
            gym = self.gy - self.gtest
            (self.stx, fxm, gxm, self.sty, fym, gym, stp, self.brackt) = _dcstep(
                self.stx, fxm, gxm, self.sty, fym, gym, stp, fm, gm, self.brackt, self.stmin, self.stmax)
            self.fx = fxm + self.stx * self.gtest
            self.fy = fym + self.sty * self.gtest
            self.gx = gxm + self.gtest
            self.gy = gym + self.gtest
        else:
            (self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, self.brackt) = _dcstep(
                self.stx, self.fx, self.gx, self.sty, self.fy, self.gy, stp, f, g, self.brackt,
                self.stmin, self.stmax)
        if self.brackt:
            if abs(self.sty - self.stx) >= 0.66 * self.width1:
                stp = self.stx + 0.5 * (self.sty - self.stx)
            self.width1 = self.width
            self.width = abs(self.sty - self.stx)
        if self.brackt:
            self.stmin = min(self.stx, self.sty)
            self.stmax = max(self.stx, self.sty)
        else:
            self.stmin = stp + 1.1 * (stp - self.stx)
            self.stmax = stp + 4.0 * (stp - self.stx)
        stp = max(stp, self.stpmin)
        stp = min(stp, self.stpmax)
        if (self.brackt and (stp <= self.stmin or stp >= self.stmax)) or \
                (self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax):
            stp = self.stx
        self.task = "FG"
        return stp, self.task


# --------------------------------------------------------------------------
# Truncated-Newton density optimizer (RhoOptN.f90, RhoDirNew.f90,
# RhoLineSearch.f90, RhoDirBFGS.f90:294-343, RhoDirCG.f90:109-142)
# --------------------------------------------------------------------------
@dataclass
class OptParams:
    max_iter: int = 500
    pot_tol: float = POT_TOL_DEFAULT
    tole: float = TOLE_DEFAULT
    conv_check: str = "ENE"      # "ENE" | "POT" | "BOTH"
    max_line: int = 200


@dataclass
class OptResult:
    rho: np.ndarray
    energy: np.ndarray
    outcome: int
    iters: int
    history: list
    num_evals: int


def _sign(x):
    """Fortran SIGN(1._DP, x): +1 for x >= +0, -1 for x < 0 or x = -0.0."""
    return np.where(np.signbit(x), -1.0, 1.0)


def newton_direction(phi, b, mu0, evaluate, dV):
    """NewtonDirection (RhoDirNew.f90:25-210): CG on H x = -b with finite-difference
    Hessian products. evaluate(rho) -> dE/dphi (optSqrt, no energy)."""
    flag = -1
    direction = np.zeros_like(phi)
    r = -b
    rr = np.sum(r ** 2)
    criteria = rr
    it = 1
    ry = 0.0
    p = None
    nevals = 0
    while True:
        y = r
        ry_last = ry
        ry = np.sum(r * y)
        p = y.copy() if it == 1 else y + ry / ry_last * p
        s1 = np.sum(phi ** 2)
        s2 = np.sum(p ** 2)
        eps = 2.0 * math.sqrt(MACH_PREC) * (1.0 + math.sqrt(s1)) / math.sqrt(s2)
        new_phi = phi + eps * p
        ap = evaluate(new_phi ** 2)
        nevals += 1
        ap = ap * _sign(new_phi)
        ap = ap - mu0 * 2.0 * new_phi
        ap = (ap - b) / eps
        pap = np.sum(p * ap)
        if pap < 0.0:
            if it == 1:
                direction = -(direction - ry / pap * p)
            flag = -2
            break
        alpha = ry / pap
        direction = direction + alpha * p
        r = r - alpha * ap
        rr_last = rr
        rr = np.sum(r ** 2)
        if rr < 0.1 * criteria:
            flag = 0
            break
        if it > 50:
            flag = 1
            break
        if abs(rr - rr_last) / rr < 0.01 and it > 9:
            flag = 2
            break
        it += 1
    return direction, it, flag, nevals


def cg_direction(it, dldphi, dldphi_old, dir_old, dV, cg_alg="HZ"):
    """CGDirection (RhoDirCG.f90:32-106), numSpin = 1, with InnerProd =
    SUM(a b) dV (:109-142): restart every 100 steps, Polak-Ribiere or
    Hager-Zhang gamma. Returns (direction, new dL/dphi_old)."""
    ip = lambda a, b: float(np.sum(a * b)) * dV   # noqa: E731
    if it == 1 or it % 100 == 1:
        gam = 0.0
    elif cg_alg[:2] == "PR":
        gam = max(0.0, ip(dldphi, dldphi - dldphi_old) / ip(dldphi_old, dldphi_old))
    else:
        y = dldphi - dldphi_old
        dty = ip(dir_old, y)
        norm_y = math.sqrt(ip(y, y) / dV)
        gam = 1.0 / dty * ip(y - 2.0 * dir_old * norm_y ** 2 / dty, dldphi)
        norm_dk = math.sqrt(ip(dir_old, dir_old) / dV)
        norm_gk = math.sqrt(ip(dldphi_old, dldphi_old) / dV)
        gam = max(gam, -1.0 / (norm_dk * min(0.01, norm_gk)))
    return -dldphi + gam * dir_old, dldphi.copy()


class BfgsMemory:
    """The RhoDirBFGS module state (RhoDirBFGS.f90:26-42) after InitializeBFGS
    (:45-92): MBFGS = 5 (s, y) pairs in a ring, the write pointer BFGSPT and
    the call counter BFGSiter; numSpin = 1. Slots are 0-based here (slot k is
    the reference's index k+1)."""

    def __init__(self, shape, m=5):
        self.m = m
        self.s = np.zeros((m,) + tuple(shape))
        self.y = np.zeros((m,) + tuple(shape))
        self.rho = np.zeros(m)
        self.alpha = np.zeros(m)
        self.pt = 0
        self.iter = 0

    def direction(self, dldphi, dV):
        """BFGSDirection (RhoDirBFGS.f90:120-218): -dL/dphi on the first call,
        else the two-loop recursion (Nocedal, Math. Comp. 35 (1980) 773-782)
        over the newest min(iter-1, MBFGS) pairs with H0 = (y.s / y.y) I from
        the newest pair; InnerProd carries dV."""
        ip = lambda a, b: float(np.sum(a * b)) * dV   # noqa: E731
        m = self.m
        self.iter += 1
        if self.iter == 1:
            return -dldphi
        bound = min(self.iter - 1, m)
        npt = (self.pt - 1) % m                        # the newest pair
        ys = ip(self.s[npt], self.y[npt])
        yy = ip(self.y[npt], self.y[npt])
        diag = ys / yy
        self.rho[npt] = 1.0 / ys
        q = -dldphi
        cp = self.pt
        for _ in range(bound):                         # newest to oldest
            cp = (cp - 1) % m
            self.alpha[cp] = self.rho[cp] * ip(self.s[cp], q)
            q = q - self.alpha[cp] * self.y[cp]
        q = q * diag
        for _ in range(bound):                         # oldest to newest
            beta = self.alpha[cp] - self.rho[cp] * ip(self.y[cp], q)
            q = q + beta * self.s[cp]
            cp = (cp + 1) % m
        return q

    def update(self, phi, temp_phi, dedphi, dldphi, tot_ele, dV):
        """UpdateBFGS (RhoDirBFGS.f90:221-290) with ChemicalPotential (:294-343):
        s = tempPhi - phi, y = (dE/dphi - 2 mu(tempPhi) tempPhi) - dL/dphi."""
        mu = dV * np.sum(dedphi * temp_phi) / (2.0 * tot_ele)
        self.s[self.pt] = temp_phi - phi
        self.y[self.pt] = (dedphi - 2.0 * mu * temp_phi) - dldphi
        self.pt = (self.pt + 1) % self.m


def rho_minimization(rho_init, evaluate_full, evaluate_pot, dV, params: OptParams = OptParams(), dir_type=1,
                     cg_alg="HZ"):
    """ProjNormRhoMinimization with dir_type = 1 (NTN), 2 (NCG) or 3 (NBF),
    numSpin = 1 (RhoOptN.f90:60-850, DensityOptimization :550-641, CheckExit
    :853-977, getNextDirection :655-687). evaluate_full(rho) -> (dE/dphi,
    eTable); evaluate_pot(rho) -> dE/dphi."""
    nr = rho_init.size
    num_points = float(nr)
    # ftol = 1e-4, gtol = 2e-1 (NTN, NBF) or 1e-2 (NCG) are default-real
    # (single precision) literals assigned to DP variables (RhoOptN.f90:255-
    # 264); xtol is 1.E-12_DP (:193)
    ftol = float(np.float32(1e-4))
    gtol = float(np.float32(1e-2 if dir_type == 2 else 2e-1))
    bfgs = BfgsMemory(rho_init.shape) if dir_type == 3 else None
    xtol = 1e-12
    tot_ele = np.sum(rho_init) * dV
    theta = 0.2
    phi = np.sqrt(np.abs(rho_init))
    dedphi, energy = evaluate_full(phi ** 2)
    nevals = 1
    dedphi = dedphi * _sign(phi)
    mu0 = dV * np.sum(dedphi * phi) / (2.0 * tot_ele)
    dldphi = dedphi - mu0 * 2.0 * phi
    pnorm0 = math.sqrt(np.sum(dldphi ** 2) / num_points)
    history = [dict(iter=0, energy=float(energy[0]), pot_norm=pnorm0, theta=0.0)]
    e = [energy[0], 0.0, 0.0]
    outcome = -1
    success = 1
    it = 1
    direction = np.zeros_like(phi)
    dir_old = np.zeros_like(phi)
    dldphi_old = np.zeros_like(phi)
    temp_phi = phi
    newton_iter, ntflag = 0, 0
    while True:
        switch_stp = False
        if dir_type == 1:
            direction, newton_iter, ntflag, ne = newton_direction(phi, dldphi, mu0, evaluate_pot, dV)
            nevals += ne
        elif dir_type == 2:
            direction, dldphi_old = cg_direction(it, dldphi, dldphi_old, dir_old, dV, cg_alg)
        else:
            direction = bfgs.direction(dldphi, dV)
        while True:
            # getNextGradient (RhoLineSearch.f90:40-139), numSpin = 1
            pd = phi + direction
            norm_phi_dir = math.sqrt(np.sum(pd * pd) * dV)
            direction = direction - phi * (np.sum(phi * direction) * dV) / tot_ele
            norm_perp = math.sqrt(np.sum(direction * direction) * dV)
            theta = min(theta, norm_perp / norm_phi_dir)
            direction = direction * math.sqrt(tot_ele) / norm_perp
            gradth = dV * np.sum(dedphi * (-phi * math.sin(0.0) + direction * math.cos(0.0)))
            # checkGradient (RhoLineSearch.f90:201-309)
            if gradth > 1.0e5:
                raise RuntimeError("The gradient is too large, the value is %r" % gradth)
            if gradth >= 0.0:
                if switch_stp:
                    outcome = 1
                    break
                switch_stp = True
                direction = -dldphi
                continue
            break
        dir_old = direction.copy()     # dirNext_old = dirNext (RhoOptN.f90:566)
        if outcome == 1:
            break
        # LineSearchTN (RhoLineSearch.f90:312-444)
        ls = Dcsrch(ftol, gtol, xtol, 0.0, PI)
        nline = 0
        f_cur, g_cur = energy[0], gradth
        while True:
            theta, task = ls.step(theta, f_cur, g_cur)
            if task.startswith("CONV"):
                success = 1
                break
            if task.startswith("ERROR"):
                success = -1
                break
            if task.startswith("WARN"):
                success = -2
                break
            temp_phi = phi * math.cos(theta) + direction * math.sin(theta)
            dedphi, energy = evaluate_full(temp_phi ** 2)
            nevals += 1
            dedphi = dedphi * _sign(temp_phi)
            g_cur = dV * np.sum(dedphi * (-phi * math.sin(theta) + direction * math.cos(theta)))
            f_cur = energy[0]
            nline += 1
            if nline > params.max_line:
                break
        if success == -1 and it > params.max_iter:
            raise RuntimeError("Fail rho optimizations")
        # StepInformation (:692-847): potNorm with the previous mu and the new phi
        pot_norm = math.sqrt(np.sum((dedphi - 2.0 * temp_phi * mu0) ** 2) / num_points)
        history.append(dict(iter=it, energy=float(energy[0]), pot_norm=pot_norm, theta=theta,
                            newton_iter=newton_iter, ntflag=ntflag, nline=nline, success=success,
                            min_den=float(np.min(temp_phi ** 2)), max_den=float(np.max(temp_phi ** 2)),
                            min_pot=float(np.min(dedphi)), max_pot=float(np.max(dedphi))))
        # CheckExit (:853-977)
        exit_flag = False
        pconv = pot_norm < params.pot_tol
        e = [energy[0], e[0], e[1]]
        econv = it >= 3 and abs(e[0] - e[1]) < params.tole and abs(e[0] - e[2]) < params.tole
        cc = params.conv_check
        if cc.startswith("ENE") and econv:
            outcome, exit_flag = 0, True
        if cc.startswith("POT") and pconv:
            outcome, exit_flag = 0, True
        if cc.startswith("BOTH") and (pconv or econv):
            outcome, exit_flag = 0, True
        if outcome == 1:
            exit_flag = True
        if it >= params.max_iter:
            outcome, exit_flag = 1, True
        if dir_type == 3:
            bfgs.update(phi, temp_phi, dedphi, dldphi, tot_ele, dV)   # RhoOptN.f90:603-610
        phi = temp_phi
        mu0 = dV * np.sum(dedphi * phi) / (2.0 * tot_ele)
        dldphi = dedphi - 2.0 * mu0 * phi
        if exit_flag:
            break
        it += 1
    return OptResult(rho=phi ** 2, energy=energy, outcome=outcome, iters=it,
                     history=history, num_evals=nevals)


def spin_theta_search(theta, energy, gradth, phi, direction, dV, evaluate_full, ftol, gtol, xtol):
    """SpinThetaSearch (RhoLineSearch.f90:447-657): at most two Polak-Ribiere
    (thetaCGDirection, :660-700) line searches in (theta_up, theta_down), each a
    Dcsrch over the step alpha along thetadir in [0, pi / max|thetadir|] with at
    most 10 trial points. theta and gradth are updated in place. Returns
    (temp_phi, dE/dphi, energy, success, evaluations)."""
    alphatol = _f32(1e-4)
    thetadir_old = np.zeros(2)
    gradth_old = np.zeros(2)
    success = 1
    alpha = 0.0
    temp_phi = np.stack([phi[k] * math.cos(theta[k]) + direction[k] * math.sin(theta[k]) for k in range(2)])
    dedphi, energy = evaluate_full(temp_phi ** 2)
    nevals = 1
    dedphi = dedphi * _sign(temp_phi)

    def grad_of_theta(th, k):         # gradOftheta (RhoLineSearch.f90:142-198)
        return dV * np.sum(dedphi[k] * (-phi[k] * math.sin(th) + direction[k] * math.cos(th)))

    j = 1
    while True:
        gam = 0.0 if j == 1 else max(0.0, float(np.sum(gradth * (gradth - gradth_old))) /
                                      float(np.sum(gradth_old * gradth_old)))
        gradth_old = gradth.copy()
        thetadir = -gradth + gam * thetadir_old
        for k in range(2):
            gradth[k] = grad_of_theta(theta[k], k)
        while True:
            dedalpha = gradth[0] * thetadir[0] + gradth[1] * thetadir[1]
            if dedalpha >= 0.0:
                thetadir = -gradth.copy()
            else:
                break
        thetadir_old = thetadir.copy()
        maxdir = max(abs(thetadir[0]), abs(thetadir[1]))
        # MIN(0.1, 0.1*pi/maxthetadir) with the single-precision literal 0.1
        alpha = min(_f32(0.1), _f32(0.1) * PI / maxdir)
        ls = Dcsrch(ftol, gtol, xtol, 0.0, PI / maxdir)
        nline = 0
        while True:
            alpha, task = ls.step(alpha, energy[0], dedalpha)
            if task.startswith("CONV"):
                success = 2
                break
            if task.startswith("ERROR"):
                success = -1
                break
            if task.startswith("WARN"):
                success = -2
                break
            tem = theta + alpha * thetadir
            temp_phi = np.stack([phi[k] * math.cos(tem[k]) + direction[k] * math.sin(tem[k]) for k in range(2)])
            dedphi, energy = evaluate_full(temp_phi ** 2)
            nevals += 1
            dedphi = dedphi * _sign(temp_phi)
            for k in range(2):
                gradth[k] = grad_of_theta(tem[k], k)
            dedalpha = gradth[0] * thetadir[0] + gradth[1] * thetadir[1]
            nline += 1
            if nline > 10:
                success = -1
                break
        theta += alpha * thetadir
        if success == -1:
            break
        if math.sqrt(gradth[0] * gradth[0] + gradth[1] * gradth[1]) < alphatol:
            break
        j += 1
        if j > 2:
            break
    return temp_phi, dedphi, energy, success, nevals


def rho_minimization_spin(rho2_init, evaluate_full, evaluate_pot, dV, params: OptParams = OptParams(),
                          tot_ele=None):
    """ProjNormRhoMinimization for numSpin = 2 with dir_type = 1 (NTN), one
    pass of the magnetization loop (RhoOptN.f90:336-420, DensityOptimization
    :550-641): per-spin electron numbers, chemical potentials and Newton
    directions (NewtonDirection perturbs one channel, RhoDirNew.f90:25-210),
    getNextGradient / checkGradient per channel (RhoLineSearch.f90:40-139,
    :201-309), the joint theta search, the potential norm over both channels
    (:745-760). evaluate_full(rho2) -> (dE/dphi (2, ...), eTable);
    evaluate_pot(rho2) -> dE/dphi. tot_ele: the electron numbers per spin
    (default: sum(rho_s) dV, :298-301)."""
    num_points = float(rho2_init[0].size)
    ftol = float(np.float32(1e-4))
    gtol = float(np.float32(2e-1))
    xtol = 1e-12
    tot = np.array([np.sum(rho2_init[s]) * dV for s in range(2)]) if tot_ele is None else np.asarray(tot_ele, float)
    theta = np.array([0.2, 0.2])
    gradth = np.zeros(2)
    phi = np.sqrt(np.abs(rho2_init))
    dedphi, energy = evaluate_full(phi ** 2)
    nevals = 1
    dedphi = dedphi * _sign(phi)
    mu0 = np.array([dV * np.sum(dedphi[s] * phi[s]) / (2.0 * tot[s]) for s in range(2)])
    dldphi = np.stack([dedphi[s] - mu0[s] * 2.0 * phi[s] for s in range(2)])
    pnorm0 = math.sqrt((np.sum(dldphi[0] ** 2) + np.sum(dldphi[1] ** 2)) / num_points / 2.0)
    history = [dict(iter=0, energy=float(energy[0]), pot_norm=pnorm0)]
    e = [energy[0], 0.0, 0.0]
    outcome = -1
    it = 1
    temp_phi = phi
    while True:
        direction = np.zeros_like(phi)
        newton_iters = []
        for s in range(2):
            other = phi[1 - s] ** 2

            def ev(rho_s, s=s, other=other):
                rr = np.empty((2,) + rho_s.shape)
                rr[s] = rho_s
                rr[1 - s] = other
                return evaluate_pot(rr)[s]
            direction[s], nit, _, ne = newton_direction(phi[s], dldphi[s], mu0[s], ev, dV)
            nevals += ne
            newton_iters.append(nit)
        sw1 = sw2 = False
        while True:
            for s in range(2):                 # getNextGradient, numSpin = 2
                pd = phi[s] + direction[s]
                _norm_phi_dir = math.sqrt(np.sum(pd * pd) * dV)   # computed, unused for numSpin = 2
                direction[s] = direction[s] - phi[s] * (np.sum(phi[s] * direction[s]) * dV) / tot[s]
                norm_perp = math.sqrt(np.sum(direction[s] * direction[s]) * dV)
                theta[s] = 0.0
                direction[s] = direction[s] * math.sqrt(tot[s]) / norm_perp
                gradth[s] = dV * np.sum(dedphi[s] * (-phi[s] * math.sin(0.0) + direction[s] * math.cos(0.0)))
            done = False                         # checkGradient, numSpin = 2
            for s in range(2):
                if gradth[s] > 1.0e5:
                    raise RuntimeError("The gradient %d is too large" % (s + 1))
                if gradth[s] >= 0.0:
                    if sw1 and sw2:
                        outcome = 1
                        done = True
                        break
                    if s == 0:
                        sw1 = True
                    else:
                        sw2 = True
                    direction[s] = -dldphi[s]
            if done or (gradth[0] < 0.0 and gradth[1] < 0.0):
                break
        if outcome == 1:
            break
        temp_phi, dedphi, energy, success, ne = spin_theta_search(theta, energy, gradth, phi, direction, dV,
                                                                  evaluate_full, ftol, gtol, xtol)
        nevals += ne
        if success == -1 and it > params.max_iter:
            raise RuntimeError("Fail rho optimizations")
        pot_norm = math.sqrt((np.sum((dedphi[0] - 2.0 * temp_phi[0] * mu0[0]) ** 2) +
                              np.sum((dedphi[1] - 2.0 * temp_phi[1] * mu0[1]) ** 2)) / num_points / 2.0)
        history.append(dict(iter=it, energy=float(energy[0]), pot_norm=pot_norm, theta=theta.copy(),
                            newton_iter=newton_iters, success=success))
        exit_flag = False
        pconv = pot_norm < params.pot_tol
        e = [energy[0], e[0], e[1]]
        econv = it >= 3 and abs(e[0] - e[1]) < params.tole and abs(e[0] - e[2]) < params.tole
        cc = params.conv_check
        if cc.startswith("ENE") and econv:
            outcome, exit_flag = 0, True
        if cc.startswith("POT") and pconv:
            outcome, exit_flag = 0, True
        if cc.startswith("BOTH") and (pconv or econv):
            outcome, exit_flag = 0, True
        if outcome == 1:
            exit_flag = True
        if it >= params.max_iter:
            outcome, exit_flag = 1, True
        phi = np.stack([phi[s] * math.cos(theta[s]) + direction[s] * math.sin(theta[s]) for s in range(2)])
        mu0 = np.array([dV * np.sum(dedphi[s] * phi[s]) / (2.0 * tot[s]) for s in range(2)])
        dldphi = np.stack([dedphi[s] - 2.0 * mu0[s] * phi[s] for s in range(2)])
        if exit_flag:
            break
        it += 1
    return OptResult(rho=phi ** 2, energy=energy, outcome=outcome, iters=it, history=history, num_evals=nevals)


@dataclass
class SpinOptResult:
    rho: np.ndarray
    energy: np.ndarray
    outcome: int
    mag_iters: int
    magnetization: float
    tot_ele: np.ndarray
    solves: list


def rho_minimization_spin_mag(rho2_init, evaluate_full, evaluate_pot, dV, params: OptParams = OptParams(),
                              fixed_mag=False, nmag=100, tolm=float(np.float32(5.0e-6))):
    """The magnetization loop around the numSpin = 2 solve (RhoOptN.f90:336-
    480): with fixed_mag (the input's FIXMAG) one solve at the initial
    magnetization; otherwise up to nmag solves, each restarted from the initial
    density rescaled to the spin populations (Ntot +- m)/2, m moved by the
    secant/parabola steps of :461-472 until |dE/dm| < tolm (or 60 steps). The
    reference reads m1 before its first assignment when it forms dE/dm after
    the first solve; here it starts at 0, which only decides whether the
    second solve's test passes (it cannot for any finite energy at m = 0)."""
    rho_r = np.array(rho2_init, dtype=float)
    tot = np.array([np.sum(rho_r[s]) * dV for s in range(2)])
    ntot = float(np.sum(tot))
    magnetization = tot[0] - tot[1]
    de = 1.0
    mag_iter = 0
    ee1 = ee2 = ee3 = 0.0
    m1 = m2 = m3 = 0.0
    solves = []
    res = None
    for _ in range(nmag):
        res = rho_minimization_spin(rho_r, evaluate_full, evaluate_pot, dV, params, tot_ele=tot.copy())
        solves.append(res)
        mag_iter += 1
        if fixed_mag:
            break
        if abs(de) < tolm:
            break
        ee3, ee2, ee1 = ee2, ee1, float(res.energy[0])
        m3, m2, m1 = m2, m1, magnetization
        with np.errstate(divide="ignore", invalid="ignore"):
            de = float(np.float64(ee1 - ee2) / np.float64(m1 - m2))   # IEEE, as the Fortran
        if mag_iter == 1:
            magnetization = m1 + _f32(0.1)
        if mag_iter == 2:
            magnetization = m1 - _f32(0.2)
        if mag_iter == 3:
            magnetization = m1 - de / ee1
        if mag_iter > 3:
            b = (m1 - m2) / (m3 - m2)
            de1 = 2.0 * (ee1 - (1 - b) * ee2 - b * ee3) / ((1.0 - 1.0 / b) * (m1 - m2) ** 2)
            if mag_iter > 60:
                break
            magnetization = m1 - de / de1
        rho_r[0] = rho_r[0] / tot[0]
        rho_r[1] = rho_r[1] / tot[1]
        tot = np.array([0.5 * (ntot + magnetization), 0.5 * (ntot - magnetization)])
        rho_r[0] = rho_r[0] * tot[0]
        rho_r[1] = rho_r[1] * tot[1]
    return SpinOptResult(rho=res.rho, energy=res.energy, outcome=res.outcome, mag_iters=mag_iter,
                         magnetization=magnetization, tot_ele=tot, solves=solves)


def make_spin_evaluators(sysm: System, lsda=1):
    """calculate_potential_plus_spin in the optimizer's two call forms."""
    def full(rho2):
        return calculate_potential_plus_spin(rho2, sysm, True, True, lsda)

    def pot_only(rho2):
        return calculate_potential_plus_spin(rho2, sysm, True, False, lsda)[0]
    return full, pot_only


def make_evaluators(sysm: System):
    """Bind calculate_potential_plus to the optimizer's two call forms
    (with eTable: RhoOptN.f90:357, RhoLineSearch.f90:402; without: RhoDirNew.f90:152)."""
    def full(rho):
        return calculate_potential_plus(rho, sysm, True, True)

    def pot_only(rho):
        return calculate_potential_plus(rho, sysm, True, False)[0]
    return full, pot_only


# --------------------------------------------------------------------------
# Ion relaxation: CG2 with LnSrch (IonOptCG2.f90, MathLNsrch.f90, Neighbor.f90)
# --------------------------------------------------------------------------
def _f32(x):
    """A default-kind (single precision) Fortran literal widened to DP."""
    return float(np.float32(x))


class LnSrch:
    """LnSrch (MathLNsrch.f90:52-296) with its module state points(2,3) and
    dsave. ``step(alp, f, g, task, c1, c2, stpmin, stpmax)`` returns
    (alp, task) as the Fortran leaves them."""

    def __init__(self):
        self.points = np.full((2, 3), -1.0)
        self.dsave = np.zeros(13)

    @staticmethod
    def _new_alp(x1, y1, x2, y2, stpmin, stpmax):
        # IEEE division as in the Fortran: a zero slope gives +-Inf, which the
        # bounds test below turns into -1
        with np.errstate(divide="ignore", invalid="ignore"):
            slope = np.float64(y1 - y2) / np.float64(x1 - x2)
            b = y1 - x1 * slope
            alp_t = float(-b / slope)
        if alp_t >= stpmax or alp_t <= stpmin:
            return -1.0
        return alp_t

    def step(self, alp, f, g, task, c1, c2, stpmin, stpmax):
        P = self.points
        na = lambda: self._new_alp(P[0, 0], P[0, 2], P[1, 0], P[1, 2], stpmin, stpmax)   # noqa: E731
        if task[:5] == "START":
            if g > 0.0:
                return alp, "ERROR: Initial G>0"
            P[:] = -1.0
            P[0] = (0.0, f, g)
            self.dsave[0], self.dsave[1] = f, g
            P[1, 0] = -1.0
            return alp, "FG" + task[2:]
        if task[:2] == "FG" and P[1, 0] > 0.0:
            if f <= self.dsave[0] + c1 * alp * self.dsave[1] and abs(g) <= c2 * abs(self.dsave[1]):
                return alp, "CONV"
        if P[1, 0] < 0.0:
            if g < P[0, 2]:
                alp = alp * 5.0 if f < P[0, 1] else (P[0, 0] + alp) / 2.0
                return alp, "WARNING: Need a better second point."
            P[1] = (alp, f, g)
            alp = na()
            if alp < 0.0 or alp < P[0, 0]:
                if P[1, 1] < P[0, 1]:
                    return 5.0 * P[1, 0], "WARNING: Interpolation error, code:-1, energy decreases, step further"
                return alp, "ERROR: Interpolation error, code:-1"
            return alp, "FG"
        if g > 0:
            if P[1, 2] > 0.0:
                if P[1, 0] > alp:
                    P[1] = (alp, f, g)
                    alp = na()
                    if alp < P[0, 0]:
                        return (P[0, 0] + P[1, 0]) / 2.0, "WARNING: Order flips, code:0, energy decreases, step further"
                    return alp, "FG"
                if f < P[1, 1]:
                    return alp * 5.0, "WARNINIG: Order flipped, code:1"
                return (P[0, 0] + P[1, 0]) / 2.0, "WARNING: Order flipped, code:2"
            if P[1, 2] < 0.0:
                P[0] = P[1]
                P[1] = (alp, f, g)
                alp = na()
                if alp < P[0, 0] or alp > P[1, 0]:
                    if P[0, 1] > P[1, 1]:
                        return P[1, 0] * 5.0, "WARNING: Interpolation error, code:1, energy decreases, step further"
                    return (P[0, 0] + P[1, 0]) / 2.0, "WARNING: Interpolation error, code:1"
                return alp, "FG"
        if g < 0.0:
            if P[1, 2] > 0.0:
                P[0] = (alp, f, g)
                alp = na()
                if alp < P[0, 0] or alp > P[1, 0]:
                    if P[0, 1] > P[1, 1]:
                        return 5.0 * P[1, 0], "WARNING: Interpolation, code:2, f decreases, step further"
                    return (P[0, 0] + P[1, 0]) / 2.0, "WARNING: Interpolation, code:4, brackets the minimum by half."
                return alp, "FG"
            if P[1, 2] < 0.0:
                P[0] = P[1]
                P[1] = (alp, f, g)
                alp = na()
                if alp < P[1, 0]:
                    if P[0, 1] > P[1, 1]:
                        return P[1, 0] * 5.0, "WARNING: Interpolation error: code:3, f decrease, step further"
                    return alp, "ERROR: Interpolation error: code:3"
                return alp, "FG"
        return alp, task


def nn_min_distance(cell: Cell, frac):
    """NNDist_bin (Neighbor.f90:68-420) for a cell of one bin (every lattice
    vector shorter than the 20 bohr bin spacing) and, in general, the minimum
    over all pairs and image shifts in {-1,0,1}^3 of the wrapped positions,
    which the binned search returns whenever the nearest distance is below the
    bin spacing."""
    fr = np.mod(np.asarray(frac, dtype=np.float64).reshape(-1, 3), 1.0)
    xyz = fr @ cell.lattice.T
    n = xyz.shape[0]
    best = 1e6
    shifts = [(i, j, k) for k in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1)]
    for a in range(n):
        for b in range(n):
            for sh in shifts:
                if a == b and sh == (0, 0, 0):
                    continue
                d = xyz[a] - (xyz[b] + cell.lattice @ np.array(sh, dtype=np.float64))
                best = min(best, float(math.sqrt(float(d @ d))))
    return best


@dataclass
class IonRelaxResult:
    frac: np.ndarray
    energy: float
    forces: np.ndarray
    steps: int
    energies: list          # total energy at every geometry evaluated, in order
    max_force: float


def ion_relax_cg2(cell: Cell, frac, evaluate, cg_type=2, force_cutoff=_f32(5e-5), max_steps=200, nn_dist=2.0,
                  max_line=10):
    """ConjugateGradient (IonOptCG2.f90:60-520). ``evaluate(frac)`` does what
    RefreshIonTerms + RhoOptimizer + CalculateForces do for the positions
    (fractional, wrapped) and returns (total energy in Ha, forces (nion, 3)
    in Ha/bohr); it is called once for the starting geometry and once per
    line-search trial."""
    ftol, gtol = _f32(1e-4), _f32(1e-1)
    frac = np.mod(np.asarray(frac, dtype=np.float64).reshape(-1, 3), 1.0)
    frac2cart = cell.lattice
    cart2frac = np.linalg.inv(cell.lattice)
    energy, forces = evaluate(frac)
    energies = [energy]
    nion = frac.shape[0]
    old_grad = np.zeros((nion, 3))
    dir_old = np.zeros((nion, 3))
    alpha = 1.0
    restart = False
    step = 0
    ls = LnSrch()
    nline = 0
    max_force = float(np.max(np.sqrt(np.sum(forces ** 2, axis=1))))
    while True:
        step += 1
        if step > max_steps:
            break
        max_force = float(np.max(np.sqrt(np.sum(forces ** 2, axis=1))))
        if max_force < force_cutoff:
            break
        old_cart = frac @ frac2cart.T
        new_grad = forces.copy()
        if step == 1 or restart or step % 30 == 0:
            gam = 0.0
            restart = False
        elif cg_type == 1:
            gam = max(0.0, float(np.sum((new_grad - old_grad) * new_grad) / np.sum(old_grad ** 2)))
        elif cg_type == 2:
            y = -new_grad + old_grad
            norm_y = math.sqrt(float(np.sum(y ** 2)))
            dy = float(np.sum(dir_old * y))
            t1 = y - 2.0 * dir_old * norm_y ** 2 / dy
            t2 = -new_grad / dy
            gam = float(np.sum(t1 * t2))
            cg_eta = 0.01
            norm_d = math.sqrt(float(np.sum(dir_old ** 2)))
            norm_g = math.sqrt(float(np.sum(old_grad ** 2)))
            gam = max(gam, -1.0 / (norm_d * min(cg_eta, norm_g)))
        else:
            y = -new_grad + old_grad
            gam = max(0.0, float(np.sum(new_grad ** 2) / np.sum(dir_old * y)))
        d = new_grad + gam * dir_old
        dir_old = d
        old_grad = new_grad
        grad_alp = -float(np.sum(forces * d))
        if nn_dist > 0:
            stpmax = abs((nn_min_distance(cell, frac) - nn_dist) / float(np.max(d)))
        else:
            stpmax = abs(1.0 / float(np.max(d)))
        stpmin = 0.0
        alpha = min(alpha, stpmax)
        task = "START"
        nline = 0
        relax_done = False
        while True:
            alpha, task = ls.step(alpha, energy, grad_alp, task, ftol, gtol, stpmin, stpmax)
            P = ls.points
            if nline >= 4 and abs(P[0, 1] - P[1, 1]) < _f32(1e-6):
                break
            if P[0, 0] > 0.0:
                bc = (P[0, 0] - P[1, 0]) * d
                if float(np.max(np.sqrt(np.sum(bc ** 2, axis=1)))) < _f32(1e-5):
                    break
            if task[:4] == "CONV":
                break
            if task[:5] == "ERROR":
                restart = True
                alpha = 1.0
                break
            if task[:2] == "FG" or task[:4] == "WARN":
                frac = np.mod((old_cart + alpha * d) @ cart2frac.T, 1.0)
                energy, forces = evaluate(frac)
                energies.append(energy)
                grad_alp = -float(np.sum(forces * d))
                max_force = float(np.max(np.sqrt(np.sum(forces ** 2, axis=1))))
                if max_force < force_cutoff:
                    relax_done = True
                    break
                nline += 1
                if nline > max_line:
                    restart = True
                    alpha = 1.0
                    break
        if relax_done:
            break
    return IonRelaxResult(frac, energy, forces, step, energies, max_force)
