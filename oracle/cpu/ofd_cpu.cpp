// CPU restatement of one WT energy + potential evaluation (CalculatePotentialPlus
// + CalculateEnergy, CalPotential.f90:41-647; numSpin = 1, LDA, Thomas-Fermi +
// von Weizsacker + Wang-Teter), FP64, on the reference's own dataflow: four
// real-to-complex and four complex-to-real 3D transforms per evaluation
// (Hartree.f90:93-169 on FFT(rho); KEDF_WT.f90:70-170 on FFT(rho^beta) and
// FFT(rho^alpha) separately, as the reference calls them; KEDF_VW.f90:282-362
// on FFT(sqrt rho)), with its transform conventions (forward scaled by 1/N,
// Fourier_v1.f90:388-393; backward unnormalized).
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the
// GPU (one core, and every core with OpenMP), pinned against the numpy oracle
// by tests/test_cpu_port.py. The FFT is this file's own mixed-radix Stockham
// (radix 4, 2, 3, 5) standing in for the reference's FFTW3, which is not
// vendored.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.141592653589793238462643;
constexpr double kCTF = 2.87123400018819;   // KEDF_TF.f90:42

// ------------------------------------------------------------------ 1D FFT
struct Fft1 {
  int n = 0;
  std::vector<int> rad;
  std::vector<cd> w;                  // w[t] = exp(-2 pi i t / n)
  std::vector<std::vector<cd>> stw;   // per stage: exp(-2 pi i k r / (ns R)), [k][r - 1]

  bool init(int n_) {
    n = n_;
    rad.clear();
    int m = n;
    while (m % 4 == 0) { rad.push_back(4); m /= 4; }
    while (m % 2 == 0) { rad.push_back(2); m /= 2; }
    while (m % 3 == 0) { rad.push_back(3); m /= 3; }
    while (m % 5 == 0) { rad.push_back(5); m /= 5; }
    if (m != 1) return false;
    w.resize(n);
    for (int t = 0; t < n; ++t) {
      const long double a = -2.0L * 3.141592653589793238462643383279L * t / n;
      w[t] = cd((double)cosl(a), (double)sinl(a));
    }
    stw.clear();
    int ns = 1;
    for (int R : rad) {
      const int step = n / (ns * R);
      std::vector<cd> t((size_t)ns * (R - 1));
      for (int k = 0; k < ns; ++k)
        for (int r = 1; r < R; ++r) t[(size_t)k * (R - 1) + r - 1] = w[(size_t)k * r * step % n];
      stw.push_back(std::move(t));
      ns *= R;
    }
    return true;
  }

  // radix-R DFT of v in place, X[k] = sum_r v[r] exp(S 2 pi i r k / R)
  template <int R, int S>
  static inline void dft(cd* v) {
    if constexpr (R == 2) {
      const cd a = v[0], b = v[1];
      v[0] = a + b;
      v[1] = a - b;
    } else if constexpr (R == 4) {
      const cd a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = v[1] - v[3];
      const cd jd = S < 0 ? cd(d.imag(), -d.real()) : cd(-d.imag(), d.real());   // S i d
      v[0] = a + c;
      v[2] = a - c;
      v[1] = b + jd;
      v[3] = b - jd;
    } else {
      cd out[R];
      for (int k = 0; k < R; ++k) {
        cd acc = 0.0;
        for (int r = 0; r < R; ++r) {
          const double a = S * 2.0 * kPi * (double)((r * k) % R) / R;
          acc += v[r] * cd(std::cos(a), std::sin(a));
        }
        out[k] = acc;
      }
      for (int k = 0; k < R; ++k) v[k] = out[k];
    }
  }

  template <int R, int S>
  void stage(const cd* __restrict a, cd* __restrict b, int ns, const cd* __restrict tws) const {
    const int nb = n / R;
    for (int j0 = 0; j0 < nb; j0 += ns) {   // blocks of ns butterflies share (j / ns)
      const int base0 = j0 * R;
      for (int k = 0; k < ns; ++k) {
        const int j = j0 + k;
        cd v[R];
        for (int r = 0; r < R; ++r) v[r] = a[j + r * nb];
        for (int r = 1; r < R; ++r) {
          const cd tw = tws[(size_t)k * (R - 1) + r - 1];
          v[r] *= S < 0 ? tw : std::conj(tw);
        }
        dft<R, S>(v);
        for (int r = 0; r < R; ++r) b[base0 + k + r * ns] = v[r];
      }
    }
  }

  // Stockham autosort transform of x (length n), y is scratch of length n;
  // sign -1 forward, +1 inverse (unnormalized)
  void run(cd* x, cd* y, int sign) const {
    cd* a = x;
    cd* b = y;
    int ns = 1;
    for (size_t st = 0; st < rad.size(); ++st) {
      const cd* tws = stw[st].data();
      switch (rad[st] * sign) {
        case -4: stage<4, -1>(a, b, ns, tws); break;
        case 4: stage<4, 1>(a, b, ns, tws); break;
        case -2: stage<2, -1>(a, b, ns, tws); break;
        case 2: stage<2, 1>(a, b, ns, tws); break;
        case -3: stage<3, -1>(a, b, ns, tws); break;
        case 3: stage<3, 1>(a, b, ns, tws); break;
        case -5: stage<5, -1>(a, b, ns, tws); break;
        default: stage<5, 1>(a, b, ns, tws); break;
      }
      std::swap(a, b);
      ns *= rad[st];
    }
    if (a != x) std::memcpy(x, a, sizeof(cd) * n);
  }
};

double lind_g(double eta, double lam, double mu) {   // LindG, MathFunctions.f90:440-528
  if (eta < 0.0) return 0.0;
  if (eta < 1e-10) return 1.0 - lam + eta * eta * (1.0 / 3.0 - 3.0 * mu);
  if (std::fabs(eta - 1.0) < 1e-10) return 2.0 - lam - 3.0 * mu + 20.0 * (eta - 1.0);
  if (eta > 3.65) {
    static const double c[15] = {-0.13714285714285712, -6.39999999999999875e-2, -3.77825602968460128e-2,
                                 -2.51824061652633074e-2, -1.80879839616166146e-2, -1.36715733124818332e-2,
                                 -1.07236045520990083e-2, -8.65192783339199453e-3, -7.1372762502456763e-3,
                                 -5.9945117538835746e-3, -5.10997527675418131e-3, -4.41060829979912465e-3,
                                 -3.84763737842981233e-3, -3.38745061493813488e-3, -3.00624946457977689e-3};
    const double e2 = eta * eta, i2 = 1.0 / e2;
    double acc = c[14];
    for (int k = 13; k >= 0; --k) acc = c[k] + i2 * acc;
    return 3.0 * (1.0 - mu) * e2 - lam - 0.6 + i2 * acc;
  }
  return 1.0 / (0.5 + 0.25 * (1.0 - eta * eta) * std::log((1.0 + eta) / std::fabs(1.0 - eta)) / eta) -
         3.0 * mu * eta * eta - lam;
}

}  // namespace

struct ofd_cpu {
  int n1, n2, n3, M, K1;
  long long nr, nc;
  double rb[9];
  double vol, dV;
  double lam, mu, alpha, beta, coef, tkf;
  Fft1 fx, fy, fz;
  std::vector<double> q2;    // |G|^2 on the half grid
  std::vector<double> wtk;   // coef * LindG(|G| / tkf)
  std::vector<cd> wsplit;   // exp(-2 pi i k / n1), k = 0..M (R2C split / C2R merge)
  std::vector<cd> spec[4];
  std::vector<double> real[8];

  // forward R2C of a real grid into spec (no 1/N)
  void forward(const double* in, std::vector<cd>& S, int nt) {
    const int n1_ = n1, M_ = M, K1_ = K1;
    (void)n1_;
#pragma omp parallel num_threads(nt)
    {
      std::vector<cd> z(std::max(M_, std::max(n2, n3)) * 4), y(z.size());
#pragma omp for schedule(static)
      for (long long row = 0; row < (long long)n2 * n3; ++row) {
        const double* x = in + row * n1_;
        for (int k = 0; k < M_; ++k) z[k] = cd(x[2 * k], x[2 * k + 1]);
        fx.run(z.data(), y.data(), -1);
        cd* out = S.data() + row * K1_;
        for (int k = 0; k <= M_; ++k) {
          const cd zk = z[k == M_ ? 0 : k], zm = std::conj(z[k == 0 ? 0 : M_ - k]);
          const cd e = 0.5 * (zk + zm), o = 0.5 * (zk - zm);   // O * i
          out[k] = e - cd(0.0, 1.0) * wsplit[k] * o;
        }
      }
      column_pass(S, n2, (long long)K1_, (long long)n2 * K1_, n3, fy, -1, z, y);
      column_pass(S, n3, (long long)n2 * K1_, (long long)K1_, n2, fz, -1, z, y);
    }
  }
  // backward C2R of spec into a real grid (unnormalized)
  void backward(std::vector<cd>& S, double* outr, int nt) {
    const int n1_ = n1, M_ = M, K1_ = K1;
#pragma omp parallel num_threads(nt)
    {
      std::vector<cd> z(std::max(M_, std::max(n2, n3)) * 4), y(z.size());
      column_pass(S, n3, (long long)n2 * K1_, (long long)K1_, n2, fz, +1, z, y);
      column_pass(S, n2, (long long)K1_, (long long)n2 * K1_, n3, fy, +1, z, y);
#pragma omp for schedule(static)
      for (long long row = 0; row < (long long)n2 * n3; ++row) {
        const cd* X = S.data() + row * K1_;
        for (int k = 0; k < M_; ++k) {
          cd xk = X[k], xm = X[M_ - k];
          if (k == 0) { xk = cd(xk.real(), 0.0); xm = cd(xm.real(), 0.0); }
          z[k] = (xk + std::conj(xm)) + cd(0.0, 1.0) * std::conj(wsplit[k]) * (xk - std::conj(xm));
        }
        fx.run(z.data(), y.data(), +1);
        double* x = outr + row * n1_;
        for (int k = 0; k < M_; ++k) {
          x[2 * k] = z[k].real();
          x[2 * k + 1] = z[k].imag();
        }
      }
    }
  }
  // transforms of length n along a column stride cs, for every (outer, k1):
  // outer index o in [0, no) at stride os. One slab (fixed o: n rows of K1
  // contiguous values) is gathered into a thread's buffer, its K1 columns
  // are transformed there (the slab stays in cache), and it is scattered back.
  void column_pass(std::vector<cd>& S, int n, long long cs, long long os, int no, const Fft1& f, int sign,
                   std::vector<cd>& z, std::vector<cd>& y) {
    thread_local std::vector<cd> slab;
    slab.resize((size_t)n * K1);
#pragma omp for schedule(static)
    for (long long o = 0; o < no; ++o) {
      cd* base = S.data() + o * os;
      for (int i = 0; i < n; ++i) std::memcpy(&slab[(size_t)i * K1], base + (size_t)i * cs, sizeof(cd) * K1);
      for (int k = 0; k < K1; ++k) {
        for (int i = 0; i < n; ++i) z[i] = slab[(size_t)i * K1 + k];
        f.run(z.data(), y.data(), sign);
        for (int i = 0; i < n; ++i) slab[(size_t)i * K1 + k] = z[i];
      }
      for (int i = 0; i < n; ++i) std::memcpy(base + (size_t)i * cs, &slab[(size_t)i * K1], sizeof(cd) * K1);
    }
  }
};

extern "C" int ofd_cpu_create(int n1, int n2, int n3, const double* cell, double lam, double mu, double alpha,
                              double beta, double rho0, ofd_cpu** out) {
  if (!cell || !out || n1 % 2 || n1 < 2 || n2 < 1 || n3 < 1) return -1;
  ofd_cpu* c = new ofd_cpu();
  c->n1 = n1; c->n2 = n2; c->n3 = n3;
  c->M = n1 / 2; c->K1 = n1 / 2 + 1;
  c->nr = (long long)n1 * n2 * n3;
  c->nc = (long long)c->K1 * n2 * n3;
  if (!c->fx.init(c->M) || !c->fy.init(n2) || !c->fz.init(n3)) {
    delete c;
    return -2;
  }
  // lattice A (columns = vectors); G = 2 pi A^-T m (CellInfo.f90:281)
  double a[3][3];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) a[i][j] = cell[3 * j + i];
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double inv[3][3];
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  for (int cc = 0; cc < 3; ++cc)
    for (int j = 0; j < 3; ++j) c->rb[3 * cc + j] = inv[j][cc] * 2.0 * kPi;
  c->vol = std::fabs(det);
  c->dV = c->vol / (double)c->nr;
  c->lam = lam; c->mu = mu;
  c->alpha = alpha < 0.0 ? 5.0 / 6.0 : alpha;
  c->beta = beta < 0.0 ? c->alpha : beta;
  // FillWT (KEDF_WTkernel.f90:81-136)
  c->coef = 5.0 / (9.0 * c->alpha * c->beta * std::pow(rho0, c->alpha + c->beta - 5.0 / 3.0));
  c->tkf = 2.0 * std::pow(3.0 * rho0 * kPi * kPi, 1.0 / 3.0);
  c->q2.resize(c->nc);
  c->wtk.resize(c->nc);
  for (int i3 = 0; i3 < n3; ++i3) {
    const int m3 = i3 > n3 / 2 ? i3 - n3 : i3;
    for (int i2 = 0; i2 < n2; ++i2) {
      const int m2 = i2 > n2 / 2 ? i2 - n2 : i2;
      for (int m1 = 0; m1 < c->K1; ++m1) {
        const double gx = c->rb[0] * m1 + c->rb[1] * m2 + c->rb[2] * m3;
        const double gy = c->rb[3] * m1 + c->rb[4] * m2 + c->rb[5] * m3;
        const double gz = c->rb[6] * m1 + c->rb[7] * m2 + c->rb[8] * m3;
        const long long g = ((long long)i3 * n2 + i2) * c->K1 + m1;
        c->q2[g] = gx * gx + gy * gy + gz * gz;
        c->wtk[g] = c->coef * lind_g(std::sqrt(c->q2[g]) / c->tkf, lam, mu);
      }
    }
  }
  c->wsplit.resize(c->K1);
  for (int k = 0; k < c->K1; ++k) {
    const long double a = -2.0L * 3.141592653589793238462643383279L * k / n1;
    c->wsplit[k] = cd((double)cosl(a), (double)sinl(a));
  }
  for (int k = 0; k < 4; ++k) c->spec[k].resize(c->nc);
  for (int k = 0; k < 8; ++k) c->real[k].resize(c->nr);
  *out = c;
  return 0;
}

extern "C" void ofd_cpu_destroy(ofd_cpu* c) { delete c; }

// One evaluation: potential (dE/dphi with opt_sqrt, else dE/drho) and eTable(9)
// (Hartree; CalculateEnergy's slots). nthreads OpenMP threads.
extern "C" int ofd_cpu_calpot(ofd_cpu* c, const double* rho, const double* vion, double e_ionion, int opt_sqrt,
                              int nthreads, double* pot, double* et) {
  if (!c || !rho || !vion || !pot) return -1;
  const int nt = nthreads > 0 ? nthreads : 1;
  const long long nr = c->nr, nc = c->nc;
  const double invn = 1.0 / (double)nr, al = c->alpha, be = c->beta;
  double* ra = c->real[0].data();
  double* rbv = c->real[1].data();
  double* ph = c->real[2].data();
  double* c13 = c->real[7].data();
  const bool wt56 = al == 5.0 / 6.0 && be == 5.0 / 6.0;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long long i = 0; i < nr; ++i) {
    const double r = rho[i];
    ph[i] = std::sqrt(r);
    c13[i] = std::cbrt(r);
    ra[i] = wt56 ? ph[i] * c13[i] : std::pow(r, al);   // rho^(5/6) = rho^(1/2) rho^(1/3)
    rbv[i] = be == al ? ra[i] : std::pow(r, be);
  }
  // the four forward transforms: rho, rho^beta, rho^alpha, sqrt rho
  c->forward(rho, c->spec[0], nt);
  c->forward(rbv, c->spec[1], nt);
  c->forward(ra, c->spec[2], nt);
  c->forward(ph, c->spec[3], nt);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long long g = 0; g < nc; ++g) {
    const double q2 = c->q2[g];
    c->spec[0][g] *= q2 > 0.0 ? 4.0 * kPi / q2 * invn : 0.0;   // Hartree, G = 0 removed
    c->spec[1][g] *= c->wtk[g] * invn;
    c->spec[2][g] *= c->wtk[g] * invn;
    c->spec[3][g] *= q2 * invn;                                 // vW: G^2 F[sqrt rho]
  }
  double* vh = c->real[3].data();
  c->backward(c->spec[0], vh, nt);
  double* p1 = c->real[4].data();
  double* p2 = c->real[5].data();
  double* tvw = c->real[6].data();
  c->backward(c->spec[1], p1, nt);
  c->backward(c->spec[2], p2, nt);
  c->backward(c->spec[3], tvw, nt);
  double s_ie = 0, s_h = 0, s_xc = 0, s_tf = 0, s_vw = 0, s_nl = 0;
  const double lam = c->lam, mu = c->mu;
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : s_ie, s_h, s_xc, s_tf, s_vw, s_nl)
  for (long long i = 0; i < nr; ++i) {
    const double r = rho[i];
    // LDA (XC_LDA.f90:51-310)
    const double r13 = c13[i];
    const double rs = 0.62035049089940009 / r13;
    double vxc, ec;
    if (rs < 1.0) {
      const double lg = std::log(rs);
      vxc = -0.98474502184269641 * r13 + lg * (3.11e-2 + 1.33333333333333333e-3 * rs) - 5.83666666666666666e-2 -
            8.4e-3 * rs;
      ec = lg * (3.11e-2 + 2e-3 * rs) - 4.8e-2 - 1.16e-2 * rs;
    } else {
      const double sq = std::sqrt(rs);
      const double den = 1.0 + 1.0529 * sq + 3.334e-1 * rs;
      vxc = -0.98474502184269641 * r13 + (-1.423e-1 - 0.174798948333333333 * sq - 6.325709333333333333e-2 * rs) /
                                             (den * den);
      ec = -1.423e-1 / den;
    }
    const double r23 = r13 * r13;
    const double vtf = 5.0 / 3.0 * kCTF * lam * r23;
    const double vnl = kCTF * (al * (ra[i] / r) * p1[i] + be * (rbv[i] / r) * p2[i]);
    double p = vion[i] + vh[i] + vxc + vtf + vnl;
    const double phi = ph[i];
    if (opt_sqrt) p = 2.0 * phi * p + mu * tvw[i];
    else p = p + mu * tvw[i] / (2.0 * phi);
    pot[i] = p;
    s_ie += r * vion[i];
    s_h += vh[i] * r;
    s_xc += -0.73855876638202234 * r * r13 + r * ec;
    s_tf += r * r23;
    s_vw += phi * tvw[i];
    s_nl += ra[i] * p1[i];
  }
  if (et) {
    double loc[9] = {0};
    loc[2] = s_ie;
    loc[3] = s_h / 2.0;
    loc[4] = s_xc;
    loc[6] = lam * kCTF * s_tf;
    loc[7] = 0.5 * s_vw * mu;
    loc[8] = kCTF * s_nl;
    loc[1] = loc[6] + loc[7] + loc[8];
    for (int k = 0; k < 9; ++k) et[k] = loc[k] * c->dV;
    et[5] = e_ionion;
    et[0] = et[1] + et[2] + et[3] + et[4] + et[5];
  }
  return 0;
}
