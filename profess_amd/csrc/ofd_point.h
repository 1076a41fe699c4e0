// Pointwise math of the evaluation passes, shared by the generic and the
// fast kernels:
//   LDA Perdew-Zunger, numSpin = 1: potential XC_LDA.f90:199-310, energy
//   density XC_LDA.f90:51-196 (the constants below are the formula's);
//   LSDA for numSpin = 2 (Perdew-Zunger and Perdew-Wang 92) and the
//   two-channel epilogue of the spin kinds;
//   LindG, MathFunctions.f90:440-528 (including its 15 Taylor terms for
//   eta > 3.65); the rho powers of the Wang-Teter terms.
// rho^(1/3), rho^(-1/3), rs and sqrt(rs) are computed once per point, all
// from one rho^(-1/6), and shared between the kinetic and the
// exchange-correlation terms.
#pragma once
#include <hip/hip_runtime.h>

#include "ofd_internal.h"

namespace ofd {

// Global-memory pointers. The buffers of a request reach the kernels through
// the descriptor array, so the compiler cannot tell their address space and
// emits flat loads and stores, which also count against the LDS wait counter
// (every wait for an LDS exchange then waits for the outstanding HBM loads
// too). Accesses through gptr compile to global_load / global_store.
template <class T> using gptr = T __attribute__((address_space(1)))*;
template <class T> __device__ __forceinline__ gptr<T> glb(T* p) { return (gptr<T>)p; }
// Complex (double2 / float2) elements through a gptr, member by member: the
// HIP vector types cannot be copy-constructed from an address-space-1
// reference, and the two members still merge into one 8- or 16-byte access.
template <class C> __device__ __forceinline__ C ldc(gptr<const C> p) {
  C r;
  r.x = p->x;
  r.y = p->y;
  return r;
}
template <class C> __device__ __forceinline__ C ldc(gptr<C> p) {
  C r;
  r.x = p->x;
  r.y = p->y;
  return r;
}
template <class C> __device__ __forceinline__ void stc(gptr<C> p, const C& v) {
  p->x = v.x;
  p->y = v.y;
}

struct PtVals {
  double phi;          // rho^(1/2)
  double r13;          // rho^(1/3)
  double ra, ram1;     // rho^alpha, rho^(alpha-1)
  double rb, rbm1;     // rho^beta,  rho^(beta-1)
  double vxc;          // LDA potential
  double exc;          // LDA energy density times rho
};

// rho^(-1/6) to FP64 accuracy: an FP32 estimate (relative error below 1e-6)
// refined by two Newton steps on y^-6 = rho, y <- y (7 - rho y^6) / 6, whose
// error maps e -> -3.5 e^2. Positive, normal rho (as densities are).
__device__ __forceinline__ double rho_m16(double rho) {
  // hardware log2 / exp2 (v_log_f32, v_exp_f32) give the seed
  const float g = __builtin_amdgcn_exp2f(-__builtin_amdgcn_logf((float)rho) * (1.0f / 6.0f));
  double y = (double)g;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double y2 = y * y;
    const double y6 = y2 * y2 * y2;
    y = y * (7.0 - rho * y6) * (1.0 / 6.0);
  }
  return y;
}

__device__ __forceinline__ void point_vals(double rho, double alpha, double beta, PtVals& p) {
  constexpr double cC = 0.62035049089940009;       // (3/(4 pi))^(1/3)
  constexpr double sqrt_cC = 0.78762331789974330;  // sqrt(cC)
  constexpr double cXp = -0.98474502184269641, cXe = -0.73855876638202234;
  // every power from one rho^(-1/6): rho^(-1/3) = y^2, rho^(1/3) = rho y^4,
  // rho^(1/2) = rho y^3, rho^(5/6) = rho y
  const double i16 = rho_m16(rho);
  const double i13 = i16 * i16;
  p.r13 = rho * (i13 * i13);
  p.phi = rho * (i13 * i16);
  const double rs = cC * i13;
  const double sq = sqrt_cC * i16;   // sqrt(rs)
  if (alpha == 5.0 / 6.0 && beta == 5.0 / 6.0) {
    p.ram1 = i16;
    p.ra = rho * i16;
    p.rb = p.ra;
    p.rbm1 = i16;
  } else {
    // one log2 for both exponents; rho^(a-1) = rho^a / rho
    const double l2 = log2(rho);
    p.ra = exp2(alpha * l2);
    p.rb = beta == alpha ? p.ra : exp2(beta * l2);
    p.ram1 = p.ra / rho;
    p.rbm1 = p.rb / rho;
  }
  double ec;
  if (rs < 1.0) {
    const double lg = log(rs);
    p.vxc = cXp * p.r13 + lg * (3.11e-2 + 1.33333333333333333e-3 * rs) - 5.83666666666666666e-2 - 8.4e-3 * rs;
    ec = lg * (3.11e-2 + 2e-3 * rs) - 4.8e-2 - 1.16e-2 * rs;
  } else {
    const double iden = 1.0 / (1.0 + 1.0529 * sq + 3.334e-1 * rs);
    p.vxc = cXp * p.r13 + (-1.423e-1 - 0.174798948333333333 * sq - 6.325709333333333333e-2 * rs) * iden * iden;
    ec = -1.423e-1 * iden;
  }
  p.exc = cXe * rho * p.r13 + rho * ec;
}

// Spin-polarized LDA at one point (numSpin = 2): both channels' potentials and
// the energy density. lsda 1: LSDAPotPZ (XC_LDA.f90:417-547, Perdew-Zunger
// correlation in Rydberg, halved with the doubled exchange); lsda 2:
// LSDAPotPW92 (XC_LDA.f90:313-414, spn_gcor :550-592). The constants are the
// reference's parametrization.
__device__ __forceinline__ void pw92_g(double a, double a1, double b1, double b2, double b3, double b4, double rs,
                                       double rs12, double& g, double& grs) {
  // p = 1: rs^p = rs, rs^(3/2) = rs12^3
  const double q0 = -2.0 * a * (1.0 + a1 * rs);
  const double q1 = 2.0 * a * (b1 * rs12 + b2 * rs + b3 * (rs12 * rs12 * rs12) + b4 * rs * rs);
  const double q2 = log(1.0 + 1.0 / q1);
  g = q0 * q2;
  const double q3 = a * (b1 / rs12 + 2.0 * b2 + 3.0 * b3 * rs12 + 4.0 * b4 * rs);
  grs = -2.0 * a * a1 * q2 - q0 * q3 / (q1 * q1 + q1);
}

__device__ __forceinline__ void lsda_point(double r1, double r2, int lsda, double& v1, double& v2, double& e) {
  constexpr double ax = -0.7385588, fzz = 1.709921, gam = 0.5198421;
  constexpr double p75vpi = 0.75 / 3.141592653589793238462643;
  const double ex1 = ax * cbrt(2.0 * r1), ex2 = ax * cbrt(2.0 * r2);
  const double vx1 = ex1 * (4.0 / 3.0), vx2 = ex2 * (4.0 / 3.0);
  const double d = r1 + r2;
  const double rs = cbrt(p75vpi / d);
  const double zet = (r1 - r2) / d;
  const double cp = cbrt(1.0 + zet), cm = cbrt(1.0 - zet);
  const double f = ((1.0 + zet) * cp + (1.0 - zet) * cm - 2.0) / gam;
  const double fz = (4.0 / 3.0) * (cp - cm) / gam;
  if (lsda == 2) {
    const double rs12 = sqrt(rs);
    double eu, eurs, ep, eprs, alfm, alfrsm;
    pw92_g(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, rs, rs12, eu, eurs);
    pw92_g(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, rs, rs12, ep, eprs);
    pw92_g(0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671, rs, rs12, alfm, alfrsm);
    const double z3 = zet * zet * zet, z4 = z3 * zet;
    const double ec = eu * (1.0 - f * z4) + ep * f * z4 - alfm * f * (1.0 - z4) / fzz;
    const double ecrs = eurs * (1.0 - f * z4) + eprs * f * z4 - alfrsm * f * (1.0 - z4) / fzz;
    const double eczet = 4.0 * z3 * f * (ep - eu + alfm / fzz) + fz * (z4 * ep - z4 * eu - (1.0 - z4) * alfm / fzz);
    const double comm = ec - rs * ecrs / 3.0 - zet * eczet;
    v1 = vx1 + comm + eczet;
    v2 = vx2 + comm - eczet;
    e = r1 * ex1 + r2 * ex2 + ec * d;
    return;
  }
  double eu, vc, es, vs;
  if (rs < 1.0) {
    const double lrs = log(rs);
    eu = 0.0622 * lrs - 0.0960 + (0.0040 * lrs - 0.0232) * rs;
    vc = eu - (0.0622 + (0.0040 * lrs + 0.0040 - 0.0232) * rs) / 3.0;
    es = 0.0311 * lrs - 0.0538 + (0.0014 * lrs - 0.0096) * rs;
    vs = es - (0.0311 + (0.0014 * lrs + 0.0014 - 0.0096) * rs) / 3.0;
  } else {
    const double srs = sqrt(rs);
    eu = -0.2846 / (1.0 + 1.0529 * srs + 0.3334 * rs);
    vc = eu * eu * (1.0 + 7.0 * 1.0529 * srs / 6.0 + 4.0 / 3.0 * 0.3334 * rs) / -0.2846;
    es = -0.1686 / (1.0 + 1.3981 * srs + 0.2611 * rs);
    vs = es * es * (1.0 + 7.0 * 1.3981 * srs / 6.0 + 4.0 / 3.0 * 0.2611 * rs) / -0.1686;
  }
  const double base = vc + f * (vs - vc);
  const double vc2 = base - (es - eu) * (1.0 + zet) * fz;
  const double vc1 = base + (es - eu) * (1.0 - zet) * fz;
  const double ec = eu + f * (es - eu);
  v1 = 0.5 * (2.0 * vx1 + vc1);
  v2 = 0.5 * (2.0 * vx2 + vc2);
  e = 0.5 * (2.0 * (r1 * ex1 + r2 * ex2) + ec * d);
}

// Spin kinds: both channels' potentials at one point and the energy sums.
// Per channel s (x_s = 2 rho_s): ion + Hartree of the total, the LSDA
// potential, TF and Wang-Teter on x_s (WTPotentialPlus(2 rho_s)), the chain
// rule 2 sqrt(rho_s) under optSqrt, and vW on sqrt(x_s) divided by sqrt(2)
// (CalVW, KEDF_VW.f90:100-140); the kinetic energies are halved sums over the
// channels (KEDF_TF.f90:66-74, CalPotential.f90:400-407).
template <int KIND, int NS>
__device__ __forceinline__ void epi_point_spin(const ReqDesc& D, double r1, double r2, double vion, const double* o,
                                               double (&acc)[NS], double& p1, double& p2) {
  constexpr double kCTF = 2.87123400018819;   // KEDF_TF.f90:42
  constexpr double kInvSqrt2 = 0.70710678118654752440;
  double vx[2], exc;
  lsda_point(r1, r2, D.lsda, vx[0], vx[1], exc);
  const double common = vion + o[0];
  const double rs[2] = {r1, r2};
  double pot[2], etf = 0.0, evw = 0.0, enl = 0.0;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double x = 2.0 * rs[s];
    PtVals p;
    point_vals(x, D.alpha, D.beta, p);
    const double r23 = p.r13 * p.r13;
    const double vtf = (5.0 / 3.0 * kCTF * D.lam) * r23;
    double vnl, t;
    if (KIND == kWT3S) {
      vnl = kCTF * (D.alpha * p.ram1 * o[1 + s] + D.beta * p.rbm1 * o[1 + s]);
      enl += p.ra * o[1 + s];
      t = o[3 + s];
    } else {
      vnl = kCTF * (D.alpha * p.ram1 * o[1 + s] + D.beta * p.rbm1 * o[3 + s]);
      enl += p.ra * o[1 + s];
      t = o[5 + s];
    }
    const double loc = common + vx[s] + vtf + vnl;
    const double tw = D.mu * t * kInvSqrt2;
    const double sq = p.phi * kInvSqrt2;   // sqrt(rho_s)
    pot[s] = D.opt_sqrt ? 2.0 * sq * loc + tw : loc + tw / (2.0 * sq);
    etf += x * r23;
    evw += p.phi * t;
  }
  const double rho = r1 + r2;
  acc[0] += rho * vion;
  acc[1] += o[0] * rho;
  acc[2] += exc;
  acc[3] += 0.5 * etf;
  acc[4] += 0.5 * evw;
  acc[5] += 0.5 * enl;
  p1 = pot[0];
  p2 = pot[1];
}

// Powers of a density with a general exponent are taken as exp2(a log2 rho)
// (positive rho): the relative error is about |a log2 rho| ulp, below 1e-14
// for the densities and exponents here, at a fraction of the cost of pow.

// rho^alpha and rho^beta for the forward fields
__device__ __forceinline__ void fwd_pows(double rho, double alpha, double beta, double& ra, double& rb) {
  if (alpha == 5.0 / 6.0 && beta == 5.0 / 6.0) {
    ra = rho * rho_m16(rho);
    rb = ra;
  } else {
    const double l2 = log2(rho);
    ra = exp2(alpha * l2);
    rb = beta == alpha ? ra : exp2(beta * l2);
  }
}

__device__ __forceinline__ double lind_g(double eta, double lam, double mu) {
  if (eta < 0.0) return 0.0;
  if (eta < 1e-10) return 1.0 - lam + eta * eta * (1.0 / 3.0 - 3.0 * mu);
  if (fabs(eta - 1.0) < 1e-10) return 2.0 - lam - 3.0 * mu + 20.0 * (eta - 1.0);
  if (eta > 3.65) {
    const double e2 = eta * eta, i2 = 1.0 / e2;
    double acc = -3.00624946457977689e-3;
    acc = -3.38745061493813488e-3 + i2 * acc;
    acc = -3.84763737842981233e-3 + i2 * acc;
    acc = -4.41060829979912465e-3 + i2 * acc;
    acc = -5.10997527675418131e-3 + i2 * acc;
    acc = -5.9945117538835746e-3 + i2 * acc;
    acc = -7.1372762502456763e-3 + i2 * acc;
    acc = -8.65192783339199453e-3 + i2 * acc;
    acc = -1.07236045520990083e-2 + i2 * acc;
    acc = -1.36715733124818332e-2 + i2 * acc;
    acc = -1.80879839616166146e-2 + i2 * acc;
    acc = -2.51824061652633074e-2 + i2 * acc;
    acc = -3.77825602968460128e-2 + i2 * acc;
    acc = -6.39999999999999875e-2 + i2 * acc;
    acc = -0.13714285714285712 + i2 * acc;
    return 3.0 * (1.0 - mu) * e2 - lam - 0.6 + i2 * acc;
  }
  return 1.0 / (0.5 + 0.25 * (1.0 - eta * eta) * log((1.0 + eta) / fabs(1.0 - eta)) / eta) - 3.0 * mu * eta * eta -
         lam;
}

__device__ __forceinline__ double fsign(double x) { return signbit(x) ? -1.0 : 1.0; }

// Workgroups are handed to the eight XCDs round-robin (id mod 8), and each
// XCD has its own L2. Renumbering so that every XCD gets one contiguous range
// of logical blocks keeps the cache lines shared by neighbouring blocks
// (line-segment boundaries of the column passes) inside one L2.
__device__ __forceinline__ int xcd_block(int w, int n) {
  const int q = n >> 3, r = n & 7, x = w & 7;
  return x * q + min(x, r) + (w >> 3);
}


}  // namespace ofd
