// Per-geometry tables on the device:
//  * WGC kernels K1..K4 on the half grid from the 1D eta table
//    (FillWGC_ReciprocalSpace, KEDF_WGCkernel.f90:109-277)
//  * the ionic potential from cardinal B-splines (IonElectronPotRecipSpline,
//    IonElectronSpline.f90:83-216) and the ion-electron forces
//    (IonElectronForcesSpline / CalculateSplineForces, :219-355, :642-753)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ofd_internal.h"
#include "ofd_point.h"

namespace ofd {

constexpr double kPi = 3.141592653589793238462643;

// rvec_bracket (MathSplines.f90:498-545): bisection, 0-based left/right
__device__ __forceinline__ void bracket(const double* __restrict__ t, int n, double x, int& left, int& right) {
  int i1 = 1, i3 = n, i2 = (i3 + i1) / 2;
  while (i2 != i1) {
    const double x2 = t[i2 - 1];
    if (x == x2) {
      left = i2 - 1;
      right = i2;
      return;
    } else if (x < x2) {
      i3 = i2;
    } else if (x > x2) {
      i1 = i2;
    }
    i2 = (i3 + i1) / 2;
  }
  if (x == t[i3 - 1]) { left = i3 - 1; right = i3; return; }
  if (x == t[i1 - 1]) { left = i1 - 1; right = i1; return; }
  left = i1 - 1;
  right = i3 - 1;
}

// spline_cubic_val (MathSplines.f90:395-470), value only
__device__ __forceinline__ double spline_val(const double* __restrict__ t, const double* __restrict__ y,
                                             const double* __restrict__ ypp, int left, int right, double x) {
  const double dt = x - t[left];
  const double h = t[right] - t[left];
  const double tmp1 = (y[right] - y[left]) / h;
  const double tmp2 = -(ypp[right] / 6.0 + ypp[left] / 3.0) * h;
  const double tmp3 = ypp[right] - ypp[left];
  return y[left] + dt * (tmp1 + tmp2 + dt * (0.5 * ypp[left] + dt * (tmp3 / (6.0 * h))));
}

__device__ __forceinline__ double gnorm2(const double* rb, int m1, int m2, int m3) {
  const double gx = rb[0] * m1 + rb[1] * m2 + rb[2] * m3;
  const double gy = rb[3] * m1 + rb[4] * m2 + rb[5] * m3;
  const double gz = rb[6] * m1 + rb[7] * m2 + rb[8] * m3;
  return gx * gx + gy * gy + gz * gz;
}

struct RB {
  double v[9];
};

__global__ void k_wgc_fill(GridDims d, RB rb, const double* __restrict__ eta, const double* __restrict__ w,
                           const double* __restrict__ iw1, const double* __restrict__ iw2,
                           const double* __restrict__ pp0, const double* __restrict__ pp1,
                           const double* __restrict__ pp2, int nk, double tkfs, double gamma, double rhoS,
                           double* __restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nc) return;
  const int k1 = (int)(idx % d.K1);
  const long long rest = idx / d.K1;
  const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const int m3 = i3 > d.n3 / 2 ? i3 - d.n3 : i3;
  const double q = sqrt(gnorm2(rb.v, k1, m2, m3));
  const double e = q / tkfs;
  int l, r;
  bracket(eta, nk, e, l, r);
  const double K1 = spline_val(eta, w, pp0, l, r, e);
  const double s1 = spline_val(eta, iw1, pp1, l, r, e);
  const double s2 = spline_val(eta, iw2, pp2, l, r, e);
  out[idx] = K1;
  out[d.nc + idx] = -s1 / (6.0 * rhoS);
  out[2 * d.nc + idx] = (s2 + (7 - gamma) * s1) / (36.0 * rhoS * rhoS);
  out[3 * d.nc + idx] = (s2 + (1 + gamma) * s1) / (36.0 * rhoS * rhoS);
}

// The Wang-Teter kernel on the half grid, coef * LindG(|G| / 2kF) / N
// (FillWT_ReciprocalSpace, KEDF_WTkernel.f90:81-208, with the forward
// transform's 1/N folded in), in the layout of the half spectra.
__global__ void k_wt_fill(GridDims d, RB rb, double coef, double tkf, double lam, double mu, double* __restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nc) return;
  const int k1 = (int)(idx % d.K1);
  const long long rest = idx / d.K1;
  const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const int m3 = i3 > d.n3 / 2 ? i3 - d.n3 : i3;
  const double q2 = gnorm2(rb.v, k1, m2, m3);
  out[idx] = coef * lind_g(sqrt(q2) / tkf, lam, mu) * (1.0 / (double)d.nr);
}

int build_wt_kernel(const ShapePlan& sp, const double* rb, double coef, double tkf, double lam, double mu,
                    double* d_out, hipStream_t st) {
  RB r;
  for (int i = 0; i < 9; ++i) r.v[i] = rb[i];
  const long long n = sp.d.nc;
  hipLaunchKernelGGL(k_wt_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sp.d, r, coef, tkf, lam, mu,
                     d_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int build_wgc_kernels(const ShapePlan& sp, const double* rb, const double* d_eta, const double* d_w,
                      const double* d_w1, const double* d_w2, const double* d_pp0, const double* d_pp1,
                      const double* d_pp2, int nknots, double tkfs, double gamma, double rhoS, double* d_out,
                      hipStream_t st) {
  RB r;
  for (int i = 0; i < 9; ++i) r.v[i] = rb[i];
  const long long n = sp.d.nc;
  hipLaunchKernelGGL(k_wgc_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sp.d, r, d_eta, d_w, d_w1, d_w2,
                     d_pp0, d_pp1, d_pp2, nknots, tkfs, gamma, rhoS, d_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------------ B-splines
// GetCardinalBSpline (CBSpline.f90:42-90): s[0..order], s[k] for k = 1..order
template <int MAXO>
__device__ __forceinline__ void card_bspline(double* s, int order, double x) {
  for (int i = 0; i <= order; ++i) s[i] = 0.0;
  s[1] = x;
  s[2] = 1.0 - x;
  for (int n = 3; n <= order; ++n) {
    for (int k = 0; k < n; ++k) {
      s[n - k] = ((x + (double)(n - k - 1)) * s[n - k] + ((double)(k + 1) - x) * s[n - k - 1]) / (double)(n - 1);
    }
  }
}

constexpr int kMaxOrder = 16;

// FillQIonTable (IonElectronSpline.f90:537-638), as a gather with a fixed
// summation order (no atomics, so the table is bitwise reproducible).
// Ion i at fractional x has, along each axis of n points, the base
// b = (-floor(n x)) mod n and covers the points (b + j) mod n, j < order,
// with weight M[j + 1] (M the cardinal B-spline of n x - floor(n x)). Its
// contribution to point l is w_i * S_1 S_2 S_3 with S = sum of M[j + 1] over
// j < order, j = (l - b) mod n (+ n, + 2n, ... when n < order).
// Prep: one thread per ion writes its three M tables (order + 1 each).
__global__ void k_qion_prep(GridDims d, const double* __restrict__ frac, const int* __restrict__ ions, int nion,
                            int order, double* __restrict__ mtab) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nion) return;
  const int ion = ions[k];
  const int n[3] = {d.n1, d.n2, d.n3};
  for (int a = 0; a < 3; ++a) {
    const double g = frac[3 * ion + a] * n[a];
    double m[kMaxOrder + 1];
    card_bspline<kMaxOrder>(m, order, g - floor(g));
    for (int j = 0; j <= order; ++j) mtab[((size_t)k * 3 + a) * (kMaxOrder + 1) + j] = m[j];
  }
}

struct QBins {
  int nb[3], bs[3];   // bins and points per bin along each axis
};

// Candidate bins along one axis for point l: the bins holding the bases
// (l - j) mod n, j < order, in the order met (at most three distinct).
__device__ __forceinline__ int qion_axis_bins(int l, int n, int order, int nb, int bs, int* out) {
  if (nb == 1) {
    out[0] = 0;
    return 1;
  }
  int c = 0;
  for (int j = 0; j < order; ++j) {
    int b = l - j;
    if (b < 0) b += n;
    const int bin = b / bs;
    if (c == 0 || out[c - 1] != bin) out[c++] = bin;
  }
  return c;
}

__device__ __forceinline__ double qion_axis_weight(int l, int b, int n, int order, const double* __restrict__ m) {
  int j = l - b;
  if (j < 0) j += n;
  double s = 0.0;
  for (; j < order; j += n) s += m[j + 1];
  return s;
}

// One thread per grid point (x fastest): bins in (z, y, x) order, ions of a
// bin in index order.
__global__ void __launch_bounds__(256) k_qion_gather(GridDims d, QBins qb, const int* __restrict__ bin_start,
                                                     const int* __restrict__ bin_ion, const int* __restrict__ base,
                                                     const double* __restrict__ weight,
                                                     const double* __restrict__ mtab, int order,
                                                     double* __restrict__ q) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nr) return;
  const int l1 = (int)(idx % d.n1);
  const long long r = idx / d.n1;
  const int l2 = (int)(r % d.n2), l3 = (int)(r / d.n2);
  int c1[3], c2[3], c3[3];
  const int n1b = qion_axis_bins(l1, d.n1, order, qb.nb[0], qb.bs[0], c1);
  const int n2b = qion_axis_bins(l2, d.n2, order, qb.nb[1], qb.bs[1], c2);
  const int n3b = qion_axis_bins(l3, d.n3, order, qb.nb[2], qb.bs[2], c3);
  constexpr int MS = kMaxOrder + 1;
  double acc = 0.0;
  for (int a3 = 0; a3 < n3b; ++a3)
    for (int a2 = 0; a2 < n2b; ++a2)
      for (int a1 = 0; a1 < n1b; ++a1) {
        const int bin = (c3[a3] * qb.nb[1] + c2[a2]) * qb.nb[0] + c1[a1];
        for (int k = bin_start[bin]; k < bin_start[bin + 1]; ++k) {
          const int i = bin_ion[k];
          const double* m = mtab + (size_t)i * 3 * MS;
          const double s1 = qion_axis_weight(l1, base[3 * i], d.n1, order, m);
          if (s1 == 0.0) continue;
          const double s2 = qion_axis_weight(l2, base[3 * i + 1], d.n2, order, m + MS);
          if (s2 == 0.0) continue;
          const double s3 = qion_axis_weight(l3, base[3 * i + 2], d.n3, order, m + 2 * MS);
          acc += weight[i] * (s1 * s2 * s3);
        }
      }
  q[idx] = acc;
}

// The B-spline table of ions ion0 .. ion0 + nion - 1 (host fractional
// coordinates h_frac, device copy d_frac) with per-ion weights h_w, written
// to every point of q. The bins are built on the host: the base of an ion is
// -floor(n x) mod n, the same IEEE product and floor the device evaluates.
int qion_table(const GridDims& d, const double* h_frac, const double* d_frac, int ion0, int nion,
               const double* h_w, int order, double* q, hipStream_t st) {
  const int n[3] = {d.n1, d.n2, d.n3};
  QBins qb;
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 2 * order) {
      qb.nb[a] = 1;
      qb.bs[a] = n[a];
    } else {
      qb.bs[a] = order;
      qb.nb[a] = (n[a] + order - 1) / order;
    }
  }
  const int nbins = qb.nb[0] * qb.nb[1] * qb.nb[2];
  std::vector<int> base(3 * (size_t)std::max(nion, 1)), binof(std::max(nion, 1)), start(nbins + 1, 0),
      order_ions(std::max(nion, 1)), ions(std::max(nion, 1));
  for (int k = 0; k < nion; ++k) {
    const int ion = ion0 + k;
    ions[k] = ion;
    int bi[3];
    for (int a = 0; a < 3; ++a) {
      const double g = h_frac[3 * (size_t)ion + a] * n[a];
      long long b = -(long long)std::floor(g);
      b %= n[a];
      if (b < 0) b += n[a];
      base[3 * k + a] = (int)b;
      bi[a] = (int)b / qb.bs[a];
    }
    binof[k] = (bi[2] * qb.nb[1] + bi[1]) * qb.nb[0] + bi[0];
    ++start[binof[k] + 1];
  }
  for (int b = 0; b < nbins; ++b) start[b + 1] += start[b];
  {
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int k = 0; k < nion; ++k) order_ions[fill[binof[k]]++] = k;   // stable: index order within a bin
  }
  std::vector<double> w(std::max(nion, 1));
  for (int k = 0; k < nion; ++k) w[k] = h_w[k];
  const size_t MS = kMaxOrder + 1;
  int* di = (int*)dev_alloc(sizeof(int) * ((size_t)nbins + 1 + 5 * (size_t)std::max(nion, 1)));
  double* dm = (double*)dev_alloc(sizeof(double) * (3 * MS + 1) * (size_t)std::max(nion, 1));
  if (!di || !dm) {
    dev_free(di);
    dev_free(dm);
    return -1;
  }
  int* d_start = di;
  int* d_binion = d_start + nbins + 1;
  int* d_base = d_binion + std::max(nion, 1);
  int* d_ions = d_base + 3 * std::max(nion, 1);
  double* d_w = dm;
  double* d_mtab = dm + std::max(nion, 1);
  hipMemcpyAsync(d_start, start.data(), sizeof(int) * (nbins + 1), hipMemcpyHostToDevice, st);
  hipMemcpyAsync(d_binion, order_ions.data(), sizeof(int) * std::max(nion, 1), hipMemcpyHostToDevice, st);
  hipMemcpyAsync(d_base, base.data(), sizeof(int) * 3 * std::max(nion, 1), hipMemcpyHostToDevice, st);
  hipMemcpyAsync(d_ions, ions.data(), sizeof(int) * std::max(nion, 1), hipMemcpyHostToDevice, st);
  hipMemcpyAsync(d_w, w.data(), sizeof(double) * std::max(nion, 1), hipMemcpyHostToDevice, st);
  if (nion > 0)
    hipLaunchKernelGGL(k_qion_prep, dim3((unsigned)((nion + 127) / 128)), dim3(128), 0, st, d, d_frac, d_ions, nion,
                       order, d_mtab);
  hipLaunchKernelGGL(k_qion_gather, dim3((unsigned)((d.nr + 255) / 256)), dim3(256), 0, st, d, qb, d_start,
                     d_binion, d_base, d_w, d_mtab, order, q);
  // the host vectors above must outlive the copies
  const hipError_t e = hipStreamSynchronize(st);
  dev_free(di);
  dev_free(dm);
  return e == hipSuccess && hipGetLastError() == hipSuccess ? 0 : -1;
}

struct PspDev {
  const double* vqs;
  const double* dd;
  double maxg, v0, spacing;
  int npts, charge;
};

// PspRecipCubicSpline (LocalPseudoPot.f90:136-198)
__device__ __forceinline__ double psp_lookup(const PspDev& p, double q) {
  if (q == 0.0) return p.v0;
  if (q >= p.maxg) return 0.0;
  const double pos = q / p.spacing + 1.0;
  int left = (int)floor(pos);
  if (left > p.npts - 1) left = p.npts - 1;   // q just below maxG rounding up
  const double dt = pos - (double)left;
  const double yl = p.vqs[left - 1], yr = p.vqs[left];
  const double pl = p.dd[left - 1], pr = p.dd[left];
  const double yval = yl + dt * (yr - yl - (pr / 6.0 + pl / 3.0) + dt * (0.5 * pl + dt * ((pr - pl) / 6.0)));
  return yval - 4.0 * kPi * (double)p.charge / pow(q, 2.0);
}

// acc(G) (+)= V_loc(|G|) * SF(G) / N  (SF = unnormalized R2C of the table)
__global__ void k_ion_accum(GridDims d, RB rb, PspDev p, const double2* __restrict__ sf, double2* __restrict__ acc,
                            int first) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nc) return;
  const int k1 = (int)(idx % d.K1);
  const long long rest = idx / d.K1;
  const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const int m3 = i3 > d.n3 / 2 ? i3 - d.n3 : i3;
  const double q = sqrt(gnorm2(rb.v, k1, m2, m3));
  const double v = psp_lookup(p, q);
  const double invn = 1.0 / (double)d.nr;
  double2 s = sf[idx];
  double2 a = first ? make_double2(0.0, 0.0) : acc[idx];
  a.x += v * (s.x * invn);
  a.y += v * (s.y * invn);
  acc[idx] = a;
}

// V(G) = conj(acc * b1 b2 b3) / vol * N   (IonElectronSpline.f90:200-203)
__global__ void k_ion_finish(GridDims d, const double2* __restrict__ b1, const double2* __restrict__ b2,
                             const double2* __restrict__ b3, double vol, double2* __restrict__ acc) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nc) return;
  const int k1 = (int)(idx % d.K1);
  const long long rest = idx / d.K1;
  const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
  double2 a = acc[idx];
  auto mul = [](double2 x, double2 y) { return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x); };
  a = mul(mul(mul(a, b1[k1]), b2[i2]), b3[i3]);
  const double sc = (double)d.nr / vol;
  acc[idx] = make_double2(a.x / vol * (double)d.nr, -a.y / vol * (double)d.nr);
  (void)sc;
}

// rhoRecip * pspRecip * b1 b2 b3 (IonElectronSpline.f90:315-331), G = 0 set to 0
__global__ void k_force_recip(GridDims d, RB rb, PspDev p, const double2* __restrict__ rho_g,
                              const double2* __restrict__ b1, const double2* __restrict__ b2,
                              const double2* __restrict__ b3, double2* __restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nc) return;
  const int k1 = (int)(idx % d.K1);
  const long long rest = idx / d.K1;
  const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const int m3 = i3 > d.n3 / 2 ? i3 - d.n3 : i3;
  double v = idx == 0 ? 0.0 : psp_lookup(p, sqrt(gnorm2(rb.v, k1, m2, m3)));
  const double invn = 1.0 / (double)d.nr;
  double2 r = rho_g[idx];
  double2 a = make_double2(r.x * invn * v, r.y * invn * v);
  auto mul = [](double2 x, double2 y) { return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x); };
  out[idx] = mul(mul(mul(a, b1[k1]), b2[i2]), b3[i3]);
}

// CalculateSplineForces (IonElectronSpline.f90:642-753): one workgroup per ion,
// fixed-order block reduction of the order^3 stencil.
__global__ void __launch_bounds__(256) k_force_gather(GridDims d, const double* __restrict__ frac, int ion0,
                                                      int order, RB cinv, const double* __restrict__ factor,
                                                      double* __restrict__ forces) {
  const int ion = ion0 + blockIdx.x;
  const double g1 = frac[3 * ion] * d.n1, g2 = frac[3 * ion + 1] * d.n2, g3 = frac[3 * ion + 2] * d.n3;
  const double f1 = floor(g1), f2 = floor(g2), f3 = floor(g3);
  double M1[kMaxOrder + 1], M2[kMaxOrder + 1], M3[kMaxOrder + 1];
  double D1[kMaxOrder + 1], D2[kMaxOrder + 1], D3[kMaxOrder + 1];
  card_bspline<kMaxOrder>(M1, order, g1 - f1);
  card_bspline<kMaxOrder>(M2, order, g2 - f2);
  card_bspline<kMaxOrder>(M3, order, g3 - f3);
  card_bspline<kMaxOrder>(D1, order - 1, g1 - f1);
  card_bspline<kMaxOrder>(D2, order - 1, g2 - f2);
  card_bspline<kMaxOrder>(D3, order - 1, g3 - f3);
  D1[order] = 0.0; D2[order] = 0.0; D3[order] = 0.0;
  auto md = [](long long a, int n) { long long r = a % n; return (int)(r < 0 ? r + n : r); };
  double acc[3] = {0.0, 0.0, 0.0};
  const int tot = order * order * order;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int iz = e / (order * order), rem = e % (order * order);
    const int iy = rem / order, ix = rem % order;
    const int l3 = md((long long)f3 - iz, d.n3);
    const int l2 = md((long long)f2 - iy, d.n2);
    const int l1 = md((long long)f1 - ix, d.n1);
    const double q1 = M2[iy + 1] * M3[iz + 1] * (double)d.n1 * (D1[ix + 1] - D1[ix]);
    const double q2 = M1[ix + 1] * M3[iz + 1] * (double)d.n2 * (D2[iy + 1] - D2[iy]);
    const double q3 = M1[ix + 1] * M2[iy + 1] * (double)d.n3 * (D3[iz + 1] - D3[iz]);
    const double fac = factor[((size_t)l3 * d.n2 + l2) * d.n1 + l1];
    // splineForces -= Vecmul(cellInv, qIonDeriv) * Factor
    for (int c = 0; c < 3; ++c) acc[c] -= (cinv.v[3 * c] * q1 + cinv.v[3 * c + 1] * q2 + cinv.v[3 * c + 2] * q3) * fac;
  }
  __shared__ double red[4][3];
  for (int c = 0; c < 3; ++c) {
    double v = acc[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[c] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int c = 0; c < 3; ++c) red[wv][c] = acc[c];
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w][threadIdx.x];
    forces[3 * ion + threadIdx.x] = s;
  }
}

// FillB (CBSpline.f90:93-196) on the host in long double
static void fill_b_host(int order, int size, int tot, std::vector<double2>& b) {
  // M = GetCardinalBSpline(order, 1.0)
  std::vector<long double> s(order + 1, 0.0L);
  const long double x = 1.0L;
  s[1] = x;
  s[2] = 1.0L - x;
  for (int n = 3; n <= order; ++n)
    for (int k = 0; k < n; ++k)
      s[n - k] = ((x + (long double)(n - k - 1)) * s[n - k] + ((long double)(k + 1) - x) * s[n - k - 1]) /
                 (long double)(n - 1);
  const long double tp = 6.283185307179586476925286766559005768L;
  b.resize(size);
  for (int ix = 0; ix < size; ++ix) {
    long double dr = 0, di = 0;
    for (int k = 0; k <= order - 2; ++k) {
      long double a = tp * (long double)ix * (long double)k / (long double)tot;
      dr += s[k + 1] * cosl(a);
      di += s[k + 1] * sinl(a);
    }
    long double a = tp * (long double)(order - 1) * (long double)ix / (long double)tot;
    long double nr = cosl(a), ni = sinl(a);
    long double den = dr * dr + di * di;
    b[ix].x = (double)((nr * dr + ni * di) / den);
    b[ix].y = (double)((ni * dr - nr * di) / den);
  }
}

struct BTables {
  double2 *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
  void make(const GridDims& d, int order, hipStream_t st) {
    std::vector<double2> h1, h2, h3;
    fill_b_host(order, d.K1, d.n1, h1);
    fill_b_host(order, d.n2, d.n2, h2);
    fill_b_host(order, d.n3, d.n3, h3);
    b1 = (double2*)dev_alloc(sizeof(double2) * h1.size());
    b2 = (double2*)dev_alloc(sizeof(double2) * h2.size());
    b3 = (double2*)dev_alloc(sizeof(double2) * h3.size());
    hipMemcpyAsync(b1, h1.data(), sizeof(double2) * h1.size(), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(b2, h2.data(), sizeof(double2) * h2.size(), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(b3, h3.data(), sizeof(double2) * h3.size(), hipMemcpyHostToDevice, st);
    hipStreamSynchronize(st);
  }
  // released after the caller has synchronized the stream
  ~BTables() {
    dev_free(b1);
    dev_free(b2);
    dev_free(b3);
  }
};

static PspDev psp_dev(int s, const double* d_vqs, const double* d_dd, const int* off, const int* npts,
                      const double* maxg, const double* v0, const int* charge) {
  PspDev p;
  p.vqs = d_vqs + off[s];
  p.dd = d_dd + off[s];
  p.maxg = maxg[s];
  p.v0 = v0[s];
  p.npts = npts[s];
  p.spacing = maxg[s] / (double)(npts[s] - 1);
  p.charge = charge[s];
  return p;
}

// scratch layout: [real grid nr doubles][complex sf nc][complex acc nc]
int build_ion_potential(const ShapePlan& sp, const double* rb, double vol, int ntypes, const int* h_type_start,
                        const double* h_frac, const double* d_frac, int nion, const double* d_psp_vqs, const double* d_psp_dd,
                        const int* h_psp_off, const int* h_psp_npts, const double* h_psp_maxg,
                        const double* h_psp_v0, const int* h_psp_charge, int order, double* d_vion,
                        void* d_scratch, hipStream_t st) {
  const GridDims& d = sp.d;
  if (order < 2 || order > kMaxOrder) return -2;
  RB r;
  for (int i = 0; i < 9; ++i) r.v[i] = rb[i];
  double* q = (double*)d_scratch;
  double2* sf = (double2*)(q + d.nr);
  double2* acc = sf + d.nc;
  ReqDesc D;
  std::memset(&D, 0, sizeof(D));
  ReqDesc* dd;
  dd = (ReqDesc*)dev_alloc(sizeof(ReqDesc));
  if (!dd) return -1;
  const unsigned gb = (unsigned)((d.nc + 255) / 256);
  for (int s = 0; s < ntypes; ++s) {
    const int i0 = h_type_start[s], n_s = h_type_start[s + 1] - i0;
    {
      const std::vector<double> ones(std::max(n_s, 1), 1.0);
      if (qion_table(d, h_frac, d_frac, i0, n_s, ones.data(), order, q, st)) {
        dev_free(dd);
        return -1;
      }
    }
    D.u = q;
    D.work = sf;
    hipMemcpyAsync(dd, &D, sizeof(D), hipMemcpyHostToDevice, st);
    run_fft3d_plain(sp, 1, dd, 1, st);
    PspDev p = psp_dev(s, d_psp_vqs, d_psp_dd, h_psp_off, h_psp_npts, h_psp_maxg, h_psp_v0, h_psp_charge);
    hipLaunchKernelGGL(k_ion_accum, dim3(gb), dim3(256), 0, st, d, r, p, sf, acc, s == 0 ? 1 : 0);
  }
  BTables bt;
  bt.make(d, order, st);
  hipLaunchKernelGGL(k_ion_finish, dim3(gb), dim3(256), 0, st, d, bt.b1, bt.b2, bt.b3, vol, acc);
  D.work = acc;
  D.out = d_vion;
  hipMemcpyAsync(dd, &D, sizeof(D), hipMemcpyHostToDevice, st);
  run_fft3d_plain(sp, 0, dd, 1, st);
  hipError_t e = hipStreamSynchronize(st);
  dev_free(dd);
  return e == hipSuccess ? 0 : -1;
}

int ion_forces_spline(const ShapePlan& sp, const double* rb, const double* cell_inv_t, int ntypes,
                      const int* h_type_start, const double* d_frac, int nion, const double* d_psp_vqs,
                      const double* d_psp_dd, const int* h_psp_off, const int* h_psp_npts, const double* h_psp_maxg,
                      const double* h_psp_v0, const int* h_psp_charge, int order, const double* d_rho,
                      double* d_forces, void* d_scratch, hipStream_t st) {
  const GridDims& d = sp.d;
  if (order < 2 || order > kMaxOrder) return -2;
  RB r, ci;
  for (int i = 0; i < 9; ++i) {
    r.v[i] = rb[i];
    ci.v[i] = cell_inv_t[i];
  }
  double* factor = (double*)d_scratch;
  double2* rg = (double2*)(factor + d.nr);
  double2* tmp = rg + d.nc;
  ReqDesc D;
  std::memset(&D, 0, sizeof(D));
  ReqDesc* dd;
  dd = (ReqDesc*)dev_alloc(sizeof(ReqDesc));
  if (!dd) return -1;
  D.u = d_rho;
  D.work = rg;
  hipMemcpyAsync(dd, &D, sizeof(D), hipMemcpyHostToDevice, st);
  run_fft3d_plain(sp, 1, dd, 1, st);
  BTables bt;
  bt.make(d, order, st);
  const unsigned gb = (unsigned)((d.nc + 255) / 256);
  for (int s = 0; s < ntypes; ++s) {
    const int i0 = h_type_start[s], n_s = h_type_start[s + 1] - i0;
    PspDev p = psp_dev(s, d_psp_vqs, d_psp_dd, h_psp_off, h_psp_npts, h_psp_maxg, h_psp_v0, h_psp_charge);
    hipLaunchKernelGGL(k_force_recip, dim3(gb), dim3(256), 0, st, d, r, p, rg, bt.b1, bt.b2, bt.b3, tmp);
    D.work = tmp;
    D.out = factor;
    hipMemcpyAsync(dd, &D, sizeof(D), hipMemcpyHostToDevice, st);
    run_fft3d_plain(sp, 0, dd, 1, st);
    if (n_s > 0)
      hipLaunchKernelGGL(k_force_gather, dim3(n_s), dim3(256), 0, st, d, d_frac, i0, order, ci, factor, d_forces);
  }
  hipError_t e = hipStreamSynchronize(st);
  dev_free(dd);
  return e == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------------ Ewald
// Real-space Ewald sum and forces (EwaldRealSpace Ewald.f90:599-775,
// EwaldRealSpaceForces :1313-1390): one workgroup per ion i, threads over the
// partners j; per pair the wrapped fractional separation (ANINT) and the
// reference's shell loop over lattice images (nr = 1, 2, ... until a shell
// adds no image within the cutoff). Writes per ion i: e[i] = sum_j q_j sum
// erfc(eta d)/d and f[3i..] = q_i eta^3 sum_j q_j sum dvec (erfc(eta d)/(eta d)^3
// + 2/sqrt(pi) exp(-(eta d)^2)/(eta d)^2); fixed-order workgroup sums.
__global__ void __launch_bounds__(256) k_ewald_real(int nion, const double* __restrict__ frac,
                                                    const double* __restrict__ charge, RB lat, RB rmet,
                                                    double eta, double rc2, double* __restrict__ e_out,
                                                    double* __restrict__ f_out) {
  const int i = blockIdx.x;
  const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
  const double two_over_sqrt_pi = 2.0 / sqrt(kPi);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < nion; j += blockDim.x) {
    double fr[3] = {fi0 - frac[3 * j], fi1 - frac[3 * j + 1], fi2 - frac[3 * j + 2]};
    for (int c = 0; c < 3; ++c) fr[c] -= round(fr[c]);   // ANINT: halves away from zero
    const double qj = charge[j];
    double se = 0.0, sf[3] = {0.0, 0.0, 0.0};
    for (int nr = 1;; ++nr) {
      bool newr = false;
      for (int ia = -nr; ia <= nr; ++ia)
        for (int ib = -nr; ib <= nr; ++ib)
          for (int ic = -nr; ic <= nr; ++ic) {
            if (!(abs(ia) == nr || abs(ib) == nr || abs(ic) == nr || nr == 1)) continue;
            const double r1 = (double)ia + fr[0], r2 = (double)ib + fr[1], r3 = (double)ic + fr[2];
            const double d2 = rmet.v[0] * r1 * r1 + rmet.v[4] * r2 * r2 + rmet.v[8] * r3 * r3 +
                              2.0 * (rmet.v[3] * r2 * r1 + rmet.v[7] * r3 * r2 + rmet.v[6] * r3 * r1);
            if (d2 <= rc2 && d2 > 1e-40) {
              newr = true;
              const double d = sqrt(d2);
              se += erfc(eta * d) / d;
              const double ed = eta * d;
              const double w = erfc(ed) / (ed * ed * ed) + two_over_sqrt_pi * exp(-(ed * ed)) / (ed * ed);
              // dvec = cellReal . r (columns of lat are the lattice vectors)
              sf[0] += (lat.v[0] * r1 + lat.v[3] * r2 + lat.v[6] * r3) * w;
              sf[1] += (lat.v[1] * r1 + lat.v[4] * r2 + lat.v[7] * r3) * w;
              sf[2] += (lat.v[2] * r1 + lat.v[5] * r2 + lat.v[8] * r3) * w;
            }
          }
      if (!newr) break;
    }
    acc[0] += qj * se;
    acc[1] += qj * sf[0];
    acc[2] += qj * sf[1];
    acc[3] += qj * sf[2];
  }
  __shared__ double red[4][4];
  for (int c = 0; c < 4; ++c) {
    double v = acc[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[c] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int c = 0; c < 4; ++c) red[wv][c] = acc[c];
  __syncthreads();
  if (threadIdx.x < 4) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[w][threadIdx.x];
    if (threadIdx.x == 0) e_out[i] = v;
    else f_out[3 * i + threadIdx.x - 1] = v * charge[i] * eta * eta * eta;
  }
}

// Reciprocal particle-mesh energy terms (EwaldRecipSpaceSpline Ewald.f90:
// 1081-1147): per half-grid point, cc = S(G)/N b1 b2 b3 with S the
// unnormalized R2C of the charge table; w |cc|^2 exp(-G^2/4eta^2)/G^2 for
// 0 < G^2 <= kc^2 (w = 1 on the k1 = 0 plane, 2 elsewhere); per-workgroup
// fixed-order partial sums.
__global__ void __launch_bounds__(256) k_ewald_recip_e(GridDims d, RB rb, const double2* __restrict__ sf,
                                                       const double2* __restrict__ b1,
                                                       const double2* __restrict__ b2,
                                                       const double2* __restrict__ b3, double eta, double kc2,
                                                       double* __restrict__ part) {
  double acc[1] = {0.0};
  const double invn = 1.0 / (double)d.nr;
  auto mul = [](double2 x, double2 y) { return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x); };
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < d.nc;
       idx += (long long)gridDim.x * blockDim.x) {
    const int k1 = (int)(idx % d.K1);
    const long long rest = idx / d.K1;
    const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
    const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
    const int m3 = i3 > d.n3 / 2 ? i3 - d.n3 : i3;
    const double qn = sqrt(gnorm2(rb.v, k1, m2, m3));
    const double q2 = qn * qn;
    if (q2 <= kc2 && q2 > 0.0) {
      double2 s = sf[idx];
      s = mul(mul(mul(make_double2(s.x * invn, s.y * invn), b1[k1]), b2[i2]), b3[i3]);
      const double s2 = s.x * s.x + s.y * s.y;
      acc[0] += (k1 == 0 ? 1.0 : 2.0) * exp(-q2 / (eta * eta) / 4.0) / q2 * s2;
    }
  }
  __shared__ double red[4];
  double v = acc[0];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

// Reciprocal particle-mesh force grid (EwaldRecipSpaceForcesSpline Ewald.f90:
// 1720-1760): conj(S/N) |b1|^2 |b2|^2 |b3|^2 exp(-G^2/4eta^2)/G^2 for
// 0 < G^2 <= kc^2, 0 elsewhere; its C2R is the force factor.
__global__ void k_ewald_recip_f(GridDims d, RB rb, const double2* __restrict__ sf,
                                const double2* __restrict__ b1, const double2* __restrict__ b2,
                                const double2* __restrict__ b3, double eta, double kc2,
                                double2* __restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= d.nc) return;
  const int k1 = (int)(idx % d.K1);
  const long long rest = idx / d.K1;
  const int i2 = (int)(rest % d.n2), i3 = (int)(rest / d.n2);
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const int m3 = i3 > d.n3 / 2 ? i3 - d.n3 : i3;
  const double q2 = gnorm2(rb.v, k1, m2, m3);
  double2 r = make_double2(0.0, 0.0);
  if (idx != 0 && q2 <= kc2) {
    const double invn = 1.0 / (double)d.nr;
    const double2 a = b1[k1], b = b2[i2], c = b3[i3];
    const double bb = (a.x * a.x + a.y * a.y) * (b.x * b.x + b.y * b.y) * (c.x * c.x + c.y * c.y);
    const double g = bb * exp(-q2 / (eta * eta) / 4.0) / q2;
    const double2 s = sf[idx];
    r = make_double2(s.x * invn * g, -s.y * invn * g);
  }
  out[idx] = r;
}

// scratch layout: [real grid nr doubles][complex sf nc][complex tmp nc]
int ewald_device(const ShapePlan& sp, const double* rb, const double* lat, double vol, int ntypes,
                 const int* h_type_start, const double* h_charge, const double* h_frac, const double* d_frac,
                 int nion, int order,
                 double eta, double rc, double kc, const double* cell_inv_t, double* comps4, double* h_forces,
                 void* d_scratch, hipStream_t st) {
  const GridDims& d = sp.d;
  if (order < 2 || order > kMaxOrder) return -2;
  RB r, L, rm, ci;
  for (int i = 0; i < 9; ++i) {
    r.v[i] = rb[i];
    L.v[i] = lat[i];
    ci.v[i] = cell_inv_t ? cell_inv_t[i] : 0.0;
  }
  // Metric3d (MathFunctions.f90:745-749): rmet(a,b) = sum_k cellReal(k,a) cellReal(k,b)
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) rm.v[3 * a + b] = lat[3 * a] * lat[3 * b] + lat[3 * a + 1] * lat[3 * b + 1] + lat[3 * a + 2] * lat[3 * b + 2];
  std::vector<double> q(nion);
  for (int s = 0; s < ntypes; ++s)
    for (int k = h_type_start[s]; k < h_type_start[s + 1]; ++k) q[k] = h_charge[s];
  double* qd = (double*)dev_alloc(sizeof(double) * nion);
  double* eo = (double*)dev_alloc(sizeof(double) * nion);
  double* fo = (double*)dev_alloc(sizeof(double) * 3 * nion);
  double* fr = (double*)dev_alloc(sizeof(double) * 3 * nion);
  const int npart = 512;
  double* part = (double*)dev_alloc(sizeof(double) * npart);
  ReqDesc* dd = (ReqDesc*)dev_alloc(sizeof(ReqDesc));
  if (!qd || !eo || !fo || !fr || !part || !dd) {
    dev_free(qd); dev_free(eo); dev_free(fo); dev_free(fr); dev_free(part); dev_free(dd);
    return -1;
  }
  hipMemcpyAsync(qd, q.data(), sizeof(double) * nion, hipMemcpyHostToDevice, st);
  hipLaunchKernelGGL(k_ewald_real, dim3(nion), dim3(256), 0, st, nion, d_frac, qd, L, rm, eta, rc * rc, eo, fo);
  // charge table, R2C
  double* tab = (double*)d_scratch;
  double2* sf = (double2*)(tab + d.nr);
  double2* tmp = sf + d.nc;
  if (qion_table(d, h_frac, d_frac, 0, nion, q.data(), order, tab, st)) {
    dev_free(qd); dev_free(eo); dev_free(fo); dev_free(fr); dev_free(part); dev_free(dd);
    return -1;
  }
  ReqDesc D;
  std::memset(&D, 0, sizeof(D));
  D.u = tab;
  D.work = sf;
  hipMemcpyAsync(dd, &D, sizeof(D), hipMemcpyHostToDevice, st);
  run_fft3d_plain(sp, 1, dd, 1, st);
  BTables bt;
  bt.make(d, order, st);
  hipLaunchKernelGGL(k_ewald_recip_e, dim3(npart), dim3(256), 0, st, d, r, sf, bt.b1, bt.b2, bt.b3, eta, kc * kc,
                     part);
  if (h_forces) {
    const unsigned gb = (unsigned)((d.nc + 255) / 256);
    hipLaunchKernelGGL(k_ewald_recip_f, dim3(gb), dim3(256), 0, st, d, r, sf, bt.b1, bt.b2, bt.b3, eta, kc * kc,
                       tmp);
    ReqDesc D2;
    std::memset(&D2, 0, sizeof(D2));
    D2.work = tmp;
    D2.out = tab;
    // the first descriptor is consumed by the forward transform above, in
    // stream order before this copy
    hipMemcpyAsync(dd, &D2, sizeof(D2), hipMemcpyHostToDevice, st);
    run_fft3d_plain(sp, 0, dd, 1, st);
    for (int s = 0; s < ntypes; ++s) {
      const int i0 = h_type_start[s], n_s = h_type_start[s + 1] - i0;
      if (n_s > 0)
        hipLaunchKernelGGL(k_force_gather, dim3(n_s), dim3(256), 0, st, d, d_frac, i0, order, ci, tab, fr);
    }
  }
  std::vector<double> he(nion), hf(3 * (size_t)nion), hr(3 * (size_t)nion), hp(npart);
  hipMemcpyAsync(he.data(), eo, sizeof(double) * nion, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(hf.data(), fo, sizeof(double) * 3 * nion, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(hp.data(), part, sizeof(double) * npart, hipMemcpyDeviceToHost, st);
  if (h_forces) hipMemcpyAsync(hr.data(), fr, sizeof(double) * 3 * nion, hipMemcpyDeviceToHost, st);
  const hipError_t e = hipStreamSynchronize(st);
  dev_free(qd); dev_free(eo); dev_free(fo); dev_free(fr); dev_free(part); dev_free(dd);
  if (e != hipSuccess) return -1;
  double real = 0.0, rec = 0.0, sq = 0.0, sumq = 0.0;
  for (int k = 0; k < nion; ++k) {
    real += q[k] * he[k];
    sq += q[k] * q[k];
    sumq += q[k];
  }
  for (int k = 0; k < npart; ++k) rec += hp[k];
  comps4[0] = real / 2.0;
  comps4[1] = -eta / std::sqrt(kPi) * sq;                                                // :778-812
  comps4[2] = rec * 2.0 * kPi / vol * (double)d.nr * (double)d.nr;
  comps4[3] = -0.5 * kPi / (eta * eta) / vol * sumq * sumq;                              // :1150-1190
  if (h_forces) {
    const double sc = 4.0 * kPi / vol * (double)d.nr;
    for (int k = 0; k < nion; ++k)
      for (int c = 0; c < 3; ++c) h_forces[3 * k + c] = hf[3 * k + c] + hr[3 * k + c] * q[k] * sc;
  }
  return 0;
}

}  // namespace ofd
