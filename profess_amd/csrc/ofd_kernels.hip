// HIP kernels for one PROFESS energy + potential evaluation on gfx950.
//
// One evaluation = five passes over the grid, batched over requests that share
// a grid shape (blockIdx.y/z indexes the request descriptor):
//   1 k_xfwd   reads rho (or phi = ca*u + cb*v and squares it), forms the NF
//              real fields of the functional pointwise, runs the x-direction
//              R2C (half-length complex FFT + split) in LDS, writes NF half rows
//   2 k_cpass  complex FFT along y, KB adjacent x-columns per workgroup
//   3 k_zgop   complex FFT along z, the G-space operator (Hartree 4 pi/G^2,
//              Wang-Teter kernel K(|G|) on the fly or the four WGC kernels,
//              vW G^2, the 1/N forward normalization), inverse FFT along z
//   4 k_cpass  inverse complex FFT along y
//   5 k_xinv   x-direction C2R of the NI fields in LDS, then the potential
//              epilogue (ion + Hartree + LDA + TF + nonlocal, chain rule
//              2 sqrt(rho), + mu*vW) and the per-workgroup FP64 energy sums
// CalPotential.f90:41-647 is the reference for what is computed; the dataflow
// (3 forward + 3 inverse transforms for WT with alpha == beta instead of 4+4)
// is this file's own.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ofd_internal.h"
#include "ofd_point.h"

namespace ofd {

constexpr double kPi = 3.141592653589793238462643;
constexpr double kCTF = 2.87123400018819;   // KEDF_TF.f90:42

template <int KIND> struct NFields;
template <> struct NFields<kWT3> { static constexpr int v = 3; };
template <> struct NFields<kWT4> { static constexpr int v = 4; };
template <> struct NFields<kWGC8> { static constexpr int v = 8; };
template <> struct NFields<kWT3S> { static constexpr int v = 5; };
template <> struct NFields<kWT4S> { static constexpr int v = 7; };
constexpr int kPlain = 9;   // one real field, no functional
template <> struct NFields<kPlain> { static constexpr int v = 1; };

template <int KIND>
__device__ __forceinline__ void gen_fields(const ReqDesc& D, double rho, double* f) {
  if (KIND == kPlain) { f[0] = rho; return; }
  double ra, rb;
  fwd_pows(rho, D.alpha, D.beta, ra, rb);
  const double phi = sqrt(rho);
  if (KIND == kWT3) {
    f[0] = rho; f[1] = rb; f[2] = phi;
  } else if (KIND == kWT4) {
    f[0] = rho; f[1] = rb; f[2] = ra; f[3] = phi;
  } else {
    double th = rho - D.rhoS;
    f[0] = rho; f[1] = phi;
    f[2] = ra; f[3] = th * ra; f[4] = th * th * ra / 2.0;
    f[5] = rb; f[6] = th * rb; f[7] = th * th * rb / 2.0;
  }
}

// Spin kinds: the fields of both channels at one point (CalPotential.f90:
// 120-160, 400-407: Hartree on rho_up + rho_dn, the kinetic terms on 2 rho_s).
template <int KIND>
__device__ __forceinline__ void gen_fields_spin(const ReqDesc& D, double r1, double r2, double* f) {
  const double x1 = 2.0 * r1, x2 = 2.0 * r2;
  double ra1, rb1, ra2, rb2;
  fwd_pows(x1, D.alpha, D.beta, ra1, rb1);
  fwd_pows(x2, D.alpha, D.beta, ra2, rb2);
  f[0] = r1 + r2;
  f[1] = rb1;
  f[2] = rb2;
  if (KIND == kWT3S) {
    f[3] = sqrt(x1); f[4] = sqrt(x2);
  } else {
    f[3] = ra1; f[4] = ra2;
    f[5] = sqrt(x1); f[6] = sqrt(x2);
  }
}

// phi_t of one channel: ca*u + cb*v (cb == 0: v may be null)
__device__ __forceinline__ double phi_t(double ca, double cb, const double* u, const double* v, size_t g) {
  return cb != 0.0 ? fma(cb, v[g], ca * u[g]) : ca * u[g];
}

// ------------------------------------------------------------ block reduce
template <int N>
__device__ __forceinline__ void block_sum_store(double (&acc)[N], double* out) {
  // wave (64 lanes) butterfly then across waves through LDS, fixed order
  __shared__ double red[16][N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[k] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) red[wv][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int nw = (blockDim.x + 63) >> 6;
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// block_sum_store with out[k * stride] (partials laid out [slot][workgroup])
template <int N>
__device__ __forceinline__ void block_sum_store_strided(double (&acc)[N], double* out, int stride) {
  __shared__ double red2[16][N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[k] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) red2[wv][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int nw = (blockDim.x + 63) >> 6;
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red2[w][threadIdx.x];
    out[(size_t)threadIdx.x * stride] = s;
  }
}

// ------------------------------------------------------------ pass 1: x R2C
// rows = n2*n3 lines of n1 reals per request; `rows_per` rows per workgroup.
// LDS: 2 buffers of rows_per*NF lines, ld = M+1 complex each.
template <typename R, int KIND, int INMODE>
__global__ void __launch_bounds__(256) k_xfwd(const ReqDesc* __restrict__ descs, GridDims d, Plan1D p,
                                              const typename V2<R>::T* __restrict__ tw,
                                              const typename V2<R>::T* __restrict__ tws, int rows_per) {
  using C = typename V2<R>::T;
  constexpr int NF = NFields<KIND>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  const ReqDesc& D = descs[blockIdx.y];
  const int M = d.M, ld = M + 1, n1 = d.n1;
  const long long nrows = (long long)d.n2 * d.n3;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const long long row0 = (long long)blk * rows_per;
  const int nrow = (int)min((long long)rows_per, nrows - row0);
  const int nl = nrow * NF;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + (size_t)rows_per * NF * ld;
  R* Ar = reinterpret_cast<R*>(A);
  // load + pointwise fields; real x at complex slot x/2, component x&1
  for (int e = threadIdx.x; e < nrow * n1; e += blockDim.x) {
    const int r = e / n1, i1 = e - r * n1;
    const long long g = (row0 + r) * n1 + i1;
    double f[NF];
    if constexpr (is_spin_kind(KIND)) {
      double r1, r2;
      if (INMODE == kInRho) {
        r1 = D.u[g];
        r2 = D.u2[g];
      } else {
        const double t1 = phi_t(D.ca, D.cb, D.u, D.v, g), t2 = phi_t(D.ca2, D.cb2, D.u2, D.v2, g);
        r1 = t1 * t1;
        r2 = t2 * t2;
      }
      gen_fields_spin<KIND>(D, r1, r2, f);
    } else {
      double rho;
      if (INMODE == kInRho) {
        rho = D.u[g];
      } else {
        const double t = D.cb != 0.0 ? fma(D.cb, D.v[g], D.ca * D.u[g]) : D.ca * D.u[g];
        rho = t * t;
      }
      gen_fields<KIND>(D, rho, f);
    }
#pragma unroll
    for (int j = 0; j < NF; ++j) Ar[2 * ((r * NF + j) * ld + (i1 >> 1)) + (i1 & 1)] = (R)f[j];
  }
  __syncthreads();
  C* res = lds_fft<-1>(A, B, p, ld, nl, tw);
  // split: X[k] = (Z[k] + conj Z[M-k])/2 - i W^k (Z[k] - conj Z[M-k])/2
  C* work = reinterpret_cast<C*>(D.work);
  const int K1 = d.K1;
  for (int e = threadIdx.x; e < nl * K1; e += blockDim.x) {
    const int l = e / K1, k = e - l * K1;
    const int r = l / NF, j = l - r * NF;
    const C* z = res + l * ld;
    C zk = z[k == M ? 0 : k];
    C zm = z[k == 0 ? 0 : M - k];
    R er = (R)0.5 * (zk.x + zm.x), ei = (R)0.5 * (zk.y - zm.y);
    R orr = (R)0.5 * (zk.y + zm.y), oi = (R)-0.5 * (zk.x - zm.x);   // O = -i(Z - conj Zm)/2
    C w = tws[k];
    C x;
    x.x = er + (w.x * orr - w.y * oi);
    x.y = ei + (w.x * oi + w.y * orr);
    work[(size_t)j * d.nc + (size_t)(row0 + r) * K1 + k] = x;
  }
}

// ------------------------------------------------------------ passes 2/4: strided complex FFT
// Transforms along an axis with element stride `es` (in complex units) for
// `kb` adjacent x-columns; blockIdx.x = outer*nchunk + chunk, blockIdx.y = field,
// blockIdx.z = request. outer index advances by `os`.
template <typename R, int S>
__global__ void __launch_bounds__(256) k_cpass(const ReqDesc* __restrict__ descs, GridDims d, Plan1D p,
                                               const typename V2<R>::T* __restrict__ tw, int kb, int nchunk,
                                               long long es, long long os) {
  using C = typename V2<R>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const ReqDesc& D = descs[blockIdx.z];
  const int bx = xcd_block(blockIdx.x, gridDim.x);
  const int chunk = bx % nchunk, outer = bx / nchunk;
  const int c0 = chunk * kb;
  const int ncol = min(kb, d.K1 - c0);
  const int n = p.n, ld = n + 1;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + (size_t)kb * ld;
  C* base = reinterpret_cast<C*>(D.work) + (size_t)blockIdx.y * d.nc + (size_t)outer * os + c0;
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    const int i = e / ncol, c = e - i * ncol;
    A[c * ld + i] = base[(size_t)i * es + c];
  }
  __syncthreads();
  C* res = lds_fft<S>(A, B, p, ld, ncol, tw);
  for (int e = threadIdx.x; e < n * ncol; e += blockDim.x) {
    const int i = e / ncol, c = e - i * ncol;
    base[(size_t)i * es + c] = res[c * ld + i];
  }
}

// ------------------------------------------------------------ pass 3: z FFT + G-space operator + z IFFT
template <int KIND>
__device__ __forceinline__ void gop(const ReqDesc& D, const double* __restrict__ wk, size_t gidx, size_t nc,
                                    double q2, double invn, double (&x)[8][2]) {
  // x[f][re/im] in, out overwrite (NI == NF for all kinds)
  const double hart = q2 > 0.0 ? 4.0 * kPi / q2 * invn : 0.0;
  if (KIND == kWT3 || KIND == kWT4) {
    const double K = D.wtk[gidx];
    const double vw = q2 * invn;
    x[0][0] *= hart; x[0][1] *= hart;
    x[1][0] *= K; x[1][1] *= K;
    if (KIND == kWT3) {
      x[2][0] *= vw; x[2][1] *= vw;
    } else {
      x[2][0] *= K; x[2][1] *= K;
      x[3][0] *= vw; x[3][1] *= vw;
    }
  } else if (KIND == kWT3S || KIND == kWT4S) {
    // Hartree on the total; the Wang-Teter kernel on each channel's x^b (and
    // x^a); vW on each channel's sqrt(x)
    constexpr int nk = KIND == kWT3S ? 2 : 4;
    const double K = D.wtk[gidx];
    const double vw = q2 * invn;
    x[0][0] *= hart; x[0][1] *= hart;
#pragma unroll
    for (int f = 1; f <= nk; ++f) { x[f][0] *= K; x[f][1] *= K; }
#pragma unroll
    for (int f = nk + 1; f <= nk + 2; ++f) { x[f][0] *= vw; x[f][1] *= vw; }
  } else if (KIND == kWGC8) {
    const double k1 = wk[gidx] * invn, k2 = wk[nc + gidx] * invn, k3 = wk[2 * nc + gidx] * invn,
                 k4 = wk[3 * nc + gidx] * invn;
    const double vw = q2 * invn;
    x[0][0] *= hart; x[0][1] *= hart;
    x[1][0] *= vw; x[1][1] *= vw;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int b = 2 + 3 * h;   // ra block at 2..4, rb block at 5..7
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        double fr = x[b][c], ft = x[b + 1][c], f2 = x[b + 2][c];
        x[b][c] = k1 * fr + k2 * ft + k3 * f2;
        x[b + 1][c] = k2 * fr + k4 * ft;
        x[b + 2][c] = k3 * fr;
      }
    }
  }
}

template <typename R, int KIND>
__global__ void __launch_bounds__(256) k_zgop(const ReqDesc* __restrict__ descs, GridDims d, Plan1D p,
                                              const typename V2<R>::T* __restrict__ tw, int kb, int nchunk) {
  using C = typename V2<R>::T;
  constexpr int NF = NFields<KIND>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  const ReqDesc& D = descs[blockIdx.y];
  const int bx = xcd_block(blockIdx.x, gridDim.x);
  const int chunk = bx % nchunk, i2 = bx / nchunk;
  const int c0 = chunk * kb;
  const int ncol = min(kb, d.K1 - c0);
  const int n3 = d.n3, ld = n3 + 1;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + (size_t)NF * kb * ld;
  C* work = reinterpret_cast<C*>(D.work);
  const size_t es = (size_t)d.n2 * d.K1;
  const size_t off = (size_t)i2 * d.K1 + c0;
  for (int e = threadIdx.x; e < NF * n3 * ncol; e += blockDim.x) {
    const int f = e / (n3 * ncol);
    const int rem = e - f * n3 * ncol;
    const int i3 = rem / ncol, c = rem - i3 * ncol;
    A[(f * ncol + c) * ld + i3] = work[(size_t)f * d.nc + (size_t)i3 * es + off + c];
  }
  __syncthreads();
  C* res = lds_fft<-1>(A, B, p, ld, NF * ncol, tw);
  C* oth = (res == A) ? B : A;
  const double invn = 1.0 / (double)d.nr;
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  for (int e = threadIdx.x; e < ncol * n3; e += blockDim.x) {
    const int c = e / n3, i3 = e - c * n3;
    const int m1 = c0 + c;
    const int m3 = i3 > n3 / 2 ? i3 - n3 : i3;
    const double gx = D.rb[0] * m1 + D.rb[1] * m2 + D.rb[2] * m3;
    const double gy = D.rb[3] * m1 + D.rb[4] * m2 + D.rb[5] * m3;
    const double gz = D.rb[6] * m1 + D.rb[7] * m2 + D.rb[8] * m3;
    const double q2 = gx * gx + gy * gy + gz * gz;
    double x[8][2];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      C v = res[(f * ncol + c) * ld + i3];
      x[f][0] = (double)v.x;
      x[f][1] = (double)v.y;
    }
    gop<KIND>(D, D.wgck, (size_t)i3 * es + off + c, (size_t)d.nc, q2, invn, x);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      C v;
      v.x = (R)x[f][0];
      v.y = (R)x[f][1];
      res[(f * ncol + c) * ld + i3] = v;
    }
  }
  __syncthreads();
  C* res2 = lds_fft<+1>(res, oth, p, ld, NF * ncol, tw);
  for (int e = threadIdx.x; e < NF * n3 * ncol; e += blockDim.x) {
    const int f = e / (n3 * ncol);
    const int rem = e - f * n3 * ncol;
    const int i3 = rem / ncol, c = rem - i3 * ncol;
    work[(size_t)f * d.nc + (size_t)i3 * es + off + c] = res2[(f * ncol + c) * ld + i3];
  }
}

// ------------------------------------------------------------ pass 5: x C2R + epilogue
template <int KIND>
__device__ __forceinline__ double epi_point(const ReqDesc& D, double rho, double vion, const double* o,
                                            double (&acc)[kNSum]) {
  PtVals p;
  point_vals(rho, D.alpha, D.beta, p);
  const double vxc = p.vxc;
  const double r23 = p.r13 * p.r13;
  const double vtf = (5.0 / 3.0 * kCTF * D.lam) * r23;
  double vnl, t, enl;
  const double vh = o[0];
  if (KIND == kWT3) {
    vnl = kCTF * (D.alpha * p.ram1 * o[1] + D.beta * p.rbm1 * o[1]);
    enl = p.ra * o[1];
    t = o[2];
  } else if (KIND == kWT4) {
    vnl = kCTF * (D.alpha * p.ram1 * o[1] + D.beta * p.rbm1 * o[2]);
    enl = p.ra * o[1];
    t = o[3];
  } else {
    const double th = rho - D.rhoS, a = D.alpha, b = D.beta;
    t = o[1];
    enl = p.rb * (o[2] + th * o[3] + th * th / 2.0 * o[4]);
    vnl = kCTF * (p.rbm1 * (b * o[2] + (rho + b * th) * o[3] + th * (rho + b * th / 2.0) * o[4]) +
                  p.ram1 * (a * o[5] + (rho + a * th) * o[6] + th * (rho + a * th / 2.0) * o[7]));
  }
  const double loc = vion + vh + vxc + vtf + vnl;
  acc[0] += rho * vion;
  acc[1] += vh * rho;
  acc[2] += p.exc;
  acc[3] += rho * r23;
  acc[4] += p.phi * t;
  acc[5] += enl;
  if (D.opt_sqrt) return 2.0 * p.phi * loc + D.mu * t;
  return loc + D.mu * t / (2.0 * p.phi);
}

template <typename R, int KIND, int INMODE, int OUTMODE>
__global__ void __launch_bounds__(256) k_xinv(const ReqDesc* __restrict__ descs, GridDims d, Plan1D p,
                                              const typename V2<R>::T* __restrict__ tw,
                                              const typename V2<R>::T* __restrict__ tws, int rows_per) {
  using C = typename V2<R>::T;
  constexpr int NF = (KIND == kPlain) ? 1 : NFields<KIND>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  const ReqDesc& D = descs[blockIdx.y];
  const int M = d.M, ld = M + 1, n1 = d.n1, K1 = d.K1;
  const long long nrows = (long long)d.n2 * d.n3;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const long long row0 = (long long)blk * rows_per;
  const int nrow = (int)min((long long)rows_per, nrows - row0);
  const int nl = nrow * NF;
  C* A = reinterpret_cast<C*>(smem);
  C* B = A + (size_t)rows_per * NF * ld;
  const C* work = reinterpret_cast<const C*>(D.work);
  for (int e = threadIdx.x; e < nl * K1; e += blockDim.x) {
    const int l = e / K1, k = e - l * K1;
    const int r = l / NF, j = l - r * NF;
    B[l * ld + k] = work[(size_t)j * d.nc + (size_t)(row0 + r) * K1 + k];
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[M-k]) + i W^{-k} (X[k] - conj X[M-k]); Im X[0] = Im X[M] = 0
  for (int e = threadIdx.x; e < nl * M; e += blockDim.x) {
    const int l = e / M, k = e - l * M;
    const C* x = B + l * ld;
    C xk = x[k], xm = x[M - k];
    if (k == 0) { xk.y = (R)0; xm.y = (R)0; }
    R er = xk.x + xm.x, ei = xk.y - xm.y;
    R dr = xk.x - xm.x, di = xk.y + xm.y;
    C w = tws[k];   // W^k; use conj
    R orr = w.x * dr + w.y * di, oi = w.x * di - w.y * dr;
    C z;
    z.x = er - oi;
    z.y = ei + orr;
    A[l * ld + k] = z;
  }
  __syncthreads();
  C* res = lds_fft<+1>(A, B, p, ld, nl, tw);
  const R* rr = reinterpret_cast<const R*>(res);
  double acc[kNSum];
#pragma unroll
  for (int k = 0; k < kNSum; ++k) acc[k] = 0.0;
  for (int e = threadIdx.x; e < nrow * n1; e += blockDim.x) {
    const int r = e / n1, i1 = e - r * n1;
    const size_t g = (size_t)(row0 + r) * n1 + i1;
    double o[8];
#pragma unroll
    for (int j = 0; j < NF; ++j) o[j] = (double)rr[2 * ((r * NF + j) * ld + (i1 >> 1)) + (i1 & 1)];
    if (KIND == kPlain) {
      D.out[g] = o[0];
      continue;
    }
    if constexpr (is_spin_kind(KIND)) {
      double t1 = 0.0, t2 = 0.0, r1, r2;
      if (INMODE == kInRho) {
        r1 = D.u[g];
        r2 = D.u2[g];
      } else {
        t1 = phi_t(D.ca, D.cb, D.u, D.v, g);
        t2 = phi_t(D.ca2, D.cb2, D.u2, D.v2, g);
        r1 = t1 * t1;
        r2 = t2 * t2;
      }
      double p1, p2;
      epi_point_spin<KIND>(D, r1, r2, D.vion[g], o, acc, p1, p2);
      if (OUTMODE == kOutPot) {
        D.out[g] = p1;
        D.out2[g] = p2;
      } else if (OUTMODE == kOutSigned) {
        const double de1 = p1 * fsign(t1), de2 = p2 * fsign(t2);
        D.out[g] = de1;
        D.out2[g] = de2;
        const double dir1 = D.s1 * D.u[g] + (D.v != nullptr ? D.s2 * D.v[g] : 0.0);
        const double dir2 = D.s1b * D.u2[g] + (D.v2 != nullptr ? D.s2b * D.v2[g] : 0.0);
        const double dl1 = de1 - 2.0 * t1 * D.mu0, dl2 = de2 - 2.0 * t2 * D.mu0b;
        acc[6] += de1 * dir1;
        acc[7] += dl1 * dl1 + dl2 * dl2;
        acc[8] += de1 * t1;
        acc[9] += de2 * dir2;
        acc[10] += de2 * t2;
      } else {
        const bool up = D.newton_ch == 0;
        const double tt = up ? t1 : t2;
        const double de = (up ? p1 : p2) * fsign(tt);
        const double ap = ((de - D.mu0 * 2.0 * tt) - D.bvec[g]) / D.eps;
        D.out[g] = ap;
        acc[6] += (up ? D.v[g] : D.v2[g]) * ap;
      }
      continue;
    }
    double tphi = 0.0, rho;
    if (INMODE == kInRho) {
      rho = D.u[g];
    } else {
      tphi = D.cb != 0.0 ? fma(D.cb, D.v[g], D.ca * D.u[g]) : D.ca * D.u[g];
      rho = tphi * tphi;
    }
    double pot = epi_point<KIND>(D, rho, D.vion[g], o, acc);
    if (OUTMODE == kOutPot) {
      D.out[g] = pot;
    } else if (OUTMODE == kOutSigned) {
      const double de = pot * fsign(tphi);
      D.out[g] = de;
      const double dir = D.s1 * D.u[g] + (D.v != nullptr ? D.s2 * D.v[g] : 0.0);
      acc[6] += de * dir;
      const double dl = de - 2.0 * tphi * D.mu0;
      acc[7] += dl * dl;
      acc[8] += de * tphi;
      acc[9] += de * de;
      acc[10] += tphi * tphi;
    } else {
      const double de = pot * fsign(tphi);
      const double ap = ((de - D.mu0 * 2.0 * tphi) - D.bvec[g]) / D.eps;
      D.out[g] = ap;
      acc[6] += D.v[g] * ap;
      if (D.w != nullptr) {
        acc[7] += D.w[g] * ap;
        acc[8] += ap * ap;
      }
    }
  }
  if (KIND != kPlain) block_sum_store_strided<kNSum>(acc, D.sums + blk, gridDim.x);
}

// ------------------------------------------------------------ finalize: sum partials
// One workgroup per (slot, request): fin[slot] = sum of the workgroup
// partials (laid out [slot][workgroup]), eight independent loads in flight
// per thread, fixed order.
__global__ void __launch_bounds__(256) k_finalize(const ReqDesc* __restrict__ descs, int nwg) {
  const ReqDesc& D = descs[blockIdx.y];
  const double* src = D.sums + (size_t)blockIdx.x * nwg;
  double acc[1] = {0.0};
  for (int base = 0; base < nwg; base += 8 * 256) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int w = base + threadIdx.x + 256 * j;
      v[j] = w < nwg ? src[w] : 0.0;
    }
    acc[0] += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  }
  __shared__ double red[4];
  double x = acc[0];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) D.fin[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ============================================================ host launchers
template <typename R>
static const typename V2<R>::T* twp(const ShapePlan& sp, int which);
template <>
const double2* twp<double>(const ShapePlan& sp, int w) {
  return w == 0 ? sp.tw_x : w == 1 ? sp.tw_y : w == 2 ? sp.tw_z : sp.tw_s;
}
template <>
const float2* twp<float>(const ShapePlan& sp, int w) {
  return w == 0 ? sp.tw_x32 : w == 1 ? sp.tw_y32 : w == 2 ? sp.tw_z32 : sp.tw_s32;
}

template <typename R, int KIND, int INMODE, int OUTMODE>
static int eval_t(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st) {
  using C = typename V2<R>::T;
  constexpr int NF = NFields<KIND>::v;
  const GridDims& d = sp.d;
  const Tiles t = tiles_for(sp, NF, (int)sizeof(C));
  size_t lds_x = 2ull * t.rows * NF * (d.M + 1) * sizeof(C);
  hipLaunchKernelGGL((k_xfwd<R, KIND, INMODE>), dim3(t.ntile, nb), dim3(256), lds_x, st, dd, d, sp.px,
                     twp<R>(sp, 0), twp<R>(sp, 3), t.rows);
  const int nchy = (d.K1 + t.kby - 1) / t.kby;
  size_t ldsy = 2ull * t.kby * (d.n2 + 1) * sizeof(C);
  hipLaunchKernelGGL((k_cpass<R, -1>), dim3(nchy * d.n3, NF, nb), dim3(256), ldsy, st, dd, d, sp.py,
                     twp<R>(sp, 1), t.kby, nchy, (long long)d.K1, (long long)d.n2 * d.K1);
  const int nchz = (d.K1 + t.kbz - 1) / t.kbz;
  size_t ldsz = 2ull * NF * t.kbz * (d.n3 + 1) * sizeof(C);
  hipLaunchKernelGGL((k_zgop<R, KIND>), dim3(nchz * d.n2, nb), dim3(256), ldsz, st, dd, d, sp.pz, twp<R>(sp, 2),
                     t.kbz, nchz);
  hipLaunchKernelGGL((k_cpass<R, +1>), dim3(nchy * d.n3, NF, nb), dim3(256), ldsy, st, dd, d, sp.py,
                     twp<R>(sp, 1), t.kby, nchy, (long long)d.K1, (long long)d.n2 * d.K1);
  hipLaunchKernelGGL((k_xinv<R, KIND, INMODE, OUTMODE>), dim3(t.ntile, nb), dim3(256), lds_x, st, dd, d, sp.px,
                     twp<R>(sp, 0), twp<R>(sp, 3), t.rows);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename R, int KIND>
static int eval_k(const ShapePlan& sp, int inmode, int outmode, const ReqDesc* dd, int nb, hipStream_t st) {
  if (inmode == kInRho) {
    if (outmode != kOutPot) return -2;
    return eval_t<R, KIND, kInRho, kOutPot>(sp, dd, nb, st);
  }
  switch (outmode) {
    case kOutPot: return eval_t<R, KIND, kInPhi, kOutPot>(sp, dd, nb, st);
    case kOutSigned: return eval_t<R, KIND, kInPhi, kOutSigned>(sp, dd, nb, st);
    default: return eval_t<R, KIND, kInPhi, kOutNewton>(sp, dd, nb, st);
  }
}

template <typename R>
static int eval_r(const ShapePlan& sp, int kind, int inmode, int outmode, const ReqDesc* dd, int nb,
                  hipStream_t st) {
  switch (kind) {
    case kWT3: return eval_k<R, kWT3>(sp, inmode, outmode, dd, nb, st);
    case kWT4: return eval_k<R, kWT4>(sp, inmode, outmode, dd, nb, st);
    case kWGC8: return eval_k<R, kWGC8>(sp, inmode, outmode, dd, nb, st);
    case kWT3S: return eval_k<R, kWT3S>(sp, inmode, outmode, dd, nb, st);
    case kWT4S: return eval_k<R, kWT4S>(sp, inmode, outmode, dd, nb, st);
    default: return -3;
  }
}

int run_eval(const ShapePlan& sp, int kind, int inmode, int outmode, int fp32, const ReqDesc* dd, int nb,
             hipStream_t st) {
  if (!sp.force_generic && fast_path_ok(sp, kind)) return fast_eval(sp, kind, inmode, outmode, fp32, dd, nb, st);
  return fp32 ? eval_r<float>(sp, kind, inmode, outmode, dd, nb, st)
              : eval_r<double>(sp, kind, inmode, outmode, dd, nb, st);
}

int run_finalize(const ShapePlan& sp, int kind, int fp32, const ReqDesc* dd, int nb, double* out, hipStream_t st) {
  if (!sp.force_generic && fast_path_ok(sp, kind)) return fast_finalize(sp, kind, fp32, dd, nb, out, st);
  const Tiles t = tiles_for(sp, func_nf(kind), fp32 ? (int)sizeof(float2) : (int)sizeof(double2));
  (void)out;
  hipLaunchKernelGGL(k_finalize, dim3(kNSum, nb), dim3(256), 0, st, dd, t.ntile);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ plain 3D transforms (per-geometry tables)
// Forward R2C of the real field D.u into D.work (no 1/N), backward C2R of
// D.work into D.out (unnormalized). Used by the ion potential and forces.
int run_fft3d_plain(const ShapePlan& sp, int forward, const ReqDesc* dd, int nb, hipStream_t st) {
  using C = double2;
  const GridDims& d = sp.d;
  const Tiles t = tiles_for(sp, 1, (int)sizeof(C));
  size_t lds_x = 2ull * t.rows * (d.M + 1) * sizeof(C);
  const int nchy = (d.K1 + t.kby - 1) / t.kby;
  size_t ldsy = 2ull * t.kby * (d.n2 + 1) * sizeof(C);
  const int nchz = (d.K1 + t.kbz - 1) / t.kbz;
  size_t ldsz = 2ull * t.kbz * (d.n3 + 1) * sizeof(C);
  if (forward) {
    hipLaunchKernelGGL((k_xfwd<double, kPlain, kInRho>), dim3(t.ntile, nb), dim3(256), lds_x, st, dd, d, sp.px,
                       sp.tw_x, sp.tw_s, t.rows);
    hipLaunchKernelGGL((k_cpass<double, -1>), dim3(nchy * d.n3, 1, nb), dim3(256), ldsy, st, dd, d, sp.py, sp.tw_y,
                       t.kby, nchy, (long long)d.K1, (long long)d.n2 * d.K1);
    hipLaunchKernelGGL((k_cpass<double, -1>), dim3(nchz * d.n2, 1, nb), dim3(256), ldsz, st, dd, d, sp.pz, sp.tw_z,
                       t.kbz, nchz, (long long)d.n2 * d.K1, (long long)d.K1);
  } else {
    hipLaunchKernelGGL((k_cpass<double, +1>), dim3(nchz * d.n2, 1, nb), dim3(256), ldsz, st, dd, d, sp.pz, sp.tw_z,
                       t.kbz, nchz, (long long)d.n2 * d.K1, (long long)d.K1);
    hipLaunchKernelGGL((k_cpass<double, +1>), dim3(nchy * d.n3, 1, nb), dim3(256), ldsy, st, dd, d, sp.py, sp.tw_y,
                       t.kby, nchy, (long long)d.K1, (long long)d.n2 * d.K1);
    hipLaunchKernelGGL((k_xinv<double, kPlain, kInRho, kOutPot>), dim3(t.ntile, nb), dim3(256), lds_x, st, dd, d,
                       sp.px, sp.tw_x, sp.tw_s, t.rows);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ============================================================ vector kernels
__global__ void k_axpby(double* y, const double* x, double a, double b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = fma(a, x[i], b * y[i]);
}
__global__ void k_lincomb(double* z, const double* x, const double* y, double a, double b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    z[i] = fma(b, y[i], a * x[i]);
}
__global__ void k_sqrt_abs(double* y, const double* x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = sqrt(fabs(x[i]));
}
__global__ void k_fill(double* y, double v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = v;
}
__global__ void k_square(double* y, const double* x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i] * x[i];
}
__global__ void k_to_f32(float* y, const double* x, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
__global__ void k_scale_sign(double* y, const double* x, const double* s, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i] * fsign(s[i]);
}
struct DotArgs {
  const double* a[4];
  const double* b[4];
};
constexpr int kDotBlocks = 512;
__global__ void __launch_bounds__(256) k_dots(DotArgs args, int np, long long n, double* partial) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < np) acc[k] += args.a[k][i] * args.b[k][i];
  }
  block_sum_store<4>(acc, partial + (size_t)blockIdx.x * 4);
}
__global__ void __launch_bounds__(256) k_dots_fin(const double* partial, int nblk, double* out) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int w = threadIdx.x; w < nblk; w += blockDim.x) {
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += partial[(size_t)w * 4 + k];
  }
  block_sum_store<4>(acc, out);
}

// One block row per operation (blockIdx.y); the kind is uniform per block row,
// so each loop is a plain streaming loop with global loads and stores.
__global__ void __launch_bounds__(256) k_vecops(const VecOp* __restrict__ ops, long long n) {
  const VecOp& o = ops[blockIdx.y];
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const gptr<double> z = glb(o.z);
  if (o.kind == kOpLincomb) {
    const gptr<const double> x = glb(o.x), y = glb(o.y);
    const double a = o.a, b = o.b;
#pragma unroll 4
    for (; i < n; i += stride) z[i] = fma(b, y[i], a * x[i]);
  } else if (o.kind == kOpAxpby) {
    const gptr<const double> x = glb(o.x);
    const double a = o.a, b = o.b;
#pragma unroll 4
    for (; i < n; i += stride) z[i] = fma(a, x[i], b * z[i]);
  } else if (o.kind == kOpCG) {
    const gptr<double> x2 = glb(o.x2), y2 = glb(o.y2);
    const gptr<const double> w = glb(o.w);
    const double a = o.a, b = o.b;
#pragma unroll 4
    for (; i < n; i += stride) {
      const double p = x2[i];
      z[i] = fma(a, p, 1.0 * z[i]);
      const double r = fma(-a, w[i], 1.0 * y2[i]);
      y2[i] = r;
      x2[i] = fma(1.0, r, b * p);
    }
  } else {
#pragma unroll 4
    for (; i < n; i += stride) z[i] = 0.0;
  }
}
__global__ void __launch_bounds__(256) k_dotsb(const DotDesc* __restrict__ dd, long long n, double* part) {
  const DotDesc& q = dd[blockIdx.y];
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc[1] = {0.0};
  const gptr<const double> a = glb(q.a);
  if (q.b != nullptr) {
    const gptr<const double> b = glb(q.b);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) acc[0] += a[i] * b[i];
  } else {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) acc[0] += a[i];
  }
  block_sum_store<1>(acc, part + (size_t)blockIdx.y * kDotBlocksBatched + blockIdx.x);
}
__global__ void __launch_bounds__(256) k_dotsb_fin(const double* part, double* res) {
  double acc[1] = {0.0};
  for (int w = threadIdx.x; w < kDotBlocksBatched; w += blockDim.x) acc[0] += part[(size_t)blockIdx.x * kDotBlocksBatched + w];
  block_sum_store<1>(acc, res + blockIdx.x);
}
void vec_ops_batched(const VecOp* d_ops, int nops, long long n, hipStream_t st) {
  if (nops <= 0) return;
  long long b = (n + 255) / 256;
  int nb = (int)(b > 512 ? 512 : (b < 1 ? 1 : b));
  hipLaunchKernelGGL(k_vecops, dim3(nb, nops), dim3(256), 0, st, d_ops, n);
}
void vec_dots_batched(const DotDesc* d_dd, int ndots, long long n, double* d_part, double* d_res, hipStream_t st) {
  if (ndots <= 0) return;
  hipLaunchKernelGGL(k_dotsb, dim3(kDotBlocksBatched, ndots), dim3(256), 0, st, d_dd, n, d_part);
  hipLaunchKernelGGL(k_dotsb_fin, dim3(ndots), dim3(256), 0, st, d_part, d_res);
}

static int nblocks(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}
void vec_axpby(double* y, const double* x, double a, double b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(256), 0, st, y, x, a, b, n);
}
void vec_lincomb(double* z, const double* x, const double* y, double a, double b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_lincomb, dim3(nblocks(n)), dim3(256), 0, st, z, x, y, a, b, n);
}
void vec_sqrt_abs(double* y, const double* x, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_sqrt_abs, dim3(nblocks(n)), dim3(256), 0, st, y, x, n);
}
void vec_fill(double* y, double v, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(256), 0, st, y, v, n);
}
void vec_square(double* y, const double* x, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_square, dim3(nblocks(n)), dim3(256), 0, st, y, x, n);
}
void vec_to_f32(float* y, const double* x, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(256), 0, st, y, x, n);
}
void vec_scale_sign(double* y, const double* x, const double* s, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_sign, dim3(nblocks(n)), dim3(256), 0, st, y, x, s, n);
}
void vec_dots(const double* const* a, const double* const* b, int np, long long n, double* d_partial,
              double* d_out, hipStream_t st) {
  DotArgs args;
  for (int k = 0; k < 4; ++k) {
    args.a[k] = k < np ? a[k] : a[0];
    args.b[k] = k < np ? b[k] : b[0];
  }
  long long want = (n + 255) / 256;
  int nb = (int)(want > kDotBlocks ? kDotBlocks : (want < 1 ? 1 : want));
  hipLaunchKernelGGL(k_dots, dim3(nb), dim3(256), 0, st, args, np, n, d_partial);
  hipLaunchKernelGGL(k_dots_fin, dim3(1), dim3(256), 0, st, d_partial, nb, d_out);
}

}  // namespace ofd
