// Register-resident two-stage FFT passes for lines of n = R1*R2 points with
// R2 dividing R1: the powers of two 16..256 and the mixed lengths 24, 40, 48,
// 80, 96, 192 (radix-12, -20 and -24 first stages).
//
// A line of n = R1*R2 complex points is owned by T = R2 threads; thread t
// holds the R1 points t + T*m (m = 0..R1-1) in registers. Stage 1 is one
// radix-R1 butterfly per thread (no twiddles), the results cross lanes once
// through a padded LDS buffer, stage 2 is R1/R2 radix-R2 butterflies per
// thread. After stage 2 thread t again holds the points t + T*m' of the
// (autosorted) transform, so a pointwise G-space operator and the inverse
// transform of the same line run without touching memory in between.
//
// Kernels (the fast path of the five evaluation passes, Wang-Teter functionals):
//   k_fx_fwd  rows along x: fields from rho, half-length complex FFT, R2C split
//   k_fy      columns along y, flattened over (i3, k1)
//   k_fz_gop  columns along z, flattened over (i2, k1): FFT, diagonal G-space
//             multiplier per field, inverse FFT
//   k_fx_inv  rows along x: C2R merge, inverse FFT, potential epilogue and
//             per-workgroup energy sums
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "ofd_internal.h"
#include "ofd_point.h"

namespace ofd {

namespace fast {

constexpr double kPi = 3.141592653589793238462643;
constexpr double kCTF = 2.87123400018819;

template <int N> struct P2;
template <> struct P2<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct P2<24> { static constexpr int R1 = 12, R2 = 2; };
template <> struct P2<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct P2<40> { static constexpr int R1 = 20, R2 = 2; };
template <> struct P2<48> { static constexpr int R1 = 12, R2 = 4; };
template <> struct P2<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct P2<80> { static constexpr int R1 = 20, R2 = 4; };
template <> struct P2<96> { static constexpr int R1 = 24, R2 = 4; };
template <> struct P2<128> { static constexpr int R1 = 16, R2 = 8; };
template <> struct P2<192> { static constexpr int R1 = 24, R2 = 8; };
template <> struct P2<256> { static constexpr int R1 = 16, R2 = 16; };

template <int N> __device__ __forceinline__ constexpr int padded(int j) { return j + j / P2<N>::R1; }
__device__ __forceinline__ int t_plus(int t, int T, int m) { return t + T * m; }
template <int N> constexpr int line_lds() { return N + N / P2<N>::R1; }
// Line stride of the column passes (y, z), in real components: their
// consecutive lanes sit on consecutive columns, and an odd stride puts the
// lanes served together by one LDS cycle on distinct banks (a multiple of
// the bank period would turn every exchange into a many-way conflict).
template <int N> constexpr int col_stride() { return line_lds<N>() | 1; }

// Two-stage FFT of one line; x[m] holds point t + T*m on entry and exit.
// `lds` is this line's padded buffer. Contains two workgroup barriers.
template <int N, int S, typename C>
__device__ __forceinline__ void fft2(C (&x)[P2<N>::R1], C* __restrict__ lds, int t, const C* __restrict__ tw,
                                     bool active) {
  constexpr int R1 = P2<N>::R1, R2 = P2<N>::R2, T = R2, Q = R1 / R2;
  if (active) {
    Bfly<R1, S, C>::run(x);
#pragma unroll
    for (int m = 0; m < R1; ++m) lds[padded<N>(t * R1 + m)] = x[m];
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b2 = t + T * q;
      C y[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) y[m] = lds[padded<N>(b2 + m * R1)];
#pragma unroll
      for (int m = 1; m < R2; ++m) {
        C w = tw[m * b2];
        if (S > 0) w.y = -w.y;
        y[m] = cmul(y[m], w);
      }
      Bfly<R2, S, C>::run(y);
#pragma unroll
      for (int m = 0; m < R2; ++m) x[q + Q * m] = y[m];
    }
  }
  __syncthreads();
}

// fft2 of one column of a group of P lines X_0 .. X_{P-1} (this line's own
// input is one of them, written to `lds` by the first stage; line p sits at
// grp + p * gstride) that a radix-P step joins first: on exit x holds the FFT
// of X_0 + sum_p w[p] X_p. The join is linear and the first stage acts on
// each line alone, so it is applied as the stage-two inputs are read from the
// group's LDS lines (no extra exchange or barrier).
template <int N, int S, int P, typename C>
__device__ __forceinline__ void fft2_join(C (&x)[P2<N>::R1], C* __restrict__ lds, const C* grp, int gstride, int t,
                                          const C* __restrict__ tw, bool active, const C (&w)[P]) {
  constexpr int R1 = P2<N>::R1, R2 = P2<N>::R2, T = R2, Q = R1 / R2;
  if (active) {
    Bfly<R1, S, C>::run(x);
#pragma unroll
    for (int m = 0; m < R1; ++m) lds[padded<N>(t * R1 + m)] = x[m];
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b2 = t + T * q;
      C y[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        const int j = padded<N>(b2 + m * R1);
        y[m] = grp[j];
#pragma unroll
        for (int p = 1; p < P; ++p) {
          const C o = cmul(grp[(size_t)p * gstride + j], w[p]);
          y[m].x += o.x;
          y[m].y += o.y;
        }
      }
#pragma unroll
      for (int m = 1; m < R2; ++m) {
        C tw1 = tw[m * b2];
        if (S > 0) tw1.y = -tw1.y;
        y[m] = cmul(y[m], tw1);
      }
      Bfly<R2, S, C>::run(y);
#pragma unroll
      for (int m = 0; m < R2; ++m) x[q + Q * m] = y[m];
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------ pointwise
// fft2 with the lane exchange done one real component at a time through a
// scalar LDS line (half the LDS of fft2, two more barriers). Used by the
// column passes, whose occupancy is bound by LDS.
template <int N, int S, typename C, typename R>
__device__ __forceinline__ void fft2s(C (&x)[P2<N>::R1], R* __restrict__ lds, int t, const C* __restrict__ tw,
                                      bool active) {
  constexpr int R1 = P2<N>::R1, R2 = P2<N>::R2, T = R2, Q = R1 / R2;
  R re[R1];   // stage-2 inputs, real parts, in (q, m) order
  if (active) {
    Bfly<R1, S, C>::run(x);
#pragma unroll
    for (int m = 0; m < R1; ++m) lds[padded<N>(t * R1 + m)] = x[m].x;
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int m = 0; m < R2; ++m) re[q * R2 + m] = lds[padded<N>(t + T * q + m * R1)];
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) lds[padded<N>(t * R1 + m)] = x[m].y;
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int b2 = t + T * q;
      C y[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        y[m].x = re[q * R2 + m];
        y[m].y = lds[padded<N>(b2 + m * R1)];
      }
#pragma unroll
      for (int m = 1; m < R2; ++m) {
        C w = tw[m * b2];
        if (S > 0) w.y = -w.y;
        y[m] = cmul(y[m], w);
      }
      Bfly<R2, S, C>::run(y);
#pragma unroll
      for (int m = 0; m < R2; ++m) x[q + Q * m] = y[m];
    }
  }
  __syncthreads();
}

template <int KIND> struct NF_;
template <> struct NF_<kWT3> { static constexpr int v = 3; };
template <> struct NF_<kWT4> { static constexpr int v = 4; };
template <> struct NF_<kWGC8> { static constexpr int v = 8; };
template <> struct NF_<kWT3S> { static constexpr int v = 5; };
template <> struct NF_<kWT4S> { static constexpr int v = 7; };

template <int KIND>
__device__ __forceinline__ void fields(const ReqDesc& D, double rho, double* f) {
  double ra, rb;
  fwd_pows(rho, D.alpha, D.beta, ra, rb);
  f[0] = rho;
  f[1] = rb;
  if (KIND == kWT3) {
    f[2] = sqrt(rho);
  } else {
    f[2] = ra;
    f[3] = sqrt(rho);
  }
}

// Fixed-order workgroup sum: wave butterflies, then the waves in order.
// out[k * stride] = sum over the workgroup of acc[k] for the first NU slots,
// 0 for the rest (slots the epilogue form never adds to).
template <int N, int NU = N>
__device__ __forceinline__ void block_sum_strided(double (&acc)[N], gptr<double> out, int stride) {
  __shared__ double red[16][NU];
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[k] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NU; ++k) red[wv][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
    if (threadIdx.x < NU) {
      const int nw = (blockDim.x + 63) >> 6;
      for (int w = 0; w < nw; ++w) s += red[w][threadIdx.x];
    }
    out[(size_t)threadIdx.x * stride] = s;
  }
}

template <int KIND>
__device__ __forceinline__ double field_of(const ReqDesc& D, int f, double rho) {
  if (KIND == kWGC8) {
    // rho, sqrt(rho), then theta^k rho^alpha / k! and theta^k rho^beta / k!
    // (k = 0, 1, 2; theta = rho - rho*) as the second-order WGC expansion
    // takes them (KEDF_WGC.f90:208-286)
    if (f == 0) return rho;
    if (f == 1) return sqrt(rho);
    const double pw = exp2((f < 5 ? D.alpha : D.beta) * log2(rho));
    const double th = rho - D.rhoS;
    const int k = (f - 2) % 3;
    return k == 0 ? pw : (k == 1 ? th * pw : th * th * pw / 2.0);
  }
  if (f == 0) return rho;
  if (f == NF_<KIND>::v - 1) return sqrt(rho);
  double ra, rb;
  fwd_pows(rho, D.alpha, D.beta, ra, rb);
  return (KIND == kWT4 && f == 2) ? ra : rb;
}

// A field value in the pass's precision. The FP32 passes of the Wang-Teter
// kinds form their powers with the hardware log2 / exp2 (relative error about
// 1e-6, as the FP32 transforms); otherwise the FP64 value, rounded.
template <int KIND, typename R>
__device__ __forceinline__ R field_r(const ReqDesc& D, int f, double rho) {
  if constexpr (sizeof(R) == 4 && (KIND == kWT3 || KIND == kWT4)) {
    const float r = (float)rho;
    if (f == 0) return r;
    if (f == NF_<KIND>::v - 1) return __builtin_amdgcn_sqrtf(r);
    const float p = (KIND == kWT4 && f == 2) ? (float)D.alpha : (float)D.beta;
    return __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(r));
  } else {
    return (R)field_of<KIND>(D, f, rho);
  }
}

// rho of one spin channel at reals g and g + 1: the density itself (kInRho)
// or phi_t^2 with phi_t = ca*u + cb*v
template <int INMODE>
__device__ __forceinline__ void rho_pair(const double* u, const double* v, double ca, double cb, size_t g, double& a,
                                         double& b) {
  const double2 uu = ldc(glb(reinterpret_cast<const double2*>(u + g)));
  if (INMODE == kInRho) {
    a = uu.x;
    b = uu.y;
    return;
  }
  double ta = ca * uu.x, tb = ca * uu.y;
  if (cb != 0.0) {
    const double2 vv = ldc(glb(reinterpret_cast<const double2*>(v + g)));
    ta = fma(cb, vv.x, ta);
    tb = fma(cb, vv.y, tb);
  }
  a = ta * ta;
  b = tb * tb;
}

// Field f of the spin kinds from both channels' densities (x = 2 rho_s):
// 0 rho_up + rho_dn; then, odd fields spin up and even fields spin down,
// x^beta (kWT3S, kWT4S), x^alpha (kWT4S), sqrt(x) (CalPotential.f90:120-160,
// 400-407; KEDF_VW.f90:100-140)
template <int KIND>
__device__ __forceinline__ double spin_field(const ReqDesc& D, int f, double r1, double r2) {
  constexpr int NF = NF_<KIND>::v;
  if (f == 0) return r1 + r2;
  const double x = 2.0 * ((f & 1) ? r1 : r2);
  if (f >= NF - 2) return sqrt(x);
  double ra, rb;
  fwd_pows(x, D.alpha, D.beta, ra, rb);
  return (KIND == kWT4S && f >= 3) ? ra : rb;
}

// The x-forward passes' complex input element at reals g, g + 1 of field f
template <int KIND, int INMODE, typename R>
__device__ __forceinline__ typename V2<R>::T field_pair(const ReqDesc& D, int f, size_t g) {
  typename V2<R>::T z;
  double a, b;
  if constexpr (is_spin_kind(KIND)) {
    double a1 = 0.0, b1 = 0.0, a2 = 0.0, b2 = 0.0;
    if (f == 0 || (f & 1)) rho_pair<INMODE>(D.u, D.v, D.ca, D.cb, g, a1, b1);
    if (f == 0 || !(f & 1)) rho_pair<INMODE>(D.u2, D.v2, D.ca2, D.cb2, g, a2, b2);
    z.x = (R)spin_field<KIND>(D, f, a1, a2);
    z.y = (R)spin_field<KIND>(D, f, b1, b2);
  } else {
    rho_pair<INMODE>(D.u, D.v, D.ca, D.cb, g, a, b);
    z.x = field_r<KIND, R>(D, f, a);
    z.y = field_r<KIND, R>(D, f, b);
  }
  return z;
}

// Rows per workgroup of the x passes: RPB * NF lines of T threads,
// whole wavefronts, at most 256 threads.
__host__ __device__ constexpr int inv_rpb(int nf, int t) {
  int r = 256 / (nf * t);
  while (r > 1 && (r * nf * t) % 64 != 0) --r;
  return r;
}

// ------------------------------------------------------------ x forward (R2C)
// LDS of the x forward pass per workgroup. For rows of 256 points and more
// whose complex lines would take more than 32 KB per workgroup (FP64), the
// lines exchange one real component at a time instead (a line of
// col_stride<M>() reals): half the LDS, for about 40 more VGPRs, which at
// these lengths lets four workgroups share a CU instead of three (shorter
// rows would lose more occupancy to the registers than they gain in LDS).
template <int M, int NF, typename C> constexpr bool fx_scalar() {
  return M >= 128 && (size_t)inv_rpb(NF, P2<M>::R2) * NF * line_lds<M>() * sizeof(C) > 32768;
}
template <int M, int NF, typename C> constexpr size_t fx_fwd_lds() {
  return fx_scalar<M, NF, C>() ? (size_t)inv_rpb(NF, P2<M>::R2) * NF * col_stride<M>() * (sizeof(C) / 2)
                               : (size_t)inv_rpb(NF, P2<M>::R2) * NF * line_lds<M>() * sizeof(C);
}

// One workgroup = RPB rows x NF fields; thread (line = row*NF + field, lane t):
// every line forms its field from rho, runs the half-length complex FFT in
// registers and LDS, and writes its R2C-split half row.
template <int M, int KIND, int INMODE, typename R>
__global__ void __launch_bounds__(256) k_fx_fwd(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw,
                                                const typename V2<R>::T* __restrict__ tws) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<M>::R1, T = P2<M>::R2, NF = NF_<KIND>::v;
  constexpr int LL = line_lds<M>();
  constexpr int RPB = inv_rpb(NF, T);
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int r = line / NF, f = line - r * NF;
  const long long nrows = (long long)d.n2 * d.n3;
  const long long row = (long long)xcd_block(blockIdx.x, gridDim.x) * RPB + r;
  const bool active = row < nrows;
  C z[R1];
  auto rho_at = [&](size_t g) {
    if (INMODE == kInRho) return (double)glb(D.u)[g];
    double tp = D.ca * glb(D.u)[g];
    if (D.cb != 0.0) tp = fma(D.cb, (double)glb(D.v)[g], tp);
    return tp * tp;
  };
  if constexpr (KIND == kWGC8) {
    // six of the eight fields are rho^alpha or rho^beta times a factor: every
    // point's two powers are computed once (one log2, two exp2) by one of the
    // row's NF*T threads and shared through LDS, which the lines' transforms
    // reuse afterwards
    double* pa = reinterpret_cast<double*>(smem) + (size_t)r * 4 * M;
    double* pb = pa + 2 * M;
    if (active) {
      const size_t base = (size_t)row * (2 * M);
      for (int q = f * T + t; q < 2 * M; q += NF * T) {
        const double l2 = log2(rho_at(base + q));
        pa[q] = exp2(D.alpha * l2);
        pb[q] = exp2(D.beta * l2);
      }
    }
    __syncthreads();
    if (active) {
      const size_t base = (size_t)row * (2 * M);
      const int kk = (f - 2) % 3;
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        const int k = t + T * m;
        double v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double rho = rho_at(base + 2 * k + h);
          if (f == 0) {
            v[h] = rho;
          } else if (f == 1) {
            v[h] = sqrt(rho);
          } else {
            const double pw = f < 5 ? pa[2 * k + h] : pb[2 * k + h];
            const double th = rho - D.rhoS;
            v[h] = kk == 0 ? pw : (kk == 1 ? th * pw : th * th * pw / 2.0);
          }
        }
        z[m].x = (R)v[0];
        z[m].y = (R)v[1];
      }
    }
    __syncthreads();
  } else if (active) {
    const size_t base = (size_t)row * (2 * M);
#pragma unroll
    for (int m = 0; m < R1; ++m) z[m] = field_pair<KIND, INMODE, R>(D, f, base + 2 * (t + T * m));
  }
  if constexpr (fx_scalar<M, NF, C>()) {
    // the split below reads Z[k] and Z[M-k] of other lanes: real parts first
    // (giving er and oi), then imaginary parts, which finish each output
    R* mr = reinterpret_cast<R*>(smem) + (size_t)line * col_stride<M>();
    fft2s<M, -1>(z, mr, t, tw, active);
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) mr[t + T * m] = z[m].x;
    }
    __syncthreads();
    R er[R1 + 1], oi[R1 + 1];
    if (active) {
#pragma unroll
      for (int m = 0; m <= R1; ++m) {
        const int k = m < R1 ? t + T * m : M;
        const R xk = mr[k == M ? 0 : k], xm = mr[k == 0 ? 0 : M - k];
        er[m] = (R)0.5 * (xk + xm);
        oi[m] = (R)-0.5 * (xk - xm);
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) mr[t + T * m] = z[m].y;
    }
    __syncthreads();
    if (active) {
      gptr<C> out = glb(reinterpret_cast<C*>(D.work)) + (size_t)f * d.nc + (size_t)row * (M + 1);
#pragma unroll
      for (int m = 0; m <= R1; ++m) {
        const int k = m < R1 ? t + T * m : M;
        if (m == R1 && t != 0) continue;
        const R yk = mr[k == M ? 0 : k], ym = mr[k == 0 ? 0 : M - k];
        const R ei = (R)0.5 * (yk - ym), orr = (R)0.5 * (yk + ym);
        const C w = tws[k];
        C x;
        x.x = er[m] + (w.x * orr - w.y * oi[m]);
        x.y = ei + (w.x * oi[m] + w.y * orr);
        stc(out + k, x);
      }
    }
  } else {
    C* mine = lds + (size_t)line * LL;
    fft2<M, -1>(z, mine, t, tw, active);
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) mine[t + T * m] = z[m];
    }
    __syncthreads();
    if (active) {
      gptr<C> out = glb(reinterpret_cast<C*>(D.work)) + (size_t)f * d.nc + (size_t)row * (M + 1);
#pragma unroll
      for (int m = 0; m <= R1; ++m) {
        const int k = m < R1 ? t + T * m : M;
        if (m == R1 && t != 0) continue;
        const C zk = mine[k == M ? 0 : k];
        const C zm = mine[k == 0 ? 0 : M - k];
        const R er = (R)0.5 * (zk.x + zm.x), ei = (R)0.5 * (zk.y - zm.y);
        const R orr = (R)0.5 * (zk.y + zm.y), oi = (R)-0.5 * (zk.x - zm.x);
        const C w = tws[k];
        C x;
        x.x = er + (w.x * orr - w.y * oi);
        x.y = ei + (w.x * oi + w.y * orr);
        stc(out + k, x);
      }
    }
  }
}

// ------------------------------------------------------------ fused x + y forward (one plane)
// One workgroup per (z-plane i3, field f, request); the NF workgroups of a
// plane are dispatched one after another on the same XCD, so the plane's rho
// is read from HBM once and from that XCD's L2 by the other fields. The plane's half
// spectrum (K1 x n2) of the field stays in LDS between the x R2C and the y
// FFT: rows of stride S >= max(M + M/R1x, M + 1) complex, n2 + n2/R1y rows so
// that the y transform can use the padded column layout as its scratch.
//
// Split form (NP row classes, N2 = n2/NP; the plans use NP = 2): one workgroup per
// (z-plane, row class p, field) takes the rows y = NP j + p only and leaves
// their N2-point y transform X_p at y-frequency rows p*N2 + k. The z pass finishes
// the y transform with its last radix-NP step (Y[k] = sum_p w^(p k) X_p[k mod
// N2], w = exp(-2 pi i / n2)) as it loads its columns. A class of rows takes
// 1/NP of the LDS (with the tightest padded column length: 81.8 KB for half a
// 128 x 128 plane in FP64), so several workgroups share a CU and the loads and
// stores of some overlap the transforms of others.
template <int M, int N2, int SPLIT = 0> struct PlaneGeo {   // SPLIT: the tight row count of the split form
  // the row stride keeps M + M/R1 (not the tight M + (M-1)/R1): at 72 complex
  // two consecutive rows fall on opposite halves of the 64 LDS banks, so the
  // 16 lanes of two rows served together do not conflict (71 measured 4x the
  // bank conflicts)
  static constexpr int LLX = M + M / P2<M>::R1;
  static constexpr int S = LLX > M + 1 ? LLX : M + 1;
  static constexpr int ROWS = SPLIT ? (N2 - 1) + (N2 - 1) / P2<N2>::R1 + 1 : N2 + N2 / P2<N2>::R1;
  static constexpr int TX = P2<M>::R2, TY = P2<N2>::R2;
  static constexpr int NTX = N2 * TX, NTY = (M + 1) * TY;
  static constexpr int THREADS = ((NTX > NTY ? NTX : NTY) + 63) / 64 * 64;
};

template <int M, int N2, int KIND, int INMODE, typename R, int NP>
__global__ void __launch_bounds__((NP > 1 ? PlaneGeo<M, N2, 1>::THREADS : 1024))
    k_plane_fwd(const ReqDesc* __restrict__ descs, GridDims d, const typename V2<R>::T* __restrict__ twx,
                const typename V2<R>::T* __restrict__ tws, const typename V2<R>::T* __restrict__ twy, int fo, int fn) {
  using C = typename V2<R>::T;
  using G = PlaneGeo<M, N2, (NP > 1)>;
  constexpr int R1X = P2<M>::R1, TX = G::TX, R1Y = P2<N2>::R1, TY = G::TY, S = G::S;
  extern __shared__ __align__(16) unsigned char smem[];
  C* P = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  // fields fo .. fo + fn - 1 of the kind
  const int pf = xcd_block(blockIdx.x, gridDim.x);
  const int hp = pf / fn, f = fo + (pf - hp * fn);
  const int i3 = hp / NP, par = hp - i3 * NP;
  // ---- x step: thread (row i2, lane t)
  const int row = threadIdx.x / TX, tx = threadIdx.x % TX;
  const bool ax = row < N2;
  C z[R1X];
  if (ax) {
    const size_t base = ((size_t)i3 * (N2 * NP) + (NP * row + par)) * (2 * M);
#pragma unroll
    for (int m = 0; m < R1X; ++m) z[m] = field_pair<KIND, INMODE, R>(D, f, base + 2 * t_plus(tx, TX, m));
  }
  C* myrow = P + (size_t)row * S;
  fft2<M, -1>(z, myrow, tx, twx, ax);
  if (ax) {
#pragma unroll
    for (int m = 0; m < R1X; ++m) myrow[tx + TX * m] = z[m];
  }
  __syncthreads();
  // split, values first into registers, then written in place
  C xo[R1X + 1];
  if (ax) {
#pragma unroll
    for (int m = 0; m <= R1X; ++m) {
      const int k = m < R1X ? tx + TX * m : M;
      const C zk = myrow[k == M ? 0 : k];
      const C zm = myrow[k == 0 ? 0 : M - k];
      const R er = (R)0.5 * (zk.x + zm.x), ei = (R)0.5 * (zk.y - zm.y);
      const R orr = (R)0.5 * (zk.y + zm.y), oi = (R)-0.5 * (zk.x - zm.x);
      const C w = tws[k];
      xo[m].x = er + (w.x * orr - w.y * oi);
      xo[m].y = ei + (w.x * oi + w.y * orr);
    }
  }
  __syncthreads();
  if (ax) {
#pragma unroll
    for (int m = 0; m < R1X; ++m) myrow[tx + TX * m] = xo[m];
    if (tx == 0) myrow[M] = xo[R1X];
  }
  __syncthreads();
  // ---- y step: thread (column c, lane t), c fastest
  const int c = threadIdx.x % (M + 1), ty = threadIdx.x / (M + 1);
  const bool ay = ty < TY;
  C y[R1Y];
  if (ay) {
#pragma unroll
    for (int m = 0; m < R1Y; ++m) y[m] = P[(size_t)(ty + TY * m) * S + c];
  }
  __syncthreads();
  // the column's padded scratch: element j at row j + j/R1Y
  {
    constexpr int R1 = R1Y, R2 = TY, Q = R1 / R2;
    if (ay) {
      Bfly<R1, -1, C>::run(y);
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        const int j = ty * R1 + m;
        P[(size_t)(j + j / R1) * S + c] = y[m];
      }
    }
    __syncthreads();
    if (ay) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int b2 = ty + TY * q;
        C w2[R2];
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          const int j = b2 + m * R1;
          w2[m] = P[(size_t)(j + j / R1) * S + c];
        }
#pragma unroll
        for (int m = 1; m < R2; ++m) w2[m] = cmul(w2[m], twy[m * b2 * NP]);
        Bfly<R2, -1, C>::run(w2);
#pragma unroll
        for (int m = 0; m < R2; ++m) y[q + Q * m] = w2[m];
      }
    }
  }
  if (ay) {
    gptr<C> out = glb(reinterpret_cast<C*>(D.work)) + (size_t)f * d.nc + (size_t)i3 * (N2 * NP) * (M + 1) +
                  (size_t)par * N2 * (M + 1) + c;
#pragma unroll
    for (int m = 0; m < R1Y; ++m) stc(out + (size_t)(ty + TY * m) * (M + 1), y[m]);
  }
}

// ------------------------------------------------------------ y columns
// columns flattened over (i3, k1); blockDim = KB * T, thread = (t, c), c fastest
template <int N, int S, typename R>
__global__ void __launch_bounds__(256) k_fy(const ReqDesc* __restrict__ descs, GridDims d,
                                            const typename V2<R>::T* __restrict__ tw, int kb, int fo) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<N>::R1, T = P2<N>::R2;
  constexpr int LL = col_stride<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  R* lds = reinterpret_cast<R*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int c = threadIdx.x % kb, t = threadIdx.x / kb;
  const long long col = (long long)xcd_block(blockIdx.x, gridDim.x) * kb + c;
  const long long ncol = (long long)d.K1 * d.n3;
  const bool active = col < ncol;
  const long long i3 = active ? col / d.K1 : 0, k1 = active ? col - i3 * d.K1 : 0;
  gptr<C> base = glb(reinterpret_cast<C*>(D.work)) + (size_t)(fo + blockIdx.y) * d.nc + (size_t)i3 * d.n2 * d.K1 + k1;
  C x[R1];
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) x[m] = ldc(base + (size_t)(t + T * m) * d.K1);
  }
  fft2s<N, S>(x, lds + (size_t)c * LL, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) stc(base + (size_t)(t + T * m) * d.K1, x[m]);
  }
}

// ------------------------------------------------------------ z columns + G-space
// One field per workgroup (blockIdx.y): for the Wang-Teter layouts the
// G-space operator is diagonal in the fields (Hartree 4 pi/G^2 on F[rho],
// K(|G|) on F[rho^b] (and F[rho^a]), G^2 on F[sqrt rho]).
// The diagonal G-space multiplier of field f at (m1, m2, m3): Hartree
// 4 pi / G^2 on F[rho], G^2 on F[sqrt rho], the Wang-Teter kernel on the
// others, with the forward 1/N. FP32 passes form it in FP32 and read the
// kernel table's FP32 copy when there is one.
template <int KIND, typename R>
__device__ __forceinline__ R gop_scale(const ReqDesc& D, int f, int m1, int m2, int m3, size_t gi, double invn) {
  constexpr int NF = NF_<KIND>::v;
  constexpr int FVW = is_spin_kind(KIND) ? NF - 2 : NF - 1;   // first vW field (one per spin channel)
  if constexpr (sizeof(R) == 4) {
    if (f != 0 && f < FVW) return D.wtk32 != nullptr ? glb(D.wtk32)[gi] : (float)glb(D.wtk)[gi];
    const float x1 = (float)m1, x2 = (float)m2, x3 = (float)m3;
    const float gx = (float)D.rb[0] * x1 + (float)D.rb[1] * x2 + (float)D.rb[2] * x3;
    const float gy = (float)D.rb[3] * x1 + (float)D.rb[4] * x2 + (float)D.rb[5] * x3;
    const float gz = (float)D.rb[6] * x1 + (float)D.rb[7] * x2 + (float)D.rb[8] * x3;
    const float q2 = gx * gx + gy * gy + gz * gz;
    if (f == 0) return q2 > 0.0f ? (float)(4.0 * kPi * invn) / q2 : 0.0f;
    return q2 * (float)invn;
  } else {
    if (f != 0 && f < FVW) return glb(D.wtk)[gi];
    const double gx = D.rb[0] * m1 + D.rb[1] * m2 + D.rb[2] * m3;
    const double gy = D.rb[3] * m1 + D.rb[4] * m2 + D.rb[5] * m3;
    const double gz = D.rb[6] * m1 + D.rb[7] * m2 + D.rb[8] * m3;
    const double q2 = gx * gx + gy * gy + gz * gz;
    if (f == 0) return q2 > 0.0 ? 4.0 * kPi / q2 * invn : 0.0;
    return q2 * invn;
  }
}

// The z pass exchanges one real component at a time (a column line of
// col_stride<N>() reals, half the LDS) for FP64 columns of 256 points and
// more, where complex lines leave two or three workgroups per CU.
template <int N, typename R, int P> constexpr bool fz_scalar() { return P == 1 && N >= 256 && sizeof(R) == 8; }

//
// P > 1 (after the split plane pass with P row classes): the kb columns of a
// workgroup are P members of kb/P groups; member q of a group holds
// y-frequency k + q n2/P (k < n2/P) at one k1, and its input is the class-q
// transform X_q[k]. Each group finishes the y transform (the radix-P step that
// joins the classes' transforms) as the first z FFT stage reads its inputs
// (fft2_join), then its members run as ordinary columns. The inverse z
// transform leaves the spectra in the ordinary layout.
template <int N, int KIND, typename R, int P>
__global__ void __launch_bounds__(256) k_fz_gop(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw, int kb,
                                                const typename V2<R>::T* __restrict__ twy, int fo) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<N>::R1, T = P2<N>::R2, NF = NF_<KIND>::v;
  constexpr int LL = col_stride<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int f = fo + blockIdx.y;
  const int c = threadIdx.x % kb, t = threadIdx.x / kb;
  const int kg = kb / P, q = c / kg, cg = c - q * kg;   // group size, member, group
  long long col;
  bool active;
  int i2, m1;
  if constexpr (P > 1) {
    const long long colh = (long long)xcd_block(blockIdx.x, gridDim.x) * kg + cg;
    active = colh < (long long)d.K1 * (d.n2 / P);
    const int i2lo = active ? (int)(colh / d.K1) : 0;
    m1 = active ? (int)(colh - (long long)i2lo * d.K1) : 0;
    i2 = i2lo + q * (d.n2 / P);
    col = (long long)i2 * d.K1 + m1;
  } else {
    col = (long long)xcd_block(blockIdx.x, gridDim.x) * kb + c;
    active = col < (long long)d.K1 * d.n2;
    i2 = active ? (int)(col / d.K1) : 0;
    m1 = active ? (int)(col - (long long)i2 * d.K1) : 0;
  }
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  // element offsets within a field in 32 bits (a field's K1 n2 n3 elements
  // stay below 2^31 on the grids this pass serves) from a base that is
  // uniform across the workgroup: scalar base + 32-bit lane offset, no 64-bit
  // index arithmetic per element
  const unsigned es = (unsigned)(d.n2 * d.K1);
  const double invn = 1.0 / (double)d.nr;
  gptr<C> fbase = glb(reinterpret_cast<C*>(D.work)) + (size_t)f * d.nc;
  const unsigned o0 = (unsigned)col + (unsigned)t * es;
  C x[R1];
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) x[m] = ldc(fbase + (o0 + (unsigned)(T * m) * es));
  }
  if constexpr (P > 1) {
    // x holds X_q; Y[i2] = sum_p w^(p i2) X_p with w^j = twy[j mod n2]
    C w[P];
#pragma unroll
    for (int p = 0; p < P; ++p) w[p] = active ? twy[(size_t)p * i2 % d.n2] : C{};
    fft2_join<N, -1, P>(x, lds + (size_t)c * LL, lds + (size_t)cg * LL, kg * LL, t, tw, active, w);
  } else if constexpr (fz_scalar<N, R, P>()) {
    fft2s<N, -1>(x, reinterpret_cast<R*>(smem) + (size_t)c * LL, t, tw, active);
  } else {
    fft2<N, -1>(x, lds + (size_t)c * LL, t, tw, active);
  }
  if (active) {
    constexpr int FVW = is_spin_kind(KIND) ? NF - 2 : NF - 1;   // first vW field
    if ((f != 0 && f < FVW) || sizeof(R) == 4) {
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        const int i3 = t + T * m;
        const int m3 = i3 > N / 2 ? i3 - N : i3;
        const R scr = gop_scale<KIND, R>(D, f, m1, m2, m3, o0 + (unsigned)(T * m) * es, invn);
        x[m].x *= scr;
        x[m].y *= scr;
      }
    } else {
      // Hartree and vW in FP64: along the column |G|^2 = a0 + m3 (a1 + m3 a2)
      const double gx = D.rb[0] * m1 + D.rb[1] * m2, gy = D.rb[3] * m1 + D.rb[4] * m2,
                   gz = D.rb[6] * m1 + D.rb[7] * m2;
      const double a0 = gx * gx + gy * gy + gz * gz;
      const double a1 = 2.0 * (gx * D.rb[2] + gy * D.rb[5] + gz * D.rb[8]);
      const double a2 = D.rb[2] * D.rb[2] + D.rb[5] * D.rb[5] + D.rb[8] * D.rb[8];
      const double ch = 4.0 * kPi * invn;
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        const int i3 = t + T * m;
        const double m3 = (double)(i3 > N / 2 ? i3 - N : i3);
        const double q2 = fma(m3, fma(m3, a2, a1), a0);
        const double scr = f == 0 ? (q2 > 0.0 ? ch / q2 : 0.0) : q2 * invn;
        x[m].x *= (R)scr;
        x[m].y *= (R)scr;
      }
    }
  }
  if constexpr (fz_scalar<N, R, P>()) fft2s<N, +1>(x, reinterpret_cast<R*>(smem) + (size_t)c * LL, t, tw, active);
  else fft2<N, +1>(x, lds + (size_t)c * LL, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) stc(fbase + (o0 + (unsigned)(T * m) * es), x[m]);
  }
}

// z columns of the WGC fields with the coupled G-space operator of the
// second-order expansion (KEDF_WGC.f90:222-231, :262-280): on each block of
// three spectra (A, B, C) = (F[rho^p], F[theta rho^p], F[theta^2 rho^p / 2])
// it makes a' = K1 A + K2 B + K3 C, b' = K2 A + K4 B, c' = K3 A.
// blockIdx.y = 0 takes F[rho] (Hartree) and F[sqrt rho] (vW), 1 the
// rho^alpha block, 2 the rho^beta block. A block is done in three phases with
// two columns per lane in registers: c' from A; then b' and the partial
// K1 A + K2 B once B is in; then a' once C is in. Each output is written to
// the slot of an input already read, so the block leaves (c', b', a') in the
// slots of (A, B, C); the x-inverse epilogue reads them in that order.
template <int N, typename R>
__global__ void __launch_bounds__(256) k_fz_wgc(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw, int kb) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<N>::R1, T = P2<N>::R2;
  constexpr int LL = col_stride<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int g = blockIdx.y;
  const int f0 = g == 0 ? 0 : 3 * g - 1;   // fields 0-1, 2-4, 5-7
  const int c = threadIdx.x % kb, t = threadIdx.x / kb;
  const long long col = (long long)xcd_block(blockIdx.x, gridDim.x) * kb + c;
  const long long ncol = (long long)d.K1 * d.n2;
  const bool active = col < ncol;
  const int i2 = active ? (int)(col / d.K1) : 0;
  const int m1 = active ? (int)(col - (long long)i2 * d.K1) : 0;
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const size_t es = (size_t)d.n2 * d.K1;
  const double invn = 1.0 / (double)d.nr;
  gptr<C> base = glb(reinterpret_cast<C*>(D.work)) + (size_t)f0 * d.nc + (size_t)col;
  C* line = lds + (size_t)c * LL;
  C xa[R1], xb[R1];
  auto load = [&](C (&x)[R1], int j) {
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) x[m] = ldc(base + (size_t)j * d.nc + (size_t)(t + T * m) * es);
    }
  };
  auto store = [&](const C (&x)[R1], int j) {
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) stc(base + (size_t)j * d.nc + (size_t)(t + T * m) * es, x[m]);
    }
  };
  if (g == 0) {
    load(xa, 0);
    load(xb, 1);
    fft2<N, -1>(xa, line, t, tw, active);
    fft2<N, -1>(xb, line, t, tw, active);
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        const int i3 = t + T * m;
        const int m3 = i3 > N / 2 ? i3 - N : i3;
        const double gx = D.rb[0] * m1 + D.rb[1] * m2 + D.rb[2] * m3;
        const double gy = D.rb[3] * m1 + D.rb[4] * m2 + D.rb[5] * m3;
        const double gz = D.rb[6] * m1 + D.rb[7] * m2 + D.rb[8] * m3;
        const double q2 = gx * gx + gy * gy + gz * gz;
        const double hart = q2 > 0.0 ? 4.0 * kPi / q2 * invn : 0.0, vw = q2 * invn;
        xa[m].x = (R)((double)xa[m].x * hart);
        xa[m].y = (R)((double)xa[m].y * hart);
        xb[m].x = (R)((double)xb[m].x * vw);
        xb[m].y = (R)((double)xb[m].y * vw);
      }
    }
    fft2<N, +1>(xa, line, t, tw, active);
    fft2<N, +1>(xb, line, t, tw, active);
    store(xa, 0);
    store(xb, 1);
    return;
  }
  const gptr<const double> wk = glb(D.wgck);
  auto kidx = [&](int m) { return (size_t)(t + T * m) * es + (size_t)col; };
  // phase 1: c' = K3 A into the slot of A. With at most 16 points per lane
  // K3 stays in registers for phase 3 (one table read fewer, within the
  // VGPRs of two waves per SIMD); longer lanes read it again.
  constexpr bool keep_k3 = R1 <= 16;
  double k3s[keep_k3 ? R1 : 1];
  load(xa, 0);
  fft2<N, -1>(xa, line, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const double k3 = wk[2 * d.nc + kidx(m)] * invn;
      if constexpr (keep_k3) k3s[m] = k3;
      xb[m].x = (R)(k3 * (double)xa[m].x);
      xb[m].y = (R)(k3 * (double)xa[m].y);
    }
  }
  fft2<N, +1>(xb, line, t, tw, active);
  store(xb, 0);
  // phase 2: b' = K2 A + K4 B into the slot of B; xa <- K1 A + K2 B
  load(xb, 1);
  fft2<N, -1>(xb, line, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const size_t gi = kidx(m);
      const double k1 = wk[gi] * invn, k2 = wk[d.nc + gi] * invn, k4 = wk[3 * d.nc + gi] * invn;
      const double ar = xa[m].x, ai = xa[m].y, br = xb[m].x, bi = xb[m].y;
      xa[m].x = (R)(k1 * ar + k2 * br);
      xa[m].y = (R)(k1 * ai + k2 * bi);
      xb[m].x = (R)(k2 * ar + k4 * br);
      xb[m].y = (R)(k2 * ai + k4 * bi);
    }
  }
  fft2<N, +1>(xb, line, t, tw, active);
  store(xb, 1);
  // phase 3: a' = (K1 A + K2 B) + K3 C into the slot of C
  load(xb, 2);
  fft2<N, -1>(xb, line, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      double k3;
      if constexpr (keep_k3) k3 = k3s[m];
      else k3 = wk[2 * d.nc + kidx(m)] * invn;
      xa[m].x = (R)((double)xa[m].x + k3 * (double)xb[m].x);
      xa[m].y = (R)((double)xa[m].y + k3 * (double)xb[m].y);
    }
  }
  fft2<N, +1>(xa, line, t, tw, active);
  store(xa, 2);
}

// ------------------------------------------------------------ x inverse (C2R) + epilogue
template <int KIND>
__device__ __forceinline__ double epi(const ReqDesc& D, double rho, double vion, const double* o, double (&acc)[kNSum]) {
  PtVals p;
  point_vals(rho, D.alpha, D.beta, p);
  const double r23 = p.r13 * p.r13;
  const double vtf = (5.0 / 3.0 * kCTF * D.lam) * r23;
  double vnl, t, enl;
  if (KIND == kWT3) {
    vnl = kCTF * (D.alpha * p.ram1 * o[1] + D.beta * p.rbm1 * o[1]);
    enl = p.ra * o[1];
    t = o[2];
  } else if (KIND == kWT4) {
    vnl = kCTF * (D.alpha * p.ram1 * o[1] + D.beta * p.rbm1 * o[2]);
    enl = p.ra * o[1];
    t = o[3];
  } else {
    // WGCPotentialPlus, second-order branch (KEDF_WGC.f90:208-286). The z
    // pass leaves each block as (c', b', a') in the slots of (A, B, C):
    // o[4], o[3], o[2] are a', b', c' of the rho^alpha block, o[7], o[6],
    // o[5] those of the rho^beta block
    const double th = rho - D.rhoS, a = D.alpha, b = D.beta;
    t = o[1];
    enl = p.rb * (o[4] + th * o[3] + th * th / 2.0 * o[2]);
    vnl = kCTF * (p.rbm1 * (b * o[4] + (rho + b * th) * o[3] + th * (rho + b * th / 2.0) * o[2]) +
                  p.ram1 * (a * o[7] + (rho + a * th) * o[6] + th * (rho + a * th / 2.0) * o[5]));
  }
  const double loc = vion + o[0] + p.vxc + vtf + vnl;
  acc[0] += rho * vion;
  acc[1] += o[0] * rho;
  acc[2] += p.exc;
  acc[3] += rho * r23;
  acc[4] += p.phi * t;
  acc[5] += enl;
  if (D.opt_sqrt) return 2.0 * p.phi * loc + D.mu * t;
  return loc + D.mu * t / (2.0 * p.phi);
}

// The epilogue of the FP32 passes (Wang-Teter kinds): the pointwise terms in
// FP32 with the hardware log2 / exp2 / sqrt (relative error about 1e-6, the
// order of the FP32 transforms' own), the energy sums in FP64.
template <int KIND>
__device__ __forceinline__ double epi32(const ReqDesc& D, double rho_d, double vion, const double* o,
                                        double (&acc)[kNSum]) {
  constexpr float cC = 0.62035049089940009f, cXp = -0.98474502184269641f, cXe = -0.73855876638202234f;
  constexpr float kLn2 = 0.69314718055994531f;
  const float rho = (float)rho_d, al = (float)D.alpha, be = (float)D.beta;
  const float l2 = __builtin_amdgcn_logf(rho);                     // log2 rho
  const float r13 = __builtin_amdgcn_exp2f(l2 * (1.0f / 3.0f));
  const float phi = __builtin_amdgcn_sqrtf(rho);
  const float ra = __builtin_amdgcn_exp2f(al * l2), ram1 = __builtin_amdgcn_exp2f((al - 1.0f) * l2);
  const float rb = KIND == kWT3 ? ra : __builtin_amdgcn_exp2f(be * l2);
  const float rbm1 = KIND == kWT3 ? ram1 : __builtin_amdgcn_exp2f((be - 1.0f) * l2);
  const float l2rs = __builtin_amdgcn_logf(cC) - l2 * (1.0f / 3.0f);  // log2 rs
  const float rs = __builtin_amdgcn_exp2f(l2rs);
  float vxc, ec;
  if (rs < 1.0f) {
    const float lg = l2rs * kLn2;
    vxc = cXp * r13 + lg * (3.11e-2f + 1.33333333333333333e-3f * rs) - 5.83666666666666666e-2f - 8.4e-3f * rs;
    ec = lg * (3.11e-2f + 2e-3f * rs) - 4.8e-2f - 1.16e-2f * rs;
  } else {
    const float sq = __builtin_amdgcn_sqrtf(rs);
    const float iden = __builtin_amdgcn_rcpf(1.0f + 1.0529f * sq + 3.334e-1f * rs);
    vxc = cXp * r13 + (-1.423e-1f - 0.174798948333333333f * sq - 6.325709333333333333e-2f * rs) * iden * iden;
    ec = -1.423e-1f * iden;
  }
  const float r23 = r13 * r13;
  const float vtf = (float)(5.0 / 3.0 * kCTF * D.lam) * r23;
  const float o1 = (float)o[1];
  const float o2 = KIND == kWT3 ? o1 : (float)o[2];
  const float vnl = (float)kCTF * (al * ram1 * o1 + be * rbm1 * o2);
  const float t = (float)o[KIND == kWT3 ? 2 : 3];
  const float loc = (float)vion + (float)o[0] + vxc + vtf + vnl;
  acc[0] += rho_d * vion;
  acc[1] += o[0] * rho_d;
  acc[2] += (double)(cXe * rho * r13 + rho * ec);
  acc[3] += rho_d * (double)r23;
  acc[4] += (double)phi * (double)t;
  acc[5] += (double)ra * o[1];
  const float mu = (float)D.mu;
  if (D.opt_sqrt) return (double)(2.0f * phi * loc + mu * t);
  return (double)(loc + mu * t / (2.0f * phi));
}

// Epilogue of the x-inverse passes over the workgroup's RPB rows, whose NF
// inverse-transformed fields sit in LDS: points IOFF .. IOFF + PPR - 1 of
// each row (PPR = 2M: the whole row), point IOFF + ii of row rr, field j at
// real ii of line rr*NF + j (LR reals per line). Consecutive threads take
// consecutive points, EPT points per thread, in chunks of PF: the global
// inputs of a whole chunk are loaded before its first point is computed, so
// they are in flight together instead of one point at a time, and the first
// chunk's loads are issued before the workgroup barrier that completes the
// lines in LDS (this function starts with that barrier). The sums go to acc.
template <int M, int KIND, int INMODE, int OUTMODE, int RPB, int NT, int LR, int PPR, typename R>
__device__ __forceinline__ void x_epi_points(const ReqDesc& D, const R* __restrict__ lr, long long row0,
                                             long long nrows, int ioff, double (&acc)[kNSum]) {
  constexpr int NF = NF_<KIND>::v;
  constexpr int EPT = (RPB * PPR + NT - 1) / NT;
  // global loads per point: u and vion, v (trial points and the optimizer
  // forms), b and w (Newton form); at most 16 doubles in flight per thread
  constexpr int NL = 2 + (INMODE == kInPhi || OUTMODE != kOutPot ? 1 : 0) + (OUTMODE == kOutNewton ? 2 : 0);
  constexpr int PF = EPT < 16 / NL ? EPT : 16 / NL;
  const int nrow = (int)min((long long)RPB, nrows - row0);
  const int npts = nrow * PPR;
  const size_t gbase = (size_t)row0 * 2 * M + ioff;
  // global index of the workgroup's point e (row e / PPR, point e % PPR + ioff)
  auto gidx = [&](int e) { return PPR == 2 * M ? gbase + e : gbase + (size_t)(e / PPR) * 2 * M + e % PPR; };
  // the second vector: the phi direction of a trial point (kInPhi), the
  // search direction of the signed form, the CG direction of the Newton form
  const bool want_v = D.v != nullptr && (INMODE == kInPhi || OUTMODE != kOutPot);
#pragma unroll
  for (int c0 = 0; c0 < EPT; c0 += PF) {
    double pu[PF], pv[PF], pion[PF], pb[PF], pw[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      // points past the workgroup's last read its last point (in bounds, no
      // branch between the loads) and are not computed
      const size_t g = gidx(min((int)threadIdx.x + (c0 + q) * NT, npts - 1));
      pv[q] = pb[q] = pw[q] = 0.0;
      pu[q] = glb(D.u)[g];
      if (want_v) pv[q] = glb(D.v)[g];
      pion[q] = glb(D.vion)[g];
      if (OUTMODE == kOutNewton) {
        pb[q] = glb(D.bvec)[g];
        if (D.w != nullptr) pw[q] = glb(D.w)[g];
      }
    }
    if (c0 == 0) __syncthreads();
    // the empty asm makes the chunk's loads complete here, together, and
    // keeps the compiler from sinking each load into its point's block
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      asm volatile("" : "+v"(pu[q]), "+v"(pion[q]));
      if (want_v) asm volatile("" : "+v"(pv[q]));
      if (OUTMODE == kOutNewton) asm volatile("" : "+v"(pb[q]), "+v"(pw[q]));
    }
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int e = threadIdx.x + (c0 + q) * NT;
      if (c0 + q >= EPT || e >= npts) break;
      const int rr = e / PPR, ii = e - rr * PPR;
      const size_t g = gidx(e);
      double o[NF];
#pragma unroll
      for (int j = 0; j < NF; ++j) o[j] = (double)lr[(size_t)(rr * NF + j) * LR + ii];
      if constexpr (is_spin_kind(KIND)) {
        // both channels: the second one's inputs are loaded here
        double t1 = 0.0, t2 = 0.0, r1, r2;
        if (INMODE == kInRho) {
          r1 = pu[q];
          r2 = glb(D.u2)[g];
        } else {
          t1 = D.cb != 0.0 ? fma(D.cb, pv[q], D.ca * pu[q]) : D.ca * pu[q];
          const double u2 = glb(D.u2)[g];
          t2 = D.cb2 != 0.0 ? fma(D.cb2, (double)glb(D.v2)[g], D.ca2 * u2) : D.ca2 * u2;
          r1 = t1 * t1;
          r2 = t2 * t2;
        }
        double p1, p2;
        epi_point_spin<KIND>(D, r1, r2, pion[q], o, acc, p1, p2);
        if (OUTMODE == kOutPot) {
          glb(D.out)[g] = p1;
          glb(D.out2)[g] = p2;
        } else if (OUTMODE == kOutSigned) {
          const double de1 = p1 * fsign(t1), de2 = p2 * fsign(t2);
          glb(D.out)[g] = de1;
          glb(D.out2)[g] = de2;
          const double dir1 = D.s1 * pu[q] + (D.v != nullptr ? D.s2 * pv[q] : 0.0);
          const double dir2 = D.s1b * glb(D.u2)[g] + (D.v2 != nullptr ? D.s2b * glb(D.v2)[g] : 0.0);
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
          const double ap = ((de - D.mu0 * 2.0 * tt) - pb[q]) / D.eps;
          glb(D.out)[g] = ap;
          acc[6] += (up ? pv[q] : (double)glb(D.v2)[g]) * ap;
        }
        continue;
      }
      double tphi, rho;
      if (INMODE == kInRho) {
        tphi = 0.0;
        rho = pu[q];
      } else {
        tphi = D.cb != 0.0 ? fma(D.cb, pv[q], D.ca * pu[q]) : D.ca * pu[q];
        rho = tphi * tphi;
      }
      double pot;
      if constexpr (sizeof(R) == 4 && (KIND == kWT3 || KIND == kWT4)) pot = epi32<KIND>(D, rho, pion[q], o, acc);
      else pot = epi<KIND>(D, rho, pion[q], o, acc);
      if (OUTMODE == kOutPot) {
        glb(D.out)[g] = pot;
      } else if (OUTMODE == kOutSigned) {
        const double de = pot * fsign(tphi);
        glb(D.out)[g] = de;
        const double dir = D.s1 * pu[q] + (D.v != nullptr ? D.s2 * pv[q] : 0.0);
        acc[6] += de * dir;
        const double dl = de - 2.0 * tphi * D.mu0;
        acc[7] += dl * dl;
        acc[8] += de * tphi;
        acc[9] += de * de;
        acc[10] += tphi * tphi;
      } else {
        const double de = pot * fsign(tphi);
        const double ap = ((de - D.mu0 * 2.0 * tphi) - pb[q]) / D.eps;
        glb(D.out)[g] = ap;
        acc[6] += pv[q] * ap;
        if (D.w != nullptr) {
          acc[7] += pw[q] * ap;
          acc[8] += ap * ap;
        }
      }
    }
  }
}

// The workgroup's partial sums, laid out [slot][workgroup]; the finalize
// kernel sums each slot in a fixed order (no device-scope fence or atomic
// here: on a multi-XCD part a release fence per workgroup writes back the
// whole L2). The energy sums (slots 0-5), plus the optimizer sums of the
// signed form (6-10) or of the Newton form (6-8).
template <int OUTMODE>
__device__ __forceinline__ void x_epi_sums(const ReqDesc& D, double (&acc)[kNSum], int blk) {
  constexpr int NU = OUTMODE == kOutPot ? 6 : (OUTMODE == kOutNewton ? 9 : kNSum);
  block_sum_strided<kNSum, NU>(acc, glb(D.sums) + blk, gridDim.x);
}

// The epilogue over whole rows held as complex lines of stride LL.
template <int M, int KIND, int INMODE, int OUTMODE, int RPB, int NT, int LL, typename R>
__device__ __forceinline__ void x_epilogue(const ReqDesc& D, const R* __restrict__ lr, long long row0,
                                           long long nrows, int blk) {
  double acc[kNSum];
#pragma unroll
  for (int k = 0; k < kNSum; ++k) acc[k] = 0.0;
  x_epi_points<M, KIND, INMODE, OUTMODE, RPB, NT, 2 * LL, 2 * M>(D, lr, row0, nrows, 0, acc);
  x_epi_sums<OUTMODE>(D, acc, blk);
}

// One workgroup = RPB rows x NF fields; thread (line = row*NF + field, lane t).
// Every line is merged (C2R), transformed and left in LDS (complex k holds the
// reals 2k, 2k+1), then all threads run the epilogue over the workgroup's
// points, consecutive threads on consecutive points.
template <int M, int KIND, int INMODE, int OUTMODE, typename R>
__global__ void __launch_bounds__(256) k_fx_inv(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw,
                                                const typename V2<R>::T* __restrict__ tws) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<M>::R1, T = P2<M>::R2, NF = NF_<KIND>::v;
  constexpr int LL = line_lds<M>();
  constexpr int RPB = inv_rpb(NF, T);
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int r = line / NF, f = line - r * NF;
  const long long nrows = (long long)d.n2 * d.n3;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const long long row0 = (long long)blk * RPB;
  const long long row = row0 + r;
  const bool active = row < nrows;
  C* mine = lds + (size_t)line * LL;
  C z[R1];
  if (active) {
    gptr<const C> X = glb(reinterpret_cast<const C*>(D.work)) + (size_t)f * d.nc + (size_t)row * (M + 1);
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int k = t + T * m;
      C xk = ldc(X + k), xm = ldc(X + (M - k));
      if (k == 0) { xk.y = (R)0; xm.y = (R)0; }
      const R er = xk.x + xm.x, ei = xk.y - xm.y;
      const R dr = xk.x - xm.x, di = xk.y + xm.y;
      const C w = tws[k];
      const R orr = w.x * dr + w.y * di, oi = w.x * di - w.y * dr;
      z[m].x = er - oi;
      z[m].y = ei + orr;
    }
  }
  if constexpr (fx_scalar<M, NF, C>()) {
    // one real component at a time through a line of LS reals, then the
    // epilogue in two halves of each row: the complex outputs t + T m with
    // m < R1/2 are the row's reals 0 .. M-1, the others M .. 2M-1
    constexpr int LS = col_stride<M>(), H = R1 / 2;
    static_assert(R1 % 2 == 0 && LS >= M, "a half row fits a line");
    R* mr = reinterpret_cast<R*>(smem) + (size_t)line * LS;
    fft2s<M, +1>(z, mr, t, tw, active);
    double acc[kNSum];
#pragma unroll
    for (int k = 0; k < kNSum; ++k) acc[k] = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h) __syncthreads();   // the first half's points are read
      if (active) {
#pragma unroll
        for (int m = 0; m < H; ++m) {
          const int i = 2 * (t + T * m);   // the half's reals i, i + 1
          mr[i] = z[h * H + m].x;
          mr[i + 1] = z[h * H + m].y;
        }
      }
      x_epi_points<M, KIND, INMODE, OUTMODE, RPB, RPB * NF * T, LS, M>(D, reinterpret_cast<const R*>(smem), row0,
                                                                      nrows, h * M, acc);
    }
    x_epi_sums<OUTMODE>(D, acc, blk);
  } else {
    fft2<M, +1>(z, mine, t, tw, active);
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) mine[t + T * m] = z[m];
    }
    x_epilogue<M, KIND, INMODE, OUTMODE, RPB, RPB * NF * T, LL>(D, reinterpret_cast<const R*>(lds), row0, nrows, blk);
  }
}

// ------------------------------------------------------------ general two-stage lengths
// n = R1*R2 where no split has R2 dividing R1 (120 = 8*15, 60 = 4*15). A line
// is owned by T threads: thread t holds the stage-1 columns c = t + T*j
// (j < J1, c < R2), i.e. the points c + R2*m (m < R1), and after the forward
// transform the stage-2 rows k1 = t + T*q (q < J2, k1 < R1), i.e. the
// outputs k1 + R1*k2. The forward transform is decimation in time (columns
// -> rows), the inverse decimation in frequency (rows -> columns), so a
// pointwise G-space operator between the two and the loads and stores
// around them need no reordering; each direction crosses lanes once
// through the line's LDS block (R2 rows of LDR complex, LDR odd).
template <int N> struct GP;
template <> struct GP<60> { static constexpr int R1 = 4, R2 = 15, T = 4; };
template <> struct GP<120> { static constexpr int R1 = 8, R2 = 15, T = 8; };
template <int N> struct IsGP { static constexpr bool v = false; };
template <> struct IsGP<60> { static constexpr bool v = true; };
template <> struct IsGP<120> { static constexpr bool v = true; };
template <int N> struct GPD {
  static constexpr int R1 = GP<N>::R1, R2 = GP<N>::R2, T = GP<N>::T;
  static constexpr int J1 = (R2 + T - 1) / T, J2 = (R1 + T - 1) / T;
  static constexpr int LDR = R1 | 1;
  static constexpr int LINE = (R2 * LDR) | 1;   // complex elements per line block
  static constexpr int NC = J1 * R1, NS = J2 * R2;
  static_assert(LINE >= N, "a line block holds the whole line in natural order");
  __device__ static constexpr bool col_ok(int t, int j) { return J1 * T == R2 || t + T * j < R2; }
  __device__ static constexpr bool row_ok(int t, int q) { return J2 * T == R1 || t + T * q < R1; }
};

// forward, decimation in time: xc (columns) -> xs (rows); two barriers
template <int N, int S, typename C>
__device__ __forceinline__ void gfwd(C (&xc)[GPD<N>::NC], C (&xs)[GPD<N>::NS], C* __restrict__ lds, int t,
                                     const C* __restrict__ tw, bool active) {
  using G = GPD<N>;
  if (active) {
#pragma unroll
    for (int j = 0; j < G::J1; ++j) {
      if (!G::col_ok(t, j)) continue;
      const int c = t + G::T * j;
      Bfly<G::R1, S, C>::run(&xc[j * G::R1]);
#pragma unroll
      for (int m = 0; m < G::R1; ++m) lds[c * G::LDR + m] = xc[j * G::R1 + m];
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int q = 0; q < G::J2; ++q) {
      if (!G::row_ok(t, q)) continue;
      const int k1 = t + G::T * q;
      C y[G::R2];
#pragma unroll
      for (int c = 0; c < G::R2; ++c) y[c] = lds[c * G::LDR + k1];
#pragma unroll
      for (int c = 1; c < G::R2; ++c) {
        C w = tw[c * k1];
        if (S > 0) w.y = -w.y;
        y[c] = cmul(y[c], w);
      }
      Bfly<G::R2, S, C>::run(y);
#pragma unroll
      for (int k2 = 0; k2 < G::R2; ++k2) xs[q * G::R2 + k2] = y[k2];
    }
  }
  __syncthreads();
}

// inverse, decimation in frequency: xs (rows) -> xc (columns); two barriers
template <int N, int S, typename C>
__device__ __forceinline__ void ginv(C (&xs)[GPD<N>::NS], C (&xc)[GPD<N>::NC], C* __restrict__ lds, int t,
                                     const C* __restrict__ tw, bool active) {
  using G = GPD<N>;
  if (active) {
#pragma unroll
    for (int q = 0; q < G::J2; ++q) {
      if (!G::row_ok(t, q)) continue;
      const int k1 = t + G::T * q;
      Bfly<G::R2, S, C>::run(&xs[q * G::R2]);
      lds[k1] = xs[q * G::R2];
#pragma unroll
      for (int c = 1; c < G::R2; ++c) {
        C w = tw[c * k1];
        if (S > 0) w.y = -w.y;
        lds[c * G::LDR + k1] = cmul(xs[q * G::R2 + c], w);
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int j = 0; j < G::J1; ++j) {
      if (!G::col_ok(t, j)) continue;
      const int c = t + G::T * j;
#pragma unroll
      for (int k1 = 0; k1 < G::R1; ++k1) xc[j * G::R1 + k1] = lds[c * G::LDR + k1];
      Bfly<G::R1, S, C>::run(&xc[j * G::R1]);
    }
  }
  __syncthreads();
}

// y columns (flattened over (i3, k1)), one field per blockIdx.y: S = -1
// loads columns and stores rows, S = +1 the reverse.
template <int N, int S, typename R>
__global__ void __launch_bounds__(256) g_fy(const ReqDesc* __restrict__ descs, GridDims d,
                                            const typename V2<R>::T* __restrict__ tw, int kb, int fo) {
  using C = typename V2<R>::T;
  using G = GPD<N>;
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int cl = threadIdx.x % kb, t = threadIdx.x / kb;
  const long long col = (long long)xcd_block(blockIdx.x, gridDim.x) * kb + cl;
  const long long ncol = (long long)d.K1 * d.n3;
  const bool active = col < ncol;
  const long long i3 = active ? col / d.K1 : 0, k1c = active ? col - i3 * d.K1 : 0;
  gptr<C> base = glb(reinterpret_cast<C*>(D.work)) + (size_t)(fo + blockIdx.y) * d.nc + (size_t)i3 * d.n2 * d.K1 + k1c;
  C xc[G::NC], xs[G::NS];
  if (S < 0) {
    if (active) {
#pragma unroll
      for (int j = 0; j < G::J1; ++j)
        if (G::col_ok(t, j)) {
#pragma unroll
          for (int m = 0; m < G::R1; ++m) xc[j * G::R1 + m] = ldc(base + (size_t)(t + G::T * j + G::R2 * m) * d.K1);
        }
    }
    gfwd<N, S>(xc, xs, lds + (size_t)cl * G::LINE, t, tw, active);
    if (active) {
#pragma unroll
      for (int q = 0; q < G::J2; ++q)
        if (G::row_ok(t, q)) {
#pragma unroll
          for (int k2 = 0; k2 < G::R2; ++k2) stc(base + (size_t)(t + G::T * q + G::R1 * k2) * d.K1, xs[q * G::R2 + k2]);
        }
    }
  } else {
    if (active) {
#pragma unroll
      for (int q = 0; q < G::J2; ++q)
        if (G::row_ok(t, q)) {
#pragma unroll
          for (int k2 = 0; k2 < G::R2; ++k2) xs[q * G::R2 + k2] = ldc(base + (size_t)(t + G::T * q + G::R1 * k2) * d.K1);
        }
    }
    ginv<N, S>(xs, xc, lds + (size_t)cl * G::LINE, t, tw, active);
    if (active) {
#pragma unroll
      for (int j = 0; j < G::J1; ++j)
        if (G::col_ok(t, j)) {
#pragma unroll
          for (int m = 0; m < G::R1; ++m) stc(base + (size_t)(t + G::T * j + G::R2 * m) * d.K1, xc[j * G::R1 + m]);
        }
    }
  }
}

// z columns + the diagonal G-space operator, one field per blockIdx.y
template <int N, int KIND, typename R>
__global__ void __launch_bounds__(256) g_fz_gop(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw, int kb, int fo) {
  using C = typename V2<R>::T;
  using G = GPD<N>;
  constexpr int NF = NF_<KIND>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int f = fo + blockIdx.y;
  const int cl = threadIdx.x % kb, t = threadIdx.x / kb;
  const long long col = (long long)xcd_block(blockIdx.x, gridDim.x) * kb + cl;
  const long long ncol = (long long)d.K1 * d.n2;
  const bool active = col < ncol;
  const int i2 = active ? (int)(col / d.K1) : 0;
  const int m1 = active ? (int)(col - (long long)i2 * d.K1) : 0;
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const size_t es = (size_t)d.n2 * d.K1;
  const double invn = 1.0 / (double)d.nr;
  gptr<C> base = glb(reinterpret_cast<C*>(D.work)) + (size_t)f * d.nc + (size_t)col;
  C xc[G::NC], xs[G::NS];
  if (active) {
#pragma unroll
    for (int j = 0; j < G::J1; ++j)
      if (G::col_ok(t, j)) {
#pragma unroll
        for (int m = 0; m < G::R1; ++m) xc[j * G::R1 + m] = ldc(base + (size_t)(t + G::T * j + G::R2 * m) * es);
      }
  }
  C* mine = lds + (size_t)cl * G::LINE;
  gfwd<N, -1>(xc, xs, mine, t, tw, active);
  if (active) {
#pragma unroll
    for (int q = 0; q < G::J2; ++q) {
      if (!G::row_ok(t, q)) continue;
#pragma unroll
      for (int k2 = 0; k2 < G::R2; ++k2) {
        const int i3 = t + G::T * q + G::R1 * k2;
        const int m3 = i3 > N / 2 ? i3 - N : i3;
        const R scr = gop_scale<KIND, R>(D, f, m1, m2, m3, (size_t)i3 * es + col, invn);
        C& v = xs[q * G::R2 + k2];
        v.x *= scr;
        v.y *= scr;
      }
    }
  }
  ginv<N, +1>(xs, xc, mine, t, tw, active);
  if (active) {
#pragma unroll
    for (int j = 0; j < G::J1; ++j)
      if (G::col_ok(t, j)) {
#pragma unroll
        for (int m = 0; m < G::R1; ++m) stc(base + (size_t)(t + G::T * j + G::R2 * m) * es, xc[j * G::R1 + m]);
      }
  }
}

// x rows, forward: fields, half-length FFT, R2C split (lines = (row, field))
template <int M, int KIND, int INMODE, typename R>
__global__ void __launch_bounds__(256) g_fx_fwd(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw,
                                                const typename V2<R>::T* __restrict__ tws) {
  using C = typename V2<R>::T;
  using G = GPD<M>;
  constexpr int NF = NF_<KIND>::v, T = G::T;
  constexpr int RPB = inv_rpb(NF, T);
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int r = line / NF, f = line - r * NF;
  const long long nrows = (long long)d.n2 * d.n3;
  const long long row = (long long)xcd_block(blockIdx.x, gridDim.x) * RPB + r;
  const bool active = row < nrows;
  C xc[G::NC], xs[G::NS];
  if (active) {
    const size_t base = (size_t)row * (2 * M);
#pragma unroll
    for (int j = 0; j < G::J1; ++j) {
      if (!G::col_ok(t, j)) continue;
#pragma unroll
      for (int m = 0; m < G::R1; ++m)
        xc[j * G::R1 + m] = field_pair<KIND, INMODE, R>(D, f, base + 2 * (t + T * j + G::R2 * m));
    }
  }
  C* mine = lds + (size_t)line * G::LINE;
  gfwd<M, -1>(xc, xs, mine, t, tw, active);
  if (active) {
#pragma unroll
    for (int q = 0; q < G::J2; ++q)
      if (G::row_ok(t, q)) {
#pragma unroll
        for (int k2 = 0; k2 < G::R2; ++k2) mine[t + T * q + G::R1 * k2] = xs[q * G::R2 + k2];
      }
  }
  __syncthreads();
  if (active) {
    gptr<C> out = glb(reinterpret_cast<C*>(D.work)) + (size_t)f * d.nc + (size_t)row * (M + 1);
    for (int k = t; k <= M; k += T) {
      const C zk = mine[k == M ? 0 : k];
      const C zm = mine[k == 0 ? 0 : M - k];
      const R er = (R)0.5 * (zk.x + zm.x), ei = (R)0.5 * (zk.y - zm.y);
      const R orr = (R)0.5 * (zk.y + zm.y), oi = (R)-0.5 * (zk.x - zm.x);
      const C w = tws[k];
      C x;
      x.x = er + (w.x * orr - w.y * oi);
      x.y = ei + (w.x * oi + w.y * orr);
      stc(out + k, x);
    }
  }
}

// x rows, inverse: C2R merge, half-length inverse FFT, epilogue and sums
template <int M, int KIND, int INMODE, int OUTMODE, typename R>
__global__ void __launch_bounds__(256) g_fx_inv(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw,
                                                const typename V2<R>::T* __restrict__ tws) {
  using C = typename V2<R>::T;
  using G = GPD<M>;
  constexpr int NF = NF_<KIND>::v, T = G::T;
  constexpr int RPB = inv_rpb(NF, T);
  constexpr int LL = G::LINE;
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  const int line = threadIdx.x / T, t = threadIdx.x % T;
  const int r = line / NF, f = line - r * NF;
  const long long nrows = (long long)d.n2 * d.n3;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const long long row0 = (long long)blk * RPB;
  const long long row = row0 + r;
  const bool active = row < nrows;
  C* mine = lds + (size_t)line * LL;
  C xc[G::NC], xs[G::NS];
  if (active) {
    gptr<const C> X = glb(reinterpret_cast<const C*>(D.work)) + (size_t)f * d.nc + (size_t)row * (M + 1);
#pragma unroll
    for (int q = 0; q < G::J2; ++q) {
      if (!G::row_ok(t, q)) continue;
#pragma unroll
      for (int k2 = 0; k2 < G::R2; ++k2) {
        const int k = t + T * q + G::R1 * k2;
        C xk = ldc(X + k), xm = ldc(X + (M - k));
        if (k == 0) { xk.y = (R)0; xm.y = (R)0; }
        const R er = xk.x + xm.x, ei = xk.y - xm.y;
        const R dr = xk.x - xm.x, di = xk.y + xm.y;
        const C w = tws[k];
        const R orr = w.x * dr + w.y * di, oi = w.x * di - w.y * dr;
        xs[q * G::R2 + k2].x = er - oi;
        xs[q * G::R2 + k2].y = ei + orr;
      }
    }
  }
  ginv<M, +1>(xs, xc, mine, t, tw, active);
  if (active) {
#pragma unroll
    for (int j = 0; j < G::J1; ++j)
      if (G::col_ok(t, j)) {
#pragma unroll
        for (int m = 0; m < G::R1; ++m) mine[t + T * j + G::R2 * m] = xc[j * G::R1 + m];
      }
  }
  x_epilogue<M, KIND, INMODE, OUTMODE, RPB, RPB * NF * T, LL>(D, reinterpret_cast<const R*>(lds), row0, nrows, blk);
}

}  // namespace fast

// ------------------------------------------------------------ dispatch
// line lengths with a two-stage plan: y and z columns, and n1/2 along x
static inline bool is_fast_len(int n) {
  switch (n) {
    case 24: case 32: case 40: case 48: case 64: case 80: case 96: case 128: case 192: case 256: return true;
    default: return false;
  }
}
static inline bool is_fast_half(int m) {
  switch (m) {
    case 16: case 24: case 32: case 40: case 48: case 64: case 96: case 128: case 256: return true;
    default: return false;
  }
}

// lengths of the general two-stage form (decimation in time forward, in
// frequency inverse)
static inline bool is_gp_len(int n) { return n == 60 || n == 120; }

inline bool fast_path_ok_impl(const ShapePlan& sp, int kind) {
  const bool xy = (is_fast_half(sp.d.M) || is_gp_len(sp.d.M)) && (is_fast_len(sp.d.n2) || is_gp_len(sp.d.n2));
  if (kind == kWGC8) return xy && is_fast_len(sp.d.n3);   // the coupled z pass has the two-stage form only
  return (kind == kWT3 || kind == kWT4 || kind == kWT3S || kind == kWT4S) && xy &&
         (is_fast_len(sp.d.n3) || is_gp_len(sp.d.n3));
}

struct FastTiles {
  int kby, kbz;
};

template <typename F> static inline int p2_of(int n, F get) {
  using namespace fast;
  switch (n) {
    case 16: return get(P2<16>());
    case 24: return get(P2<24>());
    case 32: return get(P2<32>());
    case 40: return get(P2<40>());
    case 48: return get(P2<48>());
    case 64: return get(P2<64>());
    case 80: return get(P2<80>());
    case 96: return get(P2<96>());
    case 128: return get(P2<128>());
    case 192: return get(P2<192>());
    default: return get(P2<256>());
  }
}
static inline int p2_T(int n) { return p2_of(n, [](auto p) { return decltype(p)::R2; }); }
static inline int p2_R1(int n) { return p2_of(n, [](auto p) { return decltype(p)::R1; }); }
static inline int p2_ll(int n) { return n + n / p2_R1(n); }
static inline int p2_cs(int n) { return p2_ll(n) | 1; }   // in real components (column passes)

template <typename F> static inline int gp_of(int n, F get) {
  using namespace fast;
  return n == 60 ? get(GPD<60>()) : get(GPD<120>());
}
// threads per line of the two-stage plan of length n (either form)
static inline int len_T(int n) {
  return is_gp_len(n) ? gp_of(n, [](auto g) { return decltype(g)::T; }) : p2_T(n);
}
// LDS bytes per line of a column pass (y: which = 1, exchanging one real
// component at a time in the power-of-two form; z: which = 2)
static inline long long col_line_bytes(int n, int cbytes, int which) {
  if (is_gp_len(n)) return (long long)gp_of(n, [](auto g) { return decltype(g)::LINE; }) * cbytes;
  return (long long)p2_cs(n) * (which == 1 ? cbytes / 2 : cbytes);
}

// Column tiles of the y and z passes: whole wavefronts, at most 256 threads
// and at most lds_target bytes of LDS per workgroup (a tuned kb_y/kb_z is
// bound by the LDS capacity instead).
static inline FastTiles fast_tiles(const ShapePlan& sp, int cbytes) {
  FastTiles ft;
  const long long tgt = sp.lds_target;
  const int Ty = len_T(sp.d.n2), Tz = len_T(sp.d.n3);
  const int wy = 64 / Ty, wz = 64 / Tz;
  const long long ly = col_line_bytes(sp.d.n2, cbytes, 1),
                  lz = col_line_bytes(sp.d.n3, cbytes, 2);
  const long long ty = sp.kb_y > 0 ? sp.lds_cap : tgt, tz = sp.kb_z > 0 ? sp.lds_cap : tgt;
  ft.kby = sp.kb_y > 0 ? std::max(wy, sp.kb_y / wy * wy) : 256 / Ty;
  while (ft.kby > wy && (ft.kby * ly > ty || ft.kby * Ty > 256)) ft.kby -= wy;
  ft.kbz = sp.kb_z > 0 ? std::max(wz, sp.kb_z / wz * wz) : 256 / Tz;
  while (ft.kbz > wz && (ft.kbz * lz > tz || ft.kbz * Tz > 256)) ft.kbz -= wz;
  return ft;
}

// Raises kernel K's dynamic LDS limit (once per size increase) when a launch
// needs more than 64 KiB.
template <auto K> static inline void allow_lds(size_t lds) {
  static size_t granted = 65536;
  if (lds > granted) {
    hipFuncSetAttribute((const void*)K, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    granted = lds;
  }
}

template <typename R, int M, int KIND, int INMODE, int OUTMODE>
static void launch_x(const ShapePlan& sp, const FastTiles& ft, const ReqDesc* dd, int nb, hipStream_t st, bool fwd,
                     const typename V2<R>::T* tw, const typename V2<R>::T* tws) {
  using namespace fast;
  using C = typename V2<R>::T;
  constexpr int NF = NF_<KIND>::v;
  (void)ft;
  if constexpr (IsGP<M>::v) {
    constexpr int RPB = fast::inv_rpb(NF, GPD<M>::T);
    const long long nrows = (long long)sp.d.n2 * sp.d.n3;
    const int ntile = (int)((nrows + RPB - 1) / RPB);
    const size_t lds = (size_t)RPB * NF * GPD<M>::LINE * sizeof(C);
    const dim3 grid(ntile, nb), block(RPB * NF * GPD<M>::T);
    if (fwd) {
      allow_lds<g_fx_fwd<M, KIND, INMODE, R>>(lds);
      hipLaunchKernelGGL((g_fx_fwd<M, KIND, INMODE, R>), grid, block, lds, st, dd, sp.d, tw, tws);
    } else {
      allow_lds<g_fx_inv<M, KIND, INMODE, OUTMODE, R>>(lds);
      hipLaunchKernelGGL((g_fx_inv<M, KIND, INMODE, OUTMODE, R>), grid, block, lds, st, dd, sp.d, tw, tws);
    }
    return;
  } else {
  constexpr int T = P2<M>::R2;
  if (fwd) {
    constexpr int RPB = fast::inv_rpb(NF, T);
    const long long nrows = (long long)sp.d.n2 * sp.d.n3;
    const int ntile = (int)((nrows + RPB - 1) / RPB);
    const size_t lds = fx_fwd_lds<M, NF, C>();
    allow_lds<k_fx_fwd<M, KIND, INMODE, R>>(lds);
    hipLaunchKernelGGL((k_fx_fwd<M, KIND, INMODE, R>), dim3(ntile, nb), dim3(RPB * NF * T), lds, st, dd, sp.d, tw,
                       tws);
  } else {
    constexpr int RPB = fast::inv_rpb(NF, T);
    const long long nrows = (long long)sp.d.n2 * sp.d.n3;
    const int ntile = (int)((nrows + RPB - 1) / RPB);
    const size_t lds = fx_fwd_lds<M, NF, C>();
    allow_lds<k_fx_inv<M, KIND, INMODE, OUTMODE, R>>(lds);
    hipLaunchKernelGGL((k_fx_inv<M, KIND, INMODE, OUTMODE, R>), dim3(ntile, nb), dim3(RPB * NF * T), lds, st, dd,
                       sp.d, tw, tws);
  }
  }
}

template <typename R, int N, int S>
static void launch_y(const ShapePlan& sp, int kb, int fo, int nf, const ReqDesc* dd, int nb, hipStream_t st,
                     const typename V2<R>::T* tw) {
  using namespace fast;
  using C = typename V2<R>::T;
  const long long ncol = (long long)sp.d.K1 * sp.d.n3;
  const int nblk = (int)((ncol + kb - 1) / kb);
  if constexpr (IsGP<N>::v) {
    const size_t lds = (size_t)kb * GPD<N>::LINE * sizeof(C);
    allow_lds<g_fy<N, S, R>>(lds);
    hipLaunchKernelGGL((g_fy<N, S, R>), dim3(nblk, nf, nb), dim3(kb * GPD<N>::T), lds, st, dd, sp.d, tw, kb, fo);
  } else {
    const size_t lds = (size_t)kb * col_stride<N>() * sizeof(R);
    allow_lds<k_fy<N, S, R>>(lds);
    hipLaunchKernelGGL((k_fy<N, S, R>), dim3(nblk, nf, nb), dim3(kb * P2<N>::R2), lds, st, dd, sp.d, tw, kb, fo);
  }
}

template <auto K, typename C>
static void launch_zk(dim3 grid, dim3 block, size_t lds, hipStream_t st, const ReqDesc* dd, const GridDims& d,
                      const C* tw, int kb, const C* twy, int fo) {
  allow_lds<K>(lds);
  hipLaunchKernelGGL(K, grid, block, lds, st, dd, d, tw, kb, twy, fo);
}

template <typename R, int N, int KIND>
static void launch_z(const ShapePlan& sp, int kb, const ReqDesc* dd, int nb, hipStream_t st,
                     const typename V2<R>::T* tw, int P, const typename V2<R>::T* twy, int fo, int fn) {
  using namespace fast;
  using C = typename V2<R>::T;
  const long long ncol = (long long)sp.d.K1 * sp.d.n2;
  const int nblk = (int)((ncol + kb - 1) / kb);
  if constexpr (KIND == kWGC8) {
    if constexpr (!IsGP<N>::v) {
      const size_t lds = (size_t)kb * col_stride<N>() * sizeof(C);
      allow_lds<k_fz_wgc<N, R>>(lds);
      hipLaunchKernelGGL((k_fz_wgc<N, R>), dim3(nblk, 3, nb), dim3(kb * P2<N>::R2), lds, st, dd, sp.d, tw, kb);
    }
  } else if constexpr (IsGP<N>::v) {
    const size_t lds = (size_t)kb * GPD<N>::LINE * sizeof(C);
    allow_lds<g_fz_gop<N, KIND, R>>(lds);
    hipLaunchKernelGGL((g_fz_gop<N, KIND, R>), dim3(nblk, fn, nb), dim3(kb * GPD<N>::T), lds, st, dd, sp.d, tw, kb,
                       fo);
  } else {
    // after the split plane pass: kb/P groups of P columns per workgroup (kb
    // rounded down to a multiple of P)
    // (the one-component z exchange: half the LDS per column, so a whole
    // 256-thread workgroup of columns unless the tuner picked a tile)
    const int kbj = P > 1 ? std::max(P, kb / P * P)
                          : (fz_scalar<N, R, 1>() && sp.kb_z <= 0 ? 256 / P2<N>::R2 : kb);
    const long long ngrp = (long long)sp.d.K1 * (sp.d.n2 / P);
    const dim3 grid((unsigned)((ngrp + kbj / P - 1) / (kbj / P)), fn, nb), block(kbj * P2<N>::R2);
    const size_t lds = (size_t)kbj * col_stride<N>() * sizeof(C);
    if (P == 2) launch_zk<k_fz_gop<N, KIND, R, 2>>(grid, block, lds, st, dd, sp.d, tw, kbj, twy, fo);
    else if (fz_scalar<N, R, 1>())
      launch_zk<k_fz_gop<N, KIND, R, 1>>(grid, block, lds / 2, st, dd, sp.d, tw, kbj, twy, fo);
    else launch_zk<k_fz_gop<N, KIND, R, 1>>(grid, block, lds, st, dd, sp.d, tw, kbj, twy, fo);
  }
}

// P > 1: the split form, N2 = n2 / P (the rows of one class)
template <typename R, int M, int N2, int KIND, int INMODE, int P = 1>
static bool launch_plane(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st,
                         const typename V2<R>::T* twx, const typename V2<R>::T* tws, const typename V2<R>::T* twy,
                         int fo, int fn) {
  using namespace fast;
  using C = typename V2<R>::T;
  using G = PlaneGeo<M, N2, (P > 1)>;
  const size_t lds = (size_t)G::ROWS * G::S * sizeof(C);
  if (lds > (size_t)sp.lds_cap || G::THREADS > 1024) return false;
  auto kern = k_plane_fwd<M, N2, KIND, INMODE, R, P>;
  static bool attr_set = false;
  if (!attr_set) {
    hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(sp.d.n3 * fn * P, nb), dim3(G::THREADS), lds, st, dd, sp.d, twx, tws, twy, fo, fn);
  return true;
}

// The split plane pass needs the z pass that finishes its y transform: the
// ordinary z + G-space kernel (not WGC's coupled one, not the general
// two-stage lengths).
template <int KIND> static inline bool plane_split_ok(const ShapePlan& sp) {
  return sp.plane_split > 1 && KIND != kWGC8 && !is_gp_len(sp.d.n3);
}

// The plane pass run, as its number of row classes: 0 none (separate x and y
// forward passes), 1 the whole plane, 2 the split form. (Four classes, a
// quarter of the rows per workgroup, measured slower at 128^3: the plane pass
// no faster, 26.5 us, and the z pass's radix-4 join 7.7 us dearer.)
template <typename R, int KIND, int INMODE>
static int try_plane(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st, const typename V2<R>::T* twx,
                     const typename V2<R>::T* tws, const typename V2<R>::T* twy, int fo, int fn) {
  if (sp.no_plane) return 0;
  const bool split = plane_split_ok<KIND>(sp);
  // the split form where the whole plane takes more than 64 KiB of LDS (at
  // most two workgroups per CU); smaller planes already share CUs, and
  // splitting them measured no faster
  auto go = [&](auto m, auto n2) -> int {
    constexpr int M = decltype(m)::value, N2 = decltype(n2)::value;
    using G0 = fast::PlaneGeo<M, N2>;
    constexpr size_t whole = (size_t)G0::ROWS * G0::S * sizeof(typename V2<R>::T);
    if (split && whole > 65536 && launch_plane<R, M, N2 / 2, KIND, INMODE, 2>(sp, dd, nb, st, twx, tws, twy, fo, fn))
      return 2;
    return launch_plane<R, M, N2, KIND, INMODE>(sp, dd, nb, st, twx, tws, twy, fo, fn) ? 1 : 0;
  };
  using std::integral_constant;
  if (sp.d.M == 64 && sp.d.n2 == 128) return go(integral_constant<int, 64>(), integral_constant<int, 128>());
  if (sp.d.M == 32 && sp.d.n2 == 64) return go(integral_constant<int, 32>(), integral_constant<int, 64>());
  if (sp.d.M == 16 && sp.d.n2 == 32) return go(integral_constant<int, 16>(), integral_constant<int, 32>());
  if (sp.d.M == 48 && sp.d.n2 == 96) return go(integral_constant<int, 48>(), integral_constant<int, 96>());
  if (sp.d.M == 40 && sp.d.n2 == 80) return go(integral_constant<int, 40>(), integral_constant<int, 80>());
  if (sp.d.M == 24 && sp.d.n2 == 48) return go(integral_constant<int, 24>(), integral_constant<int, 48>());
  return 0;
}

// Whether a plane pass (whole or split) serves this shape: the shapes
// try_plane knows, when the whole plane fits in LDS (the split form takes
// less), so that the branch launches it for certain.
template <typename R> static inline bool plane_for(const ShapePlan& sp) {
  using namespace fast;
  if (sp.no_plane) return false;
  auto fits = [&](auto m, auto n2) {
    using G = PlaneGeo<decltype(m)::value, decltype(n2)::value>;
    return (size_t)G::ROWS * G::S * sizeof(typename V2<R>::T) <= (size_t)sp.lds_cap && G::THREADS <= 1024;
  };
  using std::integral_constant;
  const int M = sp.d.M, n2 = sp.d.n2;
  if (M == 64 && n2 == 128) return fits(integral_constant<int, 64>(), integral_constant<int, 128>());
  if (M == 32 && n2 == 64) return fits(integral_constant<int, 32>(), integral_constant<int, 64>());
  if (M == 16 && n2 == 32) return fits(integral_constant<int, 16>(), integral_constant<int, 32>());
  if (M == 48 && n2 == 96) return fits(integral_constant<int, 48>(), integral_constant<int, 96>());
  if (M == 40 && n2 == 80) return fits(integral_constant<int, 40>(), integral_constant<int, 80>());
  if (M == 24 && n2 == 48) return fits(integral_constant<int, 24>(), integral_constant<int, 48>());
  return false;
}

// The pass chain of one launch set. (Forking the fields' independent
// passes, plane or forward y through inverse y, onto a second stream for a
// single request measured slower at 96^3 and 128^3 and 3 % faster at 256^3:
// the x-inverse pass still waits for both branches, and the smaller launches
// of each branch lose what the overlap gains.)
template <typename R, int KIND, int INMODE, int OUTMODE, int M>
static void fast_chain(const ShapePlan& sp, const FastTiles& ft, const ReqDesc* dd, int nb, hipStream_t st,
                       const typename V2<R>::T* twx, const typename V2<R>::T* tws, const typename V2<R>::T* twy,
                       const typename V2<R>::T* twz) {
  constexpr int NF = fast::NF_<KIND>::v;
  auto ylaunch = [&](auto sgn, hipStream_t s, int fo, int fn) {
    constexpr int S = decltype(sgn)::value;
    switch (sp.d.n2) {
      case 24: launch_y<R, 24, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 32: launch_y<R, 32, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 40: launch_y<R, 40, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 48: launch_y<R, 48, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 60: launch_y<R, 60, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 64: launch_y<R, 64, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 80: launch_y<R, 80, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 96: launch_y<R, 96, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 120: launch_y<R, 120, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 128: launch_y<R, 128, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      case 192: launch_y<R, 192, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
      default: launch_y<R, 256, S>(sp, ft.kby, fo, fn, dd, nb, s, twy); break;
    }
  };
  auto zlaunch = [&](hipStream_t s, int sy, int fo, int fn) {
    switch (sp.d.n3) {
      case 24: launch_z<R, 24, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 32: launch_z<R, 32, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 40: launch_z<R, 40, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 48: launch_z<R, 48, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 60: launch_z<R, 60, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 64: launch_z<R, 64, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 80: launch_z<R, 80, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 96: launch_z<R, 96, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 120: launch_z<R, 120, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 128: launch_z<R, 128, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      case 192: launch_z<R, 192, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
      default: launch_z<R, 256, KIND>(sp, ft.kbz, dd, nb, s, twz, sy, twy, fo, fn); break;
    }
  };
  const bool plane = plane_for<R>(sp);
  if (!plane) launch_x<R, M, KIND, INMODE, OUTMODE>(sp, ft, dd, nb, st, true, twx, tws);
  // one branch: the plane pass (or the forward y pass), z + G-space, inverse y
  auto branch = [&](hipStream_t s, int fo, int fn) {
    int sy = 1;
    if (plane) {
      const int pl = try_plane<R, KIND, INMODE>(sp, dd, nb, s, twx, tws, twy, fo, fn);
      sy = pl > 1 ? pl : 1;   // row classes the z pass joins
    } else {
      ylaunch(std::integral_constant<int, -1>(), s, fo, fn);
    }
    zlaunch(s, sy, fo, fn);
    ylaunch(std::integral_constant<int, +1>(), s, fo, fn);
  };
  branch(st, 0, NF);
  launch_x<R, M, KIND, INMODE, OUTMODE>(sp, ft, dd, nb, st, false, twx, tws);
}

template <typename R, int KIND, int INMODE, int OUTMODE>
static int fast_eval_t(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st) {
  const FastTiles ft = fast_tiles(sp, (int)sizeof(typename V2<R>::T));
  const bool f32 = sizeof(R) == 4;
  const typename V2<R>::T *twx, *tws, *twy, *twz;
  if constexpr (sizeof(R) == 4) {
    twx = sp.tw_x32; tws = sp.tw_s32; twy = sp.tw_y32; twz = sp.tw_z32;
  } else {
    twx = sp.tw_x; tws = sp.tw_s; twy = sp.tw_y; twz = sp.tw_z;
  }
  (void)f32;
  switch (sp.d.M) {
    case 16: fast_chain<R, KIND, INMODE, OUTMODE, 16>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 24: fast_chain<R, KIND, INMODE, OUTMODE, 24>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 32: fast_chain<R, KIND, INMODE, OUTMODE, 32>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 40: fast_chain<R, KIND, INMODE, OUTMODE, 40>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 48: fast_chain<R, KIND, INMODE, OUTMODE, 48>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 60: fast_chain<R, KIND, INMODE, OUTMODE, 60>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 96: fast_chain<R, KIND, INMODE, OUTMODE, 96>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 120: fast_chain<R, KIND, INMODE, OUTMODE, 120>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 64: fast_chain<R, KIND, INMODE, OUTMODE, 64>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 128: fast_chain<R, KIND, INMODE, OUTMODE, 128>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    default: fast_chain<R, KIND, INMODE, OUTMODE, 256>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename R, int KIND>
int fast_eval_k(const ShapePlan& sp, int inmode, int outmode, const ReqDesc* dd, int nb, hipStream_t st) {
  if (inmode == kInRho) return fast_eval_t<R, KIND, kInRho, kOutPot>(sp, dd, nb, st);
  if (outmode == kOutPot) return fast_eval_t<R, KIND, kInPhi, kOutPot>(sp, dd, nb, st);
  if (outmode == kOutSigned) return fast_eval_t<R, KIND, kInPhi, kOutSigned>(sp, dd, nb, st);
  return fast_eval_t<R, KIND, kInPhi, kOutNewton>(sp, dd, nb, st);
}

}  // namespace ofd
