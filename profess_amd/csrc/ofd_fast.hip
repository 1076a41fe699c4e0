// Register-resident two-stage FFT passes for power-of-two lines (n = 16..256).
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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "ofd_internal.h"

namespace ofd {

namespace fast {

constexpr double kPi = 3.141592653589793238462643;
constexpr double kCTF = 2.87123400018819;

template <int N> struct P2;
template <> struct P2<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct P2<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct P2<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct P2<128> { static constexpr int R1 = 16, R2 = 8; };
template <> struct P2<256> { static constexpr int R1 = 16, R2 = 16; };

template <int N> __device__ __forceinline__ constexpr int padded(int j) { return j + j / P2<N>::R1; }
__device__ __forceinline__ int t_plus(int t, int T, int m) { return t + T * m; }
template <int N> constexpr int line_lds() { return N + N / P2<N>::R1; }

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

// ------------------------------------------------------------ pointwise
__device__ __forceinline__ double lda_pot(double r13) {
  const double cX = -0.98474502184269641, cC = 0.62035049089940009;
  const double a1 = 3.11e-2, cC1 = -5.83666666666666666e-2, cC2 = 1.33333333333333333e-3, cC3 = -8.4e-3;
  const double g1 = -1.423e-1, cC4 = -0.174798948333333333, cC5 = -6.325709333333333333e-2;
  const double b1 = 1.0529, b2 = 3.334e-1;
  double rs = cC / r13;
  if (rs < 1.0) return cX * r13 + log(rs) * (a1 + cC2 * rs) + cC1 + cC3 * rs;
  double sq = sqrt(rs);
  double den = 1.0 + b1 * sq + b2 * rs;
  return cX * r13 + (g1 + cC4 * sq + cC5 * rs) / (den * den);
}
__device__ __forceinline__ double lda_edens(double rho, double r13) {
  const double cX = -0.73855876638202234, cC = 0.62035049089940009;
  const double a = 3.11e-2, b = -4.8e-2, c = 2e-3, dd = -1.16e-2, g = -1.423e-1, b1 = 1.0529, b2 = 3.334e-1;
  double rs = cC / r13;
  double ec;
  if (rs < 1.0) ec = log(rs) * (a + c * rs) + b + dd * rs;
  else ec = g / (1.0 + b1 * sqrt(rs) + b2 * rs);
  return cX * rho * r13 + rho * ec;
}
__device__ __forceinline__ double lind_g(double eta, double lam, double mu) {
  if (eta < 0.0) return 0.0;
  if (eta < 1e-10) return 1.0 - lam + eta * eta * (1.0 / 3.0 - 3.0 * mu);
  if (fabs(eta - 1.0) < 1e-10) return 2.0 - lam - 3.0 * mu + 20.0 * (eta - 1.0);
  if (eta > 3.65) {
    double e2 = eta * eta, i2 = 1.0 / e2;
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

struct Pw {
  double phi, r13, ra, ram1, rb, rbm1;
};
__device__ __forceinline__ void pows(double rho, double alpha, double beta, Pw& p) {
  p.phi = sqrt(rho);
  p.r13 = cbrt(rho);
  if (alpha == 5.0 / 6.0 && beta == 5.0 / 6.0) {
    const double r16 = sqrt(p.r13);
    p.ra = rho / r16;
    p.ram1 = 1.0 / r16;
    p.rb = p.ra;
    p.rbm1 = p.ram1;
  } else {
    p.ra = pow(rho, alpha);
    p.ram1 = pow(rho, alpha - 1.0);
    p.rb = pow(rho, beta);
    p.rbm1 = pow(rho, beta - 1.0);
  }
}

template <int KIND> struct NF_;
template <> struct NF_<kWT3> { static constexpr int v = 3; };
template <> struct NF_<kWT4> { static constexpr int v = 4; };

template <int KIND>
__device__ __forceinline__ void fields(const ReqDesc& D, double rho, double* f) {
  Pw p;
  pows(rho, D.alpha, D.beta, p);
  f[0] = rho;
  f[1] = p.rb;
  if (KIND == kWT3) {
    f[2] = p.phi;
  } else {
    f[2] = p.ra;
    f[3] = p.phi;
  }
}

template <int INMODE>
__device__ __forceinline__ double load_rho(const ReqDesc& D, size_t g, double& tphi) {
  if (INMODE == kInRho) {
    tphi = 0.0;
    return D.u[g];
  }
  tphi = D.cb != 0.0 ? fma(D.cb, D.v[g], D.ca * D.u[g]) : D.ca * D.u[g];
  return tphi * tphi;
}

template <int N>
__device__ __forceinline__ void block_sum(double (&acc)[N], double* out) {
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

// ------------------------------------------------------------ x forward (R2C)
// blockDim = RPB * T threads; thread = (row r, lane t), t fastest.
template <int M, int KIND, int INMODE, typename R>
__global__ void __launch_bounds__(256) k_fx_fwd(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw,
                                                const typename V2<R>::T* __restrict__ tws, int rpb) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<M>::R1, T = P2<M>::R2, NF = NF_<KIND>::v;
  constexpr int LL = line_lds<M>();
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const long long nrows = (long long)d.n2 * d.n3;
  const long long row = (long long)blockIdx.x * rpb + r;
  const bool active = row < nrows;
  C z[NF][R1];
  if (active) {
    const size_t base = (size_t)row * (2 * M);
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int k = t + T * m;
      double a, b;
      {
        const double2 uu = *reinterpret_cast<const double2*>(D.u + base + 2 * k);
        if (INMODE == kInRho) {
          a = uu.x;
          b = uu.y;
        } else {
          double ta = D.ca * uu.x, tb = D.ca * uu.y;
          if (D.cb != 0.0) {
            const double2 vv = *reinterpret_cast<const double2*>(D.v + base + 2 * k);
            ta = fma(D.cb, vv.x, ta);
            tb = fma(D.cb, vv.y, tb);
          }
          a = ta * ta;
          b = tb * tb;
        }
      }
      double fa[NF], fb[NF];
      fields<KIND>(D, a, fa);
      fields<KIND>(D, b, fb);
#pragma unroll
      for (int j = 0; j < NF; ++j) {
        z[j][m].x = (R)fa[j];
        z[j][m].y = (R)fb[j];
      }
    }
  }
  C* mine = lds + (size_t)r * NF * LL;
#pragma unroll
  for (int j = 0; j < NF; ++j) fft2<M, -1>(z[j], mine + j * LL, t, tw, active);
  // Z to LDS (unpadded, per field), then the split
  if (active) {
#pragma unroll
    for (int j = 0; j < NF; ++j)
#pragma unroll
      for (int m = 0; m < R1; ++m) mine[j * LL + t + T * m] = z[j][m];
  }
  __syncthreads();
  if (active) {
    C* work = reinterpret_cast<C*>(D.work);
    const size_t ob = (size_t)row * (M + 1);
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      const C* Z = mine + j * LL;
#pragma unroll
      for (int m = 0; m <= R1; ++m) {
        const int k = m < R1 ? t + T * m : M;
        if (m == R1 && t != 0) continue;
        const C zk = Z[k == M ? 0 : k];
        const C zm = Z[k == 0 ? 0 : M - k];
        const R er = (R)0.5 * (zk.x + zm.x), ei = (R)0.5 * (zk.y - zm.y);
        const R orr = (R)0.5 * (zk.y + zm.y), oi = (R)-0.5 * (zk.x - zm.x);
        const C w = tws[k];
        C x;
        x.x = er + (w.x * orr - w.y * oi);
        x.y = ei + (w.x * oi + w.y * orr);
        work[(size_t)j * d.nc + ob + k] = x;
      }
    }
  }
}

// ------------------------------------------------------------ fused x + y forward (one plane)
// One workgroup per (z-plane i3, field f, request). The plane's half
// spectrum (K1 x n2) of the field stays in LDS between the x R2C and the y
// FFT: rows of stride S >= max(M + M/R1x, M + 1) complex, n2 + n2/R1y rows so
// that the y transform can use the padded column layout as its scratch.
template <int M, int N2> struct PlaneGeo {
  static constexpr int LLX = M + M / P2<M>::R1;
  static constexpr int S = LLX > M + 1 ? LLX : M + 1;
  static constexpr int ROWS = N2 + N2 / P2<N2>::R1;
  static constexpr int TX = P2<M>::R2, TY = P2<N2>::R2;
  static constexpr int NTX = N2 * TX, NTY = (M + 1) * TY;
  static constexpr int THREADS = ((NTX > NTY ? NTX : NTY) + 63) / 64 * 64;
};

template <int KIND>
__device__ __forceinline__ double field_of(const ReqDesc& D, int f, double rho) {
  if (f == 0) return rho;
  if (KIND == kWT3) {
    if (f == 2) return sqrt(rho);
  } else {
    if (f == 3) return sqrt(rho);
    if (f == 2) {
      if (D.alpha == 5.0 / 6.0) return rho / sqrt(cbrt(rho));
      return pow(rho, D.alpha);
    }
  }
  if (D.beta == 5.0 / 6.0) return rho / sqrt(cbrt(rho));
  return pow(rho, D.beta);
}

template <int M, int N2, int KIND, int INMODE, typename R>
__global__ void __launch_bounds__(1024) k_plane_fwd(const ReqDesc* __restrict__ descs, GridDims d,
                                                    const typename V2<R>::T* __restrict__ twx,
                                                    const typename V2<R>::T* __restrict__ tws,
                                                    const typename V2<R>::T* __restrict__ twy) {
  using C = typename V2<R>::T;
  using G = PlaneGeo<M, N2>;
  constexpr int R1X = P2<M>::R1, TX = G::TX, R1Y = P2<N2>::R1, TY = G::TY, S = G::S;
  extern __shared__ __align__(16) unsigned char smem[];
  C* P = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int f = blockIdx.y;
  const int i3 = blockIdx.x;
  // ---- x step: thread (row i2, lane t)
  const int row = threadIdx.x / TX, tx = threadIdx.x % TX;
  const bool ax = row < N2;
  C z[R1X];
  if (ax) {
    const size_t base = ((size_t)i3 * N2 + row) * (2 * M);
#pragma unroll
    for (int m = 0; m < R1X; ++m) {
      const int k = t_plus(tx, TX, m);
      const double2 uu = *reinterpret_cast<const double2*>(D.u + base + 2 * k);
      double a, b;
      if (INMODE == kInRho) {
        a = uu.x;
        b = uu.y;
      } else {
        double ta = D.ca * uu.x, tb = D.ca * uu.y;
        if (D.cb != 0.0) {
          const double2 vv = *reinterpret_cast<const double2*>(D.v + base + 2 * k);
          ta = fma(D.cb, vv.x, ta);
          tb = fma(D.cb, vv.y, tb);
        }
        a = ta * ta;
        b = tb * tb;
      }
      z[m].x = (R)field_of<KIND>(D, f, a);
      z[m].y = (R)field_of<KIND>(D, f, b);
    }
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
        for (int m = 1; m < R2; ++m) w2[m] = cmul(w2[m], twy[m * b2]);
        Bfly<R2, -1, C>::run(w2);
#pragma unroll
        for (int m = 0; m < R2; ++m) y[q + Q * m] = w2[m];
      }
    }
  }
  if (ay) {
    C* out = reinterpret_cast<C*>(D.work) + (size_t)f * d.nc + (size_t)i3 * N2 * (M + 1) + c;
#pragma unroll
    for (int m = 0; m < R1Y; ++m) out[(size_t)(ty + TY * m) * (M + 1)] = y[m];
  }
}

// ------------------------------------------------------------ y columns
// columns flattened over (i3, k1); blockDim = KB * T, thread = (t, c), c fastest
template <int N, int S, typename R>
__global__ void __launch_bounds__(256) k_fy(const ReqDesc* __restrict__ descs, GridDims d,
                                            const typename V2<R>::T* __restrict__ tw, int kb) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<N>::R1, T = P2<N>::R2;
  constexpr int LL = line_lds<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int c = threadIdx.x % kb, t = threadIdx.x / kb;
  const long long col = (long long)blockIdx.x * kb + c;
  const long long ncol = (long long)d.K1 * d.n3;
  const bool active = col < ncol;
  const long long i3 = active ? col / d.K1 : 0, k1 = active ? col - i3 * d.K1 : 0;
  C* base = reinterpret_cast<C*>(D.work) + (size_t)blockIdx.y * d.nc + (size_t)i3 * d.n2 * d.K1 + k1;
  C x[R1];
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) x[m] = base[(size_t)(t + T * m) * d.K1];
  }
  fft2<N, S>(x, lds + (size_t)c * LL, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) base[(size_t)(t + T * m) * d.K1] = x[m];
  }
}

// ------------------------------------------------------------ z columns + G-space
// One field per workgroup (blockIdx.y): for the Wang-Teter layouts the
// G-space operator is diagonal in the fields (Hartree 4 pi/G^2 on F[rho],
// K(|G|) on F[rho^b] (and F[rho^a]), G^2 on F[sqrt rho]).
template <int N, int KIND, typename R>
__global__ void __launch_bounds__(256) k_fz_gop(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw, int kb) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<N>::R1, T = P2<N>::R2, NF = NF_<KIND>::v;
  constexpr int LL = line_lds<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.z];
  const int f = blockIdx.y;
  const int c = threadIdx.x % kb, t = threadIdx.x / kb;
  const long long col = (long long)blockIdx.x * kb + c;
  const long long ncol = (long long)d.K1 * d.n2;
  const bool active = col < ncol;
  const int i2 = active ? (int)(col / d.K1) : 0;
  const int m1 = active ? (int)(col - (long long)i2 * d.K1) : 0;
  const int m2 = i2 > d.n2 / 2 ? i2 - d.n2 : i2;
  const size_t es = (size_t)d.n2 * d.K1;
  const double invn = 1.0 / (double)d.nr;
  C* base = reinterpret_cast<C*>(D.work) + (size_t)f * d.nc + (size_t)col;
  C x[R1];
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) x[m] = base[(size_t)(t + T * m) * es];
  }
  fft2<N, -1>(x, lds + (size_t)c * LL, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int i3 = t + T * m;
      const int m3 = i3 > N / 2 ? i3 - N : i3;
      const double gx = D.rb[0] * m1 + D.rb[1] * m2 + D.rb[2] * m3;
      const double gy = D.rb[3] * m1 + D.rb[4] * m2 + D.rb[5] * m3;
      const double gz = D.rb[6] * m1 + D.rb[7] * m2 + D.rb[8] * m3;
      const double q2 = gx * gx + gy * gy + gz * gz;
      double sc;
      if (f == 0) sc = q2 > 0.0 ? 4.0 * kPi / q2 * invn : 0.0;
      else if (f == NF - 1) sc = q2 * invn;
      else sc = D.coef * lind_g(sqrt(q2) / D.tkf, D.lam, D.mu) * invn;
      x[m].x = (R)((double)x[m].x * sc);
      x[m].y = (R)((double)x[m].y * sc);
    }
  }
  fft2<N, +1>(x, lds + (size_t)c * LL, t, tw, active);
  if (active) {
#pragma unroll
    for (int m = 0; m < R1; ++m) base[(size_t)(t + T * m) * es] = x[m];
  }
}

// ------------------------------------------------------------ x inverse (C2R) + epilogue
template <int KIND>
__device__ __forceinline__ double epi(const ReqDesc& D, double rho, double vion, const double* o, double (&acc)[kNSum]) {
  Pw p;
  pows(rho, D.alpha, D.beta, p);
  const double vxc = lda_pot(p.r13);
  const double r23 = p.r13 * p.r13;
  const double vtf = (5.0 / 3.0 * kCTF * D.lam) * r23;
  double vnl, t, enl;
  if (KIND == kWT3) {
    vnl = kCTF * (D.alpha * p.ram1 * o[1] + D.beta * p.rbm1 * o[1]);
    enl = p.ra * o[1];
    t = o[2];
  } else {
    vnl = kCTF * (D.alpha * p.ram1 * o[1] + D.beta * p.rbm1 * o[2]);
    enl = p.ra * o[1];
    t = o[3];
  }
  const double loc = vion + o[0] + vxc + vtf + vnl;
  acc[0] += rho * vion;
  acc[1] += o[0] * rho;
  acc[2] += lda_edens(rho, p.r13);
  acc[3] += rho * r23;
  acc[4] += p.phi * t;
  acc[5] += enl;
  if (D.opt_sqrt) return 2.0 * p.phi * loc + D.mu * t;
  return loc + D.mu * t / (2.0 * p.phi);
}

// One row per T threads; thread t keeps the reals 2k, 2k+1 (k = t + T*m) of
// every field in registers, so the epilogue runs on the thread's own points
// and LDS holds only one padded line per row for the FFT's lane exchange.
template <int EPI_OUT, int KIND>
__device__ __forceinline__ void epi_store(const ReqDesc& D, size_t g, double rho, double tphi, double vion,
                                          const double* o, double bv, double dirv, double (&acc)[kNSum],
                                          double& res) {
  const double pot = epi<KIND>(D, rho, vion, o, acc);
  if (EPI_OUT == kOutPot) {
    res = pot;
  } else if (EPI_OUT == kOutSigned) {
    const double de = pot * fsign(tphi);
    res = de;
    acc[6] += de * dirv;
    const double dl = de - 2.0 * tphi * D.mu0;
    acc[7] += dl * dl;
    acc[8] += de * tphi;
  } else {
    const double de = pot * fsign(tphi);
    const double ap = ((de - D.mu0 * 2.0 * tphi) - bv) / D.eps;
    res = ap;
    acc[6] += dirv * ap;
  }
}

template <int M, int KIND, int INMODE, int OUTMODE, typename R>
__global__ void __launch_bounds__(256) k_fx_inv(const ReqDesc* __restrict__ descs, GridDims d,
                                                const typename V2<R>::T* __restrict__ tw,
                                                const typename V2<R>::T* __restrict__ tws, int rpb) {
  using C = typename V2<R>::T;
  constexpr int R1 = P2<M>::R1, T = P2<M>::R2, NF = NF_<KIND>::v;
  constexpr int LL = line_lds<M>();
  extern __shared__ __align__(16) unsigned char smem[];
  C* lds = reinterpret_cast<C*>(smem);
  const ReqDesc& D = descs[blockIdx.y];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const long long nrows = (long long)d.n2 * d.n3;
  const long long row = (long long)blockIdx.x * rpb + r;
  const bool active = row < nrows;
  C* mine = lds + (size_t)r * LL;
  double o[NF][2 * R1];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    C z[R1];
    if (active) {
      const C* X = reinterpret_cast<const C*>(D.work) + (size_t)j * d.nc + (size_t)row * (M + 1);
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        const int k = t + T * m;
        C xk = X[k], xm = X[M - k];
        if (k == 0) { xk.y = (R)0; xm.y = (R)0; }
        const R er = xk.x + xm.x, ei = xk.y - xm.y;
        const R dr = xk.x - xm.x, di = xk.y + xm.y;
        const C w = tws[k];
        const R orr = w.x * dr + w.y * di, oi = w.x * di - w.y * dr;
        z[m].x = er - oi;
        z[m].y = ei + orr;
      }
    }
    fft2<M, +1>(z, mine, t, tw, active);
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      o[j][2 * m] = (double)z[m].x;       // complex k holds reals 2k, 2k+1
      o[j][2 * m + 1] = (double)z[m].y;
    }
  }
  double acc[kNSum];
#pragma unroll
  for (int k = 0; k < kNSum; ++k) acc[k] = 0.0;
  if (active) {
    const size_t base = (size_t)row * (2 * M);
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const size_t g = base + 2 * (size_t)(t + T * m);
      double rho0, rho1, tp0 = 0.0, tp1 = 0.0;
      const double2 uu = *reinterpret_cast<const double2*>(D.u + g);
      double dir0 = 0.0, dir1 = 0.0, bv0 = 0.0, bv1 = 0.0;
      if (INMODE == kInRho) {
        rho0 = uu.x;
        rho1 = uu.y;
      } else {
        double2 vv = make_double2(0.0, 0.0);
        if (D.v != nullptr) vv = *reinterpret_cast<const double2*>(D.v + g);
        tp0 = D.cb != 0.0 ? fma(D.cb, vv.x, D.ca * uu.x) : D.ca * uu.x;
        tp1 = D.cb != 0.0 ? fma(D.cb, vv.y, D.ca * uu.y) : D.ca * uu.y;
        rho0 = tp0 * tp0;
        rho1 = tp1 * tp1;
        if (OUTMODE == kOutSigned) {
          dir0 = D.s1 * uu.x + (D.v != nullptr ? D.s2 * vv.x : 0.0);
          dir1 = D.s1 * uu.y + (D.v != nullptr ? D.s2 * vv.y : 0.0);
        } else if (OUTMODE == kOutNewton) {
          dir0 = vv.x;
          dir1 = vv.y;
          const double2 bb = *reinterpret_cast<const double2*>(D.bvec + g);
          bv0 = bb.x;
          bv1 = bb.y;
        }
      }
      const double2 vi = *reinterpret_cast<const double2*>(D.vion + g);
      double oa[NF], ob[NF];
#pragma unroll
      for (int j = 0; j < NF; ++j) {
        oa[j] = o[j][2 * m];
        ob[j] = o[j][2 * m + 1];
      }
      double2 res;
      epi_store<OUTMODE, KIND>(D, g, rho0, tp0, vi.x, oa, bv0, dir0, acc, res.x);
      epi_store<OUTMODE, KIND>(D, g + 1, rho1, tp1, vi.y, ob, bv1, dir1, acc, res.y);
      *reinterpret_cast<double2*>(D.out + g) = res;
    }
  }
  block_sum<kNSum>(acc, D.sums + (size_t)blockIdx.x * kNSum);
  // the last workgroup of this request sums the partials in a fixed order
  __shared__ unsigned int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) is_last = (atomicAdd(D.cnt, 1u) == gridDim.x - 1) ? 1u : 0u;
  __syncthreads();
  if (is_last) {
    __threadfence();
    double a2[kNSum];
#pragma unroll
    for (int k = 0; k < kNSum; ++k) a2[k] = 0.0;
    for (int w = threadIdx.x; w < (int)gridDim.x; w += blockDim.x) {
#pragma unroll
      for (int k = 0; k < kNSum; ++k) a2[k] += __builtin_nontemporal_load(D.sums + (size_t)w * kNSum + k);
    }
    block_sum<kNSum>(a2, D.fin);
    if (threadIdx.x == 0) *D.cnt = 0u;
  }
}

__global__ void __launch_bounds__(256) k_fin(const ReqDesc* __restrict__ descs, int nwg, double* out) {
  const ReqDesc& D = descs[blockIdx.y];
  const int k = blockIdx.x;   // one workgroup per sum slot
  double acc[1] = {0.0};
  for (int w = threadIdx.x; w < nwg; w += blockDim.x) acc[0] += D.sums[(size_t)w * kNSum + k];
  __shared__ double red[16];
  double v = acc[0];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    out[(size_t)blockIdx.y * kNSum + k] = s;
  }
}

}  // namespace fast

// ------------------------------------------------------------ dispatch
static bool is_fast_len(int n) { return n == 32 || n == 64 || n == 128 || n == 256; }

bool fast_path_ok(const ShapePlan& sp, int kind) {
  return (kind == kWT3 || kind == kWT4) && is_fast_len(sp.d.M) && is_fast_len(sp.d.n2) && is_fast_len(sp.d.n3);
}

struct FastTiles {
  int rpb, kby, kbz, ntile;
};

static int p2_T(int n) { return n == 16 ? 4 : n == 32 ? 4 : n == 64 ? 8 : n == 128 ? 8 : 16; }
static int p2_R1(int n) { return n == 16 ? 4 : n == 32 ? 8 : n == 64 ? 8 : 16; }
static int p2_ll(int n) { return n + n / p2_R1(n); }

// Tiles of the fast passes: at most 256 threads per workgroup and at most
// lds_target bytes of LDS (the tuner may override rows_x/kb_y/kb_z).
FastTiles fast_tiles(const ShapePlan& sp, int nf, int cbytes) {
  FastTiles ft;
  const long long tgt = sp.lds_target;
  const int M = sp.d.M;
  const int llx = std::max(p2_ll(M), M + 1);
  // whole wavefronts only: lines per workgroup are multiples of 64/T
  const int wx = 64 / p2_T(M), wy = 64 / p2_T(sp.d.n2), wz = 64 / p2_T(sp.d.n3);
  ft.rpb = sp.rows_x > 0 ? sp.rows_x : 256 / p2_T(M);
  while (ft.rpb > wx && ((long long)ft.rpb * nf * llx * cbytes > tgt || ft.rpb * p2_T(M) > 256)) ft.rpb -= wx;
  ft.kby = sp.kb_y > 0 ? sp.kb_y : 256 / p2_T(sp.d.n2);
  while (ft.kby > wy && ((long long)ft.kby * p2_ll(sp.d.n2) * cbytes > tgt || ft.kby * p2_T(sp.d.n2) > 256)) ft.kby -= wy;
  ft.kbz = sp.kb_z > 0 ? sp.kb_z : 256 / p2_T(sp.d.n3);
  while (ft.kbz > wz && ((long long)ft.kbz * p2_ll(sp.d.n3) * cbytes > tgt || ft.kbz * p2_T(sp.d.n3) > 256)) ft.kbz -= wz;
  const long long nrows = (long long)sp.d.n2 * sp.d.n3;
  ft.ntile = (int)((nrows + ft.rpb - 1) / ft.rpb);
  return ft;
}

template <typename R, int M, int KIND, int INMODE, int OUTMODE>
static void launch_x(const ShapePlan& sp, const FastTiles& ft, const ReqDesc* dd, int nb, hipStream_t st, bool fwd,
                     const typename V2<R>::T* tw, const typename V2<R>::T* tws) {
  using namespace fast;
  using C = typename V2<R>::T;
  constexpr int T = P2<M>::R2, NF = NF_<KIND>::v;
  const int threads = ft.rpb * T;
  if (fwd) {
    size_t lds = (size_t)ft.rpb * NF * line_lds<M>() * sizeof(C);
    hipLaunchKernelGGL((k_fx_fwd<M, KIND, INMODE, R>), dim3(ft.ntile, nb), dim3(threads), lds, st, dd, sp.d, tw, tws,
                       ft.rpb);
  } else {
    const int rpb = 256 / T;
    const long long nrows = (long long)sp.d.n2 * sp.d.n3;
    const int ntile = (int)((nrows + rpb - 1) / rpb);
    size_t lds = (size_t)rpb * line_lds<M>() * sizeof(C);
    hipLaunchKernelGGL((k_fx_inv<M, KIND, INMODE, OUTMODE, R>), dim3(ntile, nb), dim3(rpb * T), lds, st, dd, sp.d,
                       tw, tws, rpb);
  }
}

template <typename R, int N, int S>
static void launch_y(const ShapePlan& sp, int kb, int nf, const ReqDesc* dd, int nb, hipStream_t st,
                     const typename V2<R>::T* tw) {
  using namespace fast;
  using C = typename V2<R>::T;
  const long long ncol = (long long)sp.d.K1 * sp.d.n3;
  const int nblk = (int)((ncol + kb - 1) / kb);
  size_t lds = (size_t)kb * line_lds<N>() * sizeof(C);
  hipLaunchKernelGGL((k_fy<N, S, R>), dim3(nblk, nf, nb), dim3(kb * P2<N>::R2), lds, st, dd, sp.d, tw, kb);
}

template <typename R, int N, int KIND>
static void launch_z(const ShapePlan& sp, int kb, const ReqDesc* dd, int nb, hipStream_t st,
                     const typename V2<R>::T* tw) {
  using namespace fast;
  using C = typename V2<R>::T;
  const long long ncol = (long long)sp.d.K1 * sp.d.n2;
  const int nblk = (int)((ncol + kb - 1) / kb);
  size_t lds = (size_t)kb * line_lds<N>() * sizeof(C);
  hipLaunchKernelGGL((k_fz_gop<N, KIND, R>), dim3(nblk, NF_<KIND>::v, nb), dim3(kb * P2<N>::R2), lds, st, dd, sp.d,
                     tw, kb);
}

template <typename R, int M, int N2, int KIND, int INMODE>
static bool launch_plane(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st,
                         const typename V2<R>::T* twx, const typename V2<R>::T* tws, const typename V2<R>::T* twy) {
  using namespace fast;
  using C = typename V2<R>::T;
  using G = PlaneGeo<M, N2>;
  const size_t lds = (size_t)G::ROWS * G::S * sizeof(C);
  if (lds > (size_t)sp.lds_cap || G::THREADS > 1024) return false;
  auto kern = k_plane_fwd<M, N2, KIND, INMODE, R>;
  static bool attr_set = false;
  if (!attr_set) {
    hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(sp.d.n3, NF_<KIND>::v, nb), dim3(G::THREADS), lds, st, dd, sp.d, twx, tws, twy);
  return true;
}

template <typename R, int KIND, int INMODE>
static bool try_plane(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st, const typename V2<R>::T* twx,
                      const typename V2<R>::T* tws, const typename V2<R>::T* twy) {
  if (sp.no_plane) return false;
  if (sp.d.M == 64 && sp.d.n2 == 128) return launch_plane<R, 64, 128, KIND, INMODE>(sp, dd, nb, st, twx, tws, twy);
  if (sp.d.M == 32 && sp.d.n2 == 64) return launch_plane<R, 32, 64, KIND, INMODE>(sp, dd, nb, st, twx, tws, twy);
  return false;
}

template <typename R, int KIND, int INMODE, int OUTMODE, int M>
static void fast_chain(const ShapePlan& sp, const FastTiles& ft, const ReqDesc* dd, int nb, hipStream_t st,
                       const typename V2<R>::T* twx, const typename V2<R>::T* tws, const typename V2<R>::T* twy,
                       const typename V2<R>::T* twz) {
  constexpr int NF = fast::NF_<KIND>::v;
  auto ylaunch = [&](auto sgn) {
    constexpr int S = decltype(sgn)::value;
    switch (sp.d.n2) {
      case 32: launch_y<R, 32, S>(sp, ft.kby, NF, dd, nb, st, twy); break;
      case 64: launch_y<R, 64, S>(sp, ft.kby, NF, dd, nb, st, twy); break;
      case 128: launch_y<R, 128, S>(sp, ft.kby, NF, dd, nb, st, twy); break;
      default: launch_y<R, 256, S>(sp, ft.kby, NF, dd, nb, st, twy); break;
    }
  };
  if (!try_plane<R, KIND, INMODE>(sp, dd, nb, st, twx, tws, twy)) {
    launch_x<R, M, KIND, INMODE, OUTMODE>(sp, ft, dd, nb, st, true, twx, tws);
    ylaunch(std::integral_constant<int, -1>());
  }
  switch (sp.d.n3) {
    case 32: launch_z<R, 32, KIND>(sp, ft.kbz, dd, nb, st, twz); break;
    case 64: launch_z<R, 64, KIND>(sp, ft.kbz, dd, nb, st, twz); break;
    case 128: launch_z<R, 128, KIND>(sp, ft.kbz, dd, nb, st, twz); break;
    default: launch_z<R, 256, KIND>(sp, ft.kbz, dd, nb, st, twz); break;
  }
  ylaunch(std::integral_constant<int, +1>());
  launch_x<R, M, KIND, INMODE, OUTMODE>(sp, ft, dd, nb, st, false, twx, tws);
}

template <typename R, int KIND, int INMODE, int OUTMODE>
static int fast_eval_t(const ShapePlan& sp, const ReqDesc* dd, int nb, hipStream_t st) {
  const FastTiles ft = fast_tiles(sp, fast::NF_<KIND>::v, (int)sizeof(typename V2<R>::T));
  const bool f32 = sizeof(R) == 4;
  const typename V2<R>::T *twx, *tws, *twy, *twz;
  if constexpr (sizeof(R) == 4) {
    twx = sp.tw_x32; tws = sp.tw_s32; twy = sp.tw_y32; twz = sp.tw_z32;
  } else {
    twx = sp.tw_x; tws = sp.tw_s; twy = sp.tw_y; twz = sp.tw_z;
  }
  (void)f32;
  switch (sp.d.M) {
    case 32: fast_chain<R, KIND, INMODE, OUTMODE, 32>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 64: fast_chain<R, KIND, INMODE, OUTMODE, 64>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    case 128: fast_chain<R, KIND, INMODE, OUTMODE, 128>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
    default: fast_chain<R, KIND, INMODE, OUTMODE, 256>(sp, ft, dd, nb, st, twx, tws, twy, twz); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename R, int KIND>
static int fast_eval_k(const ShapePlan& sp, int inmode, int outmode, const ReqDesc* dd, int nb, hipStream_t st) {
  if (inmode == kInRho) return fast_eval_t<R, KIND, kInRho, kOutPot>(sp, dd, nb, st);
  if (outmode == kOutPot) return fast_eval_t<R, KIND, kInPhi, kOutPot>(sp, dd, nb, st);
  if (outmode == kOutSigned) return fast_eval_t<R, KIND, kInPhi, kOutSigned>(sp, dd, nb, st);
  return fast_eval_t<R, KIND, kInPhi, kOutNewton>(sp, dd, nb, st);
}

int fast_eval(const ShapePlan& sp, int kind, int inmode, int outmode, int fp32, const ReqDesc* dd, int nb,
              hipStream_t st) {
  if (kind == kWT3)
    return fp32 ? fast_eval_k<float, kWT3>(sp, inmode, outmode, dd, nb, st)
                : fast_eval_k<double, kWT3>(sp, inmode, outmode, dd, nb, st);
  return fp32 ? fast_eval_k<float, kWT4>(sp, inmode, outmode, dd, nb, st)
              : fast_eval_k<double, kWT4>(sp, inmode, outmode, dd, nb, st);
}

int fast_ntile(const ShapePlan& sp, int kind, int fp32) {
  return fast_tiles(sp, func_nf(kind), fp32 ? (int)sizeof(float2) : (int)sizeof(double2)).ntile;
}

// The fast x-inverse pass reduces its partials itself (last workgroup), so
// there is nothing left to launch.
int fast_finalize(const ShapePlan&, int, int, const ReqDesc*, int, double*, hipStream_t) { return 0; }

}  // namespace ofd
