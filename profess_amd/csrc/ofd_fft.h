// Device-side mixed-radix Stockham FFT engine for gfx950 (CDNA4).
//
// Transforms of length n = 2^a 3^b 5^c 7^d live in LDS, one buffer per
// "line" (a row, pencil or column of the 3D grid). Every stage reads a radix-r
// butterfly's r inputs from one LDS buffer and writes the autosorted outputs
// to the other (out-of-place Stockham), so no bit reversal pass is needed and
// a whole workgroup can run many lines of the same length in lock step: the
// butterfly loop is flat over (line, butterfly) pairs and one barrier closes
// each stage. Twiddles come from a per-length table in global memory
// (exp(-2 pi i j / n), j < n), cached in L1/L2; they are computed on the host
// in long double so both precisions get correctly rounded factors.
//
// Sign convention: S = -1 forward (e^{-i...}), S = +1 backward (e^{+i...}).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace ofd {

constexpr int kMaxStages = 12;

struct Plan1D {
  int n;
  int nst;
  int rad[kMaxStages];
};

template <typename R> struct V2;
template <> struct V2<double> { using T = double2; };
template <> struct V2<float> { using T = float2; };

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by S*i
template <int S, typename C>
__device__ __forceinline__ C mul_si(C a) {
  C r;
  if (S < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
  return r;
}

// ---------------------------------------------------------------- butterflies
template <int R, int S, typename C> struct Bfly;

template <int S, typename C> struct Bfly<2, S, C> {
  __device__ static void run(C* x) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <int S, typename C> struct Bfly<4, S, C> {
  __device__ static void run(C* x) {
    C t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
    C t2 = cadd(x[1], x[3]), t3 = mul_si<S>(csub(x[1], x[3]));
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    x[1] = cadd(t1, t3);
    x[3] = csub(t1, t3);
  }
};

template <int S, typename C> struct Bfly<8, S, C> {
  __device__ static void run(C* x) {
    using R = decltype(x[0].x);
    const R h = (R)0.70710678118654752440084436210485;
    C a[4] = {x[0], x[2], x[4], x[6]};
    C b[4] = {x[1], x[3], x[5], x[7]};
    Bfly<4, S, C>::run(a);
    Bfly<4, S, C>::run(b);
    // twiddle b[k] by w8^k, w8 = e^{S i pi/4}
    C b1 = b[1], b3 = b[3];
    C w1; w1.x = h * (b1.x - (R)S * b1.y); w1.y = h * (b1.y + (R)S * b1.x);
    C b2 = mul_si<S>(b[2]);
    // w8^3 = S*i * w8
    C w3; w3.x = h * (b3.x - (R)S * b3.y); w3.y = h * (b3.y + (R)S * b3.x);
    w3 = mul_si<S>(w3);
    x[0] = cadd(a[0], b[0]); x[4] = csub(a[0], b[0]);
    x[1] = cadd(a[1], w1);   x[5] = csub(a[1], w1);
    x[2] = cadd(a[2], b2);   x[6] = csub(a[2], b2);
    x[3] = cadd(a[3], w3);   x[7] = csub(a[3], w3);
  }
};

template <int S, typename C> struct Bfly<16, S, C> {
  __device__ static void run(C* x) {
    using R = decltype(x[0].x);
    // 4 x 4 decomposition: inputs x[4*j + m], j = 0..3
    C a[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
      for (int j = 0; j < 4; ++j) a[m][j] = x[4 * j + m];
      Bfly<4, S, C>::run(a[m]);
    }
    // twiddles w16^(m*k), m,k in 0..3
    const R c1 = (R)0.92387953251128675612818318939679;
    const R s1 = (R)0.38268343236508977172845998403040;
    const R h = (R)0.70710678118654752440084436210485;
    const R cs[4] = {(R)1, c1, h, s1};     // cos(2 pi q/16), q = 0..3
    const R sn[4] = {(R)0, s1, h, c1};     // sin(2 pi q/16)
#pragma unroll
    for (int m = 1; m < 4; ++m) {
#pragma unroll
      for (int k = 1; k < 4; ++k) {
        int q = m * k;  // 1..9
        R c, s;
        // angle = S * 2 pi q / 16
        int qq = q & 3, quad = q >> 2;
        R cq = cs[qq], sq = sn[qq];
        // rotate by quad * 90 degrees
        if (quad == 0) { c = cq; s = sq; }
        else if (quad == 1) { c = -sq; s = cq; }
        else { c = -cq; s = -sq; }
        C w; w.x = c; w.y = (R)S * s;
        a[m][k] = cmul(a[m][k], w);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      C t[4] = {a[0][k], a[1][k], a[2][k], a[3][k]};
      Bfly<4, S, C>::run(t);
#pragma unroll
      for (int q = 0; q < 4; ++q) x[k + 4 * q] = t[q];
    }
  }
};

// odd prime radices: y_q = x0 + sum_j a_j cos(2 pi j q/p) + S i b_j sin(2 pi j q/p)
template <int P> struct PrimeConst;
template <> struct PrimeConst<3> {
  __device__ static double c(int k) { const double v[2] = {1.0, -0.5}; return v[k]; }
  __device__ static double s(int k) { const double v[2] = {0.0, 0.86602540378443864676372317075294}; return v[k]; }
};
template <> struct PrimeConst<5> {
  __device__ static double c(int k) {
    const double v[3] = {1.0, 0.30901699437494742410229341718282, -0.80901699437494742410229341718282};
    return v[k];
  }
  __device__ static double s(int k) {
    const double v[3] = {0.0, 0.95105651629515357211643933337938, 0.58778525229247312916870595463907};
    return v[k];
  }
};
template <> struct PrimeConst<7> {
  __device__ static double c(int k) {
    const double v[4] = {1.0, 0.62348980185873353052500488400424, -0.22252093395631440428890256449679,
                         -0.90096886790241912623610231950745};
    return v[k];
  }
  __device__ static double s(int k) {
    const double v[4] = {0.0, 0.78183148246802980870844452667406, 0.97492791218182360701813168299393,
                         0.43388373911755812047576833284836};
    return v[k];
  }
};

template <int P, int S, typename C> struct BflyPrime {
  __device__ static void run(C* x) {
    using R = decltype(x[0].x);
    constexpr int H = (P - 1) / 2;
    C a[H + 1], b[H + 1];
    C y0 = x[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      a[j] = cadd(x[j], x[P - j]);
      b[j] = csub(x[j], x[P - j]);
      y0 = cadd(y0, a[j]);
    }
    C out[P];
    out[0] = y0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      C re = x[0];
      C im; im.x = (R)0; im.y = (R)0;
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        int k = (j * q) % P;
        // cos(2 pi k/P) = cos(2 pi (P-k)/P); sin flips sign
        int kk = k <= H ? k : P - k;
        R cc = (R)PrimeConst<P>::c(kk);
        R ss = (R)PrimeConst<P>::s(kk);
        if (k > H) ss = -ss;
        re.x += a[j].x * cc; re.y += a[j].y * cc;
        im.x += b[j].x * ss; im.y += b[j].y * ss;
      }
      // y_q = re + S i im ; y_{P-q} = re - S i im
      C sim = mul_si<S>(im);
      out[q] = cadd(re, sim);
      out[P - q] = csub(re, sim);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) x[q] = out[q];
  }
};
template <int S, typename C> struct Bfly<3, S, C> { __device__ static void run(C* x) { BflyPrime<3, S, C>::run(x); } };
template <int S, typename C> struct Bfly<5, S, C> { __device__ static void run(C* x) { BflyPrime<5, S, C>::run(x); } };
template <int S, typename C> struct Bfly<7, S, C> { __device__ static void run(C* x) { BflyPrime<7, S, C>::run(x); } };

// ---------------------------------------------------------- composite radices
// cos and sin of 2 pi j / n evaluated by the compiler: quadrant reduction to
// an angle in [0, pi/4], then Taylor series (terms to x^25, far below one
// ulp there). Used only in constant expressions.
constexpr double ct_sin_series(double x) {
  double t = x, s = x;
  for (int k = 1; k <= 12; ++k) {
    t *= -x * x / (double)((2 * k) * (2 * k + 1));
    s += t;
  }
  return s;
}
constexpr double ct_cos_series(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k <= 12; ++k) {
    t *= -x * x / (double)((2 * k - 1) * (2 * k));
    s += t;
  }
  return s;
}
// out: c = cos(2 pi j/n), s = sin(2 pi j/n)
constexpr void ct_sincos(int j, int n, double& c, double& s) {
  const double half_pi = 1.5707963267948966192313216916398;
  j = ((j % n) + n) % n;
  const int q = (4 * j) / n;        // quadrant
  const int r = 4 * j - q * n;      // angle within it = (pi/2) r/n
  double cb = 0.0, sb = 0.0;
  if (2 * r <= n) {
    const double b = half_pi * (double)r / (double)n;
    cb = ct_cos_series(b);
    sb = ct_sin_series(b);
  } else {
    const double b = half_pi * (double)(n - r) / (double)n;
    cb = ct_sin_series(b);
    sb = ct_cos_series(b);
  }
  if (r == 0) { cb = 1.0; sb = 0.0; }
  if (q == 0) { c = cb; s = sb; }
  else if (q == 1) { c = -sb; s = cb; }
  else if (q == 2) { c = -cb; s = -sb; }
  else { c = sb; s = -cb; }
}
constexpr double ct_cos(int j, int n) { double c = 0, s = 0; ct_sincos(j, n, c, s); return c; }
constexpr double ct_sin(int j, int n) { double c = 0, s = 0; ct_sincos(j, n, c, s); return s; }
template <int J, int N> struct CtW {
  static constexpr double c = ct_cos(J, N);
  static constexpr double s = ct_sin(J, N);
};

// compile-time loop: f(std::integral_constant<int, I>) for I in [I0, E)
template <int I0, int E> struct SFor {
  template <typename F> __device__ __forceinline__ static void run(F&& f) {
    if constexpr (I0 < E) {
      f(std::integral_constant<int, I0>{});
      SFor<I0 + 1, E>::run(f);
    }
  }
};

// Radix A*B in registers: B radix-A butterflies over x[B*j + m] (j < A),
// twiddles exp(S 2 pi i m k / (A B)) as compile-time constants, then A
// radix-B butterflies; the result is in natural order, x[k + A*q].
template <int A, int B, int S, typename C> struct BflyAB {
  __device__ static void run(C* x) {
    using R = decltype(x[0].x);
    constexpr int N = A * B;
    C a[B][A];
#pragma unroll
    for (int m = 0; m < B; ++m) {
#pragma unroll
      for (int j = 0; j < A; ++j) a[m][j] = x[B * j + m];
      Bfly<A, S, C>::run(a[m]);
    }
    SFor<1, B>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      SFor<1, A>::run([&](auto ki) {
        constexpr int k = decltype(ki)::value;
        C w;
        w.x = (R)CtW<m * k, N>::c;
        w.y = (R)((double)S * CtW<m * k, N>::s);
        a[m][k] = cmul(a[m][k], w);
      });
    });
#pragma unroll
    for (int k = 0; k < A; ++k) {
      C t[B];
#pragma unroll
      for (int m = 0; m < B; ++m) t[m] = a[m][k];
      Bfly<B, S, C>::run(t);
#pragma unroll
      for (int q = 0; q < B; ++q) x[k + A * q] = t[q];
    }
  }
};
template <int S, typename C> struct Bfly<12, S, C> { __device__ static void run(C* x) { BflyAB<4, 3, S, C>::run(x); } };
template <int S, typename C> struct Bfly<20, S, C> { __device__ static void run(C* x) { BflyAB<4, 5, S, C>::run(x); } };
template <int S, typename C> struct Bfly<24, S, C> { __device__ static void run(C* x) { BflyAB<8, 3, S, C>::run(x); } };
template <int S, typename C> struct Bfly<15, S, C> { __device__ static void run(C* x) { BflyAB<5, 3, S, C>::run(x); } };

// ------------------------------------------------------------------ stages
// One Stockham stage of radix RAD over `nlines` lines of length n.
// Line l occupies in[l*ld .. l*ld+n). Ns = product of the radices already done.
template <int RAD, int S, typename C>
__device__ __forceinline__ void stage_rad(const C* __restrict__ in, C* __restrict__ out, int n, int ld, int nlines,
                                          int Ns, const C* __restrict__ tw, int tid, int nthr) {
  const int nr = n / RAD;
  const int tws = n / (Ns * RAD);
  const int total = nlines * nr;
  for (int it = tid; it < total; it += nthr) {
    const int line = it / nr;
    const int b = it - line * nr;
    const int k = b % Ns;
    const C* src = in + line * ld;
    C* dst = out + line * ld;
    C x[RAD];
#pragma unroll
    for (int m = 0; m < RAD; ++m) x[m] = src[b + m * nr];
    if (Ns > 1) {
#pragma unroll
      for (int m = 1; m < RAD; ++m) {
        C w = tw[m * k * tws];
        if (S > 0) w.y = -w.y;
        x[m] = cmul(x[m], w);
      }
    }
    Bfly<RAD, S, C>::run(x);
    const int j = (b - k) * RAD + k;
#pragma unroll
    for (int m = 0; m < RAD; ++m) dst[j + m * Ns] = x[m];
  }
}

// The same stage with the length, radix and stride known at compile time
// (divisions by constants, unrolled indexing).
template <int RAD, int S, int N, int NS, typename C>
__device__ __forceinline__ void stage_rad_c(const C* __restrict__ in, C* __restrict__ out, int ld, int nlines,
                                            const C* __restrict__ tw, int tid, int nthr) {
  constexpr int nr = N / RAD;
  constexpr int tws = N / (NS * RAD);
  const int total = nlines * nr;
  for (int it = tid; it < total; it += nthr) {
    const int line = it / nr;
    const int b = it - line * nr;
    const int k = b % NS;
    const C* src = in + line * ld;
    C* dst = out + line * ld;
    C x[RAD];
#pragma unroll
    for (int m = 0; m < RAD; ++m) x[m] = src[b + m * nr];
    if (NS > 1) {
#pragma unroll
      for (int m = 1; m < RAD; ++m) {
        C w = tw[m * k * tws];
        if (S > 0) w.y = -w.y;
        x[m] = cmul(x[m], w);
      }
    }
    Bfly<RAD, S, C>::run(x);
    const int j = (b - k) * RAD + k;
#pragma unroll
    for (int m = 0; m < RAD; ++m) dst[j + m * NS] = x[m];
  }
}

template <int S, int N, int NS, typename C, int R, int... Rest>
__device__ __forceinline__ C* stages_c(C* a, C* b, int ld, int nlines, const C* __restrict__ tw) {
  stage_rad_c<R, S, N, NS>(a, b, ld, nlines, tw, threadIdx.x, blockDim.x);
  __syncthreads();
  if constexpr (sizeof...(Rest) > 0) return stages_c<S, N, NS * R, C, Rest...>(b, a, ld, nlines, tw);
  else return b;
}

// Runs all stages of `p` on nlines lines that start in buffer a (scratch b).
// Returns the buffer holding the result (a if p.nst is even, else b).
// Must be called by every thread of the workgroup (it contains barriers).
// Lengths of the serving and benchmark grids run a compile-time plan; any
// other length runs the plan in `p`.
template <int S, typename C>
__device__ C* lds_fft(C* a, C* b, const Plan1D& p, int ld, int nlines, const C* __restrict__ tw) {
  switch (p.n) {
    case 24: return stages_c<S, 24, 1, C, 8, 3>(a, b, ld, nlines, tw);
    case 40: return stages_c<S, 40, 1, C, 8, 5>(a, b, ld, nlines, tw);
    case 48: return stages_c<S, 48, 1, C, 16, 3>(a, b, ld, nlines, tw);
    case 60: return stages_c<S, 60, 1, C, 4, 3, 5>(a, b, ld, nlines, tw);
    case 80: return stages_c<S, 80, 1, C, 16, 5>(a, b, ld, nlines, tw);
    case 96: return stages_c<S, 96, 1, C, 8, 4, 3>(a, b, ld, nlines, tw);
    case 120: return stages_c<S, 120, 1, C, 8, 3, 5>(a, b, ld, nlines, tw);
    case 192: return stages_c<S, 192, 1, C, 8, 8, 3>(a, b, ld, nlines, tw);
    default: break;
  }
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int r = p.rad[s];
    switch (r) {
      case 2: stage_rad<2, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
      case 3: stage_rad<3, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
      case 4: stage_rad<4, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
      case 5: stage_rad<5, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
      case 7: stage_rad<7, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
      case 8: stage_rad<8, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
      default: stage_rad<16, S>(a, b, p.n, ld, nlines, Ns, tw, tid, nthr); break;
    }
    __syncthreads();
    C* t = a; a = b; b = t;
    Ns *= r;
  }
  return a;
}

}  // namespace ofd
