// Internal declarations shared by the HIP kernels (ofd_kernels.hip) and the
// host side of the library (ofd_api.cpp). Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ofd_fft.h"

namespace ofd {

// Functional layouts of one evaluation: number of forward (NF) and
// inverse (NI) transformed fields.
enum FuncKind : int {
  kWT3 = 0,   // Wang-Teter, alpha == beta: F[rho], F[rho^b], F[sqrt rho]
  kWT4 = 1,   // Wang-Teter, alpha != beta: F[rho], F[rho^b], F[rho^a], F[sqrt rho]
  kWGC8 = 2,  // WGC 2nd order: F[rho], F[phi], 3 x rho^a terms, 3 x rho^b terms
  // two spin channels (numSpin = 2); x_s = 2 rho_s, rho = rho_up + rho_dn
  kWT3S = 3,  // F[rho], F[x_up^b], F[x_dn^b], F[sqrt x_up], F[sqrt x_dn]
  kWT4S = 4,  // F[rho], F[x_up^b], F[x_dn^b], F[x_up^a], F[x_dn^a], F[sqrt x_up], F[sqrt x_dn]
};
constexpr bool is_spin_kind(int k) { return k == kWT3S || k == kWT4S; }
inline int func_nf(int k) {
  switch (k) {
    case kWT3: return 3;
    case kWT4: return 4;
    case kWT3S: return 5;
    case kWT4S: return 7;
    default: return 8;
  }
}

// Input forms of the density: rho directly, or rho = (ca*u + cb*v)^2 with
// u, v two phi-like vectors (optimizer trial points).
enum InMode : int { kInRho = 0, kInPhi = 1 };
// Epilogue forms.
//  kOutPot     potential = dE/dphi (opt_sqrt) or dE/drho
//  kOutSigned  potential * sign(phi_t); extra sums
//                s0 = sum dEdphi*(s1*u + s2*v), s1 = sum (dEdphi - 2 mu0 phi_t)^2,
//                s2 = sum dEdphi*phi_t
//  kOutNewton  aux = (potential*sign(phi_t) - 2 mu0 phi_t - b)/eps written to
//              `out`; extra sum s0 = sum v*aux   (v = the CG direction p)
// Spin kinds write both channels' potentials (out, out2); kOutSigned then sums
//   6 dE/dphi_up*dir_up, 7 both channels' (dE/dphi_s - 2 mu_s phi_s)^2,
//   8 dE/dphi_up*phi_up, 9 dE/dphi_dn*dir_dn, 10 dE/dphi_dn*phi_dn
// and kOutNewton acts on channel newton_ch only (its p is v or v2).
enum OutMode : int { kOutPot = 0, kOutSigned = 1, kOutNewton = 2 };

// Per-pass partial sums written by every epilogue workgroup.
constexpr int kNSum = 11;
// 0 ion-electron, 1 Hartree(x2), 2 XC, 3 TF, 4 vW, 5 nonlocal, 6..10 optimizer sums:
//   kOutSigned: 6 sum dE/dphi*dir, 7 sum (dE/dphi - 2 mu0 phi)^2, 8 sum dE/dphi*phi,
//               9 sum (dE/dphi)^2, 10 sum phi^2
//   kOutNewton: 6 sum p*Ap, 7 sum r*Ap (when w is set), 8 sum Ap^2 (when w is set)

struct ReqDesc {
  // inputs of one evaluation
  const double* u;      // rho (kInRho) or phi base (kInPhi)
  const double* v;      // second phi vector (kInPhi), may be null when cb == 0
  const double* bvec;   // dL/dphi at the base point (kOutNewton)
  const double* w;      // CG residual r (kOutNewton), null when the sums 7, 8 are not wanted
  const double* vion;   // ionic potential (real grid)
  double* out;          // potential / dEdphi / Ap
  double* sums;         // kNSum partials per epilogue workgroup
  const double* wgck;   // WGC kernels K1..K4 (4 x Nc doubles), kWGC8 only
  const double* wtk;    // WT kernel coef * LindG / N (Nc doubles), WT kinds
  double* fin;          // kNSum finalized sums of this evaluation (device)
  void* work;           // complex work buffer, NF fields x Nc
  double ca, cb;        // phi_t = ca*u + cb*v
  double s1, s2;        // kOutSigned direction weights
  double mu0, eps;      // kOutSigned / kOutNewton scalars
  double inv_eps;
  // geometry (reciprocal vectors: G_c = sum_j rb[3c+j] m_j) and functional
  double rb[9];
  double lam, mu, alpha, beta, rhoS, coef, tkf;
  int opt_sqrt;
  int pad_;
  // second spin channel (spin kinds): rho_dn (kInRho) or phi_dn = ca2*u2 + cb2*v2,
  // its potential, its kOutSigned direction weights and chemical potential
  const double* u2;
  const double* v2;
  double* out2;
  double ca2, cb2, s1b, s2b, mu0b;
  int lsda;        // exchange-correlation of the spin kinds: 1 Perdew-Zunger, 2 Perdew-Wang 92
  int newton_ch;   // kOutNewton with spin kinds: the perturbed channel (0 up, 1 down)
  const float* wtk32;   // FP32 copy of wtk for the FP32 passes (may be null)
};

struct GridDims {
  int n1, n2, n3;
  int M;    // n1/2
  int K1;   // n1/2 + 1
  long long nr;   // n1*n2*n3
  long long nc;   // K1*n2*n3
};

// Device-resident plans and twiddles for one grid shape, both precisions.
struct ShapePlan {
  GridDims d;
  Plan1D px, py, pz;
  // twiddles: forward tables exp(-2 pi i j/n) for the three 1D lengths and
  // the R2C/C2R split table exp(-2 pi i k/n1), k = 0..M
  double2* tw_x = nullptr;
  double2* tw_y = nullptr;
  double2* tw_z = nullptr;
  double2* tw_s = nullptr;
  float2* tw_x32 = nullptr;
  float2* tw_y32 = nullptr;
  float2* tw_z32 = nullptr;
  float2* tw_s32 = nullptr;
  // launch shapes: 0 = derived from the LDS budget (see tiles_for), else a
  // value picked by the tuner
  int rows_x = 0;   // rows per workgroup in the x passes
  int kb_y = 0;     // columns per workgroup in the y pass
  int kb_z = 0;     // columns per workgroup in the z pass
  int lds_cap = 65536;     // device limit per workgroup
  int lds_target = 65536;  // soft budget used for the automatic tiles
  int force_generic = 0;   // 1: never use the register-resident fast path
  int no_plane = 0;        // 1: never use the fused x+y plane kernel
  int plane_split = 2;     // 2: plane passes above 64 KiB of LDS in their even/odd-row split form (1: whole)
};

struct Tiles {
  int rows, kby, kbz, ntile;
};

// Device memory through a process-wide cache of freed blocks (ofd_plan.cpp).
// dev_free may only be called once no queued device work uses the block.
void* dev_alloc(size_t bytes);   // nullptr on failure
void dev_free(void* p);
// Tiles for NF fields with complex elements of `cbytes` bytes.
Tiles tiles_for(const ShapePlan& sp, int nf, int cbytes);

bool make_plan1d(int n, Plan1D* p);
int shape_plan_init(ShapePlan* sp, int n1, int n2, int n3, std::string* err);
void shape_plan_free(ShapePlan* sp);

// Runs one batched evaluation (all requests share `sp` and `kind`).
// descs is a device array of nb ReqDesc. `fp32` selects float FFT buffers.
int run_eval(const ShapePlan& sp, int kind, int inmode, int outmode, int fp32, const ReqDesc* d_descs, int nb,
             hipStream_t st);
// Sums the per-workgroup partials of every request: sums_out[req*kNSum + k].
int run_finalize(const ShapePlan& sp, int kind, int fp32, const ReqDesc* d_descs, int nb, double* d_sums_out,
                 hipStream_t st);
// Register-resident fast path (power-of-two WT grids; ofd_fast.hip)
bool fast_path_ok(const ShapePlan& sp, int kind);
int fast_eval(const ShapePlan& sp, int kind, int inmode, int outmode, int fp32, const ReqDesc* d_descs, int nb,
              hipStream_t st);
int fast_finalize(const ShapePlan& sp, int kind, int fp32, const ReqDesc* d_descs, int nb, double* out,
                  hipStream_t st);
// Plain 3D transforms of D.u -> D.work (forward, no 1/N) or D.work -> D.out.
int run_fft3d_plain(const ShapePlan& sp, int forward, const ReqDesc* d_descs, int nb, hipStream_t st);

// Vector kernels used by the optimizer (all FP64, device pointers).
void vec_axpby(double* y, const double* x, double a, double b, long long n, hipStream_t st);   // y = a*x + b*y
void vec_lincomb(double* z, const double* x, const double* y, double a, double b, long long n, hipStream_t st);  // z = fma(b, y, a*x)
void vec_sqrt_abs(double* y, const double* x, long long n, hipStream_t st);                   // y = sqrt(|x|)
void vec_square(double* y, const double* x, long long n, hipStream_t st);                     // y = x*x
void vec_fill(double* y, double v, long long n, hipStream_t st);                               // y = v
// dot products of up to 4 pairs into out[4] (device), deterministic order
void vec_dots(const double* const* a, const double* const* b, int npairs, long long n, double* d_partial,
              double* d_out, hipStream_t st);
void vec_scale_sign(double* y, const double* x, const double* s, long long n, hipStream_t st);  // y = x*sign(s)

// Tile candidates of the fast passes for the autotuner: columns per
// workgroup of the y pass (which = 1) or the z pass (which = 2), whole
// wavefronts within the LDS limit.
std::vector<int> fast_tile_candidates(const ShapePlan& sp, int kind, int fp32, int which);
// Whether the fused x+y plane pass exists for this shape and fits in LDS.
bool fast_plane_available(const ShapePlan& sp, int kind, int fp32);

// Batched vector operations over many requests of one grid size.
enum VecOpKind : int { kOpLincomb = 0, kOpAxpby = 1, kOpZero = 2, kOpCG = 3 };
struct VecOp {
  int kind;
  double* z;
  const double* x;
  const double* y;
  double a, b;
  double* x2;        // kOpCG: p (read and written)
  double* y2;        // kOpCG: r (read and written)
  const double* w;   // kOpCG: Ap
};
// kOpLincomb: z = fma(b, y, a*x); kOpAxpby: z = fma(a, x, b*z); kOpZero: z = 0;
// kOpCG (one CG step, the three axpby updates fused):
//   dir = fma(a, p, dir); r = fma(-a, Ap, r); p = fma(1, r, b*p)
void vec_ops_batched(const VecOp* d_ops, int nops, long long n, hipStream_t st);
struct DotDesc {
  const double* a;
  const double* b;
};
constexpr int kDotBlocksBatched = 256;
// res[k] = sum a_k . b_k (sum a_k when b_k is null), fixed-order reduction
// (kDotBlocksBatched partials each)
void vec_dots_batched(const DotDesc* d_dd, int ndots, long long n, double* d_part, double* d_res, hipStream_t st);

// Per-geometry tables.
// coef * LindG(|G|/tkf; lam, mu) / N on the half grid (WT kinds)
int build_wt_kernel(const ShapePlan& sp, const double* rb, double coef, double tkf, double lam, double mu,
                    double* d_out, hipStream_t st);
// dst = (float)src, n values (device)
void vec_to_f32(float* dst, const double* src, long long n, hipStream_t st);
int build_wgc_kernels(const ShapePlan& sp, const double* rb, const double* d_eta, const double* d_w,
                      const double* d_w1, const double* d_w2, const double* d_pp0, const double* d_pp1,
                      const double* d_pp2, int nknots, double tkfs, double gamma, double rhoS, double* d_out,
                      hipStream_t st);
int build_ion_potential(const ShapePlan& sp, const double* rb, double vol, int ntypes, const int* h_type_start,
                        const double* h_frac, const double* d_frac, int nion, const double* d_psp_vqs, const double* d_psp_dd,
                        const int* h_psp_off, const int* h_psp_npts, const double* h_psp_maxg,
                        const double* h_psp_v0, const int* h_psp_charge, int order, double* d_vion,
                        void* d_scratch, hipStream_t st);
int ion_forces_spline(const ShapePlan& sp, const double* rb, const double* cell_inv_t, int ntypes,
                      const int* h_type_start, const double* d_frac, int nion, const double* d_psp_vqs,
                      const double* d_psp_dd, const int* h_psp_off, const int* h_psp_npts,
                      const double* h_psp_maxg, const double* h_psp_v0, const int* h_psp_charge, int order,
                      const double* d_rho, double* d_forces, void* d_scratch, hipStream_t st);

// Ion-ion Ewald sum on the device (ofd_tables.hip): real-space pair sums and
// forces, particle-mesh reciprocal energy and forces on the density grid.
// comps4: real, self, reciprocal, average-potential terms (Hartree); h_forces
// (host, 3 per ion) may be NULL. Scratch as for build_ion_potential.
int ewald_device(const ShapePlan& sp, const double* rb, const double* lat, double vol, int ntypes,
                 const int* h_type_start, const double* h_charge, const double* h_frac, const double* d_frac,
                 int nion, int order,
                 double eta, double rc, double kc, const double* cell_inv_t, double* comps4, double* h_forces,
                 void* d_scratch, hipStream_t st);

}  // namespace ofd
