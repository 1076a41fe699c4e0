// C ABI of libprofess_amd.so: contexts, CalculatePotentialPlus, the NTN
// density optimizer and the batched request pool.
//
// The optimizer is a resumable state machine (NtnSolver) that stops every
// time it needs one evaluation of the potential; a single-request call
// (ofd_rho_opt) drives one solver to completion, and the pool drives many,
// batching the evaluations of all requests that share a grid shape. Vector
// state lives on the device; the host keeps only scalars (CG coefficients,
// the Moré–Thuente line-search state, convergence bookkeeping), exactly the
// data the reference keeps in scalars (RhoOptN.f90, RhoDirNew.f90,
// RhoLineSearch.f90).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ofd.h"
#include "ofd_internal.h"

using namespace ofd;

namespace {

constexpr double kPi = 3.141592653589793238462643;
constexpr double kCTF = 2.87123400018819;

#define HIPCHK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    }                                                                      \
  } while (0)

// device buffers of contexts and scratch come from the process-wide cache
template <typename T>
T* dalloc(size_t n) {
  void* p = dev_alloc((n ? n : 1) * sizeof(T));
  if (!p) throw std::runtime_error("device allocation failed");
  return (T*)p;
}

// Streams of contexts are recycled: creating and destroying a stream per
// request costs more than the request's own launches on a busy device.
std::mutex g_stream_mu;
std::vector<hipStream_t> g_streams;
hipStream_t stream_get() {
  {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (!g_streams.empty()) {
      hipStream_t s = g_streams.back();
      g_streams.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}
void stream_put(hipStream_t s) {
  if (!s) return;
  if (hipStreamSynchronize(s) != hipSuccess) {
    hipStreamDestroy(s);
    return;
  }
  std::lock_guard<std::mutex> lk(g_stream_mu);
  g_streams.push_back(s);
}

// A blocking copy ordered on `st` (never on the legacy null stream, which
// HIP refuses while another thread captures a graph).
hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

// ------------------------------------------------------------------ Moré–Thuente
// Restates MINPACK-2 dcstep/dcsrch (MathDCstep.f, MathDCsrch.f) as used by
// LineSearchTN (RhoLineSearch.f90:369-419).
struct LineSearch {
  double ftol, gtol, xtol, stpmin, stpmax;
  bool started = false;
  bool brackt = false;
  int stage = 1;
  double ginit = 0, gtest = 0, gx = 0, gy = 0, finit = 0, fx = 0, fy = 0, stx = 0, sty = 0, stmin = 0, stmax = 0,
         width = 0, width1 = 0;

  static void step(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp, double fp,
                   double dp, bool& brackt, double stpmin, double stpmax) {
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf, stpc, stpq, theta, s, gamma, p, q, r;
    if (fp > fx) {
      theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dx), std::fabs(dp)));
      gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp < stx) gamma = -gamma;
      p = (gamma - dx) + theta;
      q = ((gamma - dx) + gamma) + dp;
      r = p / q;
      stpc = stx + r * (stp - stx);
      stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
      stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2.0;
      brackt = true;
    } else if (sgnd < 0.0) {
      theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dx), std::fabs(dp)));
      gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp > stx) gamma = -gamma;
      p = (gamma - dp) + theta;
      q = ((gamma - dp) + gamma) + dx;
      r = p / q;
      stpc = stp + r * (stx - stp);
      stpq = stp + (dp / (dp - dx)) * (stx - stp);
      stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
      brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {
      theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dx), std::fabs(dp)));
      gamma = s * std::sqrt(std::fmax(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
      if (stp > stx) gamma = -gamma;
      p = (gamma - dp) + theta;
      q = (gamma + (dx - dp)) + gamma;
      r = p / q;
      if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
      else if (stp > stx) stpc = stpmax;
      else stpc = stpmin;
      stpq = stp + (dp / (dp - dx)) * (stx - stp);
      if (brackt) {
        stpf = std::fabs(stpc - stp) < std::fabs(stpq - stp) ? stpc : stpq;
        if (stp > stx) stpf = std::fmin(stp + 0.66 * (sty - stp), stpf);
        else stpf = std::fmax(stp + 0.66 * (sty - stp), stpf);
      } else {
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        stpf = std::fmin(stpmax, stpf);
        stpf = std::fmax(stpmin, stpf);
      }
    } else {
      if (brackt) {
        theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
        s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dy), std::fabs(dp)));
        gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
        if (stp > sty) gamma = -gamma;
        p = (gamma - dp) + theta;
        q = ((gamma - dp) + gamma) + dy;
        r = p / q;
        stpc = stp + r * (sty - stp);
        stpf = stpc;
      } else if (stp > stx) {
        stpf = stpmax;
      } else {
        stpf = stpmin;
      }
    }
    if (fp > fx) {
      sty = stp; fy = fp; dy = dp;
    } else {
      if (sgnd < 0.0) { sty = stx; fy = fx; dy = dx; }
      stx = stp; fx = fp; dx = dp;
    }
    stp = stpf;
  }

  // returns 1 = FG (evaluate at stp), 2 = convergence, 3 = warning, 4 = error
  int next(double& stp, double f, double g) {
    if (!started) {
      started = true;
      if (stp < stpmin || stp > stpmax || g >= 0.0 || ftol < 0.0 || gtol < 0.0 || xtol < 0.0 || stpmin < 0.0 ||
          stpmax < stpmin)
        return 4;
      brackt = false;
      stage = 1;
      finit = f;
      ginit = g;
      gtest = ftol * ginit;
      width = stpmax - stpmin;
      width1 = width / 0.5;
      stx = 0.0; fx = finit; gx = ginit;
      sty = 0.0; fy = finit; gy = ginit;
      stmin = 0.0;
      stmax = stp + 4.0 * stp;
      return 1;
    }
    const double ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
    int task = 1;
    if (brackt && (stp <= stmin || stp >= stmax)) task = 3;
    if (brackt && stmax - stmin <= xtol * stmax) task = 3;
    if (stp == stpmax && f <= ftest && g <= gtest) task = 3;
    if (stp == stpmin && (f > ftest || g >= gtest)) task = 3;
    if (f <= ftest && std::fabs(g) <= gtol * (-ginit)) task = 2;
    if (task != 1) return task;
    if (stage == 1 && f <= fx && f > ftest) {
      double fm = f - stp * gtest, fxm = fx - stx * gtest, fym = fy - sty * gtest;
      double gm = g - gtest, gxm = gx - gtest, gym = gy - gtest;
      step(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
      fx = fxm + stx * gtest;
      fy = fym + sty * gtest;
      gx = gxm + gtest;
      gy = gym + gtest;
    } else {
      step(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
    }
    if (brackt) {
      if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
      width1 = width;
      width = std::fabs(sty - stx);
    }
    if (brackt) {
      stmin = std::fmin(stx, sty);
      stmax = std::fmax(stx, sty);
    } else {
      stmin = stp + 1.1 * (stp - stx);
      stmax = stp + 4.0 * (stp - stx);
    }
    stp = std::fmax(stp, stpmin);
    stp = std::fmin(stp, stpmax);
    if ((brackt && (stp <= stmin || stp >= stmax)) || (brackt && stmax - stmin <= xtol * stmax)) stp = stx;
    return 1;
  }
};

}  // namespace

// ====================================================================== launch graphs
// The launch chain of one batched evaluation (five passes + finalize) is
// captured once per (functional kind, input mode, output mode, precision,
// descriptor array, batch width) and replayed. The ShapePlan passed with a
// key must outlive the cache (its twiddle tables are kernel arguments baked
// into the graph): the kernels read their per-request parameters
// from the descriptor array in device memory, so a replay only needs the
// descriptors uploaded first. The first use of a key runs eagerly (so that
// one-time function attributes are set outside any capture). No memset or
// copy is ever recorded. Capture is thread-local: a server's progress thread
// captures while other threads admit and collect requests. OFD_NO_GRAPH=1
// disables the capture.
struct GraphKey {
  int kind, inmode, outmode, fp32, nb;
  const void* desc;
  bool operator<(const GraphKey& o) const {
    return std::tie(kind, inmode, outmode, fp32, nb, desc) < std::tie(o.kind, o.inmode, o.outmode, o.fp32, o.nb, o.desc);
  }
};
struct GraphCache {
  std::map<GraphKey, hipGraphExec_t> exec;
  std::map<GraphKey, int> seen;
};
static void graphs_clear(GraphCache& g) {
  for (auto& kv : g.exec) hipGraphExecDestroy(kv.second);
  g.exec.clear();
  g.seen.clear();
}
// Launch graphs are opt-in (OFD_GRAPH=1; OFD_NO_GRAPH=1 always turns them
// off): on MI355X under ROCm 7 a replayed five-kernel evaluation took 4-5 us
// longer on the GPU than the same kernels launched one by one, at every grid
// from 32^3 to 256^3, and the launches of one evaluation cost the host less
// than its GPU time even at 32^3.
static bool graphs_enabled() {
  const char* off = std::getenv("OFD_NO_GRAPH");
  if (off && std::atoi(off) != 0) return false;
  const char* on = std::getenv("OFD_GRAPH");
  return on && std::atoi(on) != 0;
}
// Runs run_eval + run_finalize for `nb` descriptors at `dd` on `st`,
// through a cached graph when possible. Returns 0 on success.
static int launch_eval(GraphCache& gc, const ShapePlan& sp, int kind, int inmode, int outmode, int fp32,
                       const ReqDesc* dd, int nb, double* fin_out, hipStream_t st) {
  const GraphKey key{kind, inmode, outmode, fp32, nb, dd};
  auto direct = [&]() -> int {
    if (run_eval(sp, kind, inmode, outmode, fp32, dd, nb, st)) return -1;
    return run_finalize(sp, kind, fp32, dd, nb, fin_out, st);
  };
  if (!graphs_enabled()) return direct();
  auto it = gc.exec.find(key);
  if (it == gc.exec.end()) {
    if (gc.seen[key]++ == 0) return direct();   // first use: eager
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return direct();
    const int rc = direct();
    if (hipStreamEndCapture(st, &g) != hipSuccess || rc != 0 || g == nullptr) {
      if (g) hipGraphDestroy(g);
      (void)hipGetLastError();
      return direct();
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) {
      (void)hipGetLastError();
      return direct();
    }
    it = gc.exec.emplace(key, ex).first;
  }
  return hipGraphLaunch(it->second, st) == hipSuccess ? 0 : -1;
}

// ====================================================================== tuned launch shapes
// Tiles picked by ofd_autotune, per (grid, functional kind, precision); new
// contexts of a tuned shape start from them.
struct TuneKey {
  int n1, n2, n3, kind, fp32;
  bool operator<(const TuneKey& o) const {
    return std::tie(n1, n2, n3, kind, fp32) < std::tie(o.n1, o.n2, o.n3, o.kind, o.fp32);
  }
};
struct TuneChoice {
  int rows_x, kb_y, kb_z, no_plane;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tuned;

// ====================================================================== context
struct ofd_ctx {
  std::string err;
  ShapePlan sp;
  int kind = kWT3;
  int kinetic = 4;
  int precision = 0;
  double lat[9];      // columns = lattice vectors (bohr)
  double rb[9];       // G_c = sum_j rb[3c+j] m_j
  double vol = 0, dV = 0, num_ele = 0;
  double lam = 1, mu = 1, alpha = -100, beta = -100, gamma = -1, rho0 = -1, rhoS = -1, coef = 0, tkf = 0;
  double e_ionion = 0;
  bool have_vion = false, have_wgc = false;
  // spin (CellInfo numSpin, XC_LDA lsda, RhoOptimizers fixedmag/nmag/tolm):
  // with nspin = 2 every density/potential array holds both channels
  // (Fortran rho(n1,n2,n3,2): spin up, then spin down)
  int nspin = 1, lsda = 1, fixed_mag = 0, nmag = 100;
  double tolm = (double)5.0e-6f;
  double mag_state[3] = {0, 0, 0};   // last spin solve: magnetization, N_up, N_down
  int mag_iters = 0;
  hipStream_t st = nullptr;
  // device buffers
  double* vion = nullptr;
  double* rho = nullptr;     // staging of a host density
  double* pot = nullptr;     // staging of a host potential
  void* work = nullptr;      // NF x nc complex (FP64 size)
  double* sums = nullptr;    // partials
  double* fin = nullptr;     // kNSum finalized sums
  double* wgck = nullptr;    // 4 x nc
  double* wtk = nullptr;     // nc: WT kernel coef * LindG / N
  float* wtk32 = nullptr;    // its FP32 copy (precision 1)
  ReqDesc* desc = nullptr;   // one descriptor
  ReqDesc last_desc{};       // what `desc` holds after an enqueue
  ReqDesc* many_desc = nullptr;
  double* many_sums = nullptr;
  int many_cap = 0;
  std::vector<ReqDesc> many_last;
  GraphCache graphs;
  // optimizer vectors
  double *phi = nullptr, *dir = nullptr, *dedphi = nullptr, *dldphi = nullptr, *r = nullptr, *p = nullptr,
         *ap = nullptr, *tmp = nullptr;
  // NBF memory: the (s, y) ring of RhoDirBFGS (BFGSs, BFGSy, MBFGS = 5 pairs),
  // allocated on the first dir_type 3 solve
  static constexpr int kBfgsM = 5;
  double* bfgs_s[kBfgsM] = {};
  double* bfgs_y[kBfgsM] = {};
  double* dot_part = nullptr;
  double* dot_out = nullptr;
  // ion data kept for forces
  std::vector<int> type_start;
  std::vector<double> frac;
  std::vector<double> psp_vqs, psp_dd, psp_maxg, psp_v0;
  std::vector<int> psp_off, psp_npts, psp_charge;
  int spline_order = 10;
  double* d_frac = nullptr;
  double* d_psp_vqs = nullptr;
  double* d_psp_dd = nullptr;

  ReqDesc base_desc() const {
    ReqDesc D;
    std::memset(&D, 0, sizeof(D));
    D.vion = vion;
    D.sums = sums;
    D.wgck = wgck;
    D.wtk = wtk;
    D.wtk32 = wtk32;
    D.work = work;
    D.fin = fin;
    for (int i = 0; i < 9; ++i) D.rb[i] = rb[i];
    D.lam = lam; D.mu = mu; D.alpha = alpha; D.beta = beta; D.rhoS = rhoS; D.coef = coef; D.tkf = tkf;
    D.opt_sqrt = 1;
    D.lsda = lsda;
    return D;
  }
  long long nr() const { return sp.d.nr; }
  // points of one density array: n1*n2*n3 per spin channel
  long long nr_all() const { return sp.d.nr * nspin; }
  // the density/potential arrays of an evaluation (kInRho, kOutPot): both
  // channels from one array of nr_all() values when nspin = 2
  void set_rho_io(ReqDesc& D, const double* rho, double* pot) const {
    D.u = rho;
    D.out = pot;
    if (nspin == 2) {
      D.u2 = rho + nr();
      D.out2 = pot + nr();
    }
  }
};

static int fail(ofd_ctx* c, const std::string& m) {
  if (c) c->err = m;
  return -1;
}

// Inverse-transpose of the lattice (long double), G vectors rb = 2 pi A^{-T}
// (CellInfo.f90:281), volume |det A| (MathFunctions Volume).
static void cell_setup(const double* cell, double* rb, double* vol, double* inv_t) {
  long double a[3][3];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) a[i][j] = cell[3 * j + i];
  long double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                    a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  long double inv[3][3];
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  const long double tp = 6.283185307179586476925286766559005768L;
  // recip = inv^T * 2 pi ; rb[3c + j] = recip(c, j) = inv(j, c) * 2pi
  for (int c = 0; c < 3; ++c)
    for (int j = 0; j < 3; ++j) {
      rb[3 * c + j] = (double)(inv[j][c] * tp);
      if (inv_t) inv_t[3 * c + j] = (double)inv[j][c];
    }
  *vol = (double)fabsl(det);
}

extern "C" int ofd_create(const ofd_grid* g, const ofd_functional* f, double num_electrons, ofd_ctx** out) {
  if (!g || !f || !out) return -1;
  *out = nullptr;
  ofd_ctx* c = new ofd_ctx();
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      c->err = "no HIP device available";
      *out = c;
      return -2;
    }
    if (shape_plan_init(&c->sp, g->n1, g->n2, g->n3, &c->err)) {
      *out = c;
      return -1;
    }
    std::memcpy(c->lat, g->cell, sizeof(c->lat));
    cell_setup(g->cell, c->rb, &c->vol, nullptr);
    if (!(c->vol > 0.0)) {
      c->err = "cell volume is zero";
      *out = c;
      return -1;
    }
    c->dV = c->vol / (double)c->sp.d.nr;    // CellInfo.f90:288
    c->num_ele = num_electrons;
    c->kinetic = f->kinetic;
    c->precision = f->precision;
    c->lam = f->lambda < -99.0 ? 1.0 : f->lambda;
    c->mu = f->mu < -99.0 ? 1.0 : f->mu;
    const double tmp_rho = num_electrons / c->vol;
    if (f->kinetic == 4) {   // SetupKEDF.f90:217-251, KEDFRefresh :654-660
      c->alpha = f->alpha < 0.0 ? 5.0 / 6.0 : f->alpha;
      c->beta = f->beta < 0.0 ? c->alpha : f->beta;
      c->rho0 = f->rho0 > 0.0 ? f->rho0 : tmp_rho;
      c->coef = 5.0 / (9.0 * c->alpha * c->beta * std::pow(c->rho0, c->alpha + c->beta - 5.0 / 3.0));
      c->tkf = 2.0 * std::pow(3.0 * c->rho0 * kPi * kPi, 1.0 / 3.0);
      c->kind = c->alpha == c->beta ? kWT3 : kWT4;
    } else if (f->kinetic == 5) {   // SetupKEDF.f90:253-322, KEDFRefresh :662-665
      c->alpha = f->alpha < 0.0 ? (5.0 + std::sqrt(5.0)) / 6.0 : f->alpha;
      c->beta = f->beta < 0.0 ? (5.0 - std::sqrt(5.0)) / 6.0 : f->beta;
      c->gamma = f->gamma < 0.0 ? 2.7 : f->gamma;
      if (c->alpha < c->beta) std::swap(c->alpha, c->beta);
      c->rho0 = f->rho0 > 0.0 ? f->rho0 : tmp_rho;
      c->rhoS = f->rhoS > 0.0 ? f->rhoS : tmp_rho;
      c->kind = kWGC8;
    } else {
      c->err = "kinetic must be 4 (WT) or 5 (WGC): other KEDFs are outside this library";
      *out = c;
      return -1;
    }
    c->st = stream_get();
    const long long nr = c->sp.d.nr, nc = c->sp.d.nc;
    const int nf = func_nf(c->kind);
    c->vion = dalloc<double>(nr);
    HIPCHK(hipMemsetAsync(c->vion, 0, nr * sizeof(double), c->st));
    c->rho = dalloc<double>(nr);
    c->pot = dalloc<double>(nr);
    c->work = dalloc<double2>((size_t)nf * nc);
    const long long nrows = (long long)g->n2 * g->n3;
    c->sums = dalloc<double>((size_t)nrows * kNSum);
    c->fin = dalloc<double>(kNSum);
    if (c->kind == kWGC8) c->wgck = dalloc<double>(4 * nc);
    if (c->kind == kWT3 || c->kind == kWT4) {
      c->wtk = dalloc<double>(nc);
      if (build_wt_kernel(c->sp, c->rb, c->coef, c->tkf, c->lam, c->mu, c->wtk, c->st))
        throw std::runtime_error("WT kernel build failed");
      if (c->precision == 1) {   // the FP32 passes read the table in FP32
        c->wtk32 = dalloc<float>(nc);
        vec_to_f32(c->wtk32, c->wtk, nc, c->st);
      }
      HIPCHK(hipStreamSynchronize(c->st));
    }
    c->desc = dalloc<ReqDesc>(1);
    c->dot_part = dalloc<double>(512 * 4);
    c->dot_out = dalloc<double>(4);
    HIPCHK(hipStreamSynchronize(c->st));   // later calls may copy on other streams
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tuned.find(TuneKey{g->n1, g->n2, g->n3, c->kind, c->precision == 1 ? 1 : 0});
      if (it != g_tuned.end()) {
        c->sp.rows_x = it->second.rows_x;
        c->sp.kb_y = it->second.kb_y;
        c->sp.kb_z = it->second.kb_z;
        c->sp.no_plane = it->second.no_plane != 0;
      }
    }
  } catch (const std::exception& e) {
    c->err = e.what();
    *out = c;
    return -1;
  }
  *out = c;
  return 0;
}

static void free_opt(ofd_ctx* c) {
  double** v[] = {&c->phi, &c->dir, &c->dedphi, &c->dldphi, &c->r, &c->p, &c->ap, &c->tmp};
  for (auto pp : v) {
    dev_free(*pp);
    *pp = nullptr;
  }
  for (int k = 0; k < ofd_ctx::kBfgsM; ++k) {
    dev_free(c->bfgs_s[k]);
    dev_free(c->bfgs_y[k]);
    c->bfgs_s[k] = c->bfgs_y[k] = nullptr;
  }
}

extern "C" void ofd_destroy(ofd_ctx* c) {
  if (!c) return;
  if (c->st) hipStreamSynchronize(c->st);
  free_opt(c);
  for (void* b : {(void*)c->vion, (void*)c->rho, (void*)c->pot, c->work, (void*)c->sums, (void*)c->fin,
                  (void*)c->wgck, (void*)c->wtk, (void*)c->wtk32, (void*)c->desc, (void*)c->dot_part, (void*)c->dot_out,
                  (void*)c->many_desc, (void*)c->many_sums, (void*)c->d_frac, (void*)c->d_psp_vqs,
                  (void*)c->d_psp_dd})
    dev_free(b);
  graphs_clear(c->graphs);
  if (c->sp.tw_x) shape_plan_free(&c->sp);
  stream_put(c->st);
  delete c;
}

extern "C" const char* ofd_last_error(const ofd_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int ofd_get_params(const ofd_ctx* c, double* o) {
  if (!c || !o) return -1;
  o[0] = c->alpha; o[1] = c->beta; o[2] = c->gamma; o[3] = c->rho0; o[4] = c->rhoS; o[5] = c->lam; o[6] = c->mu;
  return 0;
}

extern "C" int ofd_get_volume(const ofd_ctx* c, double* vol) {
  if (!c || !vol) return -1;
  *vol = c->vol;
  return 0;
}

extern "C" int ofd_set_ion_terms(ofd_ctx* c, const double* vion, double eii) {
  if (!c || !vion) return fail(c, "null argument");
  if (copy_sync(c->vion, vion, c->nr() * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess)
    return fail(c, "copy of the ion potential failed");
  c->e_ionion = eii;
  c->have_vion = true;
  return 0;
}

extern "C" int ofd_get_ion_potential(ofd_ctx* c, double* vion) {
  if (!c || !vion) return fail(c, "null argument");
  if (copy_sync(vion, c->vion, c->nr() * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return fail(c, "copy of the ion potential failed");
  return 0;
}

// eTable from the finalized sums (CalculateEnergy, CalPotential.f90:610-644)
static __host__ __device__ void etable_calc(const double* s, double lam, double mu, double dV, double eii,
                                            double* et) {
  double loc[9] = {0};
  loc[2] = s[0];
  loc[3] = s[1] / 2.0;
  loc[4] = s[2];
  loc[6] = lam * kCTF * s[3];
  loc[7] = 0.5 * s[4] * mu;
  loc[8] = kCTF * s[5];
  loc[1] = loc[6] + loc[7] + loc[8];
  for (int i = 0; i < 9; ++i) et[i] = loc[i] * dV;
  et[5] = eii;
  et[0] = et[1] + et[2] + et[3] + et[4] + et[5];
}
static void etable_from_sums(const ofd_ctx* c, const double* s, double* et) {
  etable_calc(s, c->lam, c->mu, c->dV, c->e_ionion, et);
}

// One evaluation on the context's own descriptor; sums (kNSum) copied back.
static int eval_one(ofd_ctx* c, const ReqDesc& D, int inmode, int outmode, int fp32, double* h_sums) {
  if (c->kind == kWGC8 && !c->have_wgc) return fail(c, "WGC kernel table not set (ofd_set_wgc_table)");
  // stream-ordered upload from the context's own copy, which stays unchanged
  // until the synchronization below
  c->last_desc = D;
  if (hipMemcpyAsync(c->desc, &c->last_desc, sizeof(D), hipMemcpyHostToDevice, c->st) != hipSuccess)
    return fail(c, "descriptor upload failed");
  if (launch_eval(c->graphs, c->sp, c->kind, inmode, outmode, fp32, c->desc, 1, c->fin, c->st))
    return fail(c, "kernel launch failed");
  if (h_sums) {
    if (hipMemcpyAsync(h_sums, c->fin, kNSum * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess)
      return fail(c, "sum copy failed");
  }
  if (hipStreamSynchronize(c->st) != hipSuccess) return fail(c, "stream synchronize failed");
  return 0;
}

extern "C" int ofd_calpot_device(ofd_ctx* c, const double* rho, int opt_sqrt, double* pot, double* et) {
  if (!c || !rho || !pot) return fail(c, "null argument");
  ReqDesc D = c->base_desc();
  c->set_rho_io(D, rho, pot);
  D.opt_sqrt = opt_sqrt ? 1 : 0;
  double s[kNSum];
  if (eval_one(c, D, kInRho, kOutPot, c->precision == 1, s)) return -1;
  if (et) etable_from_sums(c, s, et);
  return 0;
}

extern "C" int ofd_calpot_enqueue(ofd_ctx* c, const double* rho, int opt_sqrt, double* pot, double* sums_dev) {
  if (!c || !rho || !pot) return fail(c, "null argument");
  if (c->kind == kWGC8 && !c->have_wgc) return fail(c, "WGC kernel table not set (ofd_set_wgc_table)");
  ReqDesc D = c->base_desc();
  c->set_rho_io(D, rho, pot);
  D.opt_sqrt = opt_sqrt ? 1 : 0;
  if (sums_dev) D.fin = sums_dev;
  // the descriptor is uploaded only when it changes, from the context's own
  // copy (the stream is drained first, so no earlier upload still reads it)
  if (std::memcmp(&D, &c->last_desc, sizeof(D)) != 0) {
    if (hipStreamSynchronize(c->st) != hipSuccess) return fail(c, "stream synchronize failed");
    c->last_desc = D;
    if (hipMemcpyAsync(c->desc, &c->last_desc, sizeof(D), hipMemcpyHostToDevice, c->st) != hipSuccess)
      return fail(c, "descriptor upload failed");
  }
  const int fp32 = c->precision == 1;
  if (launch_eval(c->graphs, c->sp, c->kind, kInRho, kOutPot, fp32, c->desc, 1, sums_dev ? sums_dev : c->fin,
                  c->st))
    return fail(c, "kernel launch failed");
  return 0;
}

// ====================================================================== evaluation without allocation
// ofd_calpot_noalloc: the evaluation of ofd_calpot_enqueue with its device
// scratch (descriptor, finalized sums, per-workgroup partial sums, the
// spectra of the transformed fields) in one file-scope device array instead
// of the context's buffers, launched eagerly: no allocation, no graph, no
// copy and no memset, so that it can be called inside a stream capture or a
// loop that must not allocate. Every buffer in the array is written before it
// is read within one call, so nothing needs clearing between calls; calls
// that overlap on one device share the array and race.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_put_desc(ReqDesc d, ReqDesc* dst) { *dst = d; }
__global__ void k_etable_dev(const double* s, double lam, double mu, double dV, double eii, double* et) {
  etable_calc(s, lam, mu, dV, eii, et);
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

// Bytes of scratch an evaluation of context c takes: descriptor, finalized
// sums, partial sums, spectra (FP64 size, as the context's own buffers).
static size_t noalloc_bytes(const ofd_ctx* c, size_t* off_fin, size_t* off_sums, size_t* off_work) {
  const size_t nrows = (size_t)c->sp.d.n2 * c->sp.d.n3;
  *off_fin = up256(sizeof(ReqDesc));
  *off_sums = *off_fin + up256(kNSum * sizeof(double));
  *off_work = *off_sums + up256(nrows * kNSum * sizeof(double));
  return *off_work + (size_t)func_nf(c->kind) * (size_t)c->sp.d.nc * sizeof(double2);
}

extern "C" long long ofd_noalloc_limit_bytes(void) { return (long long)kNoAllocBytes; }

extern "C" int ofd_calpot_noalloc(ofd_ctx* c, const double* rho, int opt_sqrt, double* pot, double* etable9) {
  if (!c || !rho || !pot) return fail(c, "null argument");
  if (c->kind == kWGC8 && !c->have_wgc) return fail(c, "WGC kernel table not set (ofd_set_wgc_table)");
  size_t off_fin, off_sums, off_work;
  const size_t need = noalloc_bytes(c, &off_fin, &off_sums, &off_work);
  if (need > kNoAllocBytes)
    return fail(c, "ofd_calpot_noalloc: this grid and functional need " + std::to_string(need) +
                       " bytes of scratch, above its limit of " + std::to_string(kNoAllocBytes) +
                       " bytes (use ofd_calpot_enqueue)");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return fail(c, "no current HIP device");
  unsigned char* base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return fail(c, "ofd_calpot_noalloc: device scratch address unavailable");
      g_scratch_ptr[dev] = p;
    }
    base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  }
  ReqDesc D = c->base_desc();
  c->set_rho_io(D, rho, pot);
  D.opt_sqrt = opt_sqrt ? 1 : 0;
  D.fin = reinterpret_cast<double*>(base + off_fin);
  D.sums = reinterpret_cast<double*>(base + off_sums);
  D.work = base + off_work;
  ReqDesc* dd = reinterpret_cast<ReqDesc*>(base);
  hipLaunchKernelGGL(k_put_desc, dim3(1), dim3(1), 0, c->st, D, dd);
  const int fp32 = c->precision == 1;
  if (run_eval(c->sp, c->kind, kInRho, kOutPot, fp32, dd, 1, c->st) ||
      run_finalize(c->sp, c->kind, fp32, dd, 1, D.fin, c->st))
    return fail(c, "kernel launch failed");
  if (etable9)
    hipLaunchKernelGGL(k_etable_dev, dim3(1), dim3(1), 0, c->st, D.fin, c->lam, c->mu, c->dV, c->e_ionion, etable9);
  if (hipGetLastError() != hipSuccess) return fail(c, "kernel launch failed");
  return 0;
}

extern "C" int ofd_sums_to_etable(const ofd_ctx* c, const double* s, double* et) {
  if (!c || !s || !et) return -1;
  etable_from_sums(c, s, et);
  return 0;
}

extern "C" void* ofd_get_stream(ofd_ctx* c) { return c ? (void*)c->st : nullptr; }

extern "C" int ofd_sync(ofd_ctx* c) {
  if (!c) return -1;
  return hipStreamSynchronize(c->st) == hipSuccess ? 0 : fail(c, "stream synchronize failed");
}

extern "C" int ofd_calpot_many(ofd_ctx* const* cs, int nb, const double* const* rho, double* const* pot, int opt_sqrt,
                               double* ets) {
  if (!cs || nb < 1 || !rho || !pot) return -1;
  ofd_ctx* c0 = cs[0];
  for (int i = 1; i < nb; ++i) {
    if (!cs[i] || cs[i]->sp.d.n1 != c0->sp.d.n1 || cs[i]->sp.d.n2 != c0->sp.d.n2 || cs[i]->sp.d.n3 != c0->sp.d.n3 ||
        cs[i]->kind != c0->kind || cs[i]->precision != c0->precision)
      return fail(c0, "ofd_calpot_many: contexts differ in grid, functional or precision");
  }
  try {
    if ((int)c0->many_cap < nb) {
      HIPCHK(hipStreamSynchronize(c0->st));
      graphs_clear(c0->graphs);
      dev_free(c0->many_desc);
      dev_free(c0->many_sums);
      c0->many_desc = dalloc<ReqDesc>(nb);
      c0->many_sums = dalloc<double>((size_t)nb * kNSum);
      c0->many_cap = nb;
    }
    std::vector<ReqDesc> h(nb);
    for (int i = 0; i < nb; ++i) {
      h[i] = cs[i]->base_desc();
      cs[i]->set_rho_io(h[i], rho[i], pot[i]);
      h[i].opt_sqrt = opt_sqrt ? 1 : 0;
      h[i].fin = c0->many_sums + (size_t)i * kNSum;
    }
    if (c0->many_last.size() != (size_t)nb ||
        std::memcmp(h.data(), c0->many_last.data(), sizeof(ReqDesc) * nb) != 0) {
      HIPCHK(hipStreamSynchronize(c0->st));
      c0->many_last = h;
      HIPCHK(hipMemcpyAsync(c0->many_desc, c0->many_last.data(), sizeof(ReqDesc) * nb, hipMemcpyHostToDevice,
                            c0->st));
    }
    const int fp32 = c0->precision == 1;
    if (launch_eval(c0->graphs, c0->sp, c0->kind, kInRho, kOutPot, fp32, c0->many_desc, nb, c0->many_sums, c0->st))
      return fail(c0, "kernel launch failed");
    if (ets) {
      std::vector<double> s((size_t)nb * kNSum);
      HIPCHK(hipMemcpyAsync(s.data(), c0->many_sums, sizeof(double) * s.size(), hipMemcpyDeviceToHost, c0->st));
      HIPCHK(hipStreamSynchronize(c0->st));
      for (int i = 0; i < nb; ++i) etable_from_sums(cs[i], &s[(size_t)i * kNSum], ets + 9 * (size_t)i);
    }
  } catch (const std::exception& e) {
    return fail(c0, e.what());
  }
  return 0;
}

extern "C" int ofd_calpot(ofd_ctx* c, const double* rho, int opt_sqrt, double* pot, double* et) {
  if (!c || !rho || !pot) return fail(c, "null argument");
  const size_t b = c->nr_all() * sizeof(double);
  if (hipMemcpyAsync(c->rho, rho, b, hipMemcpyHostToDevice, c->st) != hipSuccess) return fail(c, "copy failed");
  if (ofd_calpot_device(c, c->rho, opt_sqrt, c->pot, et)) return -1;
  if (copy_sync(pot, c->pot, b, hipMemcpyDeviceToHost, c->st) != hipSuccess) return fail(c, "copy failed");
  return 0;
}

// ====================================================================== NTN solver
// ProjNormRhoMinimization with dir_type = 1 (NTN), numSpin = 1, as a state
// machine. Each state either queues vector operations (run later, in order,
// by the driver) or stops on a "need": one evaluation of the potential, or a
// few dot products. The driver (ofd_rho_opt for one request, ofd_pool_step
// for many) launches what the solvers need, batched across requests, and
// hands the results back through deliver().
namespace {

struct EvalReq {
  int outmode;
  const double *u, *v, *bvec;
  double* out;
  double ca, cb, s1, s2, mu0, eps;
  bool energy;
  const double* w = nullptr;   // CG residual for the fused Newton sums
};

enum NeedKind { NEED_NONE = 0, NEED_EVAL = 1, NEED_DOTS = 2 };
struct Need {
  NeedKind kind = NEED_NONE;
  EvalReq ev{};
  int np = 0;
  const double* a[4] = {nullptr, nullptr, nullptr, nullptr};
  const double* b[4] = {nullptr, nullptr, nullptr, nullptr};
};

constexpr int kFp32MaxLine = 20;

enum SState {
  ST_INIT_EVAL, ST_INIT_DOTS, ST_NEWTON_PP, ST_NEWTON_EVAL, ST_NEWTON_RR, ST_LP1, ST_LP2, ST_LP3, ST_LINE_EVAL,
  ST_ACC_DOTS, ST_CG_DOTS1, ST_CG_DOTS2, ST_BFGS_Q0, ST_BFGS_Q, ST_BFGS_R, ST_DONE
};

struct NtnSolver {
  ofd_ctx* c = nullptr;
  ofd_opt_params prm{};
  std::vector<VecOp> ops;   // queued vector operations (device pointers)
  Need need;
  SState st = ST_INIT_EVAL;
  std::string err;
  // scalars (RhoOptN.f90:120-170)
  double tot_ele = 0, mu0 = 0, theta = 0.2, energy[9] = {0};
  double e_hist[3] = {0, 0, 0};
  int it = 1, outcome = -1, success = 1;
  // NewtonDirection (RhoDirNew.f90:54-75)
  int nit = 0, ntflag = 0;
  double rr = 0, rr_last = 0, criteria = 0, ry = 0, ry_last = 0, ss1 = 0, eps = 0;
  // direction type (RhoOptN.f90:655-687): 1 NTN, 2 NCG; NCG keeps dL/dphi_old
  // in r and the previous direction in p (the Newton-only vectors)
  int dir_type = 1;
  bool cg_pr = false;          // cg_alg 'PR' (default 'HZ', RhoDirCG.f90:26)
  bool cg_restart = false;
  // NBF (RhoDirBFGS.f90:26-42): ring pointer BFGSPT, call counter BFGSiter,
  // BFGSrho, BFGSalpha; the two-loop position (slot, step, bound) and H0
  static constexpr int kM = ofd_ctx::kBfgsM;
  int bfgs_pt = 0, bfgs_iter = 0, bfgs_cp = 0, bfgs_k = 0, bfgs_bound = 0;
  double bfgs_rho[kM] = {0}, bfgs_alpha[kM] = {0}, bfgs_diag = 0.0;
  double cg_dots[6] = {0};
  // fused_dots: |p|^2, r.p and r_new.p_old carried by the CG recurrences
  bool fused = false;
  double pp = 0, rp = 0, rp_after = 0;
  bool pending_cg = false;   // fused: dir and r updates of the last step not queued yet
  double alpha_cg = 0;
  // getNextGradient / line search
  bool switched = false;
  double norm_phi_dir = 0;
  LineSearch ls;
  int nline = 0;
  double f_cur = 0, g_cur = 0;
  double last_sums[kNSum] = {0};
  double pot_norm = 0;
  bool fp32 = false;
  bool force64 = false;   // set when an FP32 phase stalls: the rest of the run is FP64
  bool trace = false;     // OFD_TRACE=1: one stderr line per accepted step
  double mach = 2.220446049250313e-16;
  ofd_opt_stats stats{};
  std::vector<double> hist_energy, hist_potnorm;

  // ---- queued vector operations
  void lincomb(double* z, const double* x, const double* y, double a, double b) {   // z = fma(b, y, a x)
    ops.push_back(VecOp{kOpLincomb, z, x, y, a, b, nullptr, nullptr, nullptr});
  }
  void axpby(double* y, const double* x, double a, double b) {   // y = fma(a, x, b y)
    ops.push_back(VecOp{kOpAxpby, y, x, y, a, b, nullptr, nullptr, nullptr});
  }
  void zero(double* z) { ops.push_back(VecOp{kOpZero, z, nullptr, nullptr, 0.0, 0.0, nullptr, nullptr, nullptr}); }
  // dir += a p; r -= a Ap; p = r + b p, in one pass (bitwise the three axpby)
  void cg_step(double a, double b) {
    ops.push_back(VecOp{kOpCG, c->dir, nullptr, nullptr, a, b, c->p, c->r, c->ap});
  }

  void need_dots(int np, const double* a0, const double* b0, const double* a1 = nullptr, const double* b1 = nullptr,
                 const double* a2 = nullptr, const double* b2 = nullptr, const double* a3 = nullptr,
                 const double* b3 = nullptr) {
    need = Need();
    need.kind = NEED_DOTS;
    need.np = np;
    need.a[0] = a0; need.b[0] = b0;
    need.a[1] = a1; need.b[1] = b1;
    need.a[2] = a2; need.b[2] = b2;
    need.a[3] = a3; need.b[3] = b3;
  }
  void need_eval(const EvalReq& q) {
    need = Need();
    need.kind = NEED_EVAL;
    need.ev = q;
  }
  void finish() {
    need = Need();
    st = ST_DONE;
  }
  // Ends the FP32 phase for good: re-evaluate at the current phi in FP64 and
  // continue from there as the reference would from this density. Taken when
  // the potential norm reaches the switch threshold, and when an FP32 phase
  // stalls (checkGradient gives up, the line search fails or runs long, or a
  // step does not lower the energy).
  void fall_back_fp64() {
    force64 = true;
    fp32 = false;
    cg_restart = true;
    bfgs_iter = bfgs_pt = 0;   // the FP32 pairs do not describe the FP64 surface
    mach = 2.220446049250313e-16;
    e_hist[0] = e_hist[1] = e_hist[2] = 0.0;
    success = 1;
    outcome = -1;
    st = ST_INIT_EVAL;
    need_eval(EvalReq{kOutSigned, c->phi, nullptr, nullptr, c->dedphi, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, true});
  }
  void fail_with(const char* m) {
    err = m;
    outcome = 1;
    finish();
  }

  void set_precision_for(double pnorm) {
    fp32 = (c->precision == 1) && !force64 && !(prm.mixed_switch > 0.0 && pnorm < prm.mixed_switch);
    mach = fp32 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
  }

  // the first evaluation at phi = sqrt|rhoR| (RhoOptN.f90:352-357)
  void start() {
    st = ST_INIT_EVAL;
    need_eval(EvalReq{kOutSigned, c->phi, nullptr, nullptr, c->dedphi, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, true});
  }

  // getNextDirection (RhoOptN.f90:655-687)
  void direction_begin(double rr0, double s1) {
    if (dir_type == 2) cg_begin();
    else if (dir_type == 3) bfgs_begin();
    else newton_begin(rr0, s1);
  }
  // BFGSDirection (RhoDirBFGS.f90:120-218): -dL/dphi on the first call, else
  // the two-loop recursion over the newest min(BFGSiter-1, MBFGS) pairs, built
  // in place in dir; H0 = (y.s / y.y) I from the newest pair; InnerProd has dV
  void bfgs_begin() {
    ++bfgs_iter;
    lincomb(c->dir, c->dldphi, c->dldphi, -1.0, 0.0);   // q = -dL/dphi
    if (bfgs_iter == 1) {
      switched = false;
      line_prep1();
      return;
    }
    bfgs_bound = std::min(bfgs_iter - 1, kM);
    bfgs_cp = (bfgs_pt + kM - 1) % kM;   // the newest pair
    bfgs_k = 0;
    st = ST_BFGS_Q0;
    need_dots(3, c->bfgs_s[bfgs_cp], c->bfgs_y[bfgs_cp], c->bfgs_y[bfgs_cp], c->bfgs_y[bfgs_cp],
              c->bfgs_s[bfgs_cp], c->dir);
  }
  // first loop, newest to oldest: alpha = rho s.q, q -= alpha y
  void bfgs_q_step(double sq) {
    bfgs_alpha[bfgs_cp] = bfgs_rho[bfgs_cp] * (sq * c->dV);
    axpby(c->dir, c->bfgs_y[bfgs_cp], -bfgs_alpha[bfgs_cp], 1.0);
    ++bfgs_k;
    if (bfgs_k < bfgs_bound) {
      bfgs_cp = (bfgs_cp + kM - 1) % kM;
      st = ST_BFGS_Q;
      need_dots(1, c->bfgs_s[bfgs_cp], c->dir);
      return;
    }
    // r0 = H0 q0, then the second loop from the oldest pair
    axpby(c->dir, c->dir, 0.0, bfgs_diag);
    bfgs_k = 0;
    st = ST_BFGS_R;
    need_dots(1, c->bfgs_y[bfgs_cp], c->dir);
  }
  // UpdateBFGS (RhoDirBFGS.f90:221-290) with phi_new in c->phi and phi_old in
  // c->tmp: s = phi_new - phi_old; y = -dL/dphi_old here, += dL/dphi_new in
  // bfgs_update_y once dL/dphi is recomputed with the new mu
  void bfgs_update_s() {
    lincomb(c->bfgs_s[bfgs_pt], c->phi, c->tmp, 1.0, -1.0);
    lincomb(c->bfgs_y[bfgs_pt], c->dldphi, c->dldphi, -1.0, 0.0);
  }
  void bfgs_update_y() {
    axpby(c->bfgs_y[bfgs_pt], c->dldphi, 1.0, 1.0);
    bfgs_pt = (bfgs_pt + 1) % kM;
  }
  // CGDirection (RhoDirCG.f90:32-106), InnerProd = SUM(a b) dV (:109-142)
  void cg_begin() {
    // restart on the reference's schedule, and after leaving an FP32 phase
    // (dL_old and the old direction belong to the FP32 iterates)
    if (it == 1 || it % 100 == 1 || cg_restart) {
      cg_restart = false;
      cg_finish(0.0);
      return;
    }
    st = ST_CG_DOTS1;
    if (cg_pr) {
      lincomb(c->ap, c->dldphi, c->r, 1.0, -1.0);   // dL - dL_old
      need_dots(2, c->dldphi, c->ap, c->r, c->r);
      return;
    }
    lincomb(c->ap, c->dldphi, c->r, 1.0, -1.0);     // y = dL - dL_old
    need_dots(4, c->p, c->ap, c->ap, c->ap, c->ap, c->dldphi, c->p, c->dldphi);
  }
  void cg_finish(double gam) {
    // dirNext = -dL + gamma dirNext_old; a restart never reads dirNext_old,
    // which holds nothing defined before the first line search
    lincomb(c->dir, c->dldphi, gam == 0.0 ? c->dldphi : c->p, -1.0, gam);
    lincomb(c->r, c->dldphi, c->dldphi, 1.0, 0.0);   // dL_old = dL
    switched = false;
    line_prep1();
  }
  // NewtonDirection start (RhoDirNew.f90:95-115); rr0 = sum r^2 = sum b^2, s1 = sum phi^2
  void newton_begin(double rr0, double s1) {
    lincomb(c->r, c->dldphi, c->dldphi, -1.0, 0.0);   // r = -b
    zero(c->dir);
    rr = rr0;
    ss1 = s1;
    criteria = rr;
    rr_last = rr * 2.0;
    nit = 1;
    ry = 0.0;
    ntflag = -1;
    newton_prep();
  }
  // y = r; p = y (+ ry/ryLast p) (RhoDirNew.f90:116-130)
  void newton_prep() {
    ry_last = ry;
    ry = rr;
    if (nit == 1) {
      lincomb(c->p, c->r, c->r, 1.0, 0.0);   // p = r exactly: sum p^2 = rr
      pp = rr;
      rp = rr;
      newton_eval(rr);
      return;
    }
    const double beta = ry / ry_last;
    if (pending_cg) {
      cg_step(alpha_cg, beta);
      pending_cg = false;
    } else {
      axpby(c->p, c->r, 1.0, beta);
    }
    if (fused) {
      // p' = r + beta p: |p'|^2 = rr + 2 beta r.p + beta^2 |p|^2, r.p' = rr + beta r.p
      const double pp_new = rr + 2.0 * beta * rp_after + beta * beta * pp;
      if (pp_new > 1e-6 * (rr + beta * beta * pp)) {
        pp = pp_new;
        rp = rr + beta * rp_after;
        newton_eval(pp);
        return;
      }
    }
    st = ST_NEWTON_PP;
    if (fused) need_dots(2, c->p, c->p, c->r, c->p);
    else need_dots(1, c->p, c->p);
  }
  // epsilon (RhoDirNew.f90:143-147) and the finite-difference product
  void newton_eval(double pp) {
    eps = 2.0 * std::sqrt(mach) * (1.0 + std::sqrt(ss1)) / std::sqrt(pp);
    st = ST_NEWTON_EVAL;
    EvalReq q{kOutNewton, c->phi, c->p, c->dldphi, c->ap, 1.0, eps, 0.0, 0.0, mu0, eps, false};
    if (fused) q.w = c->r;
    need_eval(q);
  }
  // CG exit tests after the residual update (RhoDirNew.f90:187-200)
  void newton_after_rr(double rr_new) {
    rr = rr_new;
    bool done = false;
    if (rr < 0.1 * criteria) { ntflag = 0; done = true; }
    else if (nit > 50) { ntflag = 1; done = true; }
    else if (std::fabs(rr - rr_last) / rr < 0.01 && nit > 9) { ntflag = 2; done = true; }
    if (done) {
      newton_done();
    } else {
      ++nit;
      newton_prep();
    }
  }
  void newton_done() {
    if (pending_cg) {   // the step's direction update (r is reset by the next Newton start)
      axpby(c->dir, c->p, alpha_cg, 1.0);
      pending_cg = false;
    }
    stats.newton_iters_total += nit;
    switched = false;
    line_prep1();
  }
  // getNextGradient (RhoLineSearch.f90:40-139) in three dot rounds
  void line_prep1() {
    lincomb(c->tmp, c->phi, c->dir, 1.0, 1.0);
    st = ST_LP1;
    need_dots(2, c->tmp, c->tmp, c->phi, c->dir);
  }
  void line_start(double gradth) {
    // LineSearchTN start (RhoLineSearch.f90:369-374); ftol/gtol are
    // default-real literals (RhoOptN.f90:255-258)
    const float ftol32 = 1e-4f, gtol32 = dir_type == 2 ? 1e-2f : 2e-1f;
    ls = LineSearch();
    ls.ftol = (double)ftol32;
    ls.gtol = (double)gtol32;
    ls.xtol = 1e-12;
    ls.stpmin = 0.0;
    ls.stpmax = kPi;
    nline = 0;
    f_cur = energy[0];
    g_cur = gradth;
    line_advance();
  }
  void line_advance() {
    const int task = ls.next(theta, f_cur, g_cur);
    if (task == 1) {
      const double ct = std::cos(theta), sn = std::sin(theta);
      st = ST_LINE_EVAL;
      need_eval(EvalReq{kOutSigned, c->phi, c->dir, nullptr, c->dedphi, ct, sn, -sn, ct, mu0, 0.0, true});
      return;
    }
    success = task == 2 ? 1 : (task == 3 ? -2 : -1);
    accept();
  }
  // StepInformation + CheckExit + update (RhoOptN.f90:595-627, :692-977)
  void accept() {
    if (success == -1 && fp32) {
      fall_back_fp64();
      return;
    }
    if (success == -1 && it > prm.max_iter) {
      fail_with("Fail rho optimizations");
      return;
    }
    const long long n = c->nr();
    pot_norm = std::sqrt(last_sums[7] / (double)n);
    hist_energy.push_back(energy[0]);
    hist_potnorm.push_back(pot_norm);
    if (trace) {
      // one line per accepted step, the columns of the reference's RhoOptN
      // iteration table (RhoOptN.f90:595-627) plus the precision
      std::fprintf(stderr, "ofd-ntn %p it %4d %s E %.12f potnorm %.4e theta %.4e nline %3d newton %2d flag %2d ok %2d\n",
                   (void*)c, it, fp32 ? "fp32" : "fp64", energy[0], pot_norm, theta, nline, nit, ntflag, success);
    }
    bool exit_flag = false;
    const bool pconv = pot_norm < prm.pot_tol;
    e_hist[2] = e_hist[1];
    e_hist[1] = e_hist[0];
    e_hist[0] = energy[0];
    const bool econv =
        it >= 3 && std::fabs(e_hist[0] - e_hist[1]) < prm.tole && std::fabs(e_hist[0] - e_hist[2]) < prm.tole;
    if (prm.conv_check == 1 && econv) { outcome = 0; exit_flag = true; }
    if (prm.conv_check == 2 && pconv) { outcome = 0; exit_flag = true; }
    if (prm.conv_check == 3 && (pconv || econv)) { outcome = 0; exit_flag = true; }
    if (outcome == 1) exit_flag = true;
    if (it >= prm.max_iter) { outcome = 1; exit_flag = true; }
    // an FP32 phase never ends the run
    if (exit_flag && fp32 && outcome == 0) {
      exit_flag = false;
      outcome = -1;
    }
    // phi = tempPhi (same fma as inside the evaluation), mu, dL/dphi
    const double ct = std::cos(theta), sn = std::sin(theta);
    lincomb(c->tmp, c->phi, c->dir, ct, sn);
    std::swap(c->phi, c->tmp);
    const bool bfgs_pair = dir_type == 3 && !exit_flag;   // RhoOptN.f90:603-610
    if (bfgs_pair) bfgs_update_s();
    mu0 = c->dV * last_sums[8] / (2.0 * tot_ele);
    lincomb(c->dldphi, c->dedphi, c->phi, 1.0, -2.0 * mu0);
    if (bfgs_pair) bfgs_update_y();
    stats.iterations = it;
    if (exit_flag) {
      finish();
      return;
    }
    // an FP32 step that does not lower the energy means FP32 has reached its
    // noise floor: continue in FP64 from the accepted density
    if (fp32 && e_hist[0] >= e_hist[1]) {
      ++it;
      fall_back_fp64();
      return;
    }
    const bool was32 = fp32;
    set_precision_for(pot_norm);
    if (was32 && !fp32) {
      // FP64 refinement starts from an FP64 evaluation at the accepted
      // density: the line search must not compare FP64 energies against an
      // FP32 one (a baseline a little too low makes it return theta = 0,
      // after which theta = MIN(theta, ...) can never grow again)
      ++it;
      fall_back_fp64();
      return;
    }
    ++it;
    st = ST_ACC_DOTS;
    double bb, ff;
    if (fused && fused_bb(last_sums, bb, ff)) {
      direction_begin(bb, ff);
      return;
    }
    need_dots(2, c->dldphi, c->dldphi, c->phi, c->phi);
  }
  // |dL/dphi|^2 and |phi|^2 at the point the sums `v` were evaluated at, with
  // dL/dphi = dE/dphi - 2 mu0 phi: |b|^2 = S(de^2) - 4 mu0 S(de phi) + 4 mu0^2 S(phi^2).
  // False when the cancellation would cost more than about 8 digits.
  bool fused_bb(const double* v, double& bb, double& ff) const {
    ff = v[10];
    bb = v[9] - 4.0 * mu0 * v[8] + 4.0 * mu0 * mu0 * v[10];
    return bb > 1e-8 * v[9];
  }

  // results of the current need: kNSum sums (evaluation) or np dot products
  void deliver(const double* v) {
    const long long n = c->nr();
    switch (st) {
      case ST_INIT_EVAL: {
        std::memcpy(last_sums, v, sizeof(last_sums));
        ++stats.num_evals;
        if (fp32) ++stats.fp32_evals;
        etable_from_sums(c, v, energy);
        mu0 = c->dV * v[8] / (2.0 * tot_ele);   // ChemicalPotential (RhoDirBFGS.f90:337)
        lincomb(c->dldphi, c->dedphi, c->phi, 1.0, -2.0 * mu0);
        st = ST_INIT_DOTS;
        double bb, ff;
        if (fused && fused_bb(v, bb, ff)) {
          const double d2[2] = {bb, ff};
          deliver(d2);
          return;
        }
        need_dots(2, c->dldphi, c->dldphi, c->phi, c->phi);
        return;
      }
      case ST_INIT_DOTS: {
        pot_norm = std::sqrt(v[0] / (double)n);
        hist_energy.push_back(energy[0]);
        hist_potnorm.push_back(pot_norm);
        e_hist[0] = energy[0];
        set_precision_for(pot_norm);
        direction_begin(v[0], v[1]);
        return;
      }
      case ST_NEWTON_PP:
        pp = v[0];
        if (fused) rp = v[1];
        newton_eval(pp);
        return;
      case ST_NEWTON_EVAL: {
        std::memcpy(last_sums, v, sizeof(last_sums));
        ++stats.num_evals;
        if (fp32) ++stats.fp32_evals;
        const double pap = v[6];
        if (pap < 0.0) {
          if (nit == 1) axpby(c->dir, c->p, ry / pap, -1.0);   // dir = -(dir - ry/pAp p)
          ntflag = -2;
          newton_done();
          return;
        }
        const double alpha = ry / pap;
        rr_last = rr;
        if (fused) {
          // r' = r - alpha Ap: |r'|^2 = rr - 2 alpha r.Ap + alpha^2 |Ap|^2, r'.p = r.p - alpha pAp
          const double rr_new = rr - 2.0 * alpha * v[7] + alpha * alpha * v[8];
          rp_after = rp - alpha * pap;
          if (rr_new > 1e-6 * rr) {
            // the dir and r updates go out with the next p update (or alone at exit)
            pending_cg = true;
            alpha_cg = alpha;
            newton_after_rr(rr_new);
            return;
          }
        }
        axpby(c->dir, c->p, alpha, 1.0);
        axpby(c->r, c->ap, -alpha, 1.0);
        st = ST_NEWTON_RR;
        if (fused) need_dots(2, c->r, c->r, c->r, c->p);
        else need_dots(1, c->r, c->r);
        return;
      }
      case ST_NEWTON_RR:
        if (fused) rp_after = v[1];
        newton_after_rr(v[0]);
        return;
      case ST_LP1: {
        norm_phi_dir = std::sqrt(v[0] * c->dV);
        const double proj = v[1] * c->dV;
        axpby(c->dir, c->phi, -(proj / tot_ele), 1.0);   // dir = dir - phi <phi,dir>/Ne
        st = ST_LP2;
        need_dots(1, c->dir, c->dir);
        return;
      }
      case ST_LP2: {
        const double norm_perp = std::sqrt(v[0] * c->dV);
        theta = std::fmin(theta, norm_perp / norm_phi_dir);
        axpby(c->dir, c->dir, 0.0, std::sqrt(tot_ele) / norm_perp);
        st = ST_LP3;
        need_dots(1, c->dedphi, c->dir);
        return;
      }
      case ST_LP3: {
        const double gradth = c->dV * v[0];   // gradOftheta at theta = 0
        // checkGradient (RhoLineSearch.f90:201-309)
        if (gradth > 1.0e5) {
          fail_with("The gradient is too large");
          return;
        }
        if (gradth >= 0.0) {
          if (switched && fp32) {
            fall_back_fp64();
            return;
          }
          if (switched) {
            outcome = 1;
            finish();
            return;
          }
          switched = true;
          lincomb(c->dir, c->dldphi, c->dldphi, -1.0, 0.0);   // steepest descent
          line_prep1();
          return;
        }
        if (dir_type == 2) lincomb(c->p, c->dir, c->dir, 1.0, 0.0);   // dirNext_old = dirNext (RhoOptN.f90:566)
        line_start(gradth);
        return;
      }
      case ST_LINE_EVAL: {
        std::memcpy(last_sums, v, sizeof(last_sums));
        ++stats.num_evals;
        if (fp32) ++stats.fp32_evals;
        etable_from_sums(c, v, energy);
        f_cur = energy[0];
        g_cur = c->dV * v[6];
        ++nline;
        if (fp32 && nline > kFp32MaxLine) {
          // FP32 noise keeps the Wolfe conditions from being met: refine in FP64
          fall_back_fp64();
          return;
        }
        if (nline > prm.max_line) accept();
        else line_advance();
        return;
      }
      case ST_ACC_DOTS:
        direction_begin(v[0], v[1]);
        return;
      case ST_CG_DOTS1: {
        if (cg_pr) {
          // gamma = MAX(0, <dL, dL - dL_old> / <dL_old, dL_old>)
          cg_finish(std::max(0.0, (v[0] * c->dV) / (v[1] * c->dV)));
          return;
        }
        for (int k = 0; k < 4; ++k) cg_dots[k] = v[k];   // d.y, y.y, y.dL, d.dL
        st = ST_CG_DOTS2;
        need_dots(2, c->p, c->p, c->r, c->r);
        return;
      }
      case ST_BFGS_Q0: {
        const double ys = v[0] * c->dV, yy = v[1] * c->dV;
        bfgs_diag = ys / yy;
        bfgs_rho[bfgs_cp] = 1.0 / ys;
        bfgs_q_step(v[2]);
        return;
      }
      case ST_BFGS_Q:
        bfgs_q_step(v[0]);
        return;
      case ST_BFGS_R: {
        // beta = alpha - rho y.r; r += beta s, oldest to newest
        const double beta = bfgs_alpha[bfgs_cp] - bfgs_rho[bfgs_cp] * (v[0] * c->dV);
        axpby(c->dir, c->bfgs_s[bfgs_cp], beta, 1.0);
        bfgs_cp = (bfgs_cp + 1) % kM;
        if (++bfgs_k < bfgs_bound) {
          need_dots(1, c->bfgs_y[bfgs_cp], c->dir);
          return;
        }
        switched = false;
        line_prep1();
        return;
      }
      case ST_CG_DOTS2: {
        // Hager-Zhang (RhoDirCG.f90:79-88)
        const double dV = c->dV;
        const double dty = cg_dots[0] * dV;
        const double norm_y = std::sqrt(cg_dots[1] * dV / dV);
        // <y - 2 d |y|^2/dTy, dL> = <y, dL> - 2 |y|^2/dTy <d, dL>
        double gam = 1.0 / dty * (cg_dots[2] * dV - 2.0 * norm_y * norm_y / dty * (cg_dots[3] * dV));
        const double norm_dk = std::sqrt(v[0] * dV / dV);
        const double norm_gk = std::sqrt(v[1] * dV / dV);
        gam = std::max(gam, -1.0 / (norm_dk * std::min(0.01, norm_gk)));
        cg_finish(gam);
        return;
      }
      default:
        return;
    }
  }
};

void fill_params(const ofd_opt_params* in, ofd_opt_params* p) {
  p->max_iter = 500;
  p->pot_tol = (double)1.0e-5f;
  p->tole = (double)1.0e-6f;
  p->conv_check = 1;
  p->max_line = 200;
  p->mixed_switch = 0.0;
  p->fused_dots = 0;
  if (!in) return;
  if (in->max_iter > 0) p->max_iter = in->max_iter;
  if (in->pot_tol > 0) p->pot_tol = in->pot_tol;
  if (in->tole > 0) p->tole = in->tole;
  if (in->conv_check >= 1 && in->conv_check <= 3) p->conv_check = in->conv_check;
  if (in->max_line > 0) p->max_line = in->max_line;
  p->mixed_switch = in->mixed_switch;
  p->fused_dots = in->fused_dots != 0 ? 1 : 0;
}

// phi, dir, dE/dphi, dL/dphi and tmp hold every spin channel; the Newton
// vectors r, p, Ap one channel (NewtonDirection works on one at a time)
void alloc_opt(ofd_ctx* c) {
  if (c->phi) return;
  const long long n = c->nr(), na = c->nr_all();
  c->phi = dalloc<double>(na); c->dir = dalloc<double>(na); c->dedphi = dalloc<double>(na);
  c->dldphi = dalloc<double>(na); c->r = dalloc<double>(n); c->p = dalloc<double>(n);
  c->ap = dalloc<double>(n); c->tmp = dalloc<double>(na);
}

ReqDesc desc_for(ofd_ctx* c, const EvalReq& q) {
  ReqDesc D = c->base_desc();
  D.u = q.u; D.v = q.v; D.bvec = q.bvec; D.out = q.out;
  D.ca = q.ca; D.cb = q.cb; D.s1 = q.s1; D.s2 = q.s2; D.mu0 = q.mu0; D.eps = q.eps;
  D.w = q.w;
  D.opt_sqrt = 1;
  return D;
}

// Scratch for batched vector ops and dots, grown on demand. The host side is
// pinned so that uploads and result copies are asynchronous on the caller's
// stream; an event per staging buffer guards its reuse.
struct BatchScratch {
  VecOp* d_ops = nullptr;
  VecOp* h_ops = nullptr;
  int ops_cap = 0;
  DotDesc* d_dots = nullptr;
  DotDesc* h_dots = nullptr;
  double* d_part = nullptr;
  double* d_res = nullptr;
  double* h_res = nullptr;
  int dots_cap = 0;
  int ndots = 0;   // results pending in h_res
  hipEvent_t ev_ops = nullptr, ev_dots = nullptr;
  void free_all() {
    dev_free(d_ops); dev_free(d_dots); dev_free(d_part); dev_free(d_res);
    hipHostFree(h_ops); hipHostFree(h_dots); hipHostFree(h_res);
    if (ev_ops) hipEventDestroy(ev_ops);
    if (ev_dots) hipEventDestroy(ev_dots);
    *this = BatchScratch();
  }
  void events() {
    if (!ev_ops) HIPCHK(hipEventCreateWithFlags(&ev_ops, hipEventDisableTiming));
    if (!ev_dots) HIPCHK(hipEventCreateWithFlags(&ev_dots, hipEventDisableTiming));
  }
};

template <typename T>
static T* halloc(size_t n) {
  void* p = nullptr;
  if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::runtime_error("hipHostMalloc failed");
  return static_cast<T*>(p);
}

// Launches the queued ops of all solvers, layer by layer (each solver's ops
// stay in order), on `st`.
void flush_ops(std::vector<NtnSolver*>& ss, long long n, BatchScratch& bs, hipStream_t st) {
  size_t layers = 0, total = 0;
  for (auto* s : ss) {
    layers = std::max(layers, s->ops.size());
    total += s->ops.size();
  }
  if (layers == 0) return;
  bs.events();
  HIPCHK(hipEventSynchronize(bs.ev_ops));   // the previous upload has left the staging buffer
  if ((int)total > bs.ops_cap) {
    HIPCHK(hipStreamSynchronize(st));   // earlier vector ops may still read d_ops
    dev_free(bs.d_ops);
    hipHostFree(bs.h_ops);
    bs.ops_cap = (int)total * 2;
    bs.d_ops = dalloc<VecOp>(bs.ops_cap);
    bs.h_ops = halloc<VecOp>(bs.ops_cap);
  }
  std::vector<int> off(layers + 1, 0);
  int k = 0;
  for (size_t L = 0; L < layers; ++L) {
    off[L] = k;
    for (auto* s : ss)
      if (L < s->ops.size()) bs.h_ops[k++] = s->ops[L];
  }
  off[layers] = k;
  HIPCHK(hipMemcpyAsync(bs.d_ops, bs.h_ops, sizeof(VecOp) * k, hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(bs.ev_ops, st));
  for (size_t L = 0; L < layers; ++L) vec_ops_batched(bs.d_ops + off[L], off[L + 1] - off[L], n, st);
  for (auto* s : ss) s->ops.clear();
}

// Launches the dot needs of `ss` (all NEED_DOTS) and the copy of their
// results to pinned memory; deliver_dots hands them over after the stream
// has been synchronized.
void launch_dots(std::vector<NtnSolver*>& ss, long long n, BatchScratch& bs, hipStream_t st) {
  int cnt = 0;
  for (auto* s : ss) cnt += s->need.np;
  bs.ndots = cnt;
  if (cnt == 0) return;
  bs.events();
  HIPCHK(hipEventSynchronize(bs.ev_dots));
  if (cnt > bs.dots_cap) {
    HIPCHK(hipStreamSynchronize(st));   // earlier dot launches may still use them
    dev_free(bs.d_dots); dev_free(bs.d_part); dev_free(bs.d_res);
    hipHostFree(bs.h_dots); hipHostFree(bs.h_res);
    bs.dots_cap = cnt * 2;
    bs.d_dots = dalloc<DotDesc>(bs.dots_cap);
    bs.d_part = dalloc<double>((size_t)bs.dots_cap * kDotBlocksBatched);
    bs.d_res = dalloc<double>(bs.dots_cap);
    bs.h_dots = halloc<DotDesc>(bs.dots_cap);
    bs.h_res = halloc<double>(bs.dots_cap);
  }
  int k = 0;
  for (auto* s : ss)
    for (int j = 0; j < s->need.np; ++j) bs.h_dots[k++] = DotDesc{s->need.a[j], s->need.b[j]};
  HIPCHK(hipMemcpyAsync(bs.d_dots, bs.h_dots, sizeof(DotDesc) * cnt, hipMemcpyHostToDevice, st));
  vec_dots_batched(bs.d_dots, cnt, n, bs.d_part, bs.d_res, st);
  HIPCHK(hipMemcpyAsync(bs.h_res, bs.d_res, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(bs.ev_dots, st));
}
void deliver_dots(std::vector<NtnSolver*>& ss, BatchScratch& bs) {
  size_t k = 0;
  for (auto* s : ss) {
    const int np = s->need.np;
    s->deliver(&bs.h_res[k]);
    k += np;
  }
  bs.ndots = 0;
}
void run_dots(std::vector<NtnSolver*>& ss, long long n, BatchScratch& bs, hipStream_t st) {
  launch_dots(ss, n, bs, st);
  if (bs.ndots == 0) return;
  HIPCHK(hipStreamSynchronize(st));
  deliver_dots(ss, bs);
}

// sum(x) over a device vector, deterministic (a dot product with no second vector)
double dsum(const double* x, long long n, BatchScratch& bs, hipStream_t st) {
  NtnSolver tmp;
  tmp.need.kind = NEED_DOTS;
  tmp.need.np = 1;
  tmp.need.a[0] = x;
  tmp.need.b[0] = nullptr;
  std::vector<NtnSolver*> one{&tmp};
  launch_dots(one, n, bs, st);
  HIPCHK(hipStreamSynchronize(st));
  const double s = bs.h_res[0];
  bs.ndots = 0;
  return s;
}

// Prepares phi = sqrt|rhoR| and totEleNum = sum(rhoR) dV (RhoOptN.f90:326-352)
// from a density already on the device in c->rho.
void solver_prepare(NtnSolver& S, ofd_ctx* c, const ofd_opt_params* prm, BatchScratch& bs, hipStream_t st,
                    int dir_type = 1) {
  S = NtnSolver();
  S.c = c;
  S.dir_type = dir_type;
  S.cg_pr = prm->cg_alg == 1;
  if (dir_type == 3 && !c->bfgs_s[0]) {
    const long long n = c->nr();
    for (int k = 0; k < ofd_ctx::kBfgsM; ++k) {
      c->bfgs_s[k] = dalloc<double>(n);
      c->bfgs_y[k] = dalloc<double>(n);
    }
  }
  if (const char* e = std::getenv("OFD_TRACE")) S.trace = std::atoi(e) != 0;
  fill_params(prm, &S.prm);
  S.fused = S.prm.fused_dots != 0;
  const long long n = c->nr();
  S.tot_ele = dsum(c->rho, n, bs, st) * c->dV;
  vec_sqrt_abs(c->phi, c->rho, n, st);
  S.start();
}

}  // namespace

static int rho_opt_spin(ofd_ctx* c, double* rhoR, int dir_type, const ofd_opt_params* prm, double* energy9,
                        int* rho_outcome, ofd_opt_stats* stats);

extern "C" int ofd_rho_opt(ofd_ctx* c, double* rhoR, int dir_type, const ofd_opt_params* prm, double* energy9,
                           int* rho_outcome, ofd_opt_stats* stats) {
  if (!c || !rhoR) return fail(c, "null argument");
  if (dir_type < 1 || dir_type > 3) return fail(c, "dir_type 1 (NTN), 2 (NCG) or 3 (NBF)");
  if (c->nspin == 2) return rho_opt_spin(c, rhoR, dir_type, prm, energy9, rho_outcome, stats);
  if (c->kind == kWGC8 && !c->have_wgc) return fail(c, "WGC kernel table not set");
  auto t0 = std::chrono::steady_clock::now();
  BatchScratch bs;
  try {
    alloc_opt(c);
    const long long n = c->nr();
    HIPCHK(hipMemcpyAsync(c->rho, rhoR, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    NtnSolver S;
    solver_prepare(S, c, prm, bs, c->st, dir_type);
    std::vector<NtnSolver*> one{&S};
    double s[kNSum];
    while (S.need.kind != NEED_NONE) {
      flush_ops(one, n, bs, c->st);
      if (S.need.kind == NEED_DOTS) {
        run_dots(one, n, bs, c->st);
      } else {
        ReqDesc D = desc_for(c, S.need.ev);
        if (eval_one(c, D, kInPhi, S.need.ev.outmode, S.fp32 ? 1 : 0, s)) {
          bs.free_all();
          return -1;
        }
        S.deliver(s);
      }
    }
    flush_ops(one, n, bs, c->st);
    // rhoR = phi**2 (:495)
    vec_square(c->rho, c->phi, n, c->st);
    HIPCHK(hipMemcpyAsync(rhoR, c->rho, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    bs.free_all();
    if (energy9) std::memcpy(energy9, S.energy, sizeof(S.energy));
    if (rho_outcome) *rho_outcome = S.outcome;
    if (stats) {
      *stats = S.stats;
      stats->final_pot_norm = S.pot_norm;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (!S.err.empty()) return fail(c, S.err);
  } catch (const std::exception& e) {
    bs.free_all();
    return fail(c, e.what());
  }
  return 0;
}

// ====================================================================== spin-polarized solver
// ProjNormRhoMinimization for numSpin = 2, dir_type = 1 (RhoOptN.f90:60-850):
// per-channel electron numbers and chemical potentials, one NewtonDirection
// per channel (RhoDirNew.f90:25-210, the finite-difference Hessian product
// perturbing that channel only), getNextGradient / checkGradient per channel
// (RhoLineSearch.f90:40-139, :201-309), SpinThetaSearch's joint search of
// (theta_up, theta_down) (:447-657), and around it the magnetization loop
// (RhoOptN.f90:336-480). Each step runs on the device (evaluations, vector
// updates, deterministic dot products); the host holds the scalars and waits
// for each result, as the reference's loop does.
namespace {

struct SpinSolver {
  ofd_ctx* c = nullptr;
  hipStream_t st = nullptr;
  long long n = 0;            // points per channel
  ofd_opt_params prm{};
  BatchScratch bs;
  double* phi[2] = {};
  double* dir[2] = {};
  double* de[2] = {};
  double* dl[2] = {};
  double tot[2] = {0, 0}, mu0[2] = {0, 0}, theta[2] = {0.2, 0.2}, gradth[2] = {0, 0};
  double energy[9] = {0};
  double s[kNSum] = {0};
  double pot_norm = 0;
  int outcome = -1, it = 1, success = 1;
  ofd_opt_stats stats{};
  const double mach = 2.220446049250313e-16;

  void bind(ofd_ctx* cc) {
    c = cc;
    st = c->st;
    n = c->nr();
    for (int k = 0; k < 2; ++k) {
      phi[k] = c->phi + k * n;
      dir[k] = c->dir + k * n;
      de[k] = c->dedphi + k * n;
      dl[k] = c->dldphi + k * n;
    }
  }
  // up to four dot products in one launch, deterministic order
  void dots(int np, const double* const* a, const double* const* b, double* out) {
    vec_dots(a, b, np, n, c->dot_part, c->dot_out, st);
    HIPCHK(copy_sync(out, c->dot_out, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  double dot(const double* a, const double* b) {
    double o[4];
    const double* A[1] = {a};
    const double* B[1] = {b};
    dots(1, A, B, o);
    return o[0];
  }
  // one evaluation at phi_t,s = cos(t_s) phi_s + sin(t_s) dir_s (t = 0: phi),
  // dE/dphi (signed) into de; sums: gradth per channel at t, the potential
  // norm with the current mu, sum dE/dphi_s phi_t,s
  void eval_at(const double* t) {
    ReqDesc D = c->base_desc();
    const bool zero = t[0] == 0.0 && t[1] == 0.0;
    D.u = phi[0];
    D.v = zero ? nullptr : dir[0];
    D.ca = std::cos(t[0]);
    D.cb = zero ? 0.0 : std::sin(t[0]);
    D.s1 = -std::sin(t[0]);
    D.s2 = std::cos(t[0]);
    D.u2 = phi[1];
    D.v2 = zero ? nullptr : dir[1];
    D.ca2 = std::cos(t[1]);
    D.cb2 = zero ? 0.0 : std::sin(t[1]);
    D.s1b = -std::sin(t[1]);
    D.s2b = std::cos(t[1]);
    D.out = de[0];
    D.out2 = de[1];
    D.mu0 = mu0[0];
    D.mu0b = mu0[1];
    D.opt_sqrt = 1;
    if (eval_one(c, D, kInPhi, kOutSigned, 0, s)) throw std::runtime_error(c->err);
    etable_from_sums(c, s, energy);
    ++stats.num_evals;
  }
  // dL/dphi_s = dE/dphi_s - 2 mu_s phi_s with mu_s = dV sum(dE/dphi_s phi_s) / (2 N_s)
  // (ChemicalPotential, RhoDirBFGS.f90:294-343) from the last evaluation's sums
  void chem_pot_and_dl() {
    mu0[0] = c->dV * s[8] / (2.0 * tot[0]);
    mu0[1] = c->dV * s[10] / (2.0 * tot[1]);
    for (int k = 0; k < 2; ++k) vec_lincomb(dl[k], de[k], phi[k], 1.0, -2.0 * mu0[k], n, st);
  }

  // NewtonDirection for channel ch (RhoDirNew.f90:25-210) into dir[ch]
  int newton(int ch) {
    double* r = c->r;
    double* p = c->p;
    double* ap = c->ap;
    vec_fill(dir[ch], 0.0, n, st);
    vec_lincomb(r, dl[ch], dl[ch], -1.0, 0.0, n, st);   // r = -b
    double rr = dot(r, r);
    const double criteria = rr;
    const double s1 = dot(phi[ch], phi[ch]);
    double ry = 0.0, ry_last = 0.0, rr_last;
    int iter = 1;
    while (true) {
      ry_last = ry;
      ry = rr;   // y = r: r.y is |r|^2
      if (iter == 1) vec_lincomb(p, r, r, 1.0, 0.0, n, st);
      else vec_lincomb(p, r, p, 1.0, ry / ry_last, n, st);
      const double s2 = dot(p, p);
      const double eps = 2.0 * std::sqrt(mach) * (1.0 + std::sqrt(s1)) / std::sqrt(s2);
      ReqDesc D = c->base_desc();
      D.u = phi[0];
      D.u2 = phi[1];
      D.ca = 1.0;
      D.ca2 = 1.0;
      if (ch == 0) { D.v = p; D.cb = eps; } else { D.v2 = p; D.cb2 = eps; }
      D.bvec = dl[ch];
      D.out = ap;
      D.mu0 = mu0[ch];
      D.eps = eps;
      D.newton_ch = ch;
      D.opt_sqrt = 1;
      if (eval_one(c, D, kInPhi, kOutNewton, 0, s)) throw std::runtime_error(c->err);
      ++stats.num_evals;
      const double pap = s[6];
      if (pap < 0.0) {
        if (iter == 1) vec_lincomb(dir[ch], dir[ch], p, -1.0, ry / pap, n, st);
        break;
      }
      const double alpha = ry / pap;
      vec_axpby(dir[ch], p, alpha, 1.0, n, st);
      vec_axpby(r, ap, -alpha, 1.0, n, st);
      rr_last = rr;
      rr = dot(r, r);
      if (rr < 0.1 * criteria) break;
      if (iter > 50) break;
      if (std::fabs(rr - rr_last) / rr < 0.01 && iter > 9) break;
      ++iter;
    }
    stats.newton_iters_total += iter;
    return iter;
  }

  // getNextGradient, numSpin = 2 (RhoLineSearch.f90:108-131): project out
  // phi_s, theta_s = 0, normalize to sqrt(N_s), gradth_s at theta = 0
  void next_gradient() {
    for (int k = 0; k < 2; ++k) {
      const double pd = dot(phi[k], dir[k]);
      vec_axpby(dir[k], phi[k], -(pd * c->dV) / tot[k], 1.0, n, st);
      const double nperp = std::sqrt(dot(dir[k], dir[k]) * c->dV);
      theta[k] = 0.0;
      vec_axpby(dir[k], dir[k], 0.0, std::sqrt(tot[k]) / nperp, n, st);
      gradth[k] = c->dV * dot(de[k], dir[k]);
    }
  }

  // SpinThetaSearch (RhoLineSearch.f90:447-657). Its first evaluation, at
  // theta = 0, is the current point, whose dE/dphi, energy and gradients are
  // already in hand, and is not repeated. Returns success as the reference
  // sets it (2 converged, -1 error or more than 10 trial points, -2 warning).
  int theta_search(double ftol, double gtol, double xtol) {
    const double alphatol = (double)1e-4f;
    double td_old[2] = {0, 0}, g_old[2] = {0, 0};
    int ok = success;
    for (int j = 1;; ++j) {
      double gam = 0.0;
      if (j > 1) {
        const double num = gradth[0] * (gradth[0] - g_old[0]) + gradth[1] * (gradth[1] - g_old[1]);
        gam = std::max(0.0, num / (g_old[0] * g_old[0] + g_old[1] * g_old[1]));
      }
      g_old[0] = gradth[0];
      g_old[1] = gradth[1];
      double td[2] = {-gradth[0] + gam * td_old[0], -gradth[1] + gam * td_old[1]};
      double dedalpha;
      while (true) {
        dedalpha = gradth[0] * td[0] + gradth[1] * td[1];
        if (dedalpha >= 0.0) {
          td[0] = -gradth[0];
          td[1] = -gradth[1];
        } else {
          break;
        }
      }
      td_old[0] = td[0];
      td_old[1] = td[1];
      const double maxdir = std::max(std::fabs(td[0]), std::fabs(td[1]));
      // MIN(0.1, 0.1*pi/maxthetadir) with the single-precision literal 0.1
      double alpha = std::min((double)0.1f, (double)0.1f * kPi / maxdir);
      LineSearch ls{ftol, gtol, xtol, 0.0, kPi / maxdir};
      int nline = 0;
      while (true) {
        const int task = ls.next(alpha, energy[0], dedalpha);
        if (task == 2) { ok = 2; break; }
        if (task == 4) { ok = -1; break; }
        if (task == 3) { ok = -2; break; }
        const double tem[2] = {theta[0] + alpha * td[0], theta[1] + alpha * td[1]};
        eval_at(tem);
        gradth[0] = c->dV * s[6];
        gradth[1] = c->dV * s[9];
        dedalpha = gradth[0] * td[0] + gradth[1] * td[1];
        if (++nline > 10) {
          ok = -1;
          break;
        }
      }
      theta[0] += alpha * td[0];
      theta[1] += alpha * td[1];
      if (ok == -1) break;
      if (std::sqrt(gradth[0] * gradth[0] + gradth[1] * gradth[1]) < alphatol) break;
      if (j + 1 > 2) break;
    }
    return ok;
  }

  // One fixed-magnetization solve from the density in c->rho (both channels)
  // with electron numbers tot (RhoOptN.f90:336-420, DensityOptimization).
  void solve() {
    outcome = -1;
    success = 1;
    it = 1;
    theta[0] = theta[1] = 0.2;
    const double ftol = (double)1e-4f, gtol = (double)2e-1f, xtol = 1e-12;
    const double npts = (double)n;
    vec_sqrt_abs(c->phi, c->rho, 2 * n, st);
    const double zero[2] = {0.0, 0.0};
    mu0[0] = mu0[1] = 0.0;
    eval_at(zero);
    chem_pot_and_dl();
    double e_hist[3] = {energy[0], 0.0, 0.0};
    while (true) {
      for (int k = 0; k < 2; ++k) newton(k);
      bool sw1 = false, sw2 = false;
      while (true) {
        next_gradient();
        bool done = false;
        for (int k = 0; k < 2 && !done; ++k) {
          if (gradth[k] > 1.0e5) throw std::runtime_error(k == 0 ? "The gradient 1 is too large" : "The gradient 2 is too large");
          if (gradth[k] >= 0.0) {
            if (sw1 && sw2) {
              outcome = 1;
              done = true;
              break;
            }
            (k == 0 ? sw1 : sw2) = true;
            vec_lincomb(dir[k], dl[k], dl[k], -1.0, 0.0, n, st);
          }
        }
        if (done || (gradth[0] < 0.0 && gradth[1] < 0.0)) break;
      }
      if (outcome == 1) break;
      success = theta_search(ftol, gtol, xtol);
      if (success == -1 && it > prm.max_iter) throw std::runtime_error("Fail rho optimizations");
      // StepInformation: both channels, the previous mu (RhoOptN.f90:745-760)
      pot_norm = std::sqrt(s[7] / npts / 2.0);
      bool exit_flag = false;
      const bool pconv = pot_norm < prm.pot_tol;
      e_hist[2] = e_hist[1];
      e_hist[1] = e_hist[0];
      e_hist[0] = energy[0];
      const bool econv = it >= 3 && std::fabs(e_hist[0] - e_hist[1]) < prm.tole &&
                         std::fabs(e_hist[0] - e_hist[2]) < prm.tole;
      if ((prm.conv_check == 1 && econv) || (prm.conv_check == 2 && pconv) ||
          (prm.conv_check == 3 && (pconv || econv))) {
        outcome = 0;
        exit_flag = true;
      }
      if (outcome == 1) exit_flag = true;
      if (it >= prm.max_iter) {
        outcome = 1;
        exit_flag = true;
      }
      // phi_s = phi_s cos(theta_s) + dir_s sin(theta_s): the point of the last
      // evaluation, so its sums give the new mu
      for (int k = 0; k < 2; ++k) vec_lincomb(phi[k], phi[k], dir[k], std::cos(theta[k]), std::sin(theta[k]), n, st);
      chem_pot_and_dl();
      if (exit_flag) break;
      ++it;
    }
    stats.iterations = it;
  }
};

}  // namespace

static int rho_opt_spin(ofd_ctx* c, double* rhoR, int dir_type, const ofd_opt_params* prm, double* energy9,
                        int* rho_outcome, ofd_opt_stats* stats) {
  if (dir_type != 1) return fail(c, "spin-polarized density optimization is built for dir_type 1 (NTN)");
  auto t0 = std::chrono::steady_clock::now();
  SpinSolver S;
  try {
    alloc_opt(c);
    S.bind(c);
    fill_params(prm, &S.prm);
    const long long n = c->nr();
    HIPCHK(hipMemcpyAsync(c->rho, rhoR, 2 * n * sizeof(double), hipMemcpyHostToDevice, c->st));
    for (int k = 0; k < 2; ++k) S.tot[k] = dsum(c->rho + k * n, n, S.bs, c->st) * c->dV;
    // the magnetization loop (RhoOptN.f90:336-480): FIXMAG solves once
    const double ntot = S.tot[0] + S.tot[1];
    double mag = S.tot[0] - S.tot[1], de = 1.0;
    double ee1 = 0, ee2 = 0, ee3 = 0, m1 = 0, m2 = 0, m3 = 0;
    int mag_iter = 0;
    for (int imag = 0; imag < c->nmag; ++imag) {
      S.solve();
      ++mag_iter;
      if (c->fixed_mag) break;
      if (std::fabs(de) < c->tolm) break;
      ee3 = ee2; ee2 = ee1; ee1 = S.energy[0];
      m3 = m2; m2 = m1; m1 = mag;
      de = (ee1 - ee2) / (m1 - m2);
      if (mag_iter == 1) mag = m1 + (double)0.1f;
      if (mag_iter == 2) mag = m1 - (double)0.2f;
      if (mag_iter == 3) mag = m1 - de / ee1;
      if (mag_iter > 3) {
        const double b = (m1 - m2) / (m3 - m2);
        const double de1 = 2.0 * (ee1 - (1 - b) * ee2 - b * ee3) / ((1.0 - 1.0 / b) * (m1 - m2) * (m1 - m2));
        if (mag_iter > 60) break;
        mag = m1 - de / de1;
      }
      // each solve restarts from the initial density, rescaled to the new
      // spin populations (Ntot +- m) / 2
      const double t_new[2] = {0.5 * (ntot + mag), 0.5 * (ntot - mag)};
      for (int k = 0; k < 2; ++k) {
        vec_axpby(c->rho + k * n, c->rho + k * n, 0.0, 1.0 / S.tot[k], n, c->st);
        vec_axpby(c->rho + k * n, c->rho + k * n, 0.0, t_new[k], n, c->st);
        S.tot[k] = t_new[k];
      }
    }
    c->mag_state[0] = mag;
    c->mag_state[1] = S.tot[0];
    c->mag_state[2] = S.tot[1];
    c->mag_iters = mag_iter;
    vec_square(c->rho, c->phi, 2 * n, c->st);   // rhoR = phi**2 (:495)
    HIPCHK(hipMemcpyAsync(rhoR, c->rho, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    S.bs.free_all();
  } catch (const std::exception& e) {
    S.bs.free_all();
    return fail(c, e.what());
  }
  if (energy9) std::memcpy(energy9, S.energy, sizeof(S.energy));
  if (rho_outcome) *rho_outcome = S.outcome;
  if (stats) {
    *stats = S.stats;
    stats->iterations = S.it;
    stats->final_pot_norm = S.pot_norm;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return 0;
}

extern "C" int ofd_set_spin(ofd_ctx* c, const ofd_spin* sp) {
  if (!c || !sp) return fail(c, "null argument");
  if (sp->nspin != 1 && sp->nspin != 2) return fail(c, "nspin must be 1 or 2");
  const int lsda = sp->lsda <= 0 ? 1 : sp->lsda;
  if (lsda != 1 && lsda != 2) return fail(c, "lsda must be 1 (Perdew-Zunger) or 2 (Perdew-Wang 92)");
  if (sp->nspin == 2 && c->kinetic != 4)
    return fail(c, "spin-polarized evaluation is built for the Wang-Teter functional (kinetic 4)");
  const int base = c->kinetic == 5 ? (int)kWGC8 : (c->alpha == c->beta ? (int)kWT3 : (int)kWT4);
  const int kind = sp->nspin == 2 ? (base == kWT3 ? (int)kWT3S : (int)kWT4S) : base;
  try {
    HIPCHK(hipStreamSynchronize(c->st));
    const long long nr = c->nr(), nc = c->sp.d.nc;
    if (func_nf(kind) > func_nf(c->kind)) {
      dev_free(c->work);
      c->work = nullptr;
      c->work = dalloc<double2>((size_t)func_nf(kind) * nc);
    }
    if (sp->nspin != c->nspin) {
      free_opt(c);
      if (sp->nspin == 2) {
        dev_free(c->rho);
        dev_free(c->pot);
        c->rho = c->pot = nullptr;
        c->rho = dalloc<double>(2 * nr);
        c->pot = dalloc<double>(2 * nr);
      }
    }
    graphs_clear(c->graphs);
    std::memset(&c->last_desc, 0, sizeof(c->last_desc));
    c->many_last.clear();
    c->kind = kind;
    c->nspin = sp->nspin;
    c->lsda = lsda;
    c->fixed_mag = sp->fixed_mag ? 1 : 0;
    c->nmag = sp->nmag > 0 ? sp->nmag : 100;
    c->tolm = sp->tolm > 0.0 ? sp->tolm : (double)5.0e-6f;
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  return 0;
}

extern "C" int ofd_get_magnetization(const ofd_ctx* c, double* out3, int* mag_iters) {
  if (!c || !out3) return -1;
  for (int k = 0; k < 3; ++k) out3[k] = c->mag_state[k];
  if (mag_iters) *mag_iters = c->mag_iters;
  return 0;
}

extern "C" int ofd_spin_split(const ofd_ctx* c, const double* rho, double magmom, double* rho2) {
  if (!c || !rho || !rho2) return -1;
  const long long n = c->nr();
  for (int s = 1; s <= 2; ++s) {
    const double sgn = (s % 2 == 0) ? 1.0 : -1.0;   // (-1)^s
    for (long long i = 0; i < n; ++i) rho2[(s - 1) * n + i] = 0.5 * (rho[i] - sgn * magmom / c->vol);
  }
  return 0;
}

// ====================================================================== splines (host)
// spline_cubic_set (MathSplines.f90:178-330) with its s3_fs tridiagonal solve
static void spline_set_host(int n, const double* t, const double* y, int ibcbeg, double ybcbeg, int ibcend,
                            double ybcend, double* ypp) {
  std::vector<double> diag(n), sub(n), sup(n), b(n);
  if (ibcbeg == 0) { b[0] = 0.0; diag[0] = 1.0; sup[0] = -1.0; }
  else if (ibcbeg == 1) { b[0] = (y[1] - y[0]) / (t[1] - t[0]) - ybcbeg; diag[0] = (t[1] - t[0]) / 3.0; sup[0] = (t[1] - t[0]) / 6.0; }
  else { b[0] = ybcbeg; diag[0] = 1.0; sup[0] = 0.0; }
  for (int i = 1; i < n - 1; ++i) {
    b[i] = (y[i + 1] - y[i]) / (t[i + 1] - t[i]) - (y[i] - y[i - 1]) / (t[i] - t[i - 1]);
    sub[i] = (t[i] - t[i - 1]) / 6.0;
    diag[i] = (t[i + 1] - t[i - 1]) / 3.0;
    sup[i] = (t[i + 1] - t[i]) / 6.0;
  }
  if (ibcend == 0) { b[n - 1] = 0.0; sub[n - 1] = -1.0; diag[n - 1] = 1.0; }
  else if (ibcend == 1) { b[n - 1] = ybcend - (y[n - 1] - y[n - 2]) / (t[n - 1] - t[n - 2]); sub[n - 1] = (t[n - 1] - t[n - 2]) / 6.0; diag[n - 1] = (t[n - 1] - t[n - 2]) / 3.0; }
  else { b[n - 1] = ybcend; sub[n - 1] = 0.0; diag[n - 1] = 1.0; }
  if (n == 2 && ibcbeg == 0 && ibcend == 0) { ypp[0] = ypp[1] = 0.0; return; }
  for (int i = 1; i < n - 1; ++i) {
    double xm = sub[i] / diag[i - 1];
    diag[i] = diag[i] - xm * sup[i - 1];
    b[i] = b[i] - xm * b[i - 1];
  }
  double xm = sub[n - 1] / diag[n - 2];
  diag[n - 1] = diag[n - 1] - xm * sup[n - 2];
  ypp[n - 1] = (b[n - 1] - xm * b[n - 2]) / diag[n - 1];
  for (int i = n - 2; i >= 0; --i) ypp[i] = (b[i] - sup[i] * ypp[i + 1]) / diag[i];
}

// ====================================================================== WGC tables
extern "C" int ofd_set_wgc_table(ofd_ctx* c, int nk, const double* eta, const double* w, const double* w1,
                                 const double* w2) {
  if (!c || !eta || !w || !w1 || !w2 || nk < 3) return fail(c, "bad WGC table arguments");
  if (c->kind != kWGC8) return fail(c, "context is not a WGC functional");
  try {
    // int_w1 = eta w', int_w2 = eta^2 w'' and natural-type splines with ibc 0
    // (KEDF_WGCkernel.f90:172-179)
    std::vector<double> iw1(nk), iw2(nk), pp0(nk), pp1(nk), pp2(nk);
    for (int i = 0; i < nk; ++i) {
      iw1[i] = eta[i] * w1[i];
      iw2[i] = eta[i] * eta[i] * w2[i];
    }
    spline_set_host(nk, eta, w, 0, 0.0, 0, 0.0, pp0.data());
    spline_set_host(nk, eta, iw1.data(), 0, 0.0, 0, 0.0, pp1.data());
    spline_set_host(nk, eta, iw2.data(), 0, 0.0, 0, 0.0, pp2.data());
    double* d = dalloc<double>(7 * (size_t)nk);
    const double* src[7] = {eta, w, iw1.data(), iw2.data(), pp0.data(), pp1.data(), pp2.data()};
    for (int k = 0; k < 7; ++k)
      HIPCHK(hipMemcpyAsync(d + (size_t)k * nk, src[k], nk * sizeof(double), hipMemcpyHostToDevice, c->st));
    const double tkfs = 2.0 * std::pow(3.0 * kPi * kPi * c->rhoS, 1.0 / 3.0);   // KEDF_WGCkernel.f90:93
    int rc = build_wgc_kernels(c->sp, c->rb, d, d + nk, d + 2 * (size_t)nk, d + 3 * (size_t)nk, d + 4 * (size_t)nk,
                               d + 5 * (size_t)nk, d + 6 * (size_t)nk, nk, tkfs, c->gamma, c->rhoS, c->wgck, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    dev_free(d);
    if (rc) return fail(c, "WGC kernel fill failed");
    c->have_wgc = true;
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  return 0;
}

// The WGC kernel ODE in eta (right-hand side of KEDF_KernelODE.f90:481-537):
//   w'' = [20 (F_L(eta) - 3 eta^2 - 1) - (gamma + 1 - 10) eta w' - 36 (5/3 - beta) beta w] / eta^2
// integrated from eta = 100 (w = wInf, w' = 0) toward 0 by an adaptive
// Dormand-Prince 5(4) pair at a tight tolerance, recording w, w', w'' on the
// knots of makeKernel2 (KEDF_KernelODE.f90:316-349, zero at the last knot).
static void wgc_rhs(double eta, const double* y, double* yp, double beta, double gamma) {
  yp[0] = y[1];
  double fl;
  if (eta == 1.0) fl = 2.0;
  else if (eta == 0.0) fl = 1.0;
  else fl = 1.0 / (0.5 + (1.0 - eta * eta) / (4.0 * eta) * std::log(std::fabs((1.0 + eta) / (1.0 - eta))));
  double v = 20.0 * (fl - 3.0 * eta * eta - 1.0) - (gamma + 1.0 - 10.0) * eta * y[1] - 36.0 * (5.0 / 3.0 - beta) * beta * y[0];
  yp[1] = v / (eta * eta);
}

static int wgc_ode_table(double alpha, double beta, double gamma, std::vector<double>& eta, std::vector<double>& w,
                         std::vector<double>& w1, std::vector<double>& w2) {
  const double ts[2] = {1e-3, 3.0 + 1e-3}, te[2] = {3.0, 100.0}, dt[2] = {1e-3, 1e-2};
  int np[2];
  for (int i = 0; i < 2; ++i) np[i] = (int)std::floor((te[i] - ts[i]) / dt[i]) - 1;
  const int n = np[0] + np[1];
  eta.resize(n); w.assign(n, 0.0); w1.assign(n, 0.0); w2.assign(n, 0.0);
  for (int i = 0; i < np[0]; ++i) eta[i] = ts[0] + (double)i * dt[0];
  for (int i = np[0]; i < n; ++i) eta[i] = ts[1] + (double)(i - np[0]) * dt[1];
  // Dormand-Prince coefficients
  const double c2 = 1. / 5, c3 = 3. / 10, c4 = 4. / 5, c5 = 8. / 9;
  const double a21 = 1. / 5, a31 = 3. / 40, a32 = 9. / 40, a41 = 44. / 45, a42 = -56. / 15, a43 = 32. / 9;
  const double a51 = 19372. / 6561, a52 = -25360. / 2187, a53 = 64448. / 6561, a54 = -212. / 729;
  const double a61 = 9017. / 3168, a62 = -355. / 33, a63 = 46732. / 5247, a64 = 49. / 176, a65 = -5103. / 18656;
  const double b1 = 35. / 384, b3 = 500. / 1113, b4 = 125. / 192, b5 = -2187. / 6784, b6 = 11. / 84;
  const double e1 = 71. / 57600, e3 = -71. / 16695, e4 = 71. / 1920, e5 = -17253. / 339200, e6 = 22. / 525, e7 = -1. / 40;
  double y[2] = {-1.6 * 20.0 / (36.0 * alpha * beta), 0.0};
  double t = eta[n - 1];
  double h = -1e-3;
  const double rtol = 1e-11, atol = 1e-13;
  for (int L = n - 2; L >= 0; --L) {
    const double tw = eta[L];
    while (t > tw) {
      if (t + h < tw) h = tw - t;
      double k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], k7[2], yt[2], y5[2];
      wgc_rhs(t, y, k1, beta, gamma);
      for (int i = 0; i < 2; ++i) yt[i] = y[i] + h * a21 * k1[i];
      wgc_rhs(t + c2 * h, yt, k2, beta, gamma);
      for (int i = 0; i < 2; ++i) yt[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
      wgc_rhs(t + c3 * h, yt, k3, beta, gamma);
      for (int i = 0; i < 2; ++i) yt[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
      wgc_rhs(t + c4 * h, yt, k4, beta, gamma);
      for (int i = 0; i < 2; ++i) yt[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
      wgc_rhs(t + c5 * h, yt, k5, beta, gamma);
      for (int i = 0; i < 2; ++i) yt[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
      wgc_rhs(t + h, yt, k6, beta, gamma);
      for (int i = 0; i < 2; ++i) y5[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
      wgc_rhs(t + h, y5, k7, beta, gamma);
      double err = 0.0;
      for (int i = 0; i < 2; ++i) {
        double ei = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * k7[i]);
        double sc = atol + rtol * std::fmax(std::fabs(y[i]), std::fabs(y5[i]));
        err = std::fmax(err, std::fabs(ei) / sc);
      }
      if (err <= 1.0) {
        t += h;
        y[0] = y5[0];
        y[1] = y5[1];
      }
      double fac = err > 0.0 ? 0.9 * std::pow(err, -0.2) : 5.0;
      fac = std::fmin(5.0, std::fmax(0.2, fac));
      h *= fac;
      if (std::fabs(h) < 1e-14) return -1;
    }
    double yp[2];
    wgc_rhs(tw, y, yp, beta, gamma);
    w[L] = y[0];
    w1[L] = y[1];
    w2[L] = yp[1];
  }
  return n;
}

extern "C" int ofd_wgc_kernel_table(double alpha, double beta, double gamma, int nmax, double* eta_out,
                                    double* w_out, double* w1_out, double* w2_out) {
  if (!(alpha > 0.0) || !(beta > 0.0) || !(gamma > 0.0)) return -1;
  try {
    std::vector<double> eta, w, w1, w2;
    const int n = wgc_ode_table(alpha, beta, gamma, eta, w, w1, w2);
    if (n <= 0) return -2;
    if (nmax < n || !eta_out || !w_out || !w1_out || !w2_out) return n;
    std::copy(eta.begin(), eta.end(), eta_out);
    std::copy(w.begin(), w.end(), w_out);
    std::copy(w1.begin(), w1.end(), w1_out);
    std::copy(w2.begin(), w2.end(), w2_out);
    return n;
  } catch (const std::exception&) {
    return -3;
  }
}

extern "C" int ofd_build_wgc_table(ofd_ctx* c, int* nk_out) {
  if (!c) return -1;
  if (c->kind != kWGC8) return fail(c, "context is not a WGC functional");
  try {
    std::vector<double> eta, w, w1, w2;
    int n = wgc_ode_table(c->alpha, c->beta, c->gamma, eta, w, w1, w2);
    if (n <= 0) return fail(c, "WGC ODE integration failed");
    if (nk_out) *nk_out = n;
    return ofd_set_wgc_table(c, n, eta.data(), w.data(), w1.data(), w2.data());
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
}

// ====================================================================== ion terms on the device
extern "C" int ofd_build_ion_potential(ofd_ctx* c, int nspecies, const int* type_start, const double* frac,
                                       const double* maxg_invA, const int* npts, const int* poff,
                                       const double* values, int order, double eii) {
  if (!c || nspecies < 1 || !type_start || !frac || !maxg_invA || !npts || !poff || !values)
    return fail(c, "null argument");
  if (order <= 0) order = 10;
  if (order % 2 == 1) order += 1;   // SetupKEDF.f90:176-183
  if (order > 16) return fail(c, "spline order above 16 is not supported");
  try {
    const double bohr = 0.529177208607388;
    const double qElec = 1.602176462e-19, light = 299792458.0, mu0 = 4.0 * kPi * 1.0e-7;
    const double eps0 = 1.0 / (mu0 * light * light);
    const double h2ev = qElec * 1.0e10 / (4.0 * kPi * eps0 * bohr);
    c->type_start.assign(type_start, type_start + nspecies + 1);
    const int nion = type_start[nspecies];
    c->frac.assign(frac, frac + 3 * (size_t)nion);
    c->psp_off.clear(); c->psp_npts.clear(); c->psp_charge.clear(); c->psp_maxg.clear(); c->psp_v0.clear();
    c->psp_vqs.clear(); c->psp_dd.clear();
    for (int s = 0; s < nspecies; ++s) {
      // ReadPseudo (ReadIonFile.f90:848-976)
      const int n = npts[s];
      if (n < 3) return fail(c, "pseudopotential table too short");
      const double maxG = maxg_invA[s] * bohr;
      std::vector<double> pv(n), t(n), vqs(n), dd(n);
      for (int i = 0; i < n; ++i) pv[i] = values[poff[s] + i] / (bohr * bohr * bohr * h2ev);
      const int charge = (int)std::lround(-(pv[1] - pv[0]) * std::pow(maxG / (double)(n - 1), 2.0) / (4.0 * kPi));
      for (int i = 0; i < n; ++i) t[i] = (double)(i + 1);
      vqs[0] = pv[0];
      for (int i = 1; i < n; ++i)
        vqs[i] = pv[i] + 4.0 * kPi * charge / std::pow((t[i] - 1.0) / (double)(n - 1) * maxG, 2.0);
      spline_set_host(n, t.data(), vqs.data(), 1, 0.0, 1, 0.0, dd.data());
      c->psp_off.push_back((int)c->psp_vqs.size());
      c->psp_npts.push_back(n);
      c->psp_charge.push_back(charge);
      c->psp_maxg.push_back(maxG);
      c->psp_v0.push_back(pv[0]);
      c->psp_vqs.insert(c->psp_vqs.end(), vqs.begin(), vqs.end());
      c->psp_dd.insert(c->psp_dd.end(), dd.begin(), dd.end());
    }
    c->spline_order = order;
    HIPCHK(hipStreamSynchronize(c->st));
    dev_free(c->d_frac); dev_free(c->d_psp_vqs); dev_free(c->d_psp_dd);
    c->d_frac = dalloc<double>(3 * (size_t)nion);
    c->d_psp_vqs = dalloc<double>(c->psp_vqs.size());
    c->d_psp_dd = dalloc<double>(c->psp_dd.size());
    // the host vectors outlive the copies: build_ion_potential synchronizes the stream
    HIPCHK(hipMemcpyAsync(c->d_frac, c->frac.data(), sizeof(double) * c->frac.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_psp_vqs, c->psp_vqs.data(), sizeof(double) * c->psp_vqs.size(),
                          hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_psp_dd, c->psp_dd.data(), sizeof(double) * c->psp_dd.size(), hipMemcpyHostToDevice,
                          c->st));
    const size_t scratch = sizeof(double) * c->sp.d.nr + 2 * sizeof(double2) * c->sp.d.nc;
    void* sc = dalloc<char>(scratch);
    int rc = build_ion_potential(c->sp, c->rb, c->vol, nspecies, c->type_start.data(), c->frac.data(), c->d_frac, nion, c->d_psp_vqs,
                                 c->d_psp_dd, c->psp_off.data(), c->psp_npts.data(), c->psp_maxg.data(),
                                 c->psp_v0.data(), c->psp_charge.data(), order, c->vion, sc, c->st);
    hipStreamSynchronize(c->st);
    dev_free(sc);
    if (rc) return fail(c, "ion potential build failed");
    c->e_ionion = eii;
    c->have_vion = true;
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  return 0;
}

extern "C" int ofd_ion_forces(ofd_ctx* c, const double* rho, double* forces) {
  if (!c || !rho || !forces) return fail(c, "null argument");
  if (c->type_start.empty()) return fail(c, "ions not set (ofd_build_ion_potential)");
  try {
    const int ns = (int)c->type_start.size() - 1;
    const int nion = c->type_start[ns];
    double cinv[9], rb[9], vol;
    cell_setup(c->lat, rb, &vol, cinv);
    HIPCHK(hipMemcpyAsync(c->rho, rho, c->nr_all() * sizeof(double), hipMemcpyHostToDevice, c->st));
    // forces act on the total density, SUM(rhoR, 4) (CalForce.f90:180)
    if (c->nspin == 2) vec_axpby(c->rho, c->rho + c->nr(), 1.0, 1.0, c->nr(), c->st);
    double* df = dalloc<double>(3 * (size_t)nion);
    const size_t scratch = sizeof(double) * c->sp.d.nr + 2 * sizeof(double2) * c->sp.d.nc;
    void* sc = dalloc<char>(scratch);
    int rc = ion_forces_spline(c->sp, c->rb, cinv, ns, c->type_start.data(), c->d_frac, nion, c->d_psp_vqs,
                               c->d_psp_dd, c->psp_off.data(), c->psp_npts.data(), c->psp_maxg.data(),
                               c->psp_v0.data(), c->psp_charge.data(), c->spline_order, c->rho, df, sc, c->st);
    HIPCHK(copy_sync(forces, df, sizeof(double) * 3 * nion, hipMemcpyDeviceToHost, c->st));
    hipStreamSynchronize(c->st);
    dev_free(sc);
    dev_free(df);
    if (rc) return fail(c, "force kernels failed");
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  return 0;
}

// ====================================================================== ion-ion Ewald sum
// EwaldSetup (Ewald.f90:208-460) with the particle-mesh grid equal to the
// density grid: recipCutoff from the grid dimensions, eta lowered from 1.2
// in steps of 0.01 A until the reciprocal error bound is below 1e-10 eV, the
// real cutoff raised from 6 bohr in steps of 0.01 A until the real-space
// bound is (errorTolerance, etaIncrement, realCutoffIncrement: Ewald.f90:112-128).
static void ewald_setup(const double* lat, double vol, int n1, int n2, int n3, const std::vector<double>& q,
                        double* eta_out, double* rc_out, double* kc_out, double* real_err, double* recip_err) {
  const double bohr = 0.529177208607388;
  const double qElec = 1.602176462e-19, light = 299792458.0, mu0 = 4.0 * kPi * 1.0e-7;
  const double eps0 = 1.0 / (mu0 * light * light);
  const double h2ev = qElec * 1.0e10 / (4.0 * kPi * eps0 * bohr);
  const double tol = 1.0e-10 / h2ev, step = 0.01 * bohr;
  const int nion = (int)q.size();
  double sq = 0.0;
  for (double v : q) sq += v * v;
  const double avg = sq / (double)nion;
  auto norm = [&](int j) { return std::sqrt(lat[3 * j] * lat[3 * j] + lat[3 * j + 1] * lat[3 * j + 1] + lat[3 * j + 2] * lat[3 * j + 2]); };
  const double kc = std::min(std::min((n1 - 1) / norm(0), (n2 - 1) / norm(1)), (n3 - 1) / norm(2)) * kPi;
  const double sqrt_pi = std::sqrt(kPi);
  double eta = 1.2, re = 0.0, ke = 0.0;
  for (;;) {
    ke = (double)nion * (double)nion * avg / sqrt_pi * eta * std::erfc(kc / (2.0 * eta));
    if (ke < tol) break;
    eta -= step;
  }
  double rc = 6.0;
  for (;;) {
    re = kPi * (double)nion * (double)nion * avg / (vol * eta * eta) * std::erfc(rc * eta);
    if (re < tol) break;
    rc += step;
  }
  *eta_out = eta;
  *rc_out = rc;
  *kc_out = kc;
  *real_err = re;
  *recip_err = ke;
}

extern "C" int ofd_ewald(ofd_ctx* c, double* energy, double* forces, double* info5) {
  if (!c || !energy) return fail(c, "null argument");
  if (c->type_start.empty()) return fail(c, "ions not set (ofd_build_ion_potential)");
  try {
    const int ns = (int)c->type_start.size() - 1;
    const int nion = c->type_start[ns];
    std::vector<double> charge(ns), q(nion);
    for (int s = 0; s < ns; ++s) {
      charge[s] = (double)c->psp_charge[s];
      for (int k = c->type_start[s]; k < c->type_start[s + 1]; ++k) q[k] = charge[s];
    }
    double eta, rc, kc, re, ke;
    ewald_setup(c->lat, c->vol, c->sp.d.n1, c->sp.d.n2, c->sp.d.n3, q, &eta, &rc, &kc, &re, &ke);
    double cinv[9], rb[9], vol;
    cell_setup(c->lat, rb, &vol, cinv);
    const size_t scratch = sizeof(double) * c->sp.d.nr + 2 * sizeof(double2) * c->sp.d.nc;
    void* sc = dalloc<char>(scratch);
    double comps[4];
    const int rcode = ewald_device(c->sp, c->rb, c->lat, c->vol, ns, c->type_start.data(), charge.data(), c->frac.data(), c->d_frac,
                                   nion, c->spline_order, eta, rc, kc, cinv, comps, forces, sc, c->st);
    hipStreamSynchronize(c->st);
    dev_free(sc);
    if (rcode) return fail(c, "Ewald kernels failed");
    // EwaldEnergy = SUM(ewaldComponents) (Ewald.f90:553-580), the context's
    // ionIonEnergy from now on (RefreshIons.f90:196)
    *energy = comps[0] + comps[1] + comps[2] + comps[3];
    c->e_ionion = *energy;
    if (info5) {
      info5[0] = eta;
      info5[1] = rc;
      info5[2] = kc;
      info5[3] = re;
      info5[4] = ke;
    }
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  return 0;
}

// ====================================================================== ion relaxation (CG2)
// RefreshIonTerms (RefreshIons.f90:125-238) for positions already stored in
// c->frac: the ionic potential from the species tables kept by
// ofd_build_ion_potential, on the device.
static int refresh_ion_potential(ofd_ctx* c) {
  const int ns = (int)c->type_start.size() - 1;
  const int nion = c->type_start[ns];
  HIPCHK(hipMemcpyAsync(c->d_frac, c->frac.data(), sizeof(double) * c->frac.size(), hipMemcpyHostToDevice, c->st));
  const size_t scratch = sizeof(double) * c->sp.d.nr + 2 * sizeof(double2) * c->sp.d.nc;
  void* sc = dalloc<char>(scratch);
  const int rc = build_ion_potential(c->sp, c->rb, c->vol, ns, c->type_start.data(), c->frac.data(), c->d_frac, nion, c->d_psp_vqs,
                                     c->d_psp_dd, c->psp_off.data(), c->psp_npts.data(), c->psp_maxg.data(),
                                     c->psp_v0.data(), c->psp_charge.data(), c->spline_order, c->vion, sc, c->st);
  hipStreamSynchronize(c->st);
  dev_free(sc);
  return rc;
}

// NNDist_bin (Neighbor.f90:68-420) for positions wrapped into the cell: the
// minimum over ion pairs and image shifts in {-1,0,1}^3 (the one-bin form;
// the binned search returns the same whenever the nearest distance is below
// its 20 bohr bin spacing).
static double nn_min_distance(const double* lat, const std::vector<double>& frac) {
  const int n = (int)frac.size() / 3;
  std::vector<double> xyz(3 * (size_t)n);
  for (int a = 0; a < n; ++a) {
    double f[3];
    for (int k = 0; k < 3; ++k) {
      f[k] = std::fmod(frac[3 * a + k], 1.0);
      if (f[k] < 0) f[k] += 1.0;
    }
    for (int i = 0; i < 3; ++i) xyz[3 * a + i] = lat[i] * f[0] + lat[3 + i] * f[1] + lat[6 + i] * f[2];
  }
  double best = 1e6;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int sk = -1; sk <= 1; ++sk)
        for (int sj = -1; sj <= 1; ++sj)
          for (int si = -1; si <= 1; ++si) {
            if (a == b && si == 0 && sj == 0 && sk == 0) continue;
            double d2 = 0.0;
            for (int i = 0; i < 3; ++i) {
              const double d = xyz[3 * a + i] - (xyz[3 * b + i] + lat[i] * si + lat[3 + i] * sj + lat[6 + i] * sk);
              d2 += d * d;
            }
            best = std::min(best, std::sqrt(d2));
          }
  return best;
}

// LnSrch (MathLNsrch.f90:52-296) with its module state points(2,3) and dsave:
// two bracket points (alpha, E, dE/dalpha), secant steps on dE/dalpha. The
// reference reports its state as a text task whose first word (FG, CONV,
// WARNING, ERROR) drives the caller; here that word is an enum, and `why`
// numbers the branch that gave a warning or an error (1: no usable second
// point yet, 2-3: secant step outside the bracket with the second point
// first, 4-5: the points out of order, 6-7: secant step outside the bracket
// after a positive slope, 8-9: the same after a negative slope).
enum class LsTask { Start, FG, Conv, Warn, Error };

struct IonLnSrch {
  double P[2][3];
  double d0 = 0, d1 = 0;   // dsave(1:2)
  int why = 0;
  static double new_alp(double x1, double y1, double x2, double y2, double stpmin, double stpmax) {
    const double slope = (y1 - y2) / (x1 - x2);
    const double b = y1 - x1 * slope;
    const double a = -b / slope;
    return (a >= stpmax || a <= stpmin) ? -1.0 : a;
  }
  double na(double stpmin, double stpmax) const { return new_alp(P[0][0], P[0][2], P[1][0], P[1][2], stpmin, stpmax); }
  void set(int k, double a, double f, double g) { P[k][0] = a; P[k][1] = f; P[k][2] = g; }
  LsTask warn(int w) { why = w; return LsTask::Warn; }
  LsTask error(int w) { why = w; return LsTask::Error; }
  LsTask step(double& alp, double f, double g, LsTask task, double c1, double c2, double stpmin, double stpmax) {
    if (task == LsTask::Start) {
      if (g > 0.0) return error(0);   // not a descent direction
      for (auto& r : P) r[0] = r[1] = r[2] = -1.0;
      set(0, 0.0, f, g);
      d0 = f;
      d1 = g;
      P[1][0] = -1.0;
      return LsTask::FG;
    }
    if (task == LsTask::FG && P[1][0] > 0.0) {
      if (f <= d0 + c1 * alp * d1 && std::fabs(g) <= c2 * std::fabs(d1)) return LsTask::Conv;
    }
    if (P[1][0] < 0.0) {
      if (g < P[0][2]) {
        alp = f < P[0][1] ? alp * 5.0 : (P[0][0] + alp) / 2.0;
        return warn(1);
      }
      set(1, alp, f, g);
      alp = na(stpmin, stpmax);
      if (alp < 0.0 || alp < P[0][0]) {
        if (P[1][1] < P[0][1]) {
          alp = 5.0 * P[1][0];
          return warn(2);
        }
        return error(3);
      }
      return LsTask::FG;
    }
    if (g > 0) {
      if (P[1][2] > 0.0) {
        if (P[1][0] > alp) {
          set(1, alp, f, g);
          alp = na(stpmin, stpmax);
          if (alp < P[0][0]) {
            alp = (P[0][0] + P[1][0]) / 2.0;
            return warn(4);
          }
          return LsTask::FG;
        }
        if (f < P[1][1]) {
          alp = alp * 5.0;
          return warn(5);
        }
        alp = (P[0][0] + P[1][0]) / 2.0;
        return warn(5);
      }
      if (P[1][2] < 0.0) {
        std::memcpy(P[0], P[1], sizeof(P[0]));
        set(1, alp, f, g);
        alp = na(stpmin, stpmax);
        if (alp < P[0][0] || alp > P[1][0]) {
          if (P[0][1] > P[1][1]) {
            alp = P[1][0] * 5.0;
            return warn(6);
          }
          alp = (P[0][0] + P[1][0]) / 2.0;
          return warn(7);
        }
        return LsTask::FG;
      }
    }
    if (g < 0.0) {
      if (P[1][2] > 0.0) {
        set(0, alp, f, g);
        alp = na(stpmin, stpmax);
        if (alp < P[0][0] || alp > P[1][0]) {
          if (P[0][1] > P[1][1]) {
            alp = 5.0 * P[1][0];
            return warn(8);
          }
          alp = (P[0][0] + P[1][0]) / 2.0;
          return warn(8);
        }
        return LsTask::FG;
      }
      if (P[1][2] < 0.0) {
        std::memcpy(P[0], P[1], sizeof(P[0]));
        set(1, alp, f, g);
        alp = na(stpmin, stpmax);
        if (alp < P[1][0]) {
          if (P[0][1] > P[1][1]) {
            alp = P[1][0] * 5.0;
            return warn(9);
          }
          return error(9);
        }
        return LsTask::FG;
      }
    }
    return task;   // g == 0 with the second point set: the reference leaves task as it was
  }
};

extern "C" int ofd_ion_relax(ofd_ctx* c, double* rhoR, double* frac_xyz, const ofd_opt_params* rho_params,
                             const ofd_ion_params* ion, double* energy9, double* forces, double* energies,
                             int max_energies, ofd_ion_stats* stats) {
  if (!c || !rhoR || !frac_xyz || !rho_params) return fail(c, "null argument");
  if (c->type_start.empty()) return fail(c, "ions not set (ofd_build_ion_potential)");
  if (c->nspin != 1) return fail(c, "ion relaxation is built for numSpin = 1");
  auto t0 = std::chrono::steady_clock::now();
  // IonOptimizers.f90:29,36,41 and Neighbor.f90:33 defaults; forceCutoff is
  // a default-real literal (5e-5 in single precision)
  const int cg_type = (ion && ion->cg_type > 0) ? ion->cg_type : 2;
  const double fcut = (ion && ion->force_cutoff > 0) ? ion->force_cutoff : (double)5e-5f;
  const int max_steps = (ion && ion->max_steps > 0) ? ion->max_steps : 200;
  const double nn_dist = (ion && ion->nn_dist >= 0) ? ion->nn_dist : 2.0;
  const int rho_dir = (ion && ion->rho_dir_type > 0) ? ion->rho_dir_type : 1;
  const int max_line = 10;
  const char* tr = std::getenv("OFD_TRACE");
  const bool trace = tr && std::atoi(tr) != 0;
  // IonOptCG2.f90:134-139: ftol, gtol default-real literals, xtol 1e-12_DP
  const double ftol = (double)1e-4f, gtol = (double)1e-1f;
  if (cg_type < 1 || cg_type > 3) return fail(c, "cg_type must be 1, 2 or 3");
  if (rho_dir > 3) return fail(c, "rho_dir_type must be 1 (NTN), 2 (NCG) or 3 (NBF)");
  const int ns = (int)c->type_start.size() - 1;
  const int nion = c->type_start[ns];
  const long long n = c->nr();
  std::vector<double> frac(frac_xyz, frac_xyz + 3 * (size_t)nion);
  for (double& v : frac) {
    v = std::fmod(v, 1.0);
    if (v < 0) v += 1.0;
  }
  double et[9] = {0};
  std::vector<double> F(3 * (size_t)nion), fie(3 * (size_t)nion), fii(3 * (size_t)nion);
  int ngeo = 0, step = 0;
  double max_force = 0.0;
  std::string err;
  // one geometry: RefreshIonTerms, RhoOptimizer (warm start from rhoR),
  // CalculateForces = Ewald + ion-electron (CalForce.f90:134-194)
  auto evaluate = [&]() -> bool {
    c->frac = frac;
    if (nn_dist > 0 && nn_min_distance(c->lat, frac) < nn_dist) {
      err = "nearest ion distance below the allowed threshold (CheckNearestDistanceAtoms)";
      return false;
    }
    // the device positions first: the Ewald sum reads them too
    if (refresh_ion_potential(c)) { err = "ion potential build failed"; return false; }
    double eii = 0.0;
    if (ofd_ewald(c, &eii, fii.data(), nullptr)) return false;
    int outcome = 0;
    ofd_opt_stats st;
    if (ofd_rho_opt(c, rhoR, rho_dir, rho_params, et, &outcome, &st)) return false;
    if (ofd_ion_forces(c, rhoR, fie.data())) return false;
    for (size_t k = 0; k < F.size(); ++k) F[k] = fii[k] + fie[k];
    if (energies && ngeo < max_energies) energies[ngeo] = et[0];
    if (trace) {
      double mf = 0.0;
      for (int a = 0; a < nion; ++a)
        mf = std::max(mf, std::sqrt(F[3 * a] * F[3 * a] + F[3 * a + 1] * F[3 * a + 1] + F[3 * a + 2] * F[3 * a + 2]));
      std::fprintf(stderr, "ofd ion-relax geometry %d: E = %.12e Ha, ion-ion %.12e, max |F| = %.6e, NTN steps %d\n",
                   ngeo, et[0], eii, mf, st.iterations);
    }
    ++ngeo;
    return true;
  };
  auto maxf = [&]() {
    double m = 0.0;
    for (int a = 0; a < nion; ++a)
      m = std::max(m, std::sqrt(F[3 * a] * F[3 * a] + F[3 * a + 1] * F[3 * a + 1] + F[3 * a + 2] * F[3 * a + 2]));
    return m;
  };
  auto dot = [](const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t k = 0; k < a.size(); ++k) s += a[k] * b[k];
    return s;
  };
  try {
    if (!evaluate()) return err.empty() ? -1 : fail(c, err);
    // frac <-> cart with cellReal (columns = lattice vectors) and its inverse
    double cinv[9], rb[9], vol;
    cell_setup(c->lat, rb, &vol, cinv);   // cinv[3c+j] = inv(cellReal)(j, c)
    const size_t m = F.size();
    std::vector<double> old_grad(m, 0.0), dir_old(m, 0.0), new_grad(m), d(m), old_cart(m);
    double alpha = 1.0;
    bool restart = false;
    IonLnSrch ls;
    for (;;) {
      ++step;
      if (step > max_steps) break;
      max_force = maxf();
      if (max_force < fcut) break;
      for (int a = 0; a < nion; ++a)
        for (int i = 0; i < 3; ++i)
          old_cart[3 * a + i] = c->lat[i] * frac[3 * a] + c->lat[3 + i] * frac[3 * a + 1] + c->lat[6 + i] * frac[3 * a + 2];
      new_grad = F;
      double gam = 0.0;
      if (step == 1 || restart || step % 30 == 0) {
        gam = 0.0;
        restart = false;
      } else if (cg_type == 1) {
        double num = 0.0;
        for (size_t k = 0; k < m; ++k) num += (new_grad[k] - old_grad[k]) * new_grad[k];
        gam = std::max(0.0, num / dot(old_grad, old_grad));
      } else if (cg_type == 2) {
        // Hager-Zhang (IonOptCG2.f90:203-214)
        std::vector<double> y(m);
        for (size_t k = 0; k < m; ++k) y[k] = -new_grad[k] + old_grad[k];
        const double ny = std::sqrt(dot(y, y)), dy = dot(dir_old, y);
        double g2 = 0.0;
        for (size_t k = 0; k < m; ++k) g2 += (y[k] - 2.0 * dir_old[k] * ny * ny / dy) * (-new_grad[k] / dy);
        const double nd = std::sqrt(dot(dir_old, dir_old)), ng = std::sqrt(dot(old_grad, old_grad));
        gam = std::max(g2, -1.0 / (nd * std::min(0.01, ng)));
      } else {
        std::vector<double> y(m);
        for (size_t k = 0; k < m; ++k) y[k] = -new_grad[k] + old_grad[k];
        gam = std::max(0.0, dot(new_grad, new_grad) / dot(dir_old, y));
      }
      for (size_t k = 0; k < m; ++k) d[k] = new_grad[k] + gam * dir_old[k];
      dir_old = d;
      old_grad = new_grad;
      double grad_alp = -dot(F, d);
      double dmax = d[0];
      for (double v : d) dmax = std::max(dmax, v);
      const double stpmin = 0.0;
      const double stpmax = nn_dist > 0 ? std::fabs((nn_min_distance(c->lat, frac) - nn_dist) / dmax)
                                        : std::fabs(1.0 / dmax);
      alpha = std::min(alpha, stpmax);
      LsTask task = LsTask::Start;
      int nline = 0;
      bool done = false;
      for (;;) {
        task = ls.step(alpha, et[0], grad_alp, task, ftol, gtol, stpmin, stpmax);
        if (nline >= 4 && std::fabs(ls.P[0][1] - ls.P[1][1]) < (double)1.e-6f) break;
        if (ls.P[0][0] > 0.0) {
          double bmax = 0.0;
          for (int a = 0; a < nion; ++a) {
            double s2 = 0.0;
            for (int i = 0; i < 3; ++i) {
              const double b = (ls.P[0][0] - ls.P[1][0]) * d[3 * a + i];
              s2 += b * b;
            }
            bmax = std::max(bmax, std::sqrt(s2));
          }
          if (bmax < (double)1e-5f) break;
        }
        if (task == LsTask::Conv) break;
        if (task == LsTask::Error) {
          restart = true;
          alpha = 1.0;
          break;
        }
        if (task == LsTask::FG || task == LsTask::Warn) {
          // coord = MODULO(cart2frac (oldCart + alpha dir), 1)
          for (int a = 0; a < nion; ++a)
            for (int j = 0; j < 3; ++j) {
              double v = 0.0;
              for (int i = 0; i < 3; ++i) v += cinv[3 * i + j] * (old_cart[3 * a + i] + alpha * d[3 * a + i]);
              v = std::fmod(v, 1.0);
              if (v < 0) v += 1.0;
              frac[3 * a + j] = v;
            }
          if (!evaluate()) return err.empty() ? -1 : fail(c, err);
          grad_alp = -dot(F, d);
          max_force = maxf();
          if (max_force < fcut) {
            done = true;
            break;
          }
          if (++nline > max_line) {
            restart = true;
            alpha = 1.0;
            break;
          }
        }
      }
      if (done) break;
    }
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  std::copy(frac.begin(), frac.end(), frac_xyz);
  if (energy9) std::memcpy(energy9, et, sizeof(et));
  if (forces) std::copy(F.begin(), F.end(), forces);
  if (stats) {
    stats->steps = step;
    stats->geometries = ngeo;
    stats->max_force = max_force;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  (void)n;
  return 0;
}

// ====================================================================== autotuner
// Times one evaluation (uniform density at the context's electron count) per
// candidate tile shape on the context's stream and keeps the fastest, by
// coordinate descent over the z-pass and the y-pass columns per workgroup.
// Both knobs are numerically neutral (each line is transformed by the same
// code whatever the tile), so a tuned context gives bitwise the results of an
// untuned one.
static double time_current(ofd_ctx* c, int reps) {
  ReqDesc D = c->base_desc();
  D.u = c->rho;
  D.out = c->pot;
  D.opt_sqrt = 1;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(copy_sync(c->desc, &D, sizeof(D), hipMemcpyHostToDevice, c->st));
  c->last_desc = D;
  const int fp32 = c->precision == 1;
  auto once = [&]() {
    if (run_eval(c->sp, c->kind, kInRho, kOutPot, fp32, c->desc, 1, c->st) ||
        run_finalize(c->sp, c->kind, fp32, c->desc, 1, c->fin, c->st))
      throw std::runtime_error("evaluation launch failed while tuning");
  };
  once();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->st));
  for (int i = 0; i < reps; ++i) once();
  HIPCHK(hipEventRecord(e1, c->st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return 1e3 * (double)ms / reps;
}

extern "C" int ofd_autotune(ofd_ctx* c, int reps, ofd_tune_result* res) {
  if (!c) return -1;
  if (reps < 1) reps = 5;
  try {
    const long long n = c->nr();
    std::vector<double> h(n, c->num_ele / c->vol);
    HIPCHK(copy_sync(c->rho, h.data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
    graphs_clear(c->graphs);
    const bool fast = !c->sp.force_generic && fast_path_ok(c->sp, c->kind);
    const int fp32 = c->precision == 1;
    std::vector<int> cy, cz, cx, cp;
    if (fast) {
      cy = fast_tile_candidates(c->sp, c->kind, fp32, 1);
      cz = fast_tile_candidates(c->sp, c->kind, fp32, 2);
      cx = {0};
      // the fused plane pass stays as the plan has it: it reaches the same
      // result as the separate passes only to rounding (different contraction
      // of the split's products), and tuning must never change a result's bits
      cp = {c->sp.no_plane ? 1 : 0};
    } else {
      // the x rows per workgroup set how the energy partials are grouped, so
      // they stay fixed: tuning never changes a result's bits
      cy = {1, 2, 4, 8, 16};
      cz = {1, 2, 4, 8};
      cx = {0};
      cp = {1};
    }
    cy.insert(cy.begin(), 0);
    cz.insert(cz.begin(), 0);
    ShapePlan& sp = c->sp;
    sp.rows_x = 0; sp.kb_y = 0; sp.kb_z = 0;
    if (cp.size() == 1) sp.no_plane = cp[0] != 0;
    const double t_def = time_current(c, reps);
    double best = t_def;
    int tried = 1;
    // coordinate descent, one dimension at a time
    int cur_plane = sp.no_plane ? 1 : 0, cur_z = 0, cur_y = 0, cur_x = 0;
    auto run_dim = [&](std::vector<int>& cand, int& cur, auto set) {
      int bestv = cur;
      for (int v : cand) {
        if (v == cur) continue;
        set(v);
        const double t = time_current(c, reps);
        ++tried;
        if (t < best * 0.995) { best = t; bestv = v; }
      }
      cur = bestv;
      set(bestv);
    };
    run_dim(cp, cur_plane, [&](int v) { sp.no_plane = v != 0; });
    run_dim(cz, cur_z, [&](int v) { sp.kb_z = v; });
    run_dim(cy, cur_y, [&](int v) { sp.kb_y = v; });
    if (!fast) run_dim(cx, cur_x, [&](int v) { sp.rows_x = v; });
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[TuneKey{sp.d.n1, sp.d.n2, sp.d.n3, c->kind, fp32}] = TuneChoice{sp.rows_x, sp.kb_y, sp.kb_z,
                                                                             sp.no_plane ? 1 : 0};
    }
    graphs_clear(c->graphs);
    std::memset(&c->last_desc, 0, sizeof(c->last_desc));
    if (res) {
      res->rows_x = sp.rows_x;
      res->kb_y = sp.kb_y;
      res->kb_z = sp.kb_z;
      res->plane = sp.no_plane ? 0 : 1;
      res->configs_tried = tried;
      res->us_default = t_def;
      res->us_best = best;
    }
  } catch (const std::exception& e) {
    return fail(c, e.what());
  }
  return 0;
}

// ====================================================================== request pool
struct ofd_pool {
  std::string err;
  int n1, n2, n3, capacity, kinetic, precision;
  hipStream_t st = nullptr;
  // the pool's own plan (twiddle tables live as long as the pool and its
  // captured graphs; a request's context may be destroyed at collection)
  ShapePlan sp;
  struct Slot {
    bool used = false;    // reserved by an admission
    bool ready = false;   // prepared and handed to the pool's stream
    bool done = false;
    ofd_ctx* c = nullptr;
    hipStream_t saved = nullptr;
    NtnSolver S;
    std::chrono::steady_clock::time_point t0;
  };
  std::vector<Slot> slots;
  // descriptors and sums of one phase: launch group g (functional kind,
  // output mode, precision) owns rows [g * capacity, (g + 1) * capacity), so
  // a group's launch always reads the same descriptor array and its captured
  // graph is keyed by the batch width alone
  static constexpr int kGroups = 3 * 3 * 2;
  ReqDesc* d_desc = nullptr;
  double* d_sums = nullptr;
  ReqDesc* h_desc = nullptr;   // pinned
  double* h_sums = nullptr;    // pinned
  BatchScratch bs;
  GraphCache graphs;
  // one phase (queued vector updates, then the dot products and the
  // evaluations the solvers wait on) is launched at a time on `st`
  std::vector<NtnSolver*> act, dots;
  std::vector<int> where;      // slot -> row of h_sums in the phase, -1 if none
  int nsum = 0;
  bool inflight = false;
  hipEvent_t ev_phase = nullptr;
  // guards slots and the staging buffers; held by admission, collection and
  // whoever drives the pool (ofd_pool_step or a server's progress thread)
  mutable std::mutex mu;
  // admission prepares a request on the request's own stream with this scratch
  std::mutex admit_mu;
  BatchScratch bs_admit;
  ofd_server* server = nullptr;
  bool failed = false;
};

extern "C" long long ofd_request_bytes(int n1, int n2, int n3, int kinetic, int precision) {
  const long long nr = (long long)n1 * n2 * n3;
  const long long nc = (long long)(n1 / 2 + 1) * n2 * n3;
  const int nf = kinetic == 5 ? 8 : 3;
  long long b = 8 * nr * 3;            // vion, rho staging, pot staging
  b += 16 * nc * nf;                   // complex work (FP64 size; FP32 uses half)
  b += 8 * 8 * nr;                     // optimizer vectors
  b += 8 * (long long)n2 * n3 * kNSum; // partial sums
  if (kinetic == 5) b += 8 * 4 * nc;   // WGC kernels
  else b += 8 * nc;                    // WT kernel
  (void)precision;
  return b;
}

extern "C" int ofd_pool_create(int n1, int n2, int n3, int capacity, int kinetic, int precision, ofd_pool** out) {
  if (!out || capacity < 1) return -1;
  ofd_pool* p = new (std::nothrow) ofd_pool();
  if (!p) return -1;
  p->n1 = n1; p->n2 = n2; p->n3 = n3; p->capacity = capacity; p->kinetic = kinetic; p->precision = precision;
  try {
    p->slots.resize(capacity);
    p->where.assign(capacity, -1);
  } catch (const std::exception&) {
    p->err = "pool slot allocation failed";
    *out = p;
    return -1;
  }
  const size_t rows = (size_t)ofd_pool::kGroups * capacity;
  if (hipHostMalloc((void**)&p->h_desc, sizeof(ReqDesc) * rows, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&p->h_sums, sizeof(double) * kNSum * rows, hipHostMallocDefault) != hipSuccess) {
    p->err = "pinned host allocation failed";
    *out = p;
    return -1;
  }
  if (shape_plan_init(&p->sp, n1, n2, n3, &p->err) != 0) {
    *out = p;
    return -1;
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(TuneKey{n1, n2, n3, kWT3, precision == 1 ? 1 : 0});
    if (it != g_tuned.end()) {
      p->sp.rows_x = it->second.rows_x;
      p->sp.kb_y = it->second.kb_y;
      p->sp.kb_z = it->second.kb_z;
      p->sp.no_plane = it->second.no_plane != 0;
    }
  }
  if (hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_phase, hipEventDisableTiming) != hipSuccess ||
      hipMalloc(&p->d_desc, sizeof(ReqDesc) * rows) != hipSuccess ||
      hipMalloc(&p->d_sums, sizeof(double) * kNSum * rows) != hipSuccess) {
    p->err = "pool allocation failed";
    *out = p;
    return -1;
  }
  *out = p;
  return 0;
}

static void server_remove_pool(ofd_server* s, ofd_pool* p);

extern "C" void ofd_pool_destroy(ofd_pool* p) {
  if (!p) return;
  if (p->server) server_remove_pool(p->server, p);
  hipStreamSynchronize(p->st);
  for (auto& s : p->slots)
    if (s.used && s.c) {
      if (s.ready) s.c->st = s.saved;
      ofd_destroy(s.c);
    }
  p->bs.free_all();
  p->bs_admit.free_all();
  graphs_clear(p->graphs);
  if (p->sp.tw_x) shape_plan_free(&p->sp);
  hipHostFree(p->h_desc);
  hipHostFree(p->h_sums);
  hipFree(p->d_desc);
  hipFree(p->d_sums);
  if (p->ev_phase) hipEventDestroy(p->ev_phase);
  hipStreamDestroy(p->st);
  delete p;
}

extern "C" const char* ofd_pool_last_error(const ofd_pool* p) { return p ? p->err.c_str() : "null pool"; }

static void server_kick(ofd_server* s);

// Admission: reserves a slot, then prepares the request (density upload,
// phi = sqrt|rho|, the electron count) on the request's own stream without
// holding the pool, so a server keeps driving the pool meanwhile; the slot
// joins the pool's phases once it is ready.
extern "C" int ofd_pool_admit(ofd_pool* p, ofd_ctx* c, const double* rho, const ofd_opt_params* prm) {
  if (!p || !c) return -1;
  if (c->sp.d.n1 != p->n1 || c->sp.d.n2 != p->n2 || c->sp.d.n3 != p->n3 || c->kinetic != p->kinetic) {
    p->err = "request grid or functional does not match the pool";
    return -1;
  }
  if (c->kind == kWGC8 && !c->have_wgc) {
    p->err = "WGC table not set";
    return -1;
  }
  if (c->nspin != 1) {
    p->err = "pools serve numSpin = 1 requests";
    return -1;
  }
  int slot = -1;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    for (int i = 0; i < p->capacity; ++i)
      if (!p->slots[i].used) { slot = i; break; }
    if (slot < 0) return -2;
    p->slots[slot] = ofd_pool::Slot();
    p->slots[slot].used = true;
    p->slots[slot].c = c;
  }
  auto& s = p->slots[slot];
  try {
    std::lock_guard<std::mutex> ak(p->admit_mu);
    HIPCHK(hipStreamSynchronize(c->st));
    alloc_opt(c);
    if (rho) HIPCHK(hipMemcpyAsync(c->rho, rho, c->nr() * sizeof(double), hipMemcpyHostToDevice, c->st));
    else vec_fill(c->rho, c->num_ele / c->vol, c->nr(), c->st);   // the uniform density
    solver_prepare(s.S, c, prm, p->bs_admit, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> lk(p->mu);
    p->err = e.what();
    s = ofd_pool::Slot();
    return -1;
  }
  {
    std::lock_guard<std::mutex> lk(p->mu);
    s.t0 = std::chrono::steady_clock::now();
    s.saved = c->st;
    c->st = p->st;   // every request of the pool runs on the pool's stream
    s.ready = true;
  }
  if (p->server) server_kick(p->server);
  return slot;
}

// Launches one phase of pool `p` on its stream (the caller holds p->mu):
// the queued vector updates of every ready request, then, for the requests
// that wait on inner products (when with_dots), one batched dot-product
// launch, and for those that wait on an evaluation (when with_evals), one
// batched launch set per (functional kind, output mode, precision); the
// results are copied to pinned memory and ev_phase is recorded after them.
// Returns false (nothing launched) when no request waits on either.
static bool launch_phase(ofd_pool* p, bool with_dots, bool with_evals) {
  const long long n = (long long)p->n1 * p->n2 * p->n3;
  p->act.clear();
  for (auto& s : p->slots)
    if (s.used && s.ready && !s.done) p->act.push_back(&s.S);
  std::fill(p->where.begin(), p->where.end(), -1);
  p->dots.clear();
  p->nsum = 0;
  if (p->act.empty()) return false;
  flush_ops(p->act, n, p->bs, p->st);
  if (with_dots) {
    for (auto* s : p->act)
      if (s->need.kind == NEED_DOTS) p->dots.push_back(s);
    if (!p->dots.empty()) launch_dots(p->dots, n, p->bs, p->st);
  }
  if (with_evals) {
    std::vector<int> groups[3][3][2];
    for (int i = 0; i < p->capacity; ++i) {
      auto& s = p->slots[i];
      if (!s.used || !s.ready || s.done) continue;
      if (s.S.need.kind == NEED_EVAL) groups[s.c->kind][s.S.need.ev.outmode][s.S.fp32 ? 1 : 0].push_back(i);
    }
    int nev = 0;
    for (int kd = 0; kd < 3; ++kd)
      for (int om = 0; om < 3; ++om)
        for (int pr = 0; pr < 2; ++pr) {
          auto& g = groups[kd][om][pr];
          if (g.empty()) continue;
          const int off = ((kd * 3 + om) * 2 + pr) * p->capacity;
          for (size_t k = 0; k < g.size(); ++k) {
            p->h_desc[off + k] = desc_for(p->slots[g[k]].c, p->slots[g[k]].S.need.ev);
            p->h_desc[off + k].fin = p->d_sums + (size_t)(off + k) * kNSum;
            p->where[g[k]] = off + (int)k;
          }
          HIPCHK(hipMemcpyAsync(p->d_desc + off, p->h_desc + off, sizeof(ReqDesc) * g.size(),
                                hipMemcpyHostToDevice, p->st));
          if (launch_eval(p->graphs, p->sp, kd, kInPhi, om, pr, p->d_desc + off, (int)g.size(),
                          p->d_sums + (size_t)off * kNSum, p->st))
            throw std::runtime_error("batched evaluation launch failed");
          HIPCHK(hipMemcpyAsync(p->h_sums + (size_t)off * kNSum, p->d_sums + (size_t)off * kNSum,
                                sizeof(double) * kNSum * g.size(), hipMemcpyDeviceToHost, p->st));
          nev += (int)g.size();
        }
    p->nsum = nev;
  }
  if (p->dots.empty() && p->nsum == 0) return false;
  HIPCHK(hipEventRecord(p->ev_phase, p->st));
  p->inflight = true;
  return true;
}

// Hands the results of the completed phase to its requests (the caller holds
// p->mu and has seen ev_phase complete). Returns how many requests finished.
static int complete_phase(ofd_pool* p) {
  if (!p->dots.empty()) deliver_dots(p->dots, p->bs);
  p->dots.clear();
  int fin = 0;
  for (int i = 0; i < p->capacity; ++i) {
    auto& s = p->slots[i];
    if (!s.used || !s.ready || s.done) continue;
    if (p->where[i] >= 0) s.S.deliver(&p->h_sums[(size_t)p->where[i] * kNSum]);
    if (s.S.need.kind == NEED_NONE) {
      s.done = true;
      ++fin;
    }
  }
  std::fill(p->where.begin(), p->where.end(), -1);
  p->nsum = 0;
  p->inflight = false;
  return fin;
}

static int count_active(const ofd_pool* p) {
  int k = 0;
  for (auto& s : p->slots)
    if (s.used && s.ready && !s.done) ++k;
  return k;
}

// Advances every active request of every pool by one evaluation. Dot rounds
// first: each round launches, on every pool's own stream, the batched vector
// updates and the inner products of all requests that need them, then waits
// once for all pools; rounds repeat until every request waits on an
// evaluation. Then the evaluations, batched per (functional kind, output
// mode, precision) within a pool, launched for all pools before one wait.
// Pools on different streams run concurrently on the device.
static int pools_step(ofd_pool* const* pools, int npool) {
  std::vector<std::unique_lock<std::mutex>> locks;
  for (int ip = 0; ip < npool; ++ip) locks.emplace_back(pools[ip]->mu);
  for (int ip = 0; ip < npool; ++ip)
    if (pools[ip]->server) throw std::runtime_error("the pool is driven by a server");
  for (int guard = 0; guard < 64; ++guard) {
    bool any = false;
    for (int ip = 0; ip < npool; ++ip) any = launch_phase(pools[ip], true, false) || any;
    if (!any) break;
    for (int ip = 0; ip < npool; ++ip)
      if (pools[ip]->inflight) {
        HIPCHK(hipEventSynchronize(pools[ip]->ev_phase));
        complete_phase(pools[ip]);
      }
  }
  for (int ip = 0; ip < npool; ++ip) launch_phase(pools[ip], false, true);
  int still = 0;
  for (int ip = 0; ip < npool; ++ip) {
    ofd_pool* p = pools[ip];
    if (p->inflight) {
      HIPCHK(hipEventSynchronize(p->ev_phase));
      complete_phase(p);
    }
    still += count_active(p);
  }
  return still;
}

extern "C" int ofd_pool_step(ofd_pool* p) {
  if (!p) return -1;
  try {
    return pools_step(&p, 1);
  } catch (const std::exception& e) {
    p->err = e.what();
    return -1;
  }
}

extern "C" int ofd_pools_step(ofd_pool* const* pools, int npool) {
  if (!pools || npool < 1) return -1;
  for (int i = 0; i < npool; ++i)
    if (!pools[i]) return -1;
  try {
    return pools_step(pools, npool);
  } catch (const std::exception& e) {
    for (int i = 0; i < npool; ++i) pools[i]->err = e.what();
    return -1;
  }
}

extern "C" int ofd_pool_finished(const ofd_pool* p, int* slots, int max_slots) {
  if (!p || (!slots && max_slots > 0)) return -1;
  std::lock_guard<std::mutex> lk(p->mu);
  int k = 0;
  for (int i = 0; i < p->capacity; ++i)
    if (p->slots[i].used && p->slots[i].done) {
      if (k < max_slots) slots[k] = i;
      ++k;
    }
  return k;
}

extern "C" int ofd_pool_collect(ofd_pool* p, int slot, double* rho_out, double* energy9, int* outcome,
                                ofd_opt_stats* stats, ofd_ctx** ctx_out) {
  if (!p || slot < 0 || slot >= p->capacity) return -1;
  std::unique_lock<std::mutex> lk(p->mu);
  auto& s = p->slots[slot];
  if (!s.used) return -1;
  if (!s.done) return 0;
  ofd_ctx* c = s.c;
  // Without rho_out the final density is not formed: the request's last
  // vector updates are dropped and nothing waits on the pool's stream (no
  // launched work of a finished request is still pending).
  if (rho_out) {
    try {
      std::vector<NtnSolver*> one{&s.S};
      flush_ops(one, c->nr(), p->bs, p->st);
      const long long n = c->nr();
      vec_square(c->rho, c->phi, n, p->st);
      HIPCHK(hipMemcpyAsync(rho_out, c->rho, n * sizeof(double), hipMemcpyDeviceToHost, p->st));
      HIPCHK(hipStreamSynchronize(p->st));
    } catch (const std::exception& e) {
      p->err = e.what();
      return -1;
    }
  }
  if (energy9) std::memcpy(energy9, s.S.energy, sizeof(s.S.energy));
  if (outcome) *outcome = s.S.err.empty() ? s.S.outcome : 1;
  if (stats) {
    *stats = s.S.stats;
    stats->final_pot_norm = s.S.pot_norm;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - s.t0).count();
  }
  c->st = s.saved;
  if (!s.S.err.empty()) c->err = s.S.err;
  s = ofd_pool::Slot();
  lk.unlock();
  if (ctx_out) *ctx_out = c;
  else ofd_destroy(c);
  return 1;
}

// ====================================================================== server
// A progress thread that drives its pools without a global step: each pool
// has at most one phase in flight on its own stream; when the phase's event
// has fired, the thread hands the results to the requests and at once
// launches the pool's next phase. Pools of different grids thus never wait
// for each other, and the caller's admissions and collections (which take a
// pool's mutex only briefly) run while the device works.
struct ofd_server {
  std::mutex mu;                  // pools, counters, error
  std::mutex sweep_mu;            // held by the thread for one pass over the pools
  std::condition_variable cv_done;
  std::condition_variable cv_work;
  std::vector<ofd_pool*> pools;
  std::thread th;
  bool stop = false;
  bool kick = false;
  long long finished = 0, reported = 0;
  std::string err;
};

static void server_kick(ofd_server* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->kick = true;
  }
  s->cv_work.notify_one();
}

static void server_loop(ofd_server* s) {
  std::vector<ofd_pool*> ps;
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stop) return;
      ps = s->pools;
    }
    bool progressed = false, busy = false;
    int newly = 0;
    std::string failure;
    {
      std::lock_guard<std::mutex> sw(s->sweep_mu);
      for (ofd_pool* p : ps) {
        std::unique_lock<std::mutex> lk(p->mu, std::try_to_lock);
        if (!lk.owns_lock()) {
          busy = true;
          continue;
        }
        if (p->failed) continue;
        try {
          if (p->inflight) {
            const hipError_t q = hipEventQuery(p->ev_phase);
            if (q == hipErrorNotReady) {
              busy = true;
              continue;
            }
            if (q != hipSuccess) throw std::runtime_error(std::string("phase failed: ") + hipGetErrorString(q));
            newly += complete_phase(p);
            progressed = true;
          }
          if (launch_phase(p, true, true)) {
            busy = true;
            progressed = true;
          }
        } catch (const std::exception& e) {
          p->failed = true;
          p->err = e.what();
          failure = e.what();
        }
      }
    }
    if (newly > 0 || !failure.empty()) {
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->finished += newly;
        if (!failure.empty() && s->err.empty()) s->err = failure;
      }
      s->cv_done.notify_all();
    }
    if (progressed) continue;
    if (busy) {
      std::this_thread::yield();
    } else {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait_for(lk, std::chrono::milliseconds(1), [&] { return s->stop || s->kick; });
      s->kick = false;
    }
  }
}

static void server_remove_pool(ofd_server* s, ofd_pool* p) {
  std::lock_guard<std::mutex> sw(s->sweep_mu);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->pools.erase(std::remove(s->pools.begin(), s->pools.end(), p), s->pools.end());
  }
  std::lock_guard<std::mutex> pk(p->mu);
  if (p->inflight) {
    if (hipEventSynchronize(p->ev_phase) == hipSuccess) {
      try {
        complete_phase(p);
      } catch (const std::exception& e) {
        p->err = e.what();
        p->failed = true;
      }
    }
    p->inflight = false;
  }
  p->server = nullptr;
}

extern "C" int ofd_server_create(ofd_server** out) {
  if (!out) return -1;
  ofd_server* s = new ofd_server();
  try {
    s->th = std::thread(server_loop, s);
  } catch (const std::exception&) {
    delete s;
    *out = nullptr;
    return -1;
  }
  *out = s;
  return 0;
}

extern "C" int ofd_server_add_pool(ofd_server* s, ofd_pool* p) {
  if (!s || !p) return -1;
  {
    std::lock_guard<std::mutex> pk(p->mu);
    if (p->server) return p->server == s ? 0 : -1;
    p->server = s;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->pools.push_back(p);
  }
  server_kick(s);
  return 0;
}

extern "C" int ofd_server_wait(ofd_server* s, double timeout_s) {
  if (!s) return -1;
  std::unique_lock<std::mutex> lk(s->mu);
  const auto dt = std::chrono::duration<double>(std::max(0.0, timeout_s));
  s->cv_done.wait_for(lk, std::chrono::duration_cast<std::chrono::microseconds>(dt),
                      [&] { return s->finished != s->reported || !s->err.empty(); });
  if (!s->err.empty()) return -1;
  const long long k = s->finished - s->reported;
  s->reported = s->finished;
  return (int)k;
}

extern "C" const char* ofd_server_last_error(const ofd_server* s) { return s ? s->err.c_str() : "null server"; }

extern "C" void ofd_server_destroy(ofd_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  std::vector<ofd_pool*> ps;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    ps = s->pools;
  }
  for (ofd_pool* p : ps) server_remove_pool(s, p);
  delete s;
}
