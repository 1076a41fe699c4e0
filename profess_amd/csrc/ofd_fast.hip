// Dispatch of the register-resident fast path (two-stage line lengths,
// Wang-Teter and WGC functionals). The kernels live in ofd_fast_impl.h; each
// (precision, functional) pair is instantiated in its own translation unit
// (ofd_fast_{d,f}{3,4,5,7,8}.hip: the number is the kind's field count).
#include "ofd_fast_impl.h"

namespace ofd {

// instantiated in their own translation units (not again here)
extern template int fast_eval_k<double, kWT3>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<double, kWT4>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<float, kWT3>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<float, kWT4>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<double, kWGC8>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<float, kWGC8>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<double, kWT3S>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<float, kWT3S>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<double, kWT4S>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);
extern template int fast_eval_k<float, kWT4S>(const ShapePlan&, int, int, const ReqDesc*, int, hipStream_t);

namespace fast {
// One workgroup per (slot, request): fin[slot] = sum over the workgroups'
// partials, fixed order.
__global__ void __launch_bounds__(256) k_fin(const ReqDesc* __restrict__ descs, int nwg) {
  const ReqDesc& D = descs[blockIdx.y];
  const int k = blockIdx.x;
  gptr<const double> src = glb(D.sums) + (size_t)k * nwg;
  double acc[1] = {0.0};
  // eight independent loads in flight per thread, summed in a fixed order
  for (int base = 0; base < nwg; base += 8 * 256) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int w = base + threadIdx.x + 256 * j;
      v[j] = w < nwg ? src[w] : 0.0;
    }
    acc[0] += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  }
  block_sum_strided<1>(acc, glb(D.fin) + k, 1);
}

}  // namespace fast

bool fast_path_ok(const ShapePlan& sp, int kind) { return fast_path_ok_impl(sp, kind); }


int fast_eval(const ShapePlan& sp, int kind, int inmode, int outmode, int fp32, const ReqDesc* dd, int nb,
              hipStream_t st) {
  if (kind == kWT3)
    return fp32 ? fast_eval_k<float, kWT3>(sp, inmode, outmode, dd, nb, st)
                : fast_eval_k<double, kWT3>(sp, inmode, outmode, dd, nb, st);
  if (kind == kWGC8)
    return fp32 ? fast_eval_k<float, kWGC8>(sp, inmode, outmode, dd, nb, st)
                : fast_eval_k<double, kWGC8>(sp, inmode, outmode, dd, nb, st);
  if (kind == kWT3S)
    return fp32 ? fast_eval_k<float, kWT3S>(sp, inmode, outmode, dd, nb, st)
                : fast_eval_k<double, kWT3S>(sp, inmode, outmode, dd, nb, st);
  if (kind == kWT4S)
    return fp32 ? fast_eval_k<float, kWT4S>(sp, inmode, outmode, dd, nb, st)
                : fast_eval_k<double, kWT4S>(sp, inmode, outmode, dd, nb, st);
  return fp32 ? fast_eval_k<float, kWT4>(sp, inmode, outmode, dd, nb, st)
              : fast_eval_k<double, kWT4>(sp, inmode, outmode, dd, nb, st);
}

std::vector<int> fast_tile_candidates(const ShapePlan& sp, int kind, int fp32, int which) {
  const int n = which == 1 ? sp.d.n2 : sp.d.n3;
  const int T = len_T(n), w = 64 / T;
  const int cb = fp32 ? (int)sizeof(float2) : (int)sizeof(double2);
  std::vector<int> out;
  for (int kb = w; kb * T <= 256; kb += w)
    if (kb * col_line_bytes(n, cb, which) <= sp.lds_cap) out.push_back(kb);
  (void)kind;
  return out;
}

bool fast_plane_available(const ShapePlan& sp, int kind, int fp32) {
  (void)kind;
  const int cb = fp32 ? (int)sizeof(float2) : (int)sizeof(double2);
  auto fits = [&](int M, int N2) {
    const int S = std::max(M + M / p2_R1(M), M + 1);
    return (long long)(N2 + N2 / p2_R1(N2)) * S * cb <= sp.lds_cap;
  };
  if (sp.d.M == 64 && sp.d.n2 == 128) return fits(64, 128);
  if (sp.d.M == 32 && sp.d.n2 == 64) return fits(32, 64);
  if (sp.d.M == 16 && sp.d.n2 == 32) return fits(16, 32);
  if (sp.d.M == 48 && sp.d.n2 == 96) return fits(48, 96);
  if (sp.d.M == 40 && sp.d.n2 == 80) return fits(40, 80);
  if (sp.d.M == 24 && sp.d.n2 == 48) return fits(24, 48);
  return false;
}

// Sums the x-inverse pass's per-workgroup partials into each request's fin.
int fast_finalize(const ShapePlan& sp, int kind, int, const ReqDesc* dd, int nb, double*, hipStream_t st) {
  const int rpb = fast::inv_rpb(func_nf(kind), len_T(sp.d.M));
  const long long nrows = (long long)sp.d.n2 * sp.d.n3;
  const int nwg = (int)((nrows + rpb - 1) / rpb);
  hipLaunchKernelGGL(fast::k_fin, dim3(kNSum, nb), dim3(256), 0, st, dd, nwg);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ofd
