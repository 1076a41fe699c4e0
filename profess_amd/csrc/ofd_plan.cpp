// Shape plans: radix factorization, twiddle tables and launch tiles for one
// grid shape n1 x n2 x n3 (n1 even, every n a product of 2, 3, 5, 7).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ofd_internal.h"

namespace ofd {

// Stage radices: powers of two as 16s (8*4 for 2^5), then 3, 5, 7.
bool make_plan1d(int n, Plan1D* p) {
  std::memset(p, 0, sizeof(*p));
  p->n = n;
  if (n < 1) return false;
  int m = n, e = 0;
  while (m % 2 == 0) { m /= 2; ++e; }
  std::vector<int> rad;
  while (e >= 4 && e != 5) { rad.push_back(16); e -= 4; }
  if (e == 5) { rad.push_back(8); rad.push_back(4); e = 0; }
  if (e == 3) rad.push_back(8);
  if (e == 2) rad.push_back(4);
  if (e == 1) rad.push_back(2);
  for (int f : {3, 5, 7}) {
    while (m % f == 0) { rad.push_back(f); m /= f; }
  }
  if (m != 1 || (int)rad.size() > kMaxStages) return false;
  p->nst = (int)rad.size();
  for (int i = 0; i < p->nst; ++i) p->rad[i] = rad[i];
  return true;
}

// ------------------------------------------------------------ device memory cache
// Blocks released by dev_free wait in per-size free lists and are handed out
// again by dev_alloc, so a server that creates and destroys a context per
// request reaches a steady state with no hipMalloc or hipFree (hipFree waits
// for the whole device). Sizes round up to 4 KiB; requests of one grid shape
// ask for identical sizes. Free blocks beyond kCacheCap bytes go back to HIP.
namespace {
constexpr size_t kCacheCap = 48ull << 30;
std::mutex g_mem_mu;
std::unordered_map<void*, size_t> g_live;
std::multimap<size_t, void*> g_free;
size_t g_free_bytes = 0;
}  // namespace

void* dev_alloc(size_t bytes) {
  const size_t b = ((bytes ? bytes : 1) + 4095) / 4096 * 4096;
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_free.find(b);
    if (it != g_free.end()) {
      void* p = it->second;
      g_free.erase(it);
      g_free_bytes -= b;
      g_live[p] = b;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, b) != hipSuccess) {
    (void)hipGetLastError();
    // release the cached blocks and retry once
    std::multimap<size_t, void*> drop;
    {
      std::lock_guard<std::mutex> lk(g_mem_mu);
      drop.swap(g_free);
      g_free_bytes = 0;
    }
    for (auto& kv : drop) hipFree(kv.second);
    p = nullptr;
    if (hipMalloc(&p, b) != hipSuccess) return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_live[p] = b;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_live.find(p);
    if (it != g_live.end()) {
      const size_t b = it->second;
      g_live.erase(it);
      if (g_free_bytes + b <= kCacheCap) {
        g_free.emplace(b, p);
        g_free_bytes += b;
        return;
      }
    }
  }
  hipFree(p);
}

// ------------------------------------------------------------ twiddle tables
// exp(-2 pi i j / n), j < count, in FP64 and FP32, uploaded once per (n, count)
// and shared by every plan of the process (never freed).
namespace {
struct TwTables {
  double2* d64 = nullptr;
  float2* d32 = nullptr;
};
std::mutex g_tw_mu;
std::map<std::pair<int, int>, TwTables> g_tw;
}  // namespace

static int upload_tw(int n, int count, double2** d64, float2** d32) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find({n, count});
  if (it != g_tw.end()) {
    *d64 = it->second.d64;
    *d32 = it->second.d32;
    return 0;
  }
  std::vector<double2> h(count);
  std::vector<float2> hf(count);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int j = 0; j < count; ++j) {
    long double a = -two_pi * (long double)j / (long double)n;
    h[j].x = (double)cosl(a);
    h[j].y = (double)sinl(a);
    hf[j].x = (float)cosl(a);
    hf[j].y = (float)sinl(a);
  }
  TwTables t;
  if (hipMalloc(&t.d64, sizeof(double2) * count) != hipSuccess) return -1;
  if (hipMalloc(&t.d32, sizeof(float2) * count) != hipSuccess) return -1;
  // copies on a stream of their own: the legacy null stream is refused while
  // another thread captures a graph
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return -1;
  const bool ok = hipMemcpyAsync(t.d64, h.data(), sizeof(double2) * count, hipMemcpyHostToDevice, st) == hipSuccess &&
                  hipMemcpyAsync(t.d32, hf.data(), sizeof(float2) * count, hipMemcpyHostToDevice, st) == hipSuccess &&
                  hipStreamSynchronize(st) == hipSuccess;
  hipStreamDestroy(st);
  if (!ok) return -1;
  g_tw[{n, count}] = t;
  *d64 = t.d64;
  *d32 = t.d32;
  return 0;
}

static int lds_cap_bytes() {
  int dev = 0, v = 0;
  hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    v = 65536;
  // leave room for the static reduction scratch
  return v - 2048;
}

int shape_plan_init(ShapePlan* sp, int n1, int n2, int n3, std::string* err) {
  *sp = ShapePlan();
  if (n1 < 2 || n2 < 1 || n3 < 1 || (n1 & 1)) {
    *err = "grid: n1 must be even and >= 2 (half-length complex R2C along x)";
    return -1;
  }
  GridDims& d = sp->d;
  d.n1 = n1; d.n2 = n2; d.n3 = n3;
  d.M = n1 / 2;
  d.K1 = n1 / 2 + 1;
  d.nr = (long long)n1 * n2 * n3;
  d.nc = (long long)d.K1 * n2 * n3;
  if (!make_plan1d(d.M, &sp->px) || !make_plan1d(n2, &sp->py) || !make_plan1d(n3, &sp->pz)) {
    *err = "grid: every dimension (and n1/2) must factor into 2, 3, 5, 7";
    return -1;
  }
  if (upload_tw(d.M, d.M, &sp->tw_x, &sp->tw_x32) || upload_tw(n2, n2, &sp->tw_y, &sp->tw_y32) ||
      upload_tw(n3, n3, &sp->tw_z, &sp->tw_z32) || upload_tw(n1, d.M + 1, &sp->tw_s, &sp->tw_s32)) {
    *err = "grid: twiddle upload failed";
    return -1;
  }
  sp->lds_cap = lds_cap_bytes();
  sp->lds_target = sp->lds_cap < 65536 ? sp->lds_cap : 65536;
  if (const char* e = std::getenv("OFD_LDS_TARGET")) {
    const int v = std::atoi(e);
    if (v >= 4096) sp->lds_target = v < sp->lds_cap ? v : sp->lds_cap;
  }
  if (const char* e = std::getenv("OFD_GENERIC")) sp->force_generic = std::atoi(e) != 0;
  if (const char* e = std::getenv("OFD_NO_PLANE")) sp->no_plane = std::atoi(e) != 0;
  if (const char* e = std::getenv("OFD_PLANE_SPLIT")) sp->plane_split = std::atoi(e);
  if (const char* e = std::getenv("OFD_KB_Y")) sp->kb_y = std::atoi(e);
  if (const char* e = std::getenv("OFD_KB_Z")) sp->kb_z = std::atoi(e);
  const size_t cb = sizeof(double2);
  if (2ull * 8 * (d.M + 1) * cb > (size_t)sp->lds_cap || 2ull * 8 * (n3 + 1) * cb > (size_t)sp->lds_cap ||
      2ull * (n2 + 1) * cb > (size_t)sp->lds_cap) {
    *err = "grid: a line does not fit in LDS";
    return -1;
  }
  return 0;
}

static int clampi(long long v, int lo, int hi) { return (int)(v < lo ? lo : (v > hi ? hi : v)); }

Tiles tiles_for(const ShapePlan& sp, int nf, int cbytes) {
  const GridDims& d = sp.d;
  Tiles t;
  const long long tgt = sp.lds_target;
  t.rows = sp.rows_x > 0 ? sp.rows_x : clampi(tgt / (2ll * nf * (d.M + 1) * cbytes), 1, 16);
  t.kby = sp.kb_y > 0 ? sp.kb_y : clampi(tgt / (2ll * (d.n2 + 1) * cbytes), 1, 16);
  t.kbz = sp.kb_z > 0 ? sp.kb_z : clampi(tgt / (2ll * nf * (d.n3 + 1) * cbytes), 1, 8);
  // never exceed the hard limit
  while (t.rows > 1 && 2ll * nf * t.rows * (d.M + 1) * cbytes > sp.lds_cap) --t.rows;
  while (t.kby > 1 && 2ll * t.kby * (d.n2 + 1) * cbytes > sp.lds_cap) --t.kby;
  while (t.kbz > 1 && 2ll * nf * t.kbz * (d.n3 + 1) * cbytes > sp.lds_cap) --t.kbz;
  const long long nrows = (long long)d.n2 * d.n3;
  t.ntile = (int)((nrows + t.rows - 1) / t.rows);
  return t;
}

// The twiddle tables are shared process-wide (upload_tw): nothing to free.
void shape_plan_free(ShapePlan* sp) { *sp = ShapePlan(); }

}  // namespace ofd
