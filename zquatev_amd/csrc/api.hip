// C ABI and the ts::zquatev drop-in.
//
// The reference's driver (zquatev.cc:42-100) works in host memory: it splits
// the left half into D0 = A and D1 = B, sets Q = I in the right half, runs the
// panel loop, calls zhbev on the band-packed tridiagonal, forms U = Q0 Z and
// V = Q1 Z and fills the Kramers half.  Here the same call is one upload of
// the left half, a device-resident solve and one download:
//   load (layout.hip) -> tridiagonalise (reduce.hip) -> |e| -> divide & conquer
//   (stedc.hip) -> back-transform (backtr.hip) -> Kramers store (layout.hip).
// Workspace lives in the context and is reused across calls.
#include "zq_internal.h"
#include "zq.h"
#include "zquatev.h"

#include <complex>
#include <cstring>
#include <mutex>
#include <vector>

namespace zq {
// event-pair records of the instrumented launches of the current solve
struct ProfState {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<int> cls;
  std::vector<double> units;
  size_t used = 0;
};
static thread_local ProfState* g_prof = nullptr;

}  // namespace zq

struct zq_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  int nb_override = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* dD = nullptr;       // staging copy of D for the host-memory entry points
  size_t dD_bytes = 0;
  double* deig = nullptr;
  size_t deig_bytes = 0;
  hipEvent_t ev[6] = {};
  bool have_times = false;
  int* info_dev = nullptr;  // leaf-solver status of the last vectors solve
  zq::ProfState prof;
};


namespace zq {
static thread_local int g_pending = -1;   // slot armed by the innermost ProfScope

ProfScope::ProfScope(int c, hipStream_t st, double u) : cls(c), s(st), units(u), slot(-1) {
  ProfState* p = g_prof;
  if (!p || !p->on) return;
  if (p->pool.size() < 2 * (p->used + 1)) {
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      p->pool.push_back(e);
    }
  }
  slot = (int)p->used++;
  if (p->cls.size() < p->used) { p->cls.resize(p->used); p->units.resize(p->used); }
  p->cls[slot] = c;
  p->units[slot] = u;
  g_pending = slot;
}
ProfScope::~ProfScope() {
  if (slot < 0) return;
  if (g_pending == slot) {
    // no kernel was launched in the scope: an empty interval
    g_pending = -1;
    (void)hipEventRecord(g_prof->pool[2 * slot], s);
    (void)hipEventRecord(g_prof->pool[2 * slot + 1], s);
  }
}
void prof_take(hipEvent_t* e0, hipEvent_t* e1) {
  if (g_pending < 0 || !g_prof) return;
  *e0 = g_prof->pool[2 * g_pending];
  *e1 = g_prof->pool[2 * g_pending + 1];
  g_pending = -1;
}
}  // namespace zq

namespace {

using zq::quat;

constexpr int kNSplit = 8;

struct Plan {
  quat* M;
  quat* X;
  double* d;
  quat* e;
  double* ea;
  zq::ReduceWork rw;
  zq::StedcWork sw;
  zq::BacktrWork bw;
};

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

struct Bump {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

// Lays out (or sizes, when base is null) the workspace of one solve.
size_t layout_plan(int n, int nb, bool vectors, char* base, Plan* p) {
  Bump b{base};
  const size_t nn = (size_t)n * n;
  const int nmax = n > 1 ? n : 2;
  p->M = b.take<quat>(nn);
  p->d = b.take<double>(nmax);
  p->e = b.take<quat>(nmax);
  p->ea = b.take<double>(nmax);
  p->rw.W = b.take<quat>((size_t)n * nb);
  p->rw.ucur = b.take<quat>(nmax);
  p->rw.Prow = b.take<quat>((size_t)cdiv(nmax, 32) * n);
  p->rw.Pcol = b.take<quat>((size_t)cdiv(nmax, 64) * n);
  p->rw.TS = b.take<quat>((size_t)2 * nb);
  p->rw.KP = b.take<double>(cdiv(nmax, 16) + 16);
  p->rw.part = b.take<double>(cdiv(nmax, 16) + 16);
  p->rw.scal = b.take<double>(16);
  p->rw.Tf = b.take<quat>((size_t)cdiv(nmax, nb) * nb * nb);
  p->rw.Gm = b.take<quat>((size_t)cdiv(nmax, nb) * nb * nb);
  zq::StedcWork& s = p->sw;
  s.lam = b.take<double>(nmax);
  s.es = b.take<double>(nmax);
  s.e2 = b.take<double>(nmax);
  s.scal = b.take<double>(16);
  if (vectors) {
    s.Z = b.take<double>(nn);
    s.Qtmp = b.take<double>(nn);
    s.U = b.take<double>(nn);
    s.dk = b.take<double>(nmax);
    s.zk = b.take<double>(nmax);
    s.dkk = b.take<double>(nmax);
    s.zkk = b.take<double>(nmax);
    s.dval = b.take<double>(nmax);
    s.tau = b.take<double>(nmax);
    s.zhat = b.take<double>(nmax);
    s.rho = b.take<double>(nmax);
    s.rot = b.take<double>((size_t)4 * nmax);
    s.src = b.take<int>(nmax);
    s.keep = b.take<int>(nmax);
    s.dfl = b.take<int>(nmax);
    s.org = b.take<int>(nmax);
    s.posK = b.take<int>(nmax);
    s.posD = b.take<int>(nmax);
    s.nrot = b.take<int>(nmax);
    s.kcnt = b.take<int>(nmax);
    s.info = b.take<int>(16);
    s.int_slots = zq::stedc_int_slots(n);
    s.ints = b.take<int>(s.int_slots);
    p->X = b.take<quat>(nn);
    p->bw.phi = b.take<quat>(nmax);
    p->bw.nsplit = kNSplit;
    p->bw.ypart_slots = (size_t)kNSplit * nb * n;
    p->bw.Ypart = b.take<quat>(p->bw.ypart_slots);
    p->bw.Y = b.take<quat>((size_t)nb * n);
    p->bw.Y2 = b.take<quat>((size_t)nb * n);
  } else {
    s.Z = s.Qtmp = s.U = nullptr;
    p->X = nullptr;
  }
  return b.off + 256;
}

int ensure(void*& ptr, size_t& have, size_t need) {
  if (have >= need && ptr) return ZQ_OK;
  if (ptr) {
    if (hipFree(ptr) != hipSuccess) return ZQ_ERR_HIP;
    ptr = nullptr;
    have = 0;
  }
  hipError_t e = hipMalloc(&ptr, need);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? ZQ_ERR_NOMEM : ZQ_ERR_HIP;
  }
  have = need;
  return ZQ_OK;
}

int panel_width(const zq_context* ctx, int n) {
  int nb = ctx->nb_override > 0 ? ctx->nb_override : zq_tuned_panel_width(n);
  if (nb > 128) nb = 128;
  if (nb < 1) nb = 1;
  return nb;
}

int make_plan(zq_context* ctx, int n, int nb, bool vectors, Plan* p) {
  const size_t need = layout_plan(n, nb, vectors, nullptr, p);
  int rc = ensure(ctx->ws, ctx->ws_bytes, need);
  if (rc) return rc;
  layout_plan(n, nb, vectors, static_cast<char*>(ctx->ws), p);
  return ZQ_OK;
}

int with_device(zq_context* ctx) {
  return hipSetDevice(ctx->device) == hipSuccess ? ZQ_OK : ZQ_ERR_HIP;
}

// The device solve: dD (ld2 complex) -> dD, deig.
int solve_device(zq_context* ctx, int n, double* dD, int ld2, double* deig, bool vectors) {
  hipStream_t s = ctx->stream;
  ctx->prof.used = 0;
  zq::g_prof = &ctx->prof;
  struct Reset { ~Reset() { zq::g_prof = nullptr; } } reset_prof;
  const int nb = panel_width(ctx, n);
  Plan p;
  int rc = make_plan(ctx, n, nb, vectors, &p);
  if (rc) return rc;
  ZQ_HIP_CHECK(hipEventRecord(ctx->ev[0], s));
  if ((rc = zq::load_left_half(n, dD, ld2, p.M, n, s))) return rc;
  ZQ_HIP_CHECK(hipEventRecord(ctx->ev[1], s));
  if ((rc = zq::reduce(p.M, n, n, nb, p.d, p.e, p.rw, s))) return rc;
  if ((rc = zq::offdiag_abs(n, p.e, p.ea, s))) return rc;
  ZQ_HIP_CHECK(hipEventRecord(ctx->ev[2], s));
  if (vectors) ZQ_HIP_CHECK(hipMemsetAsync(p.sw.info, 0, sizeof(int), s));
  ctx->info_dev = vectors ? p.sw.info : nullptr;
  if ((rc = zq::stedc(n, p.d, p.ea, deig, p.sw, vectors, s))) return rc;
  ZQ_HIP_CHECK(hipEventRecord(ctx->ev[3], s));
  if (vectors) {
    if ((rc = zq::backtransform(n, nb, p.M, n, p.rw.Tf, p.e, p.sw.Z, n, p.X, n, p.bw, s))) return rc;
    ZQ_HIP_CHECK(hipEventRecord(ctx->ev[4], s));
    if ((rc = zq::store_vectors(n, p.X, n, dD, ld2, s))) return rc;
  } else {
    ZQ_HIP_CHECK(hipEventRecord(ctx->ev[4], s));
  }
  ZQ_HIP_CHECK(hipEventRecord(ctx->ev[5], s));
  ctx->have_times = true;
  return ZQ_OK;
}

int check_args(int n2, const void* D, int nld2, const void* eig) {
  if (n2 < 0 || (n2 & 1)) return -1;
  if (n2 > 0 && !D) return -2;
  if (nld2 < (n2 > 1 ? n2 : 1)) return -3;
  if (n2 > 0 && !eig) return -4;
  return 0;
}

std::mutex g_mu;
zq_context* g_ctx = nullptr;

}  // namespace

extern "C" {

int zq_version(void) { return 100; }

const char* zq_error_string(int code) {
  switch (code) {
    case ZQ_OK: return "success";
    case ZQ_ERR_HIP: return "HIP runtime error";
    case ZQ_ERR_NOMEM: return "device memory exhausted";
    case ZQ_ERR_ARG: return "unsupported internal configuration";
    case -1: return "argument 1 (n2) is negative or odd";
    case -2: return "argument 2 (D) is null";
    case -3: return "argument 3 (nld2) is smaller than n2";
    case -4: return "argument 4 (eig) is null";
    default: return code > 0 ? "tridiagonal eigensolver did not converge" : "invalid argument";
  }
}

int zq_tuned_panel_width(int n) {
  if (n <= 256) return 32;
  return 64;
}

int zq_create(zq_context** out, int device) {
  if (!out) return -1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return ZQ_ERR_HIP;
  zq_context* c = new zq_context();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ZQ_ERR_HIP;
  }
  c->stream = c->own_stream;
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return ZQ_ERR_HIP;
    }
  }
  *out = c;
  return ZQ_OK;
}

int zq_destroy(zq_context* c) {
  if (!c) return ZQ_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->ws) (void)hipFree(c->ws);
  if (c->dD) (void)hipFree(c->dD);
  if (c->deig) (void)hipFree(c->deig);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->prof.pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return ZQ_OK;
}

int zq_set_stream(zq_context* c, void* stream) {
  if (!c) return -1;
  c->stream = static_cast<hipStream_t>(stream);
  return ZQ_OK;
}

int zq_use_own_stream(zq_context* c) {
  if (!c) return -1;
  c->stream = c->own_stream;
  return ZQ_OK;
}

int zq_set_panel_width(zq_context* c, int nb) {
  if (!c) return -1;
  if (nb < 0 || nb > 128) return -2;
  c->nb_override = nb;
  return ZQ_OK;
}

int zq_zquatev_device(zq_context* c, int n2, double* dD, int nld2, double* deig, int jobz) {
  if (!c) return ZQ_ERR_ARG;
  int rc = check_args(n2, dD, nld2, deig);
  if (rc) return rc;
  if (jobz != 'V' && jobz != 'v' && jobz != 'N' && jobz != 'n') return -5;
  if (n2 == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  return solve_device(c, n2 / 2, dD, nld2, deig, jobz == 'V' || jobz == 'v');
}

int zq_zquatev_ctx(zq_context* c, int n2, double* D, int nld2, double* eig) {
  if (!c) return ZQ_ERR_ARG;
  int rc = check_args(n2, D, nld2, eig);
  if (rc) return rc;
  if (n2 == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  const int n = n2 / 2;
  const size_t colb = (size_t)n2 * 16;
  if ((rc = ensure(c->dD, c->dD_bytes, colb * n2))) return rc;
  if ((rc = ensure(reinterpret_cast<void*&>(c->deig), c->deig_bytes, (size_t)n * sizeof(double)))) return rc;
  double* dD = static_cast<double*>(c->dD);
  hipStream_t s = c->stream;
  ZQ_HIP_CHECK(hipMemcpy2DAsync(dD, colb, D, (size_t)nld2 * 16, colb, n, hipMemcpyHostToDevice, s));
  if ((rc = solve_device(c, n, dD, n2, c->deig, true))) return rc;
  ZQ_HIP_CHECK(hipMemcpy2DAsync(D, (size_t)nld2 * 16, dD, colb, colb, n2, hipMemcpyDeviceToHost, s));
  ZQ_HIP_CHECK(hipMemcpyAsync(eig, c->deig, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  int info = 0;
  if (c->info_dev) ZQ_HIP_CHECK(hipMemcpyAsync(&info, c->info_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  ZQ_HIP_CHECK(hipStreamSynchronize(s));
  return info;
}

int zq_zquatev(int n2, double* D, int nld2, double* eig) {
  int rc = check_args(n2, D, nld2, eig);
  if (rc || n2 == 0) return rc;
  // one solve at a time on the shared context
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx) {
    int rc = zq_create(&g_ctx, 0);
    if (rc) return rc;
  }
  return zq_zquatev_ctx(g_ctx, n2, D, nld2, eig);
}

int zq_tridiagonalize(zq_context* c, int n2, const double* dD, int nld2, double* dd, double* de) {
  if (!c) return ZQ_ERR_ARG;
  int rc = check_args(n2, dD, nld2, dd);
  if (rc) return rc;
  if (n2 == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  const int n = n2 / 2;
  const int nb = panel_width(c, n);
  Plan p;
  if ((rc = make_plan(c, n, nb, false, &p))) return rc;
  hipStream_t s = c->stream;
  if ((rc = zq::load_left_half(n, dD, nld2, p.M, n, s))) return rc;
  if ((rc = zq::reduce(p.M, n, n, nb, p.d, p.e, p.rw, s))) return rc;
  if ((rc = zq::offdiag_abs(n, p.e, p.ea, s))) return rc;
  ZQ_HIP_CHECK(hipMemcpyAsync(dd, p.d, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (n > 1 && de)
    ZQ_HIP_CHECK(hipMemcpyAsync(de, p.ea, (size_t)(n - 1) * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ZQ_OK;
}

int zq_tridiag_eig(zq_context* c, int n, const double* dd, const double* de, double* dw, double* dZ) {
  if (!c) return ZQ_ERR_ARG;
  if (n < 0) return -2;
  if (n > 0 && (!dd || !dw)) return -3;
  if (n == 0) return ZQ_OK;
  int rc;
  if ((rc = with_device(c))) return rc;
  Plan p;
  const bool vec = dZ != nullptr;
  if ((rc = make_plan(c, n, 32, vec, &p))) return rc;
  hipStream_t s = c->stream;
  if (vec) ZQ_HIP_CHECK(hipMemsetAsync(p.sw.info, 0, sizeof(int), s));
  if ((rc = zq::stedc(n, dd, de, dw, p.sw, vec, s))) return rc;
  if (vec) ZQ_HIP_CHECK(hipMemcpyAsync(dZ, p.sw.Z, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ZQ_OK;
}

int zq_last_timings(zq_context* c, double* ms, int count) {
  if (!c || !ms) return 0;
  if (!c->have_times) return 0;
  if (hipEventSynchronize(c->ev[5]) != hipSuccess) return 0;
  float t[5] = {0, 0, 0, 0, 0};
  (void)hipEventElapsedTime(&t[0], c->ev[1], c->ev[2]);
  (void)hipEventElapsedTime(&t[1], c->ev[2], c->ev[3]);
  (void)hipEventElapsedTime(&t[2], c->ev[3], c->ev[4]);
  float a = 0, b = 0;
  (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&b, c->ev[4], c->ev[5]);
  t[3] = a + b;
  (void)hipEventElapsedTime(&t[4], c->ev[0], c->ev[5]);
  int k = count < 5 ? count : 5;
  for (int i = 0; i < k; ++i) ms[i] = t[i];
  return k;
}

int zq_set_instrument(zq_context* c, int on) {
  if (!c) return ZQ_ERR_ARG;
  c->prof.on = on != 0;
  return ZQ_OK;
}

int zq_instrument_report(zq_context* c, int cls, double* total_ms, int* count, double* units) {
  if (!c || cls < 0 || cls >= zq::PROF_NCLASS) return ZQ_ERR_ARG;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return ZQ_ERR_HIP;
  double t = 0.0, u = 0.0;
  int k = 0;
  for (size_t i = 0; i < c->prof.used; ++i) {
    if (c->prof.cls[i] != cls) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->prof.pool[2 * i], c->prof.pool[2 * i + 1]) != hipSuccess) return ZQ_ERR_HIP;
    t += ms;
    u += c->prof.units[i];
    ++k;
  }
  if (total_ms) *total_ms = t;
  if (count) *count = k;
  if (units) *units = u;
  return ZQ_OK;
}

int zq_selftest_mfma(zq_context* c, const double* dA, const double* dB, double* dOut) {
  if (!c) return ZQ_ERR_ARG;
  int rc;
  if ((rc = with_device(c))) return rc;
  return zq::mfma_probe(dA, dB, dOut, c->stream);
}

int zq_synchronize(zq_context* c) {
  if (!c) return ZQ_ERR_ARG;
  return hipStreamSynchronize(c->stream) == hipSuccess ? ZQ_OK : ZQ_ERR_HIP;
}

}  // extern "C"

namespace ts {
int zquatev(const int n2, std::complex<double>* const D, const int nld2, double* const eig) {
  return zq_zquatev(n2, reinterpret_cast<double*>(D), nld2, eig);
}
}  // namespace ts
