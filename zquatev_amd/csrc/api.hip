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

#include <algorithm>
#include <climits>
#include <complex>
#include <cstring>
#include <mutex>
#include <new>
#include <system_error>
#include <thread>
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
  hipStream_t own_stream = nullptr;   // created on first use (see with_device)
  bool use_own = true;                // work runs on own_stream rather than a caller's stream
  int nb_override = 0;
  int tuned_nb[32] = {};      // autotuned panel width per bucket floor(log2 n) (0: none)
  // tile shapes (per bucket, autotuned; -1 / 0: the default): the matvec
  // strip switch (narrow strips up to this trailing size) and the GEMM tile
  // rows threshold (64-row tiles from this many 64 x 64 tiles per CU)
  int tuned_tsm[32] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                       -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  int tuned_t64[32] = {};
  int tsm_override = -1;
  int t64_override = 0;
  bool ws_locked = false;     // captured graphs reference ws / binfo: no reallocation until released
  char* pin[2] = {nullptr, nullptr};   // pinned chunks of the host entry points' copies
  size_t pin_cap = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  void* ws = nullptr;
  size_t ws_bytes = 0;
  size_t ws_limit = 0;        // most workspace bytes the context may hold (0: device memory decides)
  // When set, a workspace being replaced by a larger one is kept (retired)
  // instead of freed: hipFree waits for the whole device, which would stall
  // the other contexts of a server.  The owner frees them while idle.
  bool defer_free = false;
  std::vector<void*> retired;
  void* dD = nullptr;       // staging copy of D for the host-memory entry points
  size_t dD_bytes = 0;
  double* deig = nullptr;
  size_t deig_bytes = 0;
  hipEvent_t ev[6] = {};
  bool have_times = false;
  int* info_dev = nullptr;  // status word of the last single solve (device)
  int last_kind = 0;        // 0 none, 1 single solve (info_dev), 2 batch (binfo, last_count entries)
  int caps[3] = {0, 0, 0};  // iteration caps of the tridiagonal solver (0: default)
  int last_count = 0;
  zq::ProfState prof;
  zq::Staging stage;
  int* binfo = nullptr;     // per-problem status of the last batch (device)
  size_t binfo_bytes = 0;
};


namespace zq {
static thread_local int g_pending = -1;   // slot armed by the innermost ProfScope

// Events used only to time work on the device (phase marks, per-kernel
// instrumentation, the autotuner): no system-scope fence when they are
// recorded.  With the fence, a hipExtLaunchKernelGGL dispatch that carries
// the stop event ends with a cache writeback; the measured 2n = 8192 matvec
// then ran ~1.1 us (6 %) longer than without instrumentation.
constexpr unsigned kTimingEvent = hipEventDisableSystemFence;

ProfScope::ProfScope(int c, hipStream_t st, double u) : cls(c), s(st), units(u), slot(-1) {
  ProfState* p = g_prof;
  if (!p || !p->on) return;
  if (p->pool.size() < 2 * (p->used + 1)) {
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, kTimingEvent) != hipSuccess) return;
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
int stage_begin(Staging& st) {
  if (st.capture) return ZQI_OK;
  st.slot ^= 1;
  st.used = 0;
  if (st.done[st.slot] && hipEventSynchronize(st.done[st.slot]) != hipSuccess) return ZQI_ERR_HIP;
  return ZQI_OK;
}

int stage_upload(Staging& st, const void* src, size_t bytes, void* dst, hipStream_t s) {
  if (bytes == 0) return ZQI_OK;
  if (st.capture) {
    const size_t off = (st.keep_used + 255) & ~size_t(255);
    if (!st.keep || off + bytes > st.keep_cap) return ZQI_ERR_GRAPH;
    std::memcpy(st.keep + off, src, bytes);
    if (hipMemcpyAsync(dst, st.keep + off, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return ZQI_ERR_HIP;
    st.keep_used = off + bytes;
    return ZQI_OK;
  }
  const int k = st.slot;
  size_t off = (st.used + 255) & ~size_t(255);
  if (off + bytes > st.cap[k]) {
    // grow: every copy already issued from this buffer must have completed
    if (st.host[k]) {
      if (hipStreamSynchronize(s) != hipSuccess) return ZQI_ERR_HIP;
      (void)hipHostFree(st.host[k]);
      st.host[k] = nullptr;
      st.cap[k] = 0;
    }
    const size_t cap = std::max<size_t>(size_t(1) << 20, 2 * (off + bytes));
    void* h = nullptr;
    if (hipHostMalloc(&h, cap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return ZQI_ERR_NOMEM;
    }
    st.host[k] = static_cast<char*>(h);
    st.cap[k] = cap;
    off = 0;
  }
  std::memcpy(st.host[k] + off, src, bytes);
  if (hipMemcpyAsync(dst, st.host[k] + off, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return ZQI_ERR_HIP;
  st.used = off + bytes;
  return ZQI_OK;
}

int stage_end(Staging& st, hipStream_t s) {
  if (st.capture) return ZQI_OK;
  const int k = st.slot;
  if (!st.done[k] && hipEventCreateWithFlags(&st.done[k], hipEventDisableTiming) != hipSuccess) return ZQI_ERR_HIP;
  return hipEventRecord(st.done[k], s) == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

void stage_free(Staging& st) {
  for (int k = 0; k < 2; ++k) {
    if (st.done[k]) (void)hipEventSynchronize(st.done[k]);
    if (st.host[k]) (void)hipHostFree(st.host[k]);
    if (st.done[k]) (void)hipEventDestroy(st.done[k]);
    st.host[k] = nullptr;
    st.done[k] = nullptr;
    st.cap[k] = 0;
  }
  if (st.keep) (void)hipHostFree(st.keep);
  st.keep = nullptr;
  st.keep_cap = st.keep_used = 0;
}
}  // namespace zq

namespace {

using zq::quat;

// Split-K count of the back-transform's long products (V^* V, VT^* X).
// Measured (same box): 4 against 8 is 2.7% faster on the batch of 512
// problems (half the partials to write and add), equal at 2n = 8192, and
// 0.5-2.7% slower for single solves of 2n = 256 .. 2048; 16 is slower
// everywhere.
constexpr int kNSplit = 4;
constexpr size_t kGraphPool = size_t(8) << 20;   // pinned upload pool for captured graphs

// Marks the context as capturing while the stream is being captured into a
// graph (for the duration of one entry-point call); ZQ_ERR_GRAPH when the
// capture cannot be served.
struct CaptureScope {
  zq_context* c;
  int rc = ZQ_OK;
  bool active = false;
  explicit CaptureScope(zq_context* ctx);
  ~CaptureScope();
};

// reflectors per back-transform block (a multiple of the panel width)
// Reflectors per compact-WY block of the back-transform.  Batches group
// their problems by this width (as by panel width), so that a member applies
// the same blocks as its solo solve.
// Back-transform block width.  Measured: 128 from n = 449 took the batch of
// 512 problems of 2n <= 1024 from 82.7 to 74.6 ms and single solves of 2n =
// 1024 .. 8192 0.6-2.7 % faster; switching lower (n = 193, 321, 385) made the
// batch slower, and 256 made 2n = 4096 and 8192 slower.
inline int backtr_width(int n) { return n > 448 ? 128 : 64; }

struct Plan {
  quat* M;
  int ldm;       // leading dimension of M (padded off multiples of 64 quaternions)
  quat* X;
  double* d;
  quat* e;
  double* ea;
  zq::ReduceWork rw;
  zq::StedcWork sw;
  zq::BacktrWork bw;
  zq::QGemmTask* btasks;   // back-transform task tables (single solve)
  zq::SumTask* bsums;
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

// The per-problem buffers of the reduction and back-transform (a batch
// repeats this block once per problem, at a fixed stride).
// Leading dimension of the reduction's matrix: a column stride that is a
// multiple of 64 quaternions (2 KiB) puts every column of a tile on the same
// memory channels; one extra 128-byte step avoids that.
int padded_ld(int n) { return (n >= 64 && n % 64 == 0) ? n + 4 : n; }

void slab_layout(Bump& b, int n, int nb, bool vectors, Plan* p) {
  const size_t nn = (size_t)n * n;
  const int nmax = n > 1 ? n : 2;
  p->ldm = padded_ld(n);
  p->M = b.take<quat>((size_t)p->ldm * n);
  p->d = b.take<double>(nmax);
  p->e = b.take<quat>(nmax);
  p->ea = b.take<double>(nmax);
  p->rw.W = b.take<quat>((size_t)n * nb);
  p->rw.Prow = b.take<quat>((size_t)zq::matvec_row_slots(nmax) * n);
  p->rw.Pcol = b.take<quat>((size_t)zq::matvec_col_slots(nmax) * n);
  p->rw.TS = b.take<quat>((size_t)2 * nb);
  p->rw.ldtsp = 8 * nb;
  p->rw.TSP = b.take<double>((size_t)(cdiv(nmax, 16) + 1) * 8 * nb);
  p->rw.QF = b.take<double>((size_t)zq::matvec_tile_slots(nmax));
  p->rw.part = b.take<double>(cdiv(nmax, 16) + 16);
  p->rw.scal = b.take<double>(16);
  if (vectors) {
    p->X = b.take<quat>(nn);
    const int wbt = backtr_width(n);
    const int nblk = cdiv(nmax - 1, wbt);
    p->bw.phi = b.take<quat>(nmax);
    p->bw.nsplit = kNSplit;
    p->bw.wbt = wbt;
    p->bw.part_slots = (size_t)kNSplit * wbt * (n + wbt);
    p->bw.part = b.take<quat>(p->bw.part_slots);
    p->bw.Y = b.take<quat>((size_t)wbt * n);
    p->bw.G = b.take<quat>((size_t)nblk * wbt * wbt);
    p->bw.T = b.take<quat>((size_t)nblk * wbt * wbt);
    p->bw.vt_slots = zq::backtr_vt_slots(n, wbt);
    p->bw.VT = b.take<quat>(p->bw.vt_slots);
  } else {
    p->X = nullptr;
  }
}

// Divide & conquer buffers for a forest of nprob tridiagonals of total size
// ntot, the largest lq.
void stedc_layout(Bump& b, int ntot, int nprob, int lq, bool vectors, zq::StedcWork* sw) {
  zq::StedcWork& s = *sw;
  const int nt = ntot > 1 ? ntot : 2;
  s.lam = b.take<double>(nt);
  s.es = b.take<double>(nt);
  s.e2 = b.take<double>(nt);
  s.scal = b.take<double>(std::max(16, nprob));
  s.int_slots = zq::stedc_int_slots(ntot, nprob);
  s.ints = b.take<int>(s.int_slots);
  s.info = b.take<int>(std::max(16, nprob));
  s.stage = nullptr;
  if (vectors) {
    const size_t zz = (size_t)nt * lq;
    s.Z = b.take<double>(zz);
    s.Qtmp = b.take<double>(zz);
    s.U = b.take<double>(zz);
    s.dk = b.take<double>(nt);
    s.zk = b.take<double>(nt);
    s.dkk = b.take<double>(nt);
    s.zkk = b.take<double>(nt);
    s.dval = b.take<double>(nt);
    s.tau = b.take<double>(nt);
    s.zhat = b.take<double>(nt);
    s.rho = b.take<double>(nt);
    s.rot = b.take<double>((size_t)4 * nt);
    s.src = b.take<int>(nt);
    s.keep = b.take<int>(nt);
    s.dfl = b.take<int>(nt);
    s.org = b.take<int>(nt);
    s.posK = b.take<int>(nt);
    s.posD = b.take<int>(nt);
    s.nrot = b.take<int>(nt);
    s.kcnt = b.take<int>(nt);
    s.mix = b.take<int>(nt);
    s.gpos = b.take<int>(nt);
    s.ginv = b.take<int>(nt);
    s.kgrp = b.take<int>(nt);
  } else {
    s.Z = s.Qtmp = s.U = nullptr;
  }
}

// Lays out (or sizes, when base is null) the workspace of one solve.
size_t layout_plan(int n, int nb, bool vectors, char* base, Plan* p) {
  Bump b{base};
  slab_layout(b, n, nb, vectors, p);
  stedc_layout(b, n, 1, n, vectors, &p->sw);
  p->btasks = nullptr;
  p->bsums = nullptr;
  if (vectors) {
    p->btasks = b.take<zq::QGemmTask>(zq::backtransform_batched_tasks(1, &n, backtr_width(n), kNSplit));
    p->bsums = b.take<zq::SumTask>(zq::backtransform_batched_sums(1, &n, backtr_width(n)));
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

int size_bucket(int n) {
  int b = 0;
  while (b < 30 && (2 << b) <= n) ++b;
  return b;
}

// Panel width for half-dimension n: the caller's override, else the context's
// autotuned pick for n's bucket, else the built-in table.
int panel_width(const zq_context* ctx, int n) {
  int nb = ctx->nb_override;
  if (nb <= 0) nb = ctx->tuned_nb[size_bucket(n)];
  if (nb <= 0) nb = zq_tuned_panel_width(n);
  if (nb > 128) nb = 128;
  if (nb < 1) nb = 1;
  return nb;
}

constexpr int kTsmDefault = 768;   // matvec strip switch (trailing size)
constexpr int kTile64Default = 16; // 64-row GEMM tiles from this many 64 x 64 tiles per CU

int matvec_switch(const zq_context* ctx, int n) {
  if (!ctx) return kTsmDefault;
  if (ctx->tsm_override >= 0) return ctx->tsm_override;
  const int t = ctx->tuned_tsm[size_bucket(n)];
  return t >= 0 ? t : kTsmDefault;
}

int gemm64_min(const zq_context* ctx, int n) {
  if (!ctx) return kTile64Default;
  if (ctx->t64_override > 0) return ctx->t64_override;
  const int t = ctx->tuned_t64[size_bucket(n)];
  return t > 0 ? t : kTile64Default;
}

// ensure() for buffers that captured graphs may reference: while they are
// held (or during a capture) a buffer can be used but not reallocated.
int ensure_held(zq_context* ctx, void*& ptr, size_t& have, size_t need) {
  if ((ctx->ws_locked || ctx->stage.capture) && (!ptr || have < need)) return ZQ_ERR_GRAPH;
  if (ctx->defer_free && ptr && have < need) {
    ctx->retired.push_back(ptr);
    ptr = nullptr;
    have = 0;
  }
  return ensure(ptr, have, need);
}

int make_plan(zq_context* ctx, int n, int nb, bool vectors, Plan* p) {
  const size_t need = layout_plan(n, nb, vectors, nullptr, p);
  if (ctx->ws_limit && need > ctx->ws_limit) return ZQ_ERR_NOMEM;
  int rc = ensure_held(ctx, ctx->ws, ctx->ws_bytes, need);
  if (rc) return rc;
  layout_plan(n, nb, vectors, static_cast<char*>(ctx->ws), p);
  return ZQ_OK;
}

void apply_caps(const zq_context* ctx, zq::StedcWork& sw) {
  if (ctx->caps[0] > 0) sw.cap_leaf = ctx->caps[0];
  if (ctx->caps[1] > 0) sw.cap_secular = ctx->caps[1];
  if (ctx->caps[2] > 0) sw.cap_bisect = ctx->caps[2];
}

// Makes the context's device current and, when the context runs on its own
// stream, creates that stream on first use.  A stream takes one of the
// process's hardware queues (HIP maps streams onto GPU_MAX_HW_QUEUES of them,
// 4 by default, sharing queues beyond that); a context that only ever runs on
// a caller's stream never holds one, so it cannot force the streams of other
// contexts -- the lanes of a server, say -- to share queues and serialise.
int with_device(zq_context* ctx) {
  if (hipSetDevice(ctx->device) != hipSuccess) return ZQ_ERR_HIP;
  if (ctx->use_own) {
    if (!ctx->own_stream && hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
      ctx->own_stream = nullptr;
      return ZQ_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
  }
  return ZQ_OK;
}

// Nothing has been queued on this context yet (its own stream is not even created).
bool idle_unstarted(const zq_context* ctx) { return ctx->use_own && !ctx->own_stream; }

// The device solve: dD (ld2 complex) -> dD, deig.
int solve_device(zq_context* ctx, int n, double* dD, int ld2, double* deig, bool vectors) {
  hipStream_t s = ctx->stream;
  const bool cap = ctx->stage.capture;   // no timing events or instrumentation inside a graph
  ctx->prof.used = 0;
  zq::g_prof = cap ? nullptr : &ctx->prof;
  struct Reset { ~Reset() { zq::g_prof = nullptr; } } reset_prof;
  auto mark = [&](int k) { return cap ? hipSuccess : hipEventRecord(ctx->ev[k], s); };
  const int nb = panel_width(ctx, n);
  Plan p;
  int rc = make_plan(ctx, n, nb, vectors, &p);
  if (rc) return rc;
  p.rw.tsm = matvec_switch(ctx, n);
  zq::TileScope tiles(gemm64_min(ctx, n));
  ZQ_HIP_CHECK(mark(0));
  ctx->last_kind = 0;
  if ((rc = zq::fill_zero(p.sw.info, sizeof(int), s))) return rc;
  if ((rc = zq::load_left_half(n, dD, ld2, p.M, p.ldm, p.sw.info, s))) return rc;
  ZQ_HIP_CHECK(mark(1));
  if ((rc = zq::reduce(p.M, n, p.ldm, nb, p.d, p.e, p.rw, s))) return rc;
  if ((rc = zq::offdiag_abs(n, p.e, p.ea, s))) return rc;
  ZQ_HIP_CHECK(mark(2));
  ctx->info_dev = p.sw.info;
  ctx->last_kind = 1;
  if ((rc = zq::stage_begin(ctx->stage))) return rc;
  struct EndStage {
    zq_context* c;
    hipStream_t s;
    ~EndStage() { (void)zq::stage_end(c->stage, s); }
  } end_stage{ctx, s};
  const int zero = 0;
  p.sw.stage = &ctx->stage;
  apply_caps(ctx, p.sw);
  if ((rc = zq::stedc(1, &zero, &n, n, n, p.d, p.ea, deig, p.sw, vectors, s))) return rc;
  ZQ_HIP_CHECK(mark(3));
  if (vectors) {
    if ((rc = zq::backtransform_batched(1, &n, zq::ZB{nullptr, 0}, nullptr, p.M, p.ldm, p.e, p.sw.Z, n, p.X, n, p.bw,
                                        p.btasks, p.bsums, ctx->stage, s)))
      return rc;
    ZQ_HIP_CHECK(mark(4));
    if ((rc = zq::store_vectors(n, p.X, n, dD, ld2, s))) return rc;
  } else {
    ZQ_HIP_CHECK(mark(4));
  }
  ZQ_HIP_CHECK(mark(5));
  ctx->have_times = !cap;
  return ZQ_OK;
}

int check_args(int n2, const void* D, int nld2, const void* eig);

// ---------------------------------------------------------------------------
// Batches: problems grouped by panel width, each group solved in lockstep
// (sizes descending), every problem's buffers in a slab of fixed stride.
// ---------------------------------------------------------------------------
struct BatchPlan {
  Plan slab;          // problem 0's buffers
  size_t zs = 0;      // slab stride (bytes)
  int* nvec = nullptr;
  int* zrow = nullptr;
  zq::BatchProb* probs = nullptr;
  double* dall = nullptr;
  double* eall = nullptr;
  double* wall = nullptr;
  zq::StedcWork sw;
  zq::QGemmTask* rtasks = nullptr;
  int* tstart = nullptr;
  zq::QGemmTask* btasks = nullptr;
  zq::SumTask* sums = nullptr;
};

size_t batch_layout(int nprob, const int* nh, int nb, bool vectors, char* base, BatchPlan* bp) {
  const int L = nh[0];
  int ntot = 0;
  for (int z = 0; z < nprob; ++z) ntot += nh[z];
  Bump sizing{nullptr};
  Plan tmp;
  slab_layout(sizing, L, nb, vectors, &tmp);
  const size_t zs = (sizing.off + 255) & ~size_t(255);
  Bump b{base};
  char* slabs = b.take<char>(zs * nprob);
  Bump s0{slabs};
  slab_layout(s0, L, nb, vectors, &bp->slab);
  bp->zs = zs;
  bp->nvec = b.take<int>(nprob);
  bp->zrow = b.take<int>(nprob);
  bp->probs = b.take<zq::BatchProb>(nprob);
  bp->dall = b.take<double>(ntot + 1);
  bp->eall = b.take<double>(ntot + 1);
  bp->wall = b.take<double>(ntot + 1);
  stedc_layout(b, ntot, nprob, L, vectors, &bp->sw);
  bp->rtasks = b.take<zq::QGemmTask>(zq::reduce_batched_tasks(nprob, nh, nb));
  bp->tstart = b.take<int>(nprob);
  if (vectors) {
    bp->btasks = b.take<zq::QGemmTask>(zq::backtransform_batched_tasks(nprob, nh, backtr_width(L), kNSplit));
    bp->sums = b.take<zq::SumTask>(zq::backtransform_batched_sums(nprob, nh, backtr_width(L)));
  }
  return b.off + 256;
}

// One lockstep group on the context's workspace (already large enough).
// P: the group's problems, sizes descending, P[z].off the forest offsets.
int solve_group(zq_context* ctx, std::vector<zq::BatchProb>& P, int nb, int tsm, bool vectors) {
  hipStream_t s = ctx->stream;
  const int nprob = (int)P.size();
  std::vector<int> nh(nprob), zrow(nprob);
  int ntot = 0;
  for (int z = 0; z < nprob; ++z) {
    nh[z] = P[z].n;
    zrow[z] = ntot;
    P[z].off = ntot;
    ntot += P[z].n;
  }
  const int L = nh[0];
  BatchPlan bp;
  batch_layout(nprob, nh.data(), nb, vectors, static_cast<char*>(ctx->ws), &bp);
  bp.slab.rw.tsm = tsm;
  zq::TileScope tiles(gemm64_min(ctx, L));
  zq::Staging& st = ctx->stage;
  int rc;
  if ((rc = zq::stage_upload(st, nh.data(), nprob * sizeof(int), bp.nvec, s))) return rc;
  if ((rc = zq::stage_upload(st, zrow.data(), nprob * sizeof(int), bp.zrow, s))) return rc;
  if ((rc = zq::stage_upload(st, P.data(), nprob * sizeof(zq::BatchProb), bp.probs, s))) return rc;
  const zq::ZB zb{bp.nvec, bp.zs};
  if ((rc = zq::fill_zero(bp.sw.info, nprob * sizeof(int), s))) return rc;
  if ((rc = zq::load_left_batched(nprob, L, bp.probs, bp.slab.M, L, zb, bp.sw.info, s))) return rc;
  if ((rc = zq::reduce_batched(nprob, nh.data(), zb, bp.slab.M, L, nb, bp.slab.d, bp.slab.e, bp.slab.rw, bp.rtasks,
                               bp.tstart, st, s)))
    return rc;
  if ((rc = zq::gather_tridiag_batched(nprob, L, bp.probs, bp.slab.d, bp.slab.e, zb, bp.dall, bp.eall, s))) return rc;
  bp.sw.stage = &st;
  apply_caps(ctx, bp.sw);
  if ((rc = zq::stedc(nprob, zrow.data(), nh.data(), ntot, L, bp.dall, bp.eall, bp.wall, bp.sw, vectors, s)))
    return rc;
  if ((rc = zq::scatter_eig_batched(nprob, L, bp.probs, bp.wall, s))) return rc;
  if (vectors) {
    if ((rc = zq::backtransform_batched(nprob, nh.data(), zb, bp.zrow, bp.slab.M, L, bp.slab.e, bp.sw.Z, ntot,
                                        bp.slab.X, L, bp.slab.bw, bp.btasks, bp.sums, st, s)))
      return rc;
    if ((rc = zq::store_vectors_batched(nprob, L, bp.probs, bp.slab.X, L, zb, s))) return rc;
  }
  return zq::scatter_info_batched(nprob, bp.probs, bp.sw.info, ctx->binfo, s);
}

// Device-memory batch: groups and solves (arguments already validated).
// Each problem's status is gathered on the device; info (host, optional)
// receives it after the call waits for completion; info_async (pinned host,
// optional) receives it asynchronously, valid once the stream has passed the
// call.
struct Groups {
  std::vector<std::vector<zq::BatchProb>> probs;   // per group, sizes descending
  std::vector<int> nb;                             // panel width of each group
  std::vector<int> wb;                             // back-transform block width of each group
  std::vector<int> tsm;                            // matvec strip switch of each group
  size_t need = 0;                                 // workspace bytes of the largest group
};

Groups make_groups(const zq_context* ctx, int count, const int* n2, double* const* D, const int* ld2,
                   double* const* eig, bool vectors) {
  Groups G;
  for (int i = 0; i < count; ++i) {
    if (n2[i] == 0) continue;
    const int n = n2[i] / 2;
    const int nb = ctx ? panel_width(ctx, n) : zq_tuned_panel_width(n);
    const int wb = backtr_width(n);
    const int tsm = matvec_switch(ctx, n);
    size_t g = 0;
    while (g < G.nb.size() && !(G.nb[g] == nb && G.wb[g] == wb && G.tsm[g] == tsm)) ++g;
    if (g == G.nb.size()) {
      G.nb.push_back(nb);
      G.wb.push_back(wb);
      G.tsm.push_back(tsm);
      G.probs.emplace_back();
    }
    G.probs[g].push_back(zq::BatchProb{D ? D[i] : nullptr, eig ? eig[i] : nullptr, ld2 ? ld2[i] : n2[i], n, 0, i});
  }
  for (size_t g = 0; g < G.probs.size(); ++g) {
    // sizes descending, ties by position
    std::stable_sort(G.probs[g].begin(), G.probs[g].end(),
                     [](const zq::BatchProb& a, const zq::BatchProb& b) { return a.n > b.n; });
    std::vector<int> nh(G.probs[g].size());
    for (size_t k = 0; k < nh.size(); ++k) nh[k] = G.probs[g][k].n;
    BatchPlan bp;
    G.need = std::max(G.need, batch_layout((int)nh.size(), nh.data(), G.nb[g], vectors, nullptr, &bp));
  }
  return G;
}

// Workspace of problems [a, b) of one group (sizes descending).
size_t run_bytes(const std::vector<zq::BatchProb>& P, size_t a, size_t b, int nb, bool vectors) {
  std::vector<int> nh(b - a);
  for (size_t k = a; k < b; ++k) nh[k - a] = P[k].n;
  BatchPlan bp;
  return batch_layout((int)nh.size(), nh.data(), nb, vectors, nullptr, &bp);
}

// Splits every group into consecutive runs (sizes descending) whose workspace
// fits `budget` bytes; each run is then solved as a lockstep batch of its own.
// Every slab of a run is sized for the run's largest problem, so one large
// problem no longer inflates the slabs of many smaller ones.  A problem that
// does not fit even alone forms its own run (its allocation then fails).
// The results do not change: each member computes exactly its solo solve.
void split_groups(Groups& G, bool vectors, size_t budget) {
  Groups out;
  for (size_t g = 0; g < G.probs.size(); ++g) {
    const std::vector<zq::BatchProb>& P = G.probs[g];
    size_t a = 0;
    while (a < P.size()) {
      size_t lo = a + 1, hi = P.size();   // [a, lo) is always taken
      if (run_bytes(P, a, hi, G.nb[g], vectors) <= budget) {
        lo = hi;
      } else {
        while (hi - lo > 1) {   // the workspace grows with every member added
          const size_t mid = (lo + hi) / 2;
          if (run_bytes(P, a, mid, G.nb[g], vectors) <= budget) lo = mid;
          else hi = mid;
        }
      }
      out.probs.emplace_back(P.begin() + a, P.begin() + lo);
      out.nb.push_back(G.nb[g]);
      out.wb.push_back(G.wb[g]);
      out.tsm.push_back(G.tsm[g]);
      out.need = std::max(out.need, run_bytes(P, a, lo, G.nb[g], vectors));
      a = lo;
    }
  }
  G = std::move(out);
}

// Bytes of workspace a batch may use on ctx: the limit set on the context,
// and what the device can give (the workspace held now, which is reused,
// plus 90% of the free memory).  While captured graphs hold the workspace,
// exactly what is held.
size_t workspace_budget(zq_context* ctx) {
  if (ctx->ws_locked || ctx->stage.capture) return ctx->ws_bytes;
  size_t budget = SIZE_MAX;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = ctx->ws_bytes + (size_t)(0.9 * (double)fr);
  else (void)hipGetLastError();
  if (ctx->ws_limit) budget = std::min(budget, ctx->ws_limit);
  return budget;
}

// Aggregate status of a batch from the per-problem codes: ZQ_ERR_NONFINITE
// if some input held a NaN or an infinity, else 1 if some problem did not
// converge, else 0.
int batch_status(const int* info, int count) {
  int agg = 0;
  for (int i = 0; i < count; ++i) {
    if (info[i] == ZQ_ERR_NONFINITE) return ZQ_ERR_NONFINITE;
    if (info[i] > 0) agg = 1;
  }
  return agg;
}

int solve_batched(zq_context* ctx, int count, const int* n2, double* const* D, const int* ld2, double* const* eig,
                  bool vectors, int* info, int* info_async) {
  Groups G = make_groups(ctx, count, n2, D, ld2, eig, vectors);
  if (G.probs.empty()) {
    if (info)
      for (int i = 0; i < count; ++i) info[i] = 0;
    if (info_async)
      for (int i = 0; i < count; ++i) info_async[i] = 0;
    return ZQ_OK;
  }
  if (G.need > ctx->ws_bytes || (ctx->ws_limit && G.need > ctx->ws_limit)) {
    const size_t budget = workspace_budget(ctx);
    if (G.need > budget) split_groups(G, vectors, budget);
  }
  const size_t need = G.need;
  if (ctx->ws_limit && need > ctx->ws_limit) return ZQ_ERR_NOMEM;   // a single problem over the limit
  int rc = ensure_held(ctx, ctx->ws, ctx->ws_bytes, need);
  if (rc) return rc;
  if ((rc = ensure_held(ctx, reinterpret_cast<void*&>(ctx->binfo), ctx->binfo_bytes,
                        (size_t)count * sizeof(int) + 256)))
    return rc;
  ctx->have_times = false;
  ctx->prof.used = 0;
  zq::g_prof = ctx->stage.capture ? nullptr : &ctx->prof;   // per-kernel-class events when instrumented
  struct ResetProf { ~ResetProf() { zq::g_prof = nullptr; } } reset_prof;
  ctx->info_dev = nullptr;
  ctx->last_kind = 2;
  ctx->last_count = count;
  hipStream_t s = ctx->stream;
  if ((rc = zq::fill_zero(ctx->binfo, (size_t)count * sizeof(int), s))) return rc;
  if ((rc = zq::stage_begin(ctx->stage))) return rc;
  for (size_t g = 0; g < G.probs.size() && !rc; ++g) rc = solve_group(ctx, G.probs[g], G.nb[g], G.tsm[g], vectors);
  const int rc2 = zq::stage_end(ctx->stage, s);
  if (rc) return rc;
  if (rc2) return rc2;
  if (info_async)
    ZQ_HIP_CHECK(hipMemcpyAsync(info_async, ctx->binfo, (size_t)count * sizeof(int), hipMemcpyDeviceToHost, s));
  if (!info) return ZQ_OK;
  ZQ_HIP_CHECK(hipMemcpyAsync(info, ctx->binfo, (size_t)count * sizeof(int), hipMemcpyDeviceToHost, s));
  ZQ_HIP_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < count; ++i) info[i] = info_code(info[i]);
  return batch_status(info, count);
}

int check_batch(const zq_context* c, int count, const int* n2, double* const* D, const int* ld2, double* const* eig,
                int* info) {
  if (!c) return ZQ_ERR_ARG;
  if (count < 0) return -2;
  if (count == 0) return ZQ_OK;
  if (!n2) return -3;
  if (!D) return -4;
  if (!ld2) return -5;
  if (!eig) return -6;
  int bad = 0;
  for (int i = 0; i < count; ++i) {
    const int r = check_args(n2[i], D[i], ld2[i], eig[i]);
    if (info) info[i] = r;
    if (r) bad = 1;
  }
  return bad ? -8 : ZQ_OK;
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

// ---- the entry point that allocates nothing ------------------------------
// Every buffer of zq_eigvals_noalloc lies in this array, one copy per device
// (held until the process ends).  Host code never uses the name itself --
// on the host it only stands for the symbol -- but the device address that
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr size_t kScratchBytes = ZQ_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on the current device (looked up on the first
// call there), or null on failure.
void* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

// ---- host <-> device copies of the host entry points --------------------
// The caller's D is pageable memory.  Columns go through two pinned chunks of
// the context: several host threads copy one chunk while the DMA engine moves
// the other (the HIP runtime's own staging of a pageable copy is one thread).
constexpr size_t kPinChunk = size_t(32) << 20;

int pin_ready(zq_context* c) {
  if (c->pin[0]) return ZQ_OK;
  for (int k = 0; k < 2; ++k) {
    void* h = nullptr;
    if (hipHostMalloc(&h, kPinChunk, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return ZQ_ERR_NOMEM;
    }
    c->pin[k] = static_cast<char*>(h);
    if (hipEventCreateWithFlags(&c->pin_ev[k], hipEventDisableTiming) != hipSuccess) return ZQ_ERR_HIP;
  }
  c->pin_cap = kPinChunk;
  return ZQ_OK;
}

// A run of ncols columns of colb bytes: column j at host + j * hstride on the
// host side, at dev + j * colb on the device side (columns contiguous there).
struct ColRun {
  char* dev;
  char* host;
  size_t hstride, colb;
  int ncols;
  size_t poff;   // offset of its first column in the pinned chunk
};

// Greedy packing of the runs, in order, into groups of at most `cap` bytes;
// a run is split between groups at a column boundary.  Runs whose columns are
// larger than a chunk are returned in `big` (copied by the runtime).
void plan_groups(const std::vector<ColRun>& runs, size_t cap, std::vector<std::vector<ColRun>>& groups,
                 std::vector<ColRun>& big) {
  size_t used = cap;
  for (ColRun r : runs) {
    if (r.ncols <= 0) continue;
    if (r.colb > cap) {
      big.push_back(r);
      continue;
    }
    while (r.ncols > 0) {
      int fit = (int)std::min<size_t>((size_t)r.ncols, (cap - used) / r.colb);
      if (fit == 0) {
        groups.emplace_back();
        used = 0;
        continue;
      }
      ColRun q = r;
      q.ncols = fit;
      q.poff = used;
      groups.back().push_back(q);
      used += (size_t)fit * r.colb;
      r.dev += (size_t)fit * r.colb;
      r.host += (size_t)fit * r.hstride;
      r.ncols -= fit;
    }
  }
}

// host columns <-> pinned chunk for one group, split over up to 8 threads
void pack_group(const std::vector<ColRun>& g, char* pin, bool to_pinned) {
  size_t cols = 0, bytes = 0;
  for (const ColRun& r : g) {
    cols += (size_t)r.ncols;
    bytes += (size_t)r.ncols * r.colb;
  }
  const int nt = (int)std::max<size_t>(1, std::min<size_t>({(size_t)8, cols, bytes >> 22}));
  auto work = [&](int t) {
    const size_t a = cols * t / nt, b = cols * (t + 1) / nt;
    size_t base = 0;
    for (const ColRun& r : g) {
      const size_t lo = std::max(a, base), hi = std::min(b, base + (size_t)r.ncols);
      for (size_t j = lo; j < hi; ++j) {
        char* h = r.host + (j - base) * r.hstride;
        char* p = pin + r.poff + (j - base) * r.colb;
        if (to_pinned) std::memcpy(p, h, r.colb);
        else std::memcpy(h, p, r.colb);
      }
      base += (size_t)r.ncols;
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
}

// host -> device, asynchronous on s (the host buffers may be reused on return)
int h2d_runs(zq_context* c, const std::vector<ColRun>& runs, hipStream_t s) {
  int rc = pin_ready(c);
  if (rc) return rc;
  std::vector<std::vector<ColRun>> groups;
  std::vector<ColRun> big;
  plan_groups(runs, c->pin_cap, groups, big);
  for (const ColRun& r : big)
    ZQ_HIP_CHECK(hipMemcpy2DAsync(r.dev, r.colb, r.host, r.hstride, r.colb, r.ncols, hipMemcpyHostToDevice, s));
  for (size_t q = 0; q < groups.size(); ++q) {
    const int k = (int)(q & 1);
    ZQ_HIP_CHECK(hipEventSynchronize(c->pin_ev[k]));   // the chunk's previous DMA has read it
    pack_group(groups[q], c->pin[k], true);
    for (const ColRun& r : groups[q])
      ZQ_HIP_CHECK(hipMemcpyAsync(r.dev, c->pin[k] + r.poff, (size_t)r.ncols * r.colb, hipMemcpyHostToDevice, s));
    ZQ_HIP_CHECK(hipEventRecord(c->pin_ev[k], s));
  }
  return ZQ_OK;
}

// device -> host, after the work queued on s; returns when the host holds the data
int d2h_runs(zq_context* c, const std::vector<ColRun>& runs, hipStream_t s) {
  int rc = pin_ready(c);
  if (rc) return rc;
  std::vector<std::vector<ColRun>> groups;
  std::vector<ColRun> big;
  plan_groups(runs, c->pin_cap, groups, big);
  for (const ColRun& r : big)
    ZQ_HIP_CHECK(hipMemcpy2DAsync(r.host, r.hstride, r.dev, r.colb, r.colb, r.ncols, hipMemcpyDeviceToHost, s));
  auto issue = [&](size_t q) -> int {
    const int k = (int)(q & 1);
    for (const ColRun& r : groups[q])
      ZQ_HIP_CHECK(hipMemcpyAsync(c->pin[k] + r.poff, r.dev, (size_t)r.ncols * r.colb, hipMemcpyDeviceToHost, s));
    ZQ_HIP_CHECK(hipEventRecord(c->pin_ev[k], s));
    return ZQ_OK;
  };
  if (!groups.empty() && (rc = issue(0))) return rc;
  for (size_t q = 0; q < groups.size(); ++q) {
    if (q + 1 < groups.size() && (rc = issue(q + 1))) return rc;   // the other chunk, while this one drains
    ZQ_HIP_CHECK(hipEventSynchronize(c->pin_ev[q & 1]));
    pack_group(groups[q], c->pin[q & 1], false);
  }
  ZQ_HIP_CHECK(hipStreamSynchronize(s));
  return ZQ_OK;
}

int exception_status() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return ZQ_ERR_NOMEM;
  } catch (const std::system_error&) {   // e.g. no thread could be started
    return ZQ_ERR_NOMEM;
  } catch (...) {
    return ZQ_ERR_HIP;
  }
}

extern "C" {

int zq_version(void) { return 100; }

const char* zq_error_string(int code) {
  switch (code) {
    case ZQ_OK: return "success";
    case ZQ_ERR_HIP: return "HIP runtime error";
    case ZQ_ERR_NOMEM: return "device memory exhausted";
    case ZQ_ERR_ARG: return "unsupported internal configuration";
    case ZQ_ERR_NONFINITE: return "the input holds a NaN or an infinity (or overflowed the reduction)";
    case ZQ_ERR_BUDGET: return "the request needs more device memory than the server's budget";
    case ZQ_ERR_GRAPH:
      return "graph capture: this call cannot be captured, its workspace is not allocated yet (run it once "
             "outside capture), or the upload pool / workspace is held by captured graphs (zq_graph_reserve, "
             "zq_graph_release)";
    case -1: return "argument 1 (n2) is negative or odd";
    case -2: return "argument 2 (D) is null";
    case -3: return "argument 3 (nld2) is smaller than n2";
    case -4: return "argument 4 (eig) is null";
    default: return code > 0 ? "tridiagonal eigensolver did not converge" : "invalid argument";
  }
}

int zq_tuned_panel_width(int n) {
  // measured on MI355X: nb = 16 is 0.3-3 % faster than 32 for single solves of
  // n = 128 .. 768 (its column kernel holds half the panel coefficients per
  // thread, so more blocks fit a CU) and 5 % for a batch of 512 problems of
  // n <= 512; 32 is fastest from n = 1024 (0.9 % there).  zq_autotune refines
  // it per size bucket (floor(log2 n)).
  return n < 1024 ? 16 : 32;
}

int zq_create(zq_context** out, int device) try {
  if (!out) return -1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return ZQ_ERR_HIP;
  zq_context* c = new zq_context();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    zq_destroy(c);
    return ZQ_ERR_HIP;
  }
  c->stream = nullptr;   // the own stream is created on first use
  for (auto& e : c->ev) {
    if (hipEventCreateWithFlags(&e, zq::kTimingEvent) != hipSuccess) {
      zq_destroy(c);
      return ZQ_ERR_HIP;
    }
  }
  // pinned pool for uploads recorded into captured graphs (allocated here:
  // nothing may be allocated during a capture)
  void* h = nullptr;
  if (hipHostMalloc(&h, kGraphPool, hipHostMallocDefault) == hipSuccess) {
    c->stage.keep = static_cast<char*>(h);
    c->stage.keep_cap = kGraphPool;
  } else {
    (void)hipGetLastError();
  }
  *out = c;
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_graph_reserve(zq_context* c, size_t bytes) {
  if (!c) return ZQ_ERR_ARG;
  if (c->stage.capture || c->stage.keep_used > 0) return ZQ_ERR_GRAPH;
  if (bytes <= c->stage.keep_cap) return ZQ_OK;
  void* h = nullptr;
  if (hipHostMalloc(&h, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return ZQ_ERR_NOMEM;
  }
  if (c->stage.keep) (void)hipHostFree(c->stage.keep);
  c->stage.keep = static_cast<char*>(h);
  c->stage.keep_cap = bytes;
  return ZQ_OK;
}

int zq_graph_release(zq_context* c) {
  if (!c) return ZQ_ERR_ARG;
  if (c->stage.capture) return ZQ_ERR_GRAPH;
  c->stage.keep_used = 0;
  c->ws_locked = false;
  return ZQ_OK;
}

int zq_destroy(zq_context* c) {
  if (!c) return ZQ_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  zq::stage_free(c->stage);
  for (int k = 0; k < 2; ++k) {
    if (c->pin[k]) (void)hipHostFree(c->pin[k]);
    if (c->pin_ev[k]) (void)hipEventDestroy(c->pin_ev[k]);
  }
  if (c->ws) (void)hipFree(c->ws);
  for (void* p : c->retired) (void)hipFree(p);
  if (c->dD) (void)hipFree(c->dD);
  if (c->deig) (void)hipFree(c->deig);
  if (c->binfo) (void)hipFree(c->binfo);
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
  c->use_own = false;
  return ZQ_OK;
}

int zq_use_own_stream(zq_context* c) {
  if (!c) return -1;
  c->use_own = true;
  c->stream = c->own_stream;   // null until first use
  return ZQ_OK;
}

int zq_set_panel_width(zq_context* c, int nb) {
  if (!c) return -1;
  if (nb < 0 || nb > 128) return -2;
  c->nb_override = nb;
  return ZQ_OK;
}

CaptureScope::CaptureScope(zq_context* ctx) : c(ctx) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (c->stream && hipStreamIsCapturing(c->stream, &cs) != hipSuccess) {
    (void)hipGetLastError();
    rc = ZQ_ERR_HIP;
    return;
  }
  if (cs == hipStreamCaptureStatusNone) return;
  if (cs != hipStreamCaptureStatusActive || !c->stage.keep) {
    rc = ZQ_ERR_GRAPH;
    return;
  }
  active = true;
  c->stage.capture = true;
  c->ws_locked = true;
}
CaptureScope::~CaptureScope() {
  if (active) c->stage.capture = false;
}

int zq_panel_width(zq_context* c, int n) {
  if (!c || n < 1) return 0;
  return panel_width(c, n);
}

// Times the reduction (the phase the panel width and the tile shapes change)
// on a synthetic U(-1,1) input generated on the device: one warm-up and two
// timed runs per candidate, the faster kept.  First the panel width (widths
// wider than the matrix are skipped: they would repeat the previous one's
// schedule), then with it the matvec strip switch (switches at or above the
// largest trailing size all give the same schedule: the first of them
// stands for the rest), then the GEMM tile-rows threshold.  Each pick is
// recorded for the size's bucket unless an override is set.
int zq_autotune(zq_context* c, const int* sizes, int count, int* chosen, double* ms) try {
  if (!c) return ZQ_ERR_ARG;
  if (count < 0) return -3;
  if (count > 0 && !sizes) return -2;
  for (int k = 0; k < count; ++k)
    if (sizes[k] < 1) return -2;
  if (count == 0) return ZQ_OK;
  int rc;
  if ((rc = with_device(c))) return rc;
  {
    CaptureScope cap(c);
    if (cap.rc) return cap.rc;
    if (cap.active) return ZQ_ERR_GRAPH;   // it waits on its timing events
  }
  static const int cand[4] = {16, 32, 64, 128};
  static const int cand_tsm[4] = {512, 768, 1024, 1536};
  static const int cand_t64[3] = {4, 16, 64};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreateWithFlags(&e0, zq::kTimingEvent) != hipSuccess || hipEventCreateWithFlags(&e1, zq::kTimingEvent) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return ZQ_ERR_HIP;
  }
  struct Ev {
    hipEvent_t a, b;
    ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } ev{e0, e1};
  hipStream_t s = c->stream;
  // best of two timed reductions after a warm-up (ms), or < 0 with rc set
  auto time_reduce = [&](int n, int nb, int tsm, int t64, double* out) -> int {
    Plan p;
    int r = make_plan(c, n, nb, false, &p);
    if (r) return r;
    p.rw.tsm = tsm;
    zq::TileScope tiles(t64);
    double tmin = -1.0;
    for (int rep = 0; rep < 3; ++rep) {
      if ((r = zq::gen_lower_uniform(n, 1001, p.M, p.ldm, s))) return r;
      ZQ_HIP_CHECK(hipEventRecord(e0, s));
      if ((r = zq::reduce(p.M, n, p.ldm, nb, p.d, p.e, p.rw, s))) return r;
      ZQ_HIP_CHECK(hipEventRecord(e1, s));
      ZQ_HIP_CHECK(hipEventSynchronize(e1));
      float f = 0.f;
      ZQ_HIP_CHECK(hipEventElapsedTime(&f, e0, e1));
      if (rep > 0 && (tmin < 0.0 || f < tmin)) tmin = f;
    }
    *out = tmin;
    return ZQ_OK;
  };
  for (int k = 0; k < count; ++k) {
    const int n = sizes[k];
    const int bucket = size_bucket(n);
    int tsm = matvec_switch(c, n), t64 = gemm64_min(c, n);
    int best = 0;
    double best_ms = 0.0;
    for (int t = 0; t < 4; ++t) {
      double tmin = -1.0;
      if (t == 0 || cand[t - 1] < n - 1) {
        if ((rc = time_reduce(n, cand[t], tsm, t64, &tmin))) return rc;
        if (best == 0 || tmin < best_ms) {
          best = cand[t];
          best_ms = tmin;
        }
      }
      if (ms) ms[4 * k + t] = tmin;
    }
    c->tuned_nb[bucket] = best;
    if (chosen) chosen[k] = best;
    const int nb = panel_width(c, n);   // the pick, or the caller's override
    if (c->tsm_override < 0 && n - 1 > cand_tsm[0]) {
      int pick = tsm;
      double pick_ms = -1.0;
      for (int t = 0; t < 4; ++t) {
        if (t > 0 && cand_tsm[t - 1] >= n - 1) break;   // same schedule as the previous candidate
        double tm = -1.0;
        if ((rc = time_reduce(n, nb, cand_tsm[t], t64, &tm))) return rc;
        if (pick_ms < 0.0 || tm < pick_ms) {
          pick = cand_tsm[t];
          pick_ms = tm;
        }
      }
      c->tuned_tsm[bucket] = pick;
      tsm = pick;
    }
    if (c->t64_override <= 0) {
      int pick = t64;
      double pick_ms = -1.0;
      for (int t = 0; t < 3; ++t) {
        double tm = -1.0;
        if ((rc = time_reduce(n, nb, tsm, cand_t64[t], &tm))) return rc;
        if (pick_ms < 0.0 || tm < pick_ms) {
          pick = cand_t64[t];
          pick_ms = tm;
        }
      }
      c->tuned_t64[bucket] = pick;
    }
  }
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_set_tile_shape(zq_context* c, int matvec_switch_rows, int gemm64_tiles_per_cu) {
  if (!c) return ZQ_ERR_ARG;
  if (matvec_switch_rows < -1 || matvec_switch_rows > 2048) return -2;
  if (gemm64_tiles_per_cu < 0) return -3;
  c->tsm_override = matvec_switch_rows;
  c->t64_override = gemm64_tiles_per_cu;
  return ZQ_OK;
}

int zq_tile_shape(zq_context* c, int n, int* matvec_switch_rows, int* gemm64_tiles_per_cu) {
  if (!c || n < 1) return ZQ_ERR_ARG;
  if (matvec_switch_rows) *matvec_switch_rows = matvec_switch(c, n);
  if (gemm64_tiles_per_cu) *gemm64_tiles_per_cu = gemm64_min(c, n);
  return ZQ_OK;
}

int zq_zquatev_device(zq_context* c, int n2, double* dD, int nld2, double* deig, int jobz) try {
  if (!c) return ZQ_ERR_ARG;
  int rc = check_args(n2, dD, nld2, deig);
  if (rc) return rc;
  if (jobz != 'V' && jobz != 'v' && jobz != 'N' && jobz != 'n') return -5;
  if (n2 == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  CaptureScope cap(c);
  if (cap.rc) return cap.rc;
  return solve_device(c, n2 / 2, dD, nld2, deig, jobz == 'V' || jobz == 'v');
} catch (...) {
  return exception_status();
}

int zq_zquatev_ctx(zq_context* c, int n2, double* D, int nld2, double* eig) try {
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
  ColRun run{reinterpret_cast<char*>(dD), reinterpret_cast<char*>(D), (size_t)nld2 * 16, colb, n, 0};
  if ((rc = h2d_runs(c, {run}, s))) return rc;
  if ((rc = solve_device(c, n, dD, n2, c->deig, true))) return rc;
  run.ncols = n2;
  if ((rc = d2h_runs(c, {run}, s))) return rc;
  ZQ_HIP_CHECK(hipMemcpyAsync(eig, c->deig, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  int info = 0;
  if (c->info_dev) ZQ_HIP_CHECK(hipMemcpyAsync(&info, c->info_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  ZQ_HIP_CHECK(hipStreamSynchronize(s));
  return info_code(info);
} catch (...) {
  return exception_status();
}

int zq_zquatev_batched_device(zq_context* c, int count, const int* n2, double* const* D, const int* ld2,
                              double* const* eig, int jobz, int* info) try {
  int rc = check_batch(c, count, n2, D, ld2, eig, info);
  if (rc) return rc;
  if (jobz != 'V' && jobz != 'v' && jobz != 'N' && jobz != 'n') return -7;
  if (count == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  CaptureScope cap(c);
  if (cap.rc) return cap.rc;
  if (cap.active && info) return ZQ_ERR_GRAPH;   // info makes the call wait for the result
  return solve_batched(c, count, n2, D, ld2, eig, jobz == 'V' || jobz == 'v', info, nullptr);
} catch (...) {
  return exception_status();
}

int zq_zquatev_batched(zq_context* c, int count, const int* n2, double* const* D, const int* ld2, double* const* eig,
                       int* info) try {
  int rc = check_batch(c, count, n2, D, ld2, eig, info);
  if (rc) return rc;
  if (count == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  // device copies, each problem tight (ld = n2), one allocation
  std::vector<size_t> off(count + 1, 0), eoff(count + 1, 0);
  for (int i = 0; i < count; ++i) {
    off[i + 1] = off[i] + (((size_t)n2[i] * n2[i] * 16 + 255) & ~size_t(255));
    eoff[i + 1] = eoff[i] + (size_t)(n2[i] / 2);
  }
  if ((rc = ensure(c->dD, c->dD_bytes, off[count] + 256))) return rc;
  if ((rc = ensure(reinterpret_cast<void*&>(c->deig), c->deig_bytes, (eoff[count] + 1) * sizeof(double)))) return rc;
  hipStream_t s = c->stream;
  std::vector<double*> dptr(count), eptr(count);
  std::vector<int> dld(count);
  std::vector<ColRun> runs;
  for (int i = 0; i < count; ++i) {
    dptr[i] = reinterpret_cast<double*>(static_cast<char*>(c->dD) + off[i]);
    eptr[i] = c->deig + eoff[i];
    dld[i] = n2[i] > 0 ? n2[i] : 1;
    if (n2[i] > 0)
      runs.push_back({reinterpret_cast<char*>(dptr[i]), reinterpret_cast<char*>(D[i]), (size_t)ld2[i] * 16,
                      (size_t)n2[i] * 16, n2[i] / 2, 0});
  }
  if ((rc = h2d_runs(c, runs, s))) return rc;
  std::vector<int> inf(count, 0);
  const int solved = solve_batched(c, count, n2, dptr.data(), dld.data(), eptr.data(), true, inf.data(), nullptr);
  if (solved < 0) return solved;
  for (ColRun& r : runs) r.ncols *= 2;
  // eigenvalues: one run per problem of one "column" of n doubles
  for (int i = 0; i < count; ++i)
    if (n2[i] > 0)
      runs.push_back({reinterpret_cast<char*>(eptr[i]), reinterpret_cast<char*>(eig[i]), 0,
                      (size_t)(n2[i] / 2) * sizeof(double), 1, 0});
  if ((rc = d2h_runs(c, runs, s))) return rc;
  if (info)
    for (int i = 0; i < count; ++i) info[i] = inf[i];
  return solved;
} catch (...) {
  return exception_status();
}

int zq_zquatev(int n2, double* D, int nld2, double* eig) try {
  int rc = check_args(n2, D, nld2, eig);
  if (rc || n2 == 0) return rc;
  // one solve at a time on the shared context
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx) {
    int rc = zq_create(&g_ctx, 0);
    if (rc) return rc;
  }
  return zq_zquatev_ctx(g_ctx, n2, D, nld2, eig);
} catch (...) {
  return exception_status();
}

int zq_eigvals_noalloc(int n2, const double* dD, int nld2, double* deig, int* dinfo, void* stream) try {
  int rc = check_args(n2, dD, nld2, deig);
  if (rc) return rc;
  if (n2 > ZQ_NOALLOC_MAX_N2) return -5;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n2 == 0) return dinfo ? zq::fill_zero(dinfo, sizeof(int), s) : ZQ_OK;
  const int n = n2 / 2;
  const int nb = zq_tuned_panel_width(n);
  Plan p;
  if (layout_plan(n, nb, false, nullptr, &p) > kScratchBytes) return -5;   // checked before any device call
  char* base = static_cast<char*>(scratch_on_current_device());
  if (!base) return ZQ_ERR_HIP;
  layout_plan(n, nb, false, base, &p);
  // the kernels of zq_zquatev_device(jobz = 'N') with the built-in panel width
  if ((rc = zq::fill_zero(p.sw.info, sizeof(int), s))) return rc;   // the status word starts at 0
  if ((rc = zq::load_left_half(n, dD, nld2, p.M, p.ldm, p.sw.info, s))) return rc;
  if ((rc = zq::reduce(p.M, n, p.ldm, nb, p.d, p.e, p.rw, s))) return rc;
  if ((rc = zq::offdiag_abs(n, p.e, p.ea, s))) return rc;
  if ((rc = zq::eigvals_one(n, p.d, p.ea, deig, p.sw, s))) return rc;
  if (dinfo && (rc = zq::status_out(p.sw.info, dinfo, s))) return rc;
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_tridiagonalize(zq_context* c, int n2, const double* dD, int nld2, double* dd, double* de) try {
  if (!c) return ZQ_ERR_ARG;
  int rc = check_args(n2, dD, nld2, dd);
  if (rc) return rc;
  if (n2 == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  const int n = n2 / 2;
  const int nb = panel_width(c, n);
  Plan p;
  if ((rc = make_plan(c, n, nb, false, &p))) return rc;
  p.rw.tsm = matvec_switch(c, n);
  zq::TileScope tiles(gemm64_min(c, n));
  hipStream_t s = c->stream;
  if ((rc = zq::fill_zero(p.sw.info, sizeof(int), s))) return rc;
  if ((rc = zq::load_left_half(n, dD, nld2, p.M, p.ldm, p.sw.info, s))) return rc;
  c->info_dev = p.sw.info;
  c->last_kind = 1;
  if ((rc = zq::reduce(p.M, n, p.ldm, nb, p.d, p.e, p.rw, s))) return rc;
  if ((rc = zq::offdiag_abs(n, p.e, p.ea, s))) return rc;
  ZQ_HIP_CHECK(hipMemcpyAsync(dd, p.d, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (n > 1 && de)
    ZQ_HIP_CHECK(hipMemcpyAsync(de, p.ea, (size_t)(n - 1) * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_qgemm(zq_context* c, int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
             const double* B, int ldb, double beta, double* C, int ldc, int lower_only) try {
  if (!c) return ZQ_ERR_ARG;
  const bool ta = transa == 'C' || transa == 'c', tb = transb == 'C' || transb == 'c';
  if (!ta && transa != 'N' && transa != 'n') return -2;
  if (!tb && transb != 'N' && transb != 'n') return -3;
  if (m < 0 || n < 0 || k < 0) return -4;
  if (m == 0 || n == 0) return ZQ_OK;
  if (k > 0 && (!A || !B)) return -5;
  if (!C) return -6;
  if (lda < std::max(1, ta ? k : m) || ldb < std::max(1, tb ? n : k) || ldc < std::max(1, m)) return -7;
  int rc;
  if ((rc = with_device(c))) return rc;
  const quat* qa = reinterpret_cast<const quat*>(A);
  const quat* qb = reinterpret_cast<const quat*>(B);
  const zq::QOperand oa{qa, qa, lda, lda, INT_MAX, ta, INT_MIN / 2};
  const zq::QOperand ob{qb, qb, ldb, ldb, INT_MAX, tb, INT_MIN / 2};
  // tile rows by the context's GEMM threshold for trailing size m (zq_set_tile_shape)
  zq::TileScope tiles(gemm64_min(c, m));
  // k == 0 runs no K stage: C = beta C
  return zq::qgemm(m, n, k, alpha, oa, ob, beta, reinterpret_cast<quat*>(C), ldc, lower_only != 0, c->stream);
} catch (...) {
  return exception_status();
}

int zq_tridiag_eig(zq_context* c, int n, const double* dd, const double* de, double* dw, double* dZ) try {
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
  if ((rc = zq::fill_zero(p.sw.info, sizeof(int), s))) return rc;
  c->info_dev = p.sw.info;
  c->last_kind = 1;
  const int zero = 0;
  p.sw.stage = &c->stage;
  apply_caps(c, p.sw);
  if ((rc = zq::stage_begin(c->stage))) return rc;
  if ((rc = zq::stedc(1, &zero, &n, n, n, dd, de, dw, p.sw, vec, s))) return rc;
  if ((rc = zq::stage_end(c->stage, s))) return rc;
  if (vec) ZQ_HIP_CHECK(hipMemcpyAsync(dZ, p.sw.Z, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ZQ_OK;
} catch (...) {
  return exception_status();
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
  if (!idle_unstarted(c) && hipStreamSynchronize(c->stream) != hipSuccess) return ZQ_ERR_HIP;
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
  if (idle_unstarted(c)) return ZQ_OK;
  return hipStreamSynchronize(c->stream) == hipSuccess ? ZQ_OK : ZQ_ERR_HIP;
}

int zq_last_info(zq_context* c) try {
  if (!c) return ZQ_ERR_ARG;
  if (idle_unstarted(c)) return ZQ_OK;
  if (with_device(c)) return ZQ_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return ZQ_ERR_HIP;
  if (c->last_kind == 1 && c->info_dev) {
    int w = 0;
    ZQ_HIP_CHECK(hipMemcpy(&w, c->info_dev, sizeof(int), hipMemcpyDeviceToHost));
    return info_code(w);
  }
  if (c->last_kind == 2 && c->binfo && c->last_count > 0) {
    std::vector<int> w(c->last_count);
    ZQ_HIP_CHECK(hipMemcpy(w.data(), c->binfo, w.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int& v : w) v = info_code(v);
    return batch_status(w.data(), c->last_count);
  }
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_set_iteration_caps(zq_context* c, int leaf, int secular, int bisect) {
  if (!c) return ZQ_ERR_ARG;
  if (leaf < 0 || secular < 0 || bisect < 0) return -2;
  c->caps[0] = leaf;
  c->caps[1] = secular;
  c->caps[2] = bisect;
  return ZQ_OK;
}

int zq_set_workspace_limit(zq_context* c, size_t bytes) {
  if (!c) return ZQ_ERR_ARG;
  c->ws_limit = bytes;
  return ZQ_OK;
}

int zq_batch_workspace(zq_context* c, int count, const int* n2, int jobz, size_t budget, size_t* bytes,
                       int* run) try {
  if (count < 0) return -2;
  if (count > 0 && !n2) return -3;
  if (jobz != 'V' && jobz != 'v' && jobz != 'N' && jobz != 'n') return -4;
  for (int i = 0; i < count; ++i)
    if (n2[i] < 0 || (n2[i] & 1)) return -3;
  const bool vectors = jobz == 'V' || jobz == 'v';
  Groups G = make_groups(c, count, n2, nullptr, nullptr, nullptr, vectors);
  if (budget > 0 && G.need > budget) split_groups(G, vectors, budget);
  if (bytes) *bytes = G.need;
  if (run) {
    for (int i = 0; i < count; ++i) run[i] = -1;
    for (size_t g = 0; g < G.probs.size(); ++g)
      for (const zq::BatchProb& b : G.probs[g]) run[b.idx] = (int)g;
  }
  return (int)G.probs.size();
} catch (...) {
  return exception_status();
}

}  // extern "C"

namespace ts {
int zquatev(const int n2, std::complex<double>* const D, const int nld2, double* const eig) {
  return zq_zquatev(n2, reinterpret_cast<double*>(D), nld2, eig);
}
}  // namespace ts

namespace zq {
int batched_async(zq_context* c, int count, const int* n2, double* const* D, const int* ld2, double* const* eig,
                  int jobz, int* info_async) {
  int rc = check_batch(c, count, n2, D, ld2, eig, nullptr);
  if (rc) return rc;
  if (jobz != 'V' && jobz != 'v' && jobz != 'N' && jobz != 'n') return -7;
  if (count == 0) return ZQ_OK;
  if ((rc = with_device(c))) return rc;
  return solve_batched(c, count, n2, D, ld2, eig, jobz == 'V' || jobz == 'v', nullptr, info_async);
}

size_t batched_workspace(const zq_context* c, int count, const int* n2, int jobz) {
  return make_groups(c, count, n2, nullptr, nullptr, nullptr, jobz == 'V' || jobz == 'v').need;
}

hipStream_t context_stream(const zq_context* c) { return c->stream; }

int context_reserve(zq_context* c, int max_count, size_t stage_bytes) {
  int rc;
  if ((rc = with_device(c))) return rc;
  if ((rc = ensure(reinterpret_cast<void*&>(c->binfo), c->binfo_bytes, (size_t)max_count * sizeof(int) + 256)))
    return rc;
  for (int k = 0; k < 2; ++k) {
    if (c->stage.cap[k] >= stage_bytes) continue;
    if (c->stage.host[k]) {
      if (hipStreamSynchronize(c->stream) != hipSuccess) return ZQ_ERR_HIP;
      (void)hipHostFree(c->stage.host[k]);
      c->stage.host[k] = nullptr;
      c->stage.cap[k] = 0;
    }
    void* h = nullptr;
    if (hipHostMalloc(&h, stage_bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return ZQ_ERR_NOMEM;
    }
    c->stage.host[k] = static_cast<char*>(h);
    c->stage.cap[k] = stage_bytes;
  }
  return ZQ_OK;
}

void context_set_defer_free(zq_context* c, bool on) { c->defer_free = on; }

int context_free_retired(zq_context* c) {
  if (c->retired.empty()) return ZQ_OK;
  if (with_device(c)) return ZQ_ERR_HIP;
  int rc = ZQ_OK;
  for (void* p : c->retired)
    if (hipFree(p) != hipSuccess) rc = ZQ_ERR_HIP;
  c->retired.clear();
  return rc;
}

int context_release_workspace(zq_context* c) {
  if (c->ws_locked || !c->ws) return ZQ_OK;
  if (with_device(c)) return ZQ_ERR_HIP;
  const hipError_t e = hipFree(c->ws);
  c->ws = nullptr;
  c->ws_bytes = 0;
  return e == hipSuccess ? ZQ_OK : ZQ_ERR_HIP;
}
}  // namespace zq
