// Internal declarations shared by the HIP translation units of libzquatev_amd.
//
// Storage conventions (all device memory, column-major):
//   * A quaternion q = (a, b) with a, b complex is the 2x2 complex block
//     [[a, -conj b], [b, conj a]]; it is stored as four doubles
//     (Re a, Im a, Re b, Im b).  The quaternion-Hermitian matrix
//     H = [[A, -conj B], [B, conj A]] of zquatev.h:41-52 is then the n x n
//     quaternion matrix M with M_ij = (A_ij, B_ij), and the eigenvector
//     matrix [[U, -conj V], [V, conj U]] is the quaternion matrix X_ij = (U_ij, V_ij).
//   * Products: (a1,b1)(a2,b2) = (a1 a2 - conj(b1) b2, b1 a2 + conj(a1) b2);
//     conjugate (a,b)^* = (conj a, -b); |q|^2 = |a|^2 + |b|^2.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <hip/hip_ext.h>
#include <cstdint>
#include <cstddef>

namespace zq {

struct __align__(16) quat {
  double ar, ai, br, bi;
};

__device__ __forceinline__ quat qzero() { return quat{0.0, 0.0, 0.0, 0.0}; }

__device__ __forceinline__ quat qload(const quat* p) {
  const double2* q = reinterpret_cast<const double2*>(p);
  double2 x = q[0];
  double2 y = q[1];
  return quat{x.x, x.y, y.x, y.y};
}

__device__ __forceinline__ void qstore(quat* p, const quat& v) {
  double2* q = reinterpret_cast<double2*>(p);
  q[0] = make_double2(v.ar, v.ai);
  q[1] = make_double2(v.br, v.bi);
}

// Loads and stores through pointers read from tables in device memory (task
// and problem tables): the compiler cannot tell their address space and would
// emit flat instructions; every one of them points to global memory.  The
// accesses go through global-space pointers to scalars or native vectors (a
// struct copy would become a generic memcpy).
#define ZQ_GLOBAL __attribute__((address_space(1)))
typedef double zq_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 d2load_g(const double2* p) {
  const zq_d2 v = *(const ZQ_GLOBAL zq_d2*)p;
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void d2store_g(double2* p, double x, double y) {
  const zq_d2 v = {x, y};
  *(ZQ_GLOBAL zq_d2*)p = v;
}
__device__ __forceinline__ void dstore_g(double* p, double x) { *(ZQ_GLOBAL double*)p = x; }
__device__ __forceinline__ quat qload_g(const quat* p) {
  const ZQ_GLOBAL zq_d2* g = (const ZQ_GLOBAL zq_d2*)p;
  const zq_d2 x = g[0], y = g[1];
  return quat{x.x, x.y, y.x, y.y};
}
__device__ __forceinline__ void qstore_g(quat* p, const quat& v) {
  ZQ_GLOBAL zq_d2* g = (ZQ_GLOBAL zq_d2*)p;
  const zq_d2 x = {v.ar, v.ai}, y = {v.br, v.bi};
  g[0] = x;
  g[1] = y;
}

__device__ __forceinline__ quat qadd(const quat& x, const quat& y) {
  return quat{x.ar + y.ar, x.ai + y.ai, x.br + y.br, x.bi + y.bi};
}
__device__ __forceinline__ quat qsub(const quat& x, const quat& y) {
  return quat{x.ar - y.ar, x.ai - y.ai, x.br - y.br, x.bi - y.bi};
}
__device__ __forceinline__ quat qscale(const quat& x, double s) {
  return quat{x.ar * s, x.ai * s, x.br * s, x.bi * s};
}
__device__ __forceinline__ quat qconj(const quat& x) { return quat{x.ar, -x.ai, -x.br, -x.bi}; }
__device__ __forceinline__ double qnorm2(const quat& x) {
  return x.ar * x.ar + x.ai * x.ai + x.br * x.br + x.bi * x.bi;
}

// x * y
__device__ __forceinline__ quat qmul(const quat& x, const quat& y) {
  quat r;
  r.ar = x.ar * y.ar - x.ai * y.ai - x.br * y.br - x.bi * y.bi;
  r.ai = x.ar * y.ai + x.ai * y.ar + x.bi * y.br - x.br * y.bi;
  r.br = x.br * y.ar - x.bi * y.ai + x.ar * y.br + x.ai * y.bi;
  r.bi = x.br * y.ai + x.bi * y.ar + x.ar * y.bi - x.ai * y.br;
  return r;
}
// acc += x * y
__device__ __forceinline__ void qfma(quat& acc, const quat& x, const quat& y) {
  acc.ar += x.ar * y.ar - x.ai * y.ai - x.br * y.br - x.bi * y.bi;
  acc.ai += x.ar * y.ai + x.ai * y.ar + x.bi * y.br - x.br * y.bi;
  acc.br += x.br * y.ar - x.bi * y.ai + x.ar * y.br + x.ai * y.bi;
  acc.bi += x.br * y.ai + x.bi * y.ar + x.ar * y.bi - x.ai * y.br;
}
// acc += conj(x) * y
__device__ __forceinline__ void qfma_cl(quat& acc, const quat& x, const quat& y) {
  qfma(acc, qconj(x), y);
}
// acc += x * conj(y)
__device__ __forceinline__ void qfma_cr(quat& acc, const quat& x, const quat& y) {
  qfma(acc, x, qconj(y));
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
  return __shfl_xor(v, m, 64);
}

// ---- cross-lane exchanges without the LDS crossbar (gfx950) ----------------
// xor 1, 2, 4, 8 inside a 16-lane row through DPP; the row-crossing steps of a
// butterfly through v_permlane16_swap / v_permlane32_swap.
__device__ __forceinline__ unsigned lo32(double v) { return (unsigned)(unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ unsigned hi32(double v) {
  return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32);
}
__device__ __forceinline__ double mkd(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// value of lane k (wave-uniform k) of a per-lane double, in scalar registers
__device__ __forceinline__ double lane_d(double v, int k) {
  return mkd((unsigned)__builtin_amdgcn_readlane((int)lo32(v), k), (unsigned)__builtin_amdgcn_readlane((int)hi32(v), k));
}
template <int CTRL>
__device__ __forceinline__ double dppd(double v) {
  return mkd(__builtin_amdgcn_update_dpp(0u, lo32(v), CTRL, 0xF, 0xF, false),
             __builtin_amdgcn_update_dpp(0u, hi32(v), CTRL, 0xF, 0xF, false));
}
// xor 4: row_shl:4 on banks 0 and 2, row_shr:4 on banks 1 and 3
__device__ __forceinline__ double dpp_xor4(double v) {
  unsigned l = __builtin_amdgcn_update_dpp(0u, lo32(v), 0x104, 0xF, 0x5, false);
  l = __builtin_amdgcn_update_dpp(l, lo32(v), 0x114, 0xF, 0xA, false);
  unsigned h = __builtin_amdgcn_update_dpp(0u, hi32(v), 0x104, 0xF, 0x5, false);
  h = __builtin_amdgcn_update_dpp(h, hi32(v), 0x114, 0xF, 0xA, false);
  return mkd(l, h);
}
// value of v in lane (lane ^ M), M in {1, 2, 4, 8}
template <int M>
__device__ __forceinline__ double xlane(double v) {
  if (M == 8) return dppd<0x128>(v);   // row_ror:8
  if (M == 4) return dpp_xor4(v);
  if (M == 2) return dppd<0x4E>(v);    // quad_perm [2,3,0,1]
  return dppd<0xB1>(v);                // quad_perm [1,0,3,2]
}
// Butterfly exchange across 32 lanes: lanes 0..31 return A + A(lane^32),
// lanes 32..63 return B(lane^32) + B.
__device__ __forceinline__ double bf32(double A, double B) {
  auto l = __builtin_amdgcn_permlane32_swap(lo32(A), lo32(B), false, false);
  auto h = __builtin_amdgcn_permlane32_swap(hi32(A), hi32(B), false, false);
  return mkd(l[0], h[0]) + mkd(l[1], h[1]);
}
// Same across 16 lanes (even rows keep A, odd rows keep B).
__device__ __forceinline__ double bf16(double A, double B) {
  auto l = __builtin_amdgcn_permlane16_swap(lo32(A), lo32(B), false, false);
  auto h = __builtin_amdgcn_permlane16_swap(hi32(A), hi32(B), false, false);
  return mkd(l[0], h[0]) + mkd(l[1], h[1]);
}
template <int M, int CNT>
__device__ __forceinline__ void bstep(double* v, int lane) {
  if (M == 32) {
#pragma unroll
    for (int k = 0; k < CNT; ++k) v[k] = bf32(v[k], v[k + CNT]);
  } else if (M == 16) {
#pragma unroll
    for (int k = 0; k < CNT; ++k) v[k] = bf16(v[k], v[k + CNT]);
  } else {
    const bool hi = (lane & M) != 0;
#pragma unroll
    for (int k = 0; k < CNT; ++k) {
      const double send = hi ? v[k] : v[k + CNT];
      const double keep = hi ? v[k + CNT] : v[k];
      v[k] = keep + xlane<M>(send);
    }
  }
}
// Halving butterfly over a wave for N (power of two <= 64) values per lane:
// lane l returns value (l >> (6 - log2 N)) summed over all 64 lanes.
template <int N>
__device__ __forceinline__ double butterfly_sum(double* v) {
  const int lane = threadIdx.x & 63;
  if (N >= 2) bstep<32, (N >= 2 ? N / 2 : 1)>(v, lane);
  if (N >= 4) bstep<16, (N >= 4 ? N / 4 : 1)>(v, lane);
  if (N >= 8) bstep<8, (N >= 8 ? N / 8 : 1)>(v, lane);
  if (N >= 16) bstep<4, (N >= 16 ? N / 16 : 1)>(v, lane);
  if (N >= 32) bstep<2, (N >= 32 ? N / 32 : 1)>(v, lane);
  if (N >= 64) bstep<1, 1>(v, lane);
  // lane bits not consumed by the halving steps still hold separate partial sums
  double r = v[0];
  if (N <= 2) r = bf16(r, r);
  if (N <= 4) r += xlane<8>(r);
  if (N <= 8) r += xlane<4>(r);
  if (N <= 16) r += xlane<2>(r);
  if (N <= 32) r += xlane<1>(r);
  if (N == 1) r = bf32(r, r);
  return r;
}
// Halving butterfly inside each 16-lane row for N (power of two, >= 16)
// values per lane: afterwards v[t], t < N/16, of lane l holds the row sum of
// value (l & 15) * (N/16) + t.
template <int N>
__device__ __forceinline__ void row16_halving(double* v) {
  const int lane = threadIdx.x & 63;
  bstep<8, N / 2>(v, lane);
  bstep<4, N / 4>(v, lane);
  bstep<2, N / 8>(v, lane);
  bstep<1, N / 16>(v, lane);
}
// sum over all 64 lanes, result in every lane
__device__ __forceinline__ double wave_sum_d(double v) {
  v += xlane<1>(v);
  v += xlane<2>(v);
  v += xlane<4>(v);
  v += xlane<8>(v);
  v = bf16(v, v);
  return bf32(v, v);
}
__device__ __forceinline__ quat wave_sum(quat v) {
  return quat{wave_sum_d(v.ar), wave_sum_d(v.ai), wave_sum_d(v.br), wave_sum_d(v.bi)};
}

}  // namespace zq

// ---------------------------------------------------------------------------
// Host-side plumbing
// ---------------------------------------------------------------------------
// status codes (mirrored in include/zq.h)
#define ZQI_OK 0
#define ZQI_ERR_HIP (-1001)
#define ZQI_ERR_NOMEM (-1002)
#define ZQI_ERR_ARG (-1003)
#define ZQI_ERR_GRAPH (-1004)
#define ZQI_ERR_NONFINITE (-1005)
// Per-problem device status word: the low bits hold a positive
// non-convergence code (atomicMax), this bit marks a NaN / infinity met in
// the input or in the reduced tridiagonal (atomicOr).  info_code() turns a
// word into the code the entry points return.
#define ZQI_INFO_NONFINITE (1 << 30)
__host__ __device__ inline int info_code(int word) {
  return (word & ZQI_INFO_NONFINITE) ? ZQI_ERR_NONFINITE : word;
}

#define ZQ_HIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return (_e == hipErrorOutOfMemory) ? ZQI_ERR_NOMEM : ZQI_ERR_HIP; \
  } while (0)

// Status for an exception caught at the C boundary (host containers, threads):
// allocation and resource failures are ZQI_ERR_NOMEM, anything else
// ZQI_ERR_HIP.  Called only from inside a catch handler.
int exception_status() noexcept;

namespace zq {

// Optional per-kernel-class timing with HIP events (enabled per context).
// Classes: symv (panel matvec), trail (trailing MFMA update), backtr (MFMA
// back-transform products), merge (D&C merge GEMM), other.  A ProfScope arms
// one event pair; the first launch_k inside the scope hands the pair to
// hipExtLaunchKernelGGL, so the recorded interval is exactly that kernel's
// execution (no launch gaps).
enum ProfClass { PROF_SYMV = 0, PROF_TRAIL = 1, PROF_BACKTR = 2, PROF_MERGE = 3, PROF_NCLASS = 4 };
struct ProfScope {
  int cls;
  hipStream_t s;
  double units;
  int slot;
  ProfScope(int cls, hipStream_t s, double units);
  ~ProfScope();
};
// events of the armed scope (nullptr when none); disarms it
void prof_take(hipEvent_t* e0, hipEvent_t* e1);

template <typename F, typename... Args>
inline void launch_k(F kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  prof_take(&e0, &e1);
  hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, e0, e1, 0u, args...);
}

struct Staging;

// Batch addressing of the per-column and per-problem kernels: problem
// z = blockIdx.z has half-dimension nvec[z] and its buffers at byte offset
// z * zs from the pointers passed (problem 0's).  The single solve passes
// {nullptr, 0}, which leaves every pointer and size as given.
struct ZB {
  const int* nvec;
  size_t zs;
};
// Byte arithmetic on the pointer itself (no round trip through an integer),
// so the compiler keeps kernel arguments in the global address space and
// emits global rather than flat memory instructions.
template <class T>
__device__ __forceinline__ T* zoff(T* p, const ZB& zb) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<B*>(p) + (size_t)blockIdx.z * zb.zs);
}

// reduce.hip: quaternion Householder tridiagonalisation of the lower triangle
// of M (n x n, ld).  Writes d (n), e (n-1 quaternions: the subdiagonal) and
// leaves the reflector u_c in M[c+1:, c].
struct ReduceWork {
  quat* W;        // n x nb (global row index)
  quat* Prow;     // matvec row partials: ceil(n/32) x n
  quat* Pcol;     // matvec column partials: ceil(n/64) x n
  quat* TS;       // tall-skinny products [W^* u | V^* u]: 2 x nb
  double* TSP;    // their per-block partials: ceil(n/16) x ldtsp
  int ldtsp;      // 8 nb
  double* QF;     // per-tile partials of x^* M x: ceil(n/64) (ceil(n/64) + 1)
  double* part;   // ceil(n/16)
  double* scal;   // 16
  int tsm = 768;  // matvec strip switch: narrow strips up to this trailing size (0 .. 2048)
};
int reduce(quat* M, int n, int ld, int nb, double* d, quat* e, const ReduceWork& w, hipStream_t s);
struct QGemmTask;
// Lockstep reduction of a batch of nprob problems sorted by n descending (nh:
// host copy of the sizes, zb.nvec the device copy); every buffer of problem z
// lies zb.zs bytes after problem z-1's (the pointers given are problem 0's,
// ld = nh[0]).  tasks: device room for the trailing-update task table
// (reduce_batched_tasks entries); tstart: device room for nprob ints.  Each
// problem's result equals its solo reduce().
size_t reduce_batched_tasks(int nprob, const int* nh, int nb);
int reduce_batched(int nprob, const int* nh, ZB zb, quat* M, int ld, int nb, double* d, quat* e, const ReduceWork& w,
                   QGemmTask* tasks, int* tstart, Staging& st, hipStream_t s);

// stedc.hip.  Several independent tridiagonals (a forest) are solved
// together: problem z owns the global index range [o_z, o_z + n_z) of every
// 1-D array, and its eigenvector block Z[o_z:o_z+n_z, 0:n_z] (ld = total size).
// A merge of global rows [lo, hi) with split mid belongs to the problem with
// base ob.
struct MergeDesc {
  int lo, mid, hi, ob, pr;   // rows [lo, hi) split at mid; problem base row ob and index pr
};
struct Staging;
struct StedcWork {
  double* Z;      // n x n eigenvectors (ld n)
  double* Qtmp;   // n x n
  double* U;      // n x n
  double *lam, *es, *e2, *dk, *zk, *dkk, *zkk, *dval, *tau, *zhat, *rho;  // n each
  double* rot;    // 4 n
  double* scal;   // 16
  int *src, *keep, *dfl, *org, *posK, *posD;                             // n each
  int *nrot, *kcnt;                                                      // n each
  // merge GEMM structure: mixed flag per sorted position, GEMM position of
  // each kept column and its inverse, and per merge the ends of the groups
  // of kept columns from the first half and from both halves (2 per merge)
  int *mix, *gpos, *ginv, *kgrp;                                         // n each
  int* info;      // one status word per problem (ZQI_INFO_NONFINITE | non-convergence: failing local row + 1)
  int* ints;      // tree descriptors
  size_t int_slots;
  Staging* stage; // host staging of the descriptors
  // iteration caps: QL sweeps per leaf eigenvalue, secular-equation steps
  // per root, multisection steps per eigenvalue (reaching one is reported
  // as non-convergence, info > 0)
  int cap_leaf = 60, cap_secular = 200, cap_bisect = 64;
};
size_t stedc_int_slots(int ntot, int nprob);
// offs/ns (host): problem offsets and sizes; ntot = sum of ns; lq >= max ns.
// d, e: concatenated diagonals and |subdiagonals| (e[o_z + i], i < n_z - 1);
// w (ntot): eigenvalues ascending per problem.  Z (when vectors) is ld ntot.
int stedc(int nprob, const int* offs, const int* ns, int ntot, int lq, const double* d, const double* e, double* w,
          const StedcWork& sw, bool vectors, hipStream_t s);
// Eigenvalues of one tridiagonal of order n, the kernels stedc runs for it
// without vectors, with nothing staged from the host (sw.ints holds the
// problem's (offset, size), written by a kernel).
int eigvals_one(int n, const double* d, const double* e, double* w, const StedcWork& sw, hipStream_t s);
// *out = info_code(*word) on the device
int status_out(const int* word, int* out, hipStream_t s);

// Pinned host staging for host -> device descriptor tables.  Two buffers
// alternate between solves; stage_begin waits until the solve that last used
// the buffer (two solves back) has completed on its stream, so the host can
// run one solve ahead without blocking.
struct Staging {
  char* host[2] = {nullptr, nullptr};
  size_t cap[2] = {0, 0};
  hipEvent_t done[2] = {nullptr, nullptr};
  int slot = 0;
  size_t used = 0;
  // While the stream is being captured into a graph, uploads are recorded as
  // copies from this pinned pool instead (the graph re-reads them at every
  // replay, so they must stay as written); nothing is allocated or waited on.
  bool capture = false;
  char* keep = nullptr;
  size_t keep_cap = 0, keep_used = 0;
};
int stage_begin(Staging& st);
int stage_upload(Staging& st, const void* src, size_t bytes, void* dst, hipStream_t s);
int stage_end(Staging& st, hipStream_t s);
void stage_free(Staging& st);

// backtr.hip: X = Q Phi Z with the reflectors u_c kept in M, applied in
// blocks of wbt consecutive reflectors (compact WY, I - V T V^*).
struct BacktrWork {
  quat* phi;          // n
  quat* part;         // split-K partials: part_slots
  size_t part_slots;  // >= nsplit * wbt * n
  int nsplit;
  int wbt;
  quat* Y;            // wbt x n
  quat* G;            // nblk x wbt x wbt: V^* V per block
  quat* T;            // nblk x wbt x wbt: T = (I + triu(G, 1))^{-1}
  quat* VT;           // V T^* of every block, block b at vt_offset(b), ld = its row count
  size_t vt_slots;
};
size_t backtr_vt_slots(int n, int wbt);
struct SumTask {
  const quat* parts;
  quat* out;
  int total;
  int nz;
};
// task-table sizes for a set of problems (sizes nh descending)
size_t backtransform_batched_tasks(int nprob, const int* nh, int wbt, int nsplit);
size_t backtransform_batched_sums(int nprob, const int* nh, int wbt);
// Problems as for reduce_batched (zb; a single solve passes nprob = 1 and
// ZB{nullptr, 0}); problem z's real eigenvectors are Z + zrow[z] (ld ldz),
// zrow_d the device copy (null for a single solve).  tasks / sums: device
// room for the task tables.
int backtransform_batched(int nprob, const int* nh, ZB zb, const int* zrow_d, const quat* M, int ld, const quat* e,
                          const double* Z, int ldz, quat* X, int ldx, const BacktrWork& w, QGemmTask* tasks,
                          SumTask* sums, Staging& st, hipStream_t s);

// layout.hip
// One problem of a batch: the caller's D (in and out, ld2 complex), its
// eigenvalue output, half-dimension n, and its offset in the concatenated
// tridiagonal (forest) arrays.
struct BatchProb {
  double* D;
  double* eig;
  int ld2;
  int n;
  int off;
  int idx;   // position in the caller's arrays
};
// The loads replace a NaN or infinite input entry by 0 and set
// ZQI_INFO_NONFINITE in the problem's status word (info[z]), so that every
// later kernel runs on finite data.
int load_left_batched(int nprob, int L, const BatchProb* P, quat* M, int ld, ZB zb, int* info, hipStream_t s);
int store_vectors_batched(int nprob, int L, const BatchProb* P, const quat* X, int ldx, ZB zb, hipStream_t s);
// dall[off + i] = d_z[i], eall[off + i] = |e_z[i]|
int gather_tridiag_batched(int nprob, int L, const BatchProb* P, const double* d, const quat* e, ZB zb, double* dall,
                           double* eall, hipStream_t s);
int scatter_eig_batched(int nprob, int L, const BatchProb* P, const double* wall, hipStream_t s);
// info_all[P[z].idx] = info[z]
int scatter_info_batched(int nprob, const BatchProb* P, const int* info, int* info_all, hipStream_t s);
int load_left_half(int n, const double* D, int ld2, quat* M, int ld, int* info, hipStream_t s);
int store_vectors(int n, const quat* X, int ldx, double* D, int ld2, hipStream_t s);
int offdiag_abs(int n, const quat* e, double* ea, hipStream_t s);
// zero `bytes` (a multiple of 4) at p with a kernel (no memset node in graphs)
int fill_zero(void* p, size_t bytes, hipStream_t s);
// synthetic U(-1,1) lower triangle (real diagonal) for timing runs
int gen_lower_uniform(int n, unsigned long long seed, quat* M, int ld, hipStream_t s);
// partial-sum slots of the reduction's matvec for trailing sizes up to nmax:
// row partials (per column strip), column partials (per row group), and the
// per-tile x^* M x partials
int matvec_row_slots(int nmax);
int matvec_col_slots(int nmax);
int matvec_tile_slots(int nmax);

// gemm.hip: C = beta C + alpha opA(A) opB(B) for quaternion operands (FP64 MFMA).
// An operand is the column concatenation [p0 | p1] split at ksplit along K;
// conj_trans selects conj-transpose use of the storage; storage element
// (r, c) is read only when r > c + mask_off (INT_MIN/2 disables the mask).
struct QOperand {
  const quat* p0;
  const quat* p1;
  int ld0, ld1;
  int ksplit;
  bool conj_trans;
  int mask_off;
};
int qgemm(int m, int nn, int k, double alpha, const QOperand& A, const QOperand& B, double beta, quat* C, int ldc,
          bool lower_only, hipStream_t s);
// One quaternion GEMM piece: C = beta C + alpha opA(A) opB(B) summed over
// logical k in [kbeg, kend); K bounds the operand reads; lower_only writes
// only i >= j.
struct QGemmTask {
  int m, nn, K, kbeg, kend;
  int ldc, lower_only;
  double alpha, beta;
  QOperand A, B;
  quat* C;
};
// ntasks independent tasks (device array) in one launch; maxm / maxnn bound
// their shapes; every task has the given operand conjugation flags.  Each task
// computes exactly what qgemm would for it.
int qgemm_tasks(const QGemmTask* tasks_dev, int ntasks, int maxm, int maxnn, bool a_conj_trans, bool b_conj_trans,
                bool lower_only, hipStream_t s);
// GEMM tile rows: 64-row tiles for launches of at least `g_tile64_per_cu`
// 64 x 64 tiles per compute unit, else 32-row tiles.  A tuning parameter of
// the calling thread (set per solve from its context; 16 by default): the
// tile shape does not change any result, only the time.
extern thread_local int g_tile64_per_cu;
struct TileScope {
  int saved;
  explicit TileScope(int v) : saved(g_tile64_per_cu) { if (v > 0) g_tile64_per_cu = v; }
  ~TileScope() { g_tile64_per_cu = saved; }
};
// number of K chunks (and the chunk length) a split-K product over K uses
int splitk_count(int K, int nsplit, int* kchunk_out);
// real merge products of one divide & conquer tree level (see stedc.hip);
// kgrp: per merge the GEMM positions ending the first-half and both-halves
// groups of kept columns
int dgemm_merges(const MergeDesc* descs_dev, int nmerge, int nsmax, const int* kcnt, const int* kgrp,
                 const double* Qtmp, const double* U, int lq, double* Z, int ldz, const int* posK, hipStream_t s);
int mfma_probe(const double* a, const double* b, double* out, hipStream_t s);

// api.hip hooks for the serving loop (serve.hip)
}  // namespace zq
struct zq_context;
namespace zq {
// zq_zquatev_batched_device without waiting: each problem's status is copied
// into info_async (pinned host, count ints), valid once the context's stream
// has passed the call.
int batched_async(zq_context* c, int count, const int* n2, double* const* D, const int* ld2, double* const* eig,
                  int jobz, int* info_async);
// device workspace bytes such a batch needs on c
size_t batched_workspace(const zq_context* c, int count, const int* n2, int jobz);
hipStream_t context_stream(const zq_context* c);
// Allocates ahead what a batch of up to max_count problems needs besides the
// workspace (status words, pinned descriptor staging of stage_bytes per
// buffer), so that later batches never free and reallocate them.
int context_reserve(zq_context* c, int max_count, size_t stage_bytes);
// Frees the workspace (the context must have no work in flight).
int context_release_workspace(zq_context* c);
// With defer-free on, a workspace replaced by a larger one is retired rather
// than freed (hipFree waits for the whole device); context_free_retired frees
// the retired ones (call while no work is in flight on the device).
void context_set_defer_free(zq_context* c, bool on);
int context_free_retired(zq_context* c);
}  // namespace zq
