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
__device__ __forceinline__ quat wave_sum(quat v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    v.ar += shfl_xor_d(v.ar, m);
    v.ai += shfl_xor_d(v.ai, m);
    v.br += shfl_xor_d(v.br, m);
    v.bi += shfl_xor_d(v.bi, m);
  }
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_d(v, m);
  return v;
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

#define ZQ_HIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return (_e == hipErrorOutOfMemory) ? ZQI_ERR_NOMEM : ZQI_ERR_HIP; \
  } while (0)

namespace zq {

// Optional per-kernel-class timing with HIP events (enabled per context).
// Classes: symv (panel matvec), trail (trailing MFMA update), backtr (MFMA
// back-transform products), merge (D&C merge GEMM), other.
enum ProfClass { PROF_SYMV = 0, PROF_TRAIL = 1, PROF_BACKTR = 2, PROF_MERGE = 3, PROF_NCLASS = 4 };
struct ProfScope {
  int cls;
  hipStream_t s;
  double units;
  int slot;
  ProfScope(int cls, hipStream_t s, double units);
  ~ProfScope();
};

// reduce.hip: quaternion Householder tridiagonalisation of the lower triangle
// of M (n x n, ld).  Writes d (n), e (n-1 quaternions: the subdiagonal),
// Tf (per panel nb x nb compact-WY factors) and leaves u_c in M[c+1:, c].
struct ReduceWork {
  quat* W;        // n x nb (global row index)
  quat* ucur;     // n
  quat* Prow;     // matvec row partials: ceil(n/32) x n
  quat* Pcol;     // matvec column partials: ceil(n/128) x n
  quat* TS;       // tall-skinny partials: ceil(n/128) x 2 x nb
  double* KP;     // ceil(n/64)
  double* part;   // ceil(n/64)
  double* scal;   // 16
  quat* Tf;       // ceil((n-1)/nb) x nb x nb
  quat* Gm;       // same shape: strict upper part of V^* V per panel
};
int reduce(quat* M, int n, int ld, int nb, double* d, quat* e, const ReduceWork& w, hipStream_t s);

// stedc.hip
struct StedcWork {
  double* Z;      // n x n eigenvectors (ld n)
  double* Qtmp;   // n x n
  double* U;      // n x n
  double *lam, *es, *e2, *dk, *zk, *dkk, *zkk, *dval, *tau, *zhat, *rho;  // n each
  double* rot;    // 4 n
  double* scal;   // 16
  int *src, *keep, *dfl, *org, *posK, *posD;                             // n each
  int *nrot, *kcnt;                                                      // n each
  int* info;      // 1
  int* ints;      // tree descriptors
  size_t int_slots;
};
size_t stedc_int_slots(int n);
int stedc(int n, const double* d, const double* e, double* w, const StedcWork& sw, bool vectors, hipStream_t s);

// backtr.hip
struct BacktrWork {
  quat* phi;          // n
  quat* Ypart;        // ypart_slots
  size_t ypart_slots;
  int nsplit;
  quat* Y;            // nb x n
  quat* Y2;           // nb x n
};
int backtransform(int n, int nb, const quat* M, int ld, const quat* Tf, const quat* e, const double* Z, int ldz,
                  quat* X, int ldx, const BacktrWork& w, hipStream_t s);

// layout.hip
int load_left_half(int n, const double* D, int ld2, quat* M, int ld, hipStream_t s);
int store_vectors(int n, const quat* X, int ldx, double* D, int ld2, hipStream_t s);
int offdiag_abs(int n, const quat* e, double* ea, hipStream_t s);

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
// number of K chunks qgemm_cn_splitk uses for a given request
int splitk_count(int K, int nsplit, int* kchunk_out);
// Ypart[z] (m x nn, ld m) = A B over the z-th K chunk (A logical m x K, B K x nn)
int qgemm_cn_splitk(int m, int nn, int K, const QOperand& A, const quat* B, int ldb, quat* Ypart, int nsplit,
                    hipStream_t s);
// real: C[:, cmap[j]] = A (m x K) B (K x K) with K read on the device (<= k_max)
int dgemm_nn(int m, int unused, const int* k_dev, int k_max, const double* A, int lda, const double* B, int ldb,
             double* C, int ldc, const int* cmap, hipStream_t s);
int mfma_probe(const double* a, const double* b, double* out, hipStream_t s);

}  // namespace zq
