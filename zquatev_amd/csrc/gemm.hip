// FP64 MFMA products for the quaternion panel updates and the divide & conquer.
//
// Quaternion GEMM as one real GEMM: with q = (a, b) stored as the 4-vector
// (Re a, Im a, Re b, Im b), the product q p is L(q) p with
//
//   L(q) = [ a_r  -a_i  -b_r  -b_i ]
//          [ a_i   a_r   b_i  -b_r ]
//          [ b_r  -b_i   a_r   a_i ]
//          [ b_i   b_r  -a_i   a_r ]      L(q)[r][k] = sign[r][k] * q[r ^ k].
//
// So C (m x n quaternions) += A B is the real product of L(A) (4m x 4k) with
// B viewed as a 4k x n real matrix, and each v_mfma_f64_16x16x4_f64 consumes
// 4 quaternion rows x 1 quaternion k of A (expanded on the fly from LDS) and
// 1 k x 16 quaternion columns of B.  Register layout of that instruction
// (pinned by zq_selftest_mfma): A[i][k] is in lane i + 16k, B[k][j] in lane
// j + 16k, and accumulator register r of lane l holds D[4r + l/16][l%16].
// Real row 4r + g is taken to be component r of quaternion row g, so each
// lane's four accumulators are one whole quaternion of C.  No flops are spent
// on zeros: 32 real flops per quaternion multiply-add, the count the
// roofline uses.
//
// These kernels replace the complex zgemm3m_ calls of the reference's panel
// update and Q accumulation (blocked.cc:416-544) and back-transform
// (zquatev.cc:86-90).
#include "zq_internal.h"
#include <climits>
#include <cstdlib>

namespace zq {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Tiles of 64 x 64 quaternions per workgroup (2 x 2 waves, 32 x 32
// quaternions = 16 MFMA accumulators per wave), K staged 8 quaternions at a
// time through double-buffered LDS.  The next stage's operands are fetched
// into registers before the current stage's MFMAs, so global latency hides
// behind the wave's own work, and one barrier per stage suffices (a buffer is
// rewritten only after the barrier that follows its last read).  The
// conjugate-transpose flags are template parameters (branch-free address
// math); the sign of the L(A) expansion is applied by flipping the sign bit.
// ---------------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 8;

template <bool CT>
__device__ __forceinline__ quat ld_opA(const QOperand& A, int i, int k, int m, int K) {
  if (i >= m || k >= K) return qzero();
  const quat* p = A.p0;
  int ldp = A.ld0, kk = k;
  if (k >= A.ksplit) { p = A.p1; ldp = A.ld1; kk = k - A.ksplit; }
  const int sr = CT ? kk : i, sc = CT ? i : kk;
  if (!(sr > sc + A.mask_off)) return qzero();
  const quat v = qload(p + sr + (size_t)sc * ldp);
  return CT ? qconj(v) : v;
}

template <bool CT>
__device__ __forceinline__ quat ld_opB(const QOperand& B, int k, int j, int K, int nn) {
  if (k >= K || j >= nn) return qzero();
  const quat* p = B.p0;
  int ldp = B.ld0, kk = k;
  if (k >= B.ksplit) { p = B.p1; ldp = B.ld1; kk = k - B.ksplit; }
  const int sr = CT ? j : kk, sc = CT ? kk : j;
  if (!(sr > sc + B.mask_off)) return qzero();
  const quat v = qload(p + sr + (size_t)sc * ldp);
  return CT ? qconj(v) : v;
}

__device__ __forceinline__ double flip_sign(double v, unsigned mask) {
  return mkd(lo32(v), hi32(v) ^ mask);
}

template <bool ACT, bool BCT>
__device__ __forceinline__ void qgemm2_tile(const QGemmTask& t, int row0, int col0, int kbeg, int kend, quat* C) {
  const int m = t.m, nn = t.nn, K = t.K;
  __shared__ double As[2][GBK][GBM][4];
  __shared__ double Bs[2][GBK][GBN][4];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;   // wave tile: rows wr*32..+32, cols wc*32..+32 (quaternions)
  const int cr = (lane & 15) >> 2, ck = lane >> 4;
  const int acomp = cr ^ ck;
  const unsigned smask = ((0x428E >> (cr * 4 + ck)) & 1) ? 0x80000000u : 0u;
  const int aq = lane & 3;
  const int bkc = lane >> 4, bj = lane & 15;
  // this thread's two A and two B stage elements
  int ai[2], ak[2], bk[2], bjj[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (!ACT) { ai[h] = tid & 63; ak[h] = (tid >> 6) + 4 * h; }
    else { ak[h] = tid & 7; ai[h] = (tid >> 3) + 32 * h; }
    if (!BCT) { bk[h] = tid & 7; bjj[h] = (tid >> 3) + 32 * h; }
    else { bjj[h] = tid & 63; bk[h] = (tid >> 6) + 4 * h; }
  }
  quat ra[2], rb[2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int ka = k0 + ak[h], kb = k0 + bk[h];
      ra[h] = (ka < kend) ? ld_opA<ACT>(t.A, row0 + ai[h], ka, m, K) : qzero();
      rb[h] = (kb < kend) ? ld_opB<BCT>(t.B, kb, col0 + bjj[h], K, nn) : qzero();
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double2* pa = reinterpret_cast<double2*>(&As[buf][ak[h]][ai[h]][0]);
      pa[0] = make_double2(ra[h].ar, ra[h].ai);
      pa[1] = make_double2(ra[h].br, ra[h].bi);
      double2* pb = reinterpret_cast<double2*>(&Bs[buf][bk[h]][bjj[h]][0]);
      pb[0] = make_double2(rb[h].ar, rb[h].ai);
      pb[1] = make_double2(rb[h].br, rb[h].bi);
    }
  };
  d4 acc[8][2];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  int cur = 0;
  if (kbeg < kend) {
    fetch(kbeg);
    stash(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += GBK) {
    const bool more = k0 + GBK < kend;
    if (more) fetch(k0 + GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; ++kk) {
      double af[8], bf[2];
#pragma unroll
      for (int a = 0; a < 8; ++a) af[a] = flip_sign(As[cur][kk][wr * 32 + 4 * a + aq][acomp], smask);
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[cur][kk][wc * 32 + 16 * b + bj][bkc];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    if (more) stash(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // epilogue: lane holds quaternion (row wr*32 + 4a + lane/16, col wc*32 + 16b + lane%16)
  const double alpha = t.alpha, beta = t.beta;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int i = row0 + wr * 32 + 4 * a + (lane >> 4);
      const int jj = col0 + wc * 32 + 16 * b + (lane & 15);
      if (i < m && jj < nn && (!t.lower_only || i >= jj)) {
        quat* cp = C + i + (size_t)jj * t.ldc;
        quat r = quat{alpha * acc[a][b][0], alpha * acc[a][b][1], alpha * acc[a][b][2], alpha * acc[a][b][3]};
        if (beta != 0.0) r = qadd(r, qscale(qload(cp), beta));
        qstore(cp, r);
      }
    }
  }
}

template <bool ACT, bool BCT>
__global__ void __launch_bounds__(256, 2) k_qgemm2(QGemmTask t, int kchunk, size_t cstride) {
  const int row0 = blockIdx.x * GBM;
  const int col0 = blockIdx.y * GBN;
  if (t.lower_only && row0 + GBM - 1 < col0) return;
  int kbeg = t.kbeg, kend = t.kend;
  quat* C = t.C;
  if (gridDim.z > 1) {
    kbeg = t.kbeg + blockIdx.z * kchunk;
    kend = min(t.kend, kbeg + kchunk);
    C += (size_t)blockIdx.z * cstride;
  }
  qgemm2_tile<ACT, BCT>(t, row0, col0, kbeg, kend, C);
}

template <bool ACT, bool BCT>
__global__ void __launch_bounds__(256, 2) k_qgemm2_batch(const QGemmTask* __restrict__ tasks) {
  const QGemmTask t = tasks[blockIdx.z];
  const int row0 = blockIdx.x * GBM;
  const int col0 = blockIdx.y * GBN;
  if (row0 >= t.m || col0 >= t.nn) return;
  if (t.lower_only && row0 + GBM - 1 < col0) return;
  qgemm2_tile<ACT, BCT>(t, row0, col0, t.kbeg, t.kend, t.C);
}

static void launch_gemm2(bool act, bool bct, dim3 grid, hipStream_t s, const QGemmTask& t, int kchunk,
                         size_t cstride) {
  if (!act && !bct) launch_k(k_qgemm2<false, false>, grid, dim3(256), 0, s, t, kchunk, cstride);
  else if (act && !bct) launch_k(k_qgemm2<true, false>, grid, dim3(256), 0, s, t, kchunk, cstride);
  else if (!act && bct) launch_k(k_qgemm2<false, true>, grid, dim3(256), 0, s, t, kchunk, cstride);
  else launch_k(k_qgemm2<true, true>, grid, dim3(256), 0, s, t, kchunk, cstride);
}

static void launch_gemm2_batch(bool act, bool bct, dim3 grid, hipStream_t s, const QGemmTask* tasks) {
  if (!act && !bct) launch_k(k_qgemm2_batch<false, false>, grid, dim3(256), 0, s, tasks);
  else if (act && !bct) launch_k(k_qgemm2_batch<true, false>, grid, dim3(256), 0, s, tasks);
  else if (!act && bct) launch_k(k_qgemm2_batch<false, true>, grid, dim3(256), 0, s, tasks);
  else launch_k(k_qgemm2_batch<true, true>, grid, dim3(256), 0, s, tasks);
}

int qgemm(int m, int nn, int k, double alpha, const QOperand& A, const QOperand& B, double beta, quat* C, int ldc,
          bool lower_only, hipStream_t s) {
  if (m <= 0 || nn <= 0) return ZQI_OK;
  QGemmTask t{m, nn, k, 0, k, ldc, lower_only ? 1 : 0, alpha, beta, A, B, C};
  launch_gemm2(A.conj_trans, B.conj_trans, dim3((m + GBM - 1) / GBM, (nn + GBN - 1) / GBN, 1), s, t, k, (size_t)0);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

int qgemm_tasks(const QGemmTask* tasks_dev, int ntasks, int maxm, int maxnn, bool act, bool bct, hipStream_t s) {
  if (ntasks <= 0 || maxm <= 0 || maxnn <= 0) return ZQI_OK;
  launch_gemm2_batch(act, bct, dim3((maxm + GBM - 1) / GBM, (maxnn + GBN - 1) / GBN, ntasks), s, tasks_dev);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

int splitk_count(int K, int nsplit, int* kchunk_out) {
  if (nsplit < 1) nsplit = 1;
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + GBK - 1) / GBK) * GBK;
  if (kchunk < GBK) kchunk = GBK;
  if (kchunk_out) *kchunk_out = kchunk;
  const int nz = (K + kchunk - 1) / kchunk;
  return nz > 0 ? nz : 1;
}

int qgemm_cn_splitk(int m, int nn, int K, const QOperand& A, const quat* B, int ldb, quat* Ypart, int nsplit,
                    hipStream_t s) {
  if (m <= 0 || nn <= 0) return ZQI_OK;
  int kchunk = 0;
  const int nz = splitk_count(K, nsplit, &kchunk);
  QOperand Bo{B, B, ldb, ldb, INT_MAX, false, INT_MIN / 2};
  QGemmTask t{m, nn, K, 0, K, m, 0, 1.0, 0.0, A, Bo, Ypart};
  launch_gemm2(A.conj_trans, false, dim3((m + GBM - 1) / GBM, (nn + GBN - 1) / GBN, nz), s, t, kchunk,
               (size_t)m * nn);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Real FP64 GEMM for the divide & conquer merges of one tree level, one merge
// per blockIdx.z: for the merge of rows [lo, hi) (ns = hi - lo) of problem
// base ob,  Z[lo:hi, lo - ob + posK[j]] = Qtmp_m (ns x K) U_m (K x K), with
// K = kcnt[z] read on the device (data-dependent after deflation) and the
// merge's Qtmp / U blocks at lo * lq (ld ns).
// ---------------------------------------------------------------------------
constexpr int DBM = 64, DBN = 64, DBK = 16;

__global__ void __launch_bounds__(256) k_dgemm_merge(const MergeDesc* __restrict__ descs, const int* __restrict__ kcnt,
                                                     const double* __restrict__ Qtmp, const double* __restrict__ U,
                                                     int lq, double* __restrict__ Z, int ldz,
                                                     const int* __restrict__ posK) {
  const MergeDesc md = descs[blockIdx.z];
  const int K = kcnt[blockIdx.z];
  const int m = md.hi - md.lo;
  const int row0 = blockIdx.x * DBM, col0 = blockIdx.y * DBN;
  if (col0 >= K || row0 >= m) return;
  const double* A = Qtmp + (size_t)md.lo * lq;
  const double* B = U + (size_t)md.lo * lq;
  const int lda = m, ldb = m;
  double* C = Z + md.lo + (size_t)(md.lo - md.ob) * ldz;
  const int* cmap = posK + md.lo;
  __shared__ double As[DBK][DBM + 1];
  __shared__ double Bs[DBK][DBN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;  // 32 x 32 per wave
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += DBK) {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int idx = tid + 256 * h;      // 1024 = 64 x 16
      const int i = idx & (DBM - 1), k = idx >> 6;
      const int gi = row0 + i, gk = k0 + k;
      As[k][i] = (gi < m && gk < K) ? A[gi + (size_t)gk * lda] : 0.0;
      const int kb = idx & (DBK - 1), jb = idx >> 4;
      const int gkb = k0 + kb, gj = col0 + jb;
      Bs[kb][jb] = (gkb < K && gj < K) ? B[gkb + (size_t)gj * ldb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < DBK; kk += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[kk + (lane >> 4)][wr * 32 + 16 * a + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[kk + (lane >> 4)][wc * 32 + 16 * b + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = col0 + wc * 32 + 16 * b + (lane & 15);
      if (j >= K) continue;
      const int cj = cmap[j];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr * 32 + 16 * a + 4 * r + (lane >> 4);
        if (i < m) C[i + (size_t)cj * ldz] = acc[a][b][r];
      }
    }
}

int dgemm_merges(const MergeDesc* descs_dev, int nmerge, int nsmax, const int* kcnt, const double* Qtmp,
                 const double* U, int lq, double* Z, int ldz, const int* posK, hipStream_t s) {
  if (nmerge <= 0 || nsmax <= 0) return ZQI_OK;
  dim3 grid((nsmax + DBM - 1) / DBM, (nsmax + DBN - 1) / DBN, nmerge);
  launch_k(k_dgemm_merge, grid, dim3(256), 0, s, descs_dev, kcnt, Qtmp, U, lq, Z, ldz, posK);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

// Raw MFMA probe: D = A (16x4) B (4x16) with lane-indexed inputs; used by the
// self-test that pins the register layout the kernels above assume.
__global__ void k_mfma_probe(const double* a, const double* b, double* out) {
  const int l = threadIdx.x;
  d4 acc = d4{0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[l * 4 + r] = acc[r];
}

int mfma_probe(const double* a, const double* b, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, a, b, out);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
