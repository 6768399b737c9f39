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

namespace zq {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int QBM = 32;   // quaternion rows per workgroup tile
constexpr int QBN = 64;   // quaternion cols per workgroup tile
constexpr int QBK = 8;    // quaternion k per LDS stage

__device__ __forceinline__ quat load_opA(const QOperand& A, int i, int k, int m, int K) {
  // logical A[i][k]; storage either (i, k) or (k, i) conj
  if (i >= m || k >= K) return qzero();
  const quat* p;
  int ldp, kk = k;
  if (k < A.ksplit) { p = A.p0; ldp = A.ld0; } else { p = A.p1; ldp = A.ld1; kk = k - A.ksplit; }
  int sr, sc;
  if (!A.conj_trans) { sr = i; sc = kk; } else { sr = kk; sc = i; }
  if (!(sr > sc + A.mask_off)) return qzero();
  quat v = qload(p + sr + (size_t)sc * ldp);
  return A.conj_trans ? qconj(v) : v;
}

__device__ __forceinline__ quat load_opB(const QOperand& B, int k, int jcol, int K, int nn) {
  if (k >= K || jcol >= nn) return qzero();
  const quat* p;
  int ldp, kk = k;
  if (k < B.ksplit) { p = B.p0; ldp = B.ld0; } else { p = B.p1; ldp = B.ld1; kk = k - B.ksplit; }
  int sr, sc;
  if (!B.conj_trans) { sr = kk; sc = jcol; } else { sr = jcol; sc = kk; }
  if (!(sr > sc + B.mask_off)) return qzero();
  quat v = qload(p + sr + (size_t)sc * ldp);
  return B.conj_trans ? qconj(v) : v;
}

// C = beta C + alpha A B over quaternions.  gridDim.z > 1 splits K into
// gridDim.z chunks of kchunk; chunk z writes C + z * cstride with beta = 0.
__global__ void __launch_bounds__(256) k_qgemm(int m, int nn, int K, double alpha, QOperand A, QOperand B, double beta,
                                               quat* __restrict__ C, int ldc, int lower_only, int kchunk,
                                               size_t cstride) {
  const int row0 = blockIdx.x * QBM;
  const int col0 = blockIdx.y * QBN;
  if (lower_only && row0 + QBM - 1 < col0) return;
  int kbeg = 0, kend = K;
  if (gridDim.z > 1) {
    kbeg = blockIdx.z * kchunk;
    kend = min(K, kbeg + kchunk);
    C += (size_t)blockIdx.z * cstride;
  }
  __shared__ double As[QBK][QBM][4];
  __shared__ double Bs[QBK][QBN][4];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;   // wave tile: rows wr*16..+16, cols wc*32..+32

  // per-lane fragment constants: A-operand real row (lane & 15) is component
  // cr = (lane & 15) >> 2 of quaternion row aq = lane & 3; real k = lane >> 4
  const int cr = (lane & 15) >> 2, ck = lane >> 4;
  const int acomp = cr ^ ck;
  // sign[cr][ck]: rows [+,-,-,-],[+,+,+,-],[+,-,+,+],[+,+,-,+]; nibble cr of
  // 0x428E has bit ck set where the sign is negative
  const double asign = ((0x428E >> (cr * 4 + ck)) & 1) ? -1.0 : 1.0;
  const int aq = lane & 3;           // quaternion row within the 4-row MFMA tile
  const int bkc = lane >> 4;         // component of B
  const int bj = lane & 15;

  d4 acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  for (int k0 = kbeg; k0 < kend; k0 += QBK) {
    // A tile: QBM x QBK = 256 quaternions, one per thread
    {
      int i, k;
      if (!A.conj_trans) { i = tid & (QBM - 1); k = tid >> 5; }
      else { k = tid & (QBK - 1); i = tid >> 3; }
      const int kg = k0 + k;
      quat v = (kg < kend) ? load_opA(A, row0 + i, kg, m, K) : qzero();
      As[k][i][0] = v.ar; As[k][i][1] = v.ai; As[k][i][2] = v.br; As[k][i][3] = v.bi;
    }
    // B tile: QBK x QBN = 512 quaternions, two per thread
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int k, jj;
      if (!B.conj_trans) { k = tid & (QBK - 1); jj = (tid >> 3) + 32 * h; }
      else { jj = tid & (QBN - 1); k = (tid >> 6) + 4 * h; }
      const int kg = k0 + k;
      quat v = (kg < kend) ? load_opB(B, kg, col0 + jj, K, nn) : qzero();
      Bs[k][jj][0] = v.ar; Bs[k][jj][1] = v.ai; Bs[k][jj][2] = v.br; Bs[k][jj][3] = v.bi;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < QBK; ++kk) {
      double af[4], bf[2];
#pragma unroll
      for (int a = 0; a < 4; ++a) af[a] = asign * As[kk][wr * 16 + 4 * a + aq][acomp];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[kk][wc * 32 + 16 * b + bj][bkc];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: lane holds quaternion (row wr*16 + 4a + lane/16, col wc*32 + 16b + lane%16)
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int i = row0 + wr * 16 + 4 * a + (lane >> 4);
      const int jj = col0 + wc * 32 + 16 * b + (lane & 15);
      if (i < m && jj < nn && (!lower_only || i >= jj)) {
        quat* cp = C + i + (size_t)jj * ldc;
        quat r = quat{alpha * acc[a][b][0], alpha * acc[a][b][1], alpha * acc[a][b][2], alpha * acc[a][b][3]};
        if (beta != 0.0) r = qadd(r, qscale(qload(cp), beta));
        qstore(cp, r);
      }
    }
  }
}

int qgemm(int m, int nn, int k, double alpha, const QOperand& A, const QOperand& B, double beta, quat* C, int ldc,
          bool lower_only, hipStream_t s) {
  if (m <= 0 || nn <= 0) return ZQI_OK;
  dim3 grid((m + QBM - 1) / QBM, (nn + QBN - 1) / QBN, 1);
  launch_k(k_qgemm, grid, dim3(256), 0, s, m, nn, k, alpha, A, B, beta, C, ldc, lower_only ? 1 : 0, k, 0);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

int splitk_count(int K, int nsplit, int* kchunk_out) {
  if (nsplit < 1) nsplit = 1;
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + QBK - 1) / QBK) * QBK;
  if (kchunk < QBK) kchunk = QBK;
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
  dim3 grid((m + QBM - 1) / QBM, (nn + QBN - 1) / QBN, nz);
  launch_k(k_qgemm, grid, dim3(256), 0, s, m, nn, K, 1.0, A, Bo, 0.0, Ypart, m, 0, kchunk,
                     (size_t)m * nn);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Real FP64 GEMM for the divide & conquer merges: C[:, cmap[j]] = A (m x K) B (K x K),
// K read from the device (data-dependent after deflation).  Batched over merges
// through gridDim.z with per-merge offsets.
// ---------------------------------------------------------------------------
constexpr int DBM = 64, DBN = 64, DBK = 16;

__global__ void __launch_bounds__(256) k_dgemm_nn(int m, const int* __restrict__ kdev, const double* __restrict__ A,
                                                  int lda, const double* __restrict__ B, int ldb,
                                                  double* __restrict__ C, int ldc, const int* __restrict__ cmap) {
  const int K = *kdev;
  const int row0 = blockIdx.x * DBM, col0 = blockIdx.y * DBN;
  if (col0 >= K || row0 >= m) return;
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
      const int cj = cmap ? cmap[j] : j;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr * 32 + 16 * a + 4 * r + (lane >> 4);
        if (i < m) C[i + (size_t)cj * ldc] = acc[a][b][r];
      }
    }
}

int dgemm_nn(int m, int, const int* k_dev, int k_max, const double* A, int lda, const double* B, int ldb, double* C,
             int ldc, const int* cmap, hipStream_t s) {
  if (m <= 0 || k_max <= 0) return ZQI_OK;
  dim3 grid((m + DBM - 1) / DBM, (k_max + DBN - 1) / DBN, 1);
  launch_k(k_dgemm_nn, grid, dim3(256), 0, s, m, k_dev, A, lda, B, ldb, C, ldc, cmap);
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
