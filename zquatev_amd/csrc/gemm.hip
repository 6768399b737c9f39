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
// Tiles of TBM x 64 quaternions per workgroup, TBM = 64 or 32 (2 x 2 waves,
// TBM/2 x 32 quaternions = TBM/4 MFMA accumulators per wave; the smaller tile
// doubles the workgroups of products too small to fill the device), K staged 8 quaternions at a
// time through double-buffered LDS.  The next stage's operands are fetched
// into registers before the current stage's MFMAs, so global latency hides
// behind the wave's own work, and one barrier per stage suffices (a buffer is
// rewritten only after the barrier that follows its last read).  The
// conjugate-transpose flags are template parameters (branch-free address
// math); the sign of the L(A) expansion is applied by flipping the sign bit.
// ---------------------------------------------------------------------------
constexpr int GBN = 64, GBK = 8;
// LDS layouts (doubles).  A: [k][row][component], B: [k][component][column],
// each k slice padded by 2 doubles, so that fragment reads (16 consecutive
// doubles per component group) and the stage writes of either operand
// orientation (8 consecutive k per 8 threads) fall in distinct banks.
constexpr int BSK = GBN * 4 + 2;   // B k-slice stride

// Operand elements as stored.  The conjugation of a CT operand is applied
// when the stage is stashed in LDS, after the current stage's MFMAs: sign
// flips here, on the freshly loaded registers, would make every prefetch
// wait for its own data before those MFMAs could start.
template <bool CT>
__device__ __forceinline__ quat ld_opA(const QOperand& A, int i, int k, int m, int K) {
  if (i >= m || k >= K) return qzero();
  const quat* p = A.p0;
  int ldp = A.ld0, kk = k;
  if (k >= A.ksplit) { p = A.p1; ldp = A.ld1; kk = k - A.ksplit; }
  const int sr = CT ? kk : i, sc = CT ? i : kk;
  if (!(sr > sc + A.mask_off)) return qzero();
  return qload_g(p + sr + (size_t)sc * ldp);
}

template <bool CT>
__device__ __forceinline__ quat ld_opB(const QOperand& B, int k, int j, int K, int nn) {
  if (k >= K || j >= nn) return qzero();
  const quat* p = B.p0;
  int ldp = B.ld0, kk = k;
  if (k >= B.ksplit) { p = B.p1; ldp = B.ld1; kk = k - B.ksplit; }
  const int sr = CT ? j : kk, sc = CT ? kk : j;
  if (!(sr > sc + B.mask_off)) return qzero();
  return qload_g(p + sr + (size_t)sc * ldp);
}

__device__ __forceinline__ double flip_sign(double v, unsigned mask) {
  return mkd(lo32(v), hi32(v) ^ mask);
}

template <int TBM, bool ACT, bool BCT>
__device__ __forceinline__ void qgemm2_tile(const QGemmTask& t, int row0, int col0, int kbeg, int kend, quat* C) {
  constexpr int AH = TBM / 32;          // A stage elements per thread
  constexpr int AK = TBM * 4 + 2;       // A k-slice stride (doubles)
  constexpr int WR = TBM / 2;           // quaternion rows per wave
  constexpr int NA = WR / 4;            // 4-row MFMA tiles per wave
  const int m = t.m, nn = t.nn, K = t.K;
  __shared__ double As[2][GBK * AK];
  __shared__ double Bs[2][GBK * BSK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;   // wave tile: rows wr*WR..+WR, cols wc*32..+32 (quaternions)
  const int cr = (lane & 15) >> 2, ck = lane >> 4;
  const int acomp = cr ^ ck;
  const unsigned smask = ((0x428E >> (cr * 4 + ck)) & 1) ? 0x80000000u : 0u;
  const int aq = lane & 3;
  const int bkc = lane >> 4, bj = lane & 15;
  // this thread's A and B stage elements
  int ai[AH], ak[AH], bk[2], bjj[2];
#pragma unroll
  for (int h = 0; h < AH; ++h) {
    if (!ACT) { ai[h] = tid % TBM; ak[h] = tid / TBM + (256 / TBM) * h; }
    else { ak[h] = tid & 7; ai[h] = (tid >> 3) + 32 * h; }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (!BCT) { bk[h] = tid & 7; bjj[h] = (tid >> 3) + 32 * h; }
    else { bjj[h] = tid & 63; bk[h] = (tid >> 6) + 4 * h; }
  }
  quat ra[AH], rb[2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < AH; ++h) {
      const int ka = k0 + ak[h];
      ra[h] = (ka < kend) ? ld_opA<ACT>(t.A, row0 + ai[h], ka, m, K) : qzero();
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int kb = k0 + bk[h];
      rb[h] = (kb < kend) ? ld_opB<BCT>(t.B, kb, col0 + bjj[h], K, nn) : qzero();
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int h = 0; h < AH; ++h) {
      const quat a = ACT ? qconj(ra[h]) : ra[h];
      double2* pa = reinterpret_cast<double2*>(&As[buf][ak[h] * AK + ai[h] * 4]);
      pa[0] = make_double2(a.ar, a.ai);
      pa[1] = make_double2(a.br, a.bi);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const quat bv = BCT ? qconj(rb[h]) : rb[h];
      double* pb = &Bs[buf][bk[h] * BSK + bjj[h]];
      pb[0] = bv.ar;
      pb[GBN] = bv.ai;
      pb[2 * GBN] = bv.br;
      pb[3 * GBN] = bv.bi;
    }
  };
  d4 acc[NA][2];
#pragma unroll
  for (int a = 0; a < NA; ++a)
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
      double af[NA], bf[2];
#pragma unroll
      for (int a = 0; a < NA; ++a) af[a] = flip_sign(As[cur][kk * AK + (wr * WR + 4 * a + aq) * 4 + acomp], smask);
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[cur][kk * BSK + bkc * GBN + wc * 32 + 16 * b + bj];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    if (more) stash(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // epilogue: lane holds quaternion (row wr*WR + 4a + lane/16, col wc*32 + 16b + lane%16)
  const double alpha = t.alpha, beta = t.beta;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int i = row0 + wr * WR + 4 * a + (lane >> 4);
      const int jj = col0 + wc * 32 + 16 * b + (lane & 15);
      if (i < m && jj < nn && (!t.lower_only || i >= jj)) {
        quat* cp = C + i + (size_t)jj * t.ldc;
        quat r = quat{alpha * acc[a][b][0], alpha * acc[a][b][1], alpha * acc[a][b][2], alpha * acc[a][b][3]};
        if (beta != 0.0) r = qadd(r, qscale(qload_g(cp), beta));
        qstore_g(cp, r);
      }
    }
  }
}

// Workgroups are dealt to the 8 XCDs round-robin by launch order.  The tile
// index is remapped so that each XCD works through a contiguous range of
// tiles (column-major over the tile grid), which share operand tiles in its
// L2: XCD k = t mod 8 owns floor(T/8) (+1 for k < T mod 8) consecutive ones.
__device__ __forceinline__ int xcd_remap(int t, int T) {
  const int k = t & 7;
  return k * (T >> 3) + min(k, T & 7) + (t >> 3);
}

// Lower-triangle products enumerate only the tiles that meet the lower
// triangle: column block y (64 wide) has row tiles x >= (64 y) / TBM.
__host__ __device__ inline int lower_tiles(int m, int tbm) {
  const int gx = (m + tbm - 1) / tbm, gy = (m + 63) / 64;
  int cnt = 0;
  for (int y = 0; y < gy; ++y) cnt += gx - (64 * y) / tbm;
  return cnt;
}

template <int TBM>
__device__ __forceinline__ bool tile_of(const QGemmTask& t, int u, int gx_rect, int& tx, int& ty) {
  if (!t.lower_only) {
    tx = u % gx_rect;
    ty = u / gx_rect;
    return true;
  }
  const int gx = (t.m + TBM - 1) / TBM, gy = (t.nn + 63) / 64;
  int acc = 0;
  for (int y = 0; y < gy; ++y) {
    const int x0 = (64 * y) / TBM;
    const int cnt = gx - x0;
    if (u < acc + cnt) {
      tx = x0 + (u - acc);
      ty = y;
      return true;
    }
    acc += cnt;
  }
  return false;
}

// grid: x = tiles (rectangular: gx * gy of them, gx given; lower: lower_tiles), z = K chunks
template <int TBM, bool ACT, bool BCT>
__global__ void __launch_bounds__(256, 2) k_qgemm2(QGemmTask t, int kchunk, size_t cstride, int gx) {
  int tx, ty;
  if (!tile_of<TBM>(t, xcd_remap(blockIdx.x, gridDim.x), gx, tx, ty)) return;
  const int row0 = tx * TBM;
  const int col0 = ty * GBN;
  int kbeg = t.kbeg, kend = t.kend;
  quat* C = t.C;
  if (gridDim.z > 1) {
    kbeg = t.kbeg + blockIdx.z * kchunk;
    kend = min(t.kend, kbeg + kchunk);
    C += (size_t)blockIdx.z * cstride;
  }
  qgemm2_tile<TBM, ACT, BCT>(t, row0, col0, kbeg, kend, C);
}

// grid: x = tiles of the largest task, z = tasks
template <int TBM, bool ACT, bool BCT>
__global__ void __launch_bounds__(256, 2) k_qgemm2_batch(const QGemmTask* __restrict__ tasks) {
  const QGemmTask t = tasks[blockIdx.z];
  // tiles of this task: the launch covers the largest task's; gx from the task
  const int gx = (t.m + TBM - 1) / TBM;
  const int ntile = t.lower_only ? lower_tiles(t.m, TBM) : gx * ((t.nn + GBN - 1) / GBN);
  const int u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= ntile) return;
  int tx, ty;
  if (!tile_of<TBM>(t, u, gx, tx, ty)) return;
  const int row0 = tx * TBM;
  const int col0 = ty * GBN;
  qgemm2_tile<TBM, ACT, BCT>(t, row0, col0, t.kbeg, t.kend, t.C);
}

// Tile rows: 64 only for launches with at least 16 64 x 64 tiles per compute
// unit, else 32 (twice the workgroups: better balance and register use).
// Measured on the 2n = 8192 solve, each step against the previous one on
// one box: switch at 2 -> 4 tiles per CU -0.6 ms, 4 -> 8 -1.5 ms, 8 -> 16
// -0.9 ms; 32 rows always is equal to 16 within noise; the batch of 512
// problems stays within 0.5%.
static int device_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (!cached[dev]) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

thread_local int g_tile64_per_cu = 16;

static int tile_rows(long tiles64) { return tiles64 >= (long)g_tile64_per_cu * device_cus() ? 64 : 32; }

template <int TBM>
static void launch_gemm2_t(bool act, bool bct, hipStream_t s, const QGemmTask& t, int nz, int kchunk,
                           size_t cstride) {
  const int gx = (t.m + TBM - 1) / TBM;
  const int nt = t.lower_only ? lower_tiles(t.m, TBM) : gx * ((t.nn + GBN - 1) / GBN);
  const dim3 grid(nt, 1, nz);
  if (!act && !bct) launch_k(k_qgemm2<TBM, false, false>, grid, dim3(256), 0, s, t, kchunk, cstride, gx);
  else if (act && !bct) launch_k(k_qgemm2<TBM, true, false>, grid, dim3(256), 0, s, t, kchunk, cstride, gx);
  else if (!act && bct) launch_k(k_qgemm2<TBM, false, true>, grid, dim3(256), 0, s, t, kchunk, cstride, gx);
  else launch_k(k_qgemm2<TBM, true, true>, grid, dim3(256), 0, s, t, kchunk, cstride, gx);
}

static void launch_gemm2(bool act, bool bct, hipStream_t s, const QGemmTask& t, int nz, int kchunk, size_t cstride) {
  const long t64 = (long)(t.lower_only ? lower_tiles(t.m, 64) : ((t.m + 63) / 64) * ((t.nn + GBN - 1) / GBN)) * nz;
  if (tile_rows(t64) == 64) launch_gemm2_t<64>(act, bct, s, t, nz, kchunk, cstride);
  else launch_gemm2_t<32>(act, bct, s, t, nz, kchunk, cstride);
}

template <int TBM>
static void launch_gemm2_batch_t(bool act, bool bct, dim3 grid, hipStream_t s, const QGemmTask* tasks) {
  if (!act && !bct) launch_k(k_qgemm2_batch<TBM, false, false>, grid, dim3(256), 0, s, tasks);
  else if (act && !bct) launch_k(k_qgemm2_batch<TBM, true, false>, grid, dim3(256), 0, s, tasks);
  else if (!act && bct) launch_k(k_qgemm2_batch<TBM, false, true>, grid, dim3(256), 0, s, tasks);
  else launch_k(k_qgemm2_batch<TBM, true, true>, grid, dim3(256), 0, s, tasks);
}

int qgemm(int m, int nn, int k, double alpha, const QOperand& A, const QOperand& B, double beta, quat* C, int ldc,
          bool lower_only, hipStream_t s) {
  if (m <= 0 || nn <= 0) return ZQI_OK;
  QGemmTask t{m, nn, k, 0, k, ldc, lower_only ? 1 : 0, alpha, beta, A, B, C};
  launch_gemm2(A.conj_trans, B.conj_trans, s, t, 1, k, (size_t)0);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

int qgemm_tasks(const QGemmTask* tasks_dev, int ntasks, int maxm, int maxnn, bool act, bool bct, bool lower,
                hipStream_t s) {
  if (ntasks <= 0 || maxm <= 0 || maxnn <= 0) return ZQI_OK;
  auto ntiles = [&](int tbm) {
    return lower ? lower_tiles(maxm, tbm) : ((maxm + tbm - 1) / tbm) * ((maxnn + GBN - 1) / GBN);
  };
  if (tile_rows((long)ntiles(64) * ntasks) == 64)
    launch_gemm2_batch_t<64>(act, bct, dim3(ntiles(64), 1, ntasks), s, tasks_dev);
  else
    launch_gemm2_batch_t<32>(act, bct, dim3(ntiles(32), 1, ntasks), s, tasks_dev);
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

// ---------------------------------------------------------------------------
// Real FP64 GEMM for the divide & conquer merges of one tree level, one merge
// per blockIdx.z: for the merge of rows [lo, hi) (ns = hi - lo) of problem
// base ob,  Z[lo:hi, lo - ob + posK[j]] = Qtmp_m (ns x K) U_m (K x K), with
// K = kcnt[z] read on the device (data-dependent after deflation) and the
// merge's Qtmp / U blocks at lo * lq (ld ns).  Qtmp's columns come in three
// groups: [0, g0) hold rows of the first half (rows < mid - lo) only, [g0,
// g1) both halves, [g1, K) the second half only (kgrp[2z], kgrp[2z + 1]), so
// a tile of first-half rows sums over [0, g1) and one of second-half rows
// over [g0, K), skipping exact zeros (as LAPACK's dlaed3 splits the product).
// ---------------------------------------------------------------------------
constexpr int DBM = 64, DBN = 64, DBK = 16;

// 64 x 64 tile per workgroup (2 x 2 waves of 32 x 32), K staged 16 at a time
// through double-buffered LDS; the next stage's operands are fetched into
// registers before the current stage's MFMAs (one barrier per stage).
__global__ void __launch_bounds__(256) k_dgemm_merge(const MergeDesc* __restrict__ descs, const int* __restrict__ kcnt,
                                                     const int* __restrict__ kgrp, const double* __restrict__ Qtmp,
                                                     const double* __restrict__ U, int lq, double* __restrict__ Z,
                                                     int ldz, const int* __restrict__ posK) {
  const MergeDesc md = descs[blockIdx.z];
  const int K = kcnt[blockIdx.z];
  const int m = md.hi - md.lo, n1 = md.mid - md.lo;
  const int row0 = blockIdx.x * DBM, col0 = blockIdx.y * DBN;
  if (col0 >= K || row0 >= m) return;
  // the sum's range for this tile's rows (the full range for a tile across the halves)
  const int kb = row0 >= n1 ? kgrp[2 * blockIdx.z] : 0;
  const int ke = row0 + DBM <= n1 ? kgrp[2 * blockIdx.z + 1] : K;
  const double* A = Qtmp + (size_t)md.lo * lq;
  const double* B = U + (size_t)md.lo * lq;
  const int lda = m, ldb = m;
  double* C = Z + md.lo + (size_t)(md.lo - md.ob) * ldz;
  const int* cmap = posK + md.lo;
  __shared__ double As[2][DBK][DBM + 1];
  __shared__ double Bs[2][DBK][DBN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv >> 1, wc = wv & 1;  // 32 x 32 per wave
  constexpr int NE = DBK * DBM / 256;   // stage elements per thread and operand
  double ra[NE], rb[NE];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < NE; ++h) {
      const int idx = tid + 256 * h;
      const int i = idx & (DBM - 1), k = idx >> 6;
      const int gi = row0 + i, gk = k0 + k;
      ra[h] = (gi < m && gk < ke) ? A[gi + (size_t)gk * lda] : 0.0;
      const int kr = idx & (DBK - 1), jb = idx / DBK;
      const int gkb = k0 + kr, gj = col0 + jb;
      rb[h] = (gkb < ke && gj < K) ? B[gkb + (size_t)gj * ldb] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int h = 0; h < NE; ++h) {
      const int idx = tid + 256 * h;
      As[buf][idx >> 6][idx & (DBM - 1)] = ra[h];
      Bs[buf][idx & (DBK - 1)][idx / DBK] = rb[h];
    }
  };
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  fetch(kb);
  stash(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = kb; k0 < ke; k0 += DBK) {
    const bool more = k0 + DBK < ke;
    if (more) fetch(k0 + DBK);
#pragma unroll
    for (int kk = 0; kk < DBK; kk += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[cur][kk + (lane >> 4)][wr * 32 + 16 * a + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[cur][kk + (lane >> 4)][wc * 32 + 16 * b + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    if (more) stash(cur ^ 1);
    __syncthreads();
    cur ^= 1;
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

int dgemm_merges(const MergeDesc* descs_dev, int nmerge, int nsmax, const int* kcnt, const int* kgrp,
                 const double* Qtmp, const double* U, int lq, double* Z, int ldz, const int* posK, hipStream_t s) {
  if (nmerge <= 0 || nsmax <= 0) return ZQI_OK;
  dim3 grid((nsmax + DBM - 1) / DBM, (nsmax + DBN - 1) / DBN, nmerge);
  launch_k(k_dgemm_merge, grid, dim3(256), 0, s, descs_dev, kcnt, kgrp, Qtmp, U, lq, Z, ldz, posK);
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
