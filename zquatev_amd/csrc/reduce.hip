// Quaternion Householder tridiagonalisation (the panel and trailing update).
//
// Replaces the reference's symplectic reduction: impl::panel_update
// (blocked.cc:45-546) and impl::unblocked_update (unblocked.cc:44-131), driven
// by the panel loop at zquatev.cc:68-72.  The reference applies three complex
// symplectic transforms per column (a B-column Householder, a Givens rotation
// between the A and B blocks, an A-column Householder).  Here one quaternion
// Householder reflector H = I - u u^* (u^* u = 2) clears column c of the
// quaternion matrix M below its subdiagonal in one step; by the implicit-Q
// theorem the real tridiagonal obtained after the diagonal phase scaling is
// the same one the reference hands to zhbev (zquatev.cc:74-84).
//
// Blocked (latrd-style) form: within a panel of nb columns starting at p the
// trailing matrix is M - V W^* - W V^*, where V holds the panel's reflectors
// (stored in place below the subdiagonal of M) and
//   w_j = p_j - (u_j^* p_j / 2) u_j,   p_j = M u_j - V (W^* u_j) - W (V^* u_j).
// After the panel, the trailing lower triangle gets the rank-2nb update on
// FP64 MFMA (gemm.hip).  Per column the only pass over the trailing matrix is
// the Hermitian matvec M u_j, which reads the lower triangle once (16 m^2
// bytes for an m x m trailing block).
#include "zq_internal.h"
#include <climits>

namespace zq {

constexpr int RB = 256;      // rows per block for row-parallel kernels
constexpr int TB = 32;       // symv tile edge (quaternions)
constexpr int TSR = 256;     // rows per tall-skinny chunk
constexpr int NBMAX = 128;   // largest supported panel width

__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// Column c of the panel: apply the panel's pending update
//   M[i,c] -= sum_{l<j} V(i,l) W(c,l)^* + W(i,l) V(c,l)^*     for i in [c, n)
// and accumulate |M[i,c]|^2 over i >= c+2 (the reflector's tail norm).
__global__ void __launch_bounds__(RB) k_col_update(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                                   const quat* __restrict__ W, int ldw, double* __restrict__ part) {
  __shared__ quat cw[NBMAX], cv[NBMAX];
  __shared__ double sh[RB / 64];
  for (int l = threadIdx.x; l < j; l += RB) {
    cw[l] = qconj(qload(W + c + (size_t)l * ldw));
    cv[l] = qconj(qload(M + c + (size_t)(p + l) * ld));
  }
  __syncthreads();
  const int i = c + blockIdx.x * RB + threadIdx.x;
  double ss = 0.0;
  if (i < n) {
    quat m = qload(M + i + (size_t)c * ld);
    if (j > 0) {
      quat acc = qzero();
      for (int l = 0; l < j; ++l) {
        qfma(acc, qload(M + i + (size_t)(p + l) * ld), cw[l]);
        qfma(acc, qload(W + i + (size_t)l * ldw), cv[l]);
      }
      m = qsub(m, acc);
      qstore(M + i + (size_t)c * ld, m);
    }
    if (i >= c + 2) ss = qnorm2(m);
  }
  ss = block_sum(ss, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = ss;
}

// Reflector for column c: x = M[c+1:, c].  H x = alpha e_1 with
// alpha = -q |x|, q = x_1/|x_1|; u = gamma (x - alpha e_1), gamma = 1/sqrt(|x| (|x| + |x_1|)).
// scal = [gamma, u_1 (4 doubles)].
__global__ void __launch_bounds__(RB) k_reflector(const quat* __restrict__ M, int ld, int n, int c, const double* part,
                                                  int nparts, double* __restrict__ scal, double* __restrict__ d,
                                                  quat* __restrict__ e) {
  __shared__ double sh[RB / 64];
  double t = 0.0;
  for (int i = threadIdx.x; i < nparts; i += RB) t += part[i];
  t = block_sum(t, sh);
  if (threadIdx.x == 0) {
    const quat x1 = qload(M + (c + 1) + (size_t)c * ld);
    d[c] = M[c + (size_t)c * ld].ar;
    const double ax1 = sqrt(qnorm2(x1));
    if (!(t > 0.0)) {
      scal[0] = 0.0;
      scal[1] = scal[2] = scal[3] = scal[4] = 0.0;
      qstore(e + c, x1);
    } else {
      const double nx = sqrt(ax1 * ax1 + t);
      quat q = (ax1 > 0.0) ? qscale(x1, 1.0 / ax1) : quat{1.0, 0.0, 0.0, 0.0};
      const double s = ax1 + nx;
      const double g = 1.0 / sqrt(nx * s);
      qstore(e + c, qscale(q, -nx));
      const quat u1 = qscale(q, s * g);
      scal[0] = g;
      scal[1] = u1.ar; scal[2] = u1.ai; scal[3] = u1.br; scal[4] = u1.bi;
    }
  }
}

__device__ __forceinline__ quat u_at(const quat* M, int ld, int c, int i, double g, const quat& u1) {
  return (i == c + 1) ? u1 : qscale(qload(M + i + (size_t)c * ld), g);
}

// Hermitian matvec y = M[c+1:, c+1:] u over the lower triangle, tiled TB x TB.
// Tile (I, J), I >= J, writes its row sums to slot J of rows in block I and its
// column sums to slot I of rows in block J (I == J: both into slot I), so each
// (slot, row) pair is written exactly once and y(row) = sum over slots.
// Blocks past the tiles compute tall-skinny partials of W^* u and V^* u.
__global__ void __launch_bounds__(256) k_symv(const quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                              const quat* __restrict__ W, int ldw, const double* __restrict__ scal,
                                              quat* __restrict__ P, int ldp, quat* __restrict__ ucur,
                                              quat* __restrict__ TS, int nb, int ntiles) {
  const int s0 = c + 1;
  const double g = scal[0];
  const quat u1 = quat{scal[1], scal[2], scal[3], scal[4]};
  const int t = blockIdx.x;
  if (t >= ntiles) {
    // tall-skinny partials over rows [s0 + h*TSR, ...)
    const int h = t - ntiles;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = s0 + h * TSR;
    quat uu[TSR / 64];
#pragma unroll
    for (int q = 0; q < TSR / 64; ++q) {
      const int i = r0 + lane + 64 * q;
      uu[q] = (i < n) ? u_at(M, ld, c, i, g, u1) : qzero();
    }
    for (int l = w; l < j; l += 4) {
      quat aw = qzero(), av = qzero();
#pragma unroll
      for (int q = 0; q < TSR / 64; ++q) {
        const int i = r0 + lane + 64 * q;
        if (i < n) {
          qfma_cl(aw, qload(W + i + (size_t)l * ldw), uu[q]);
          qfma_cl(av, qload(M + i + (size_t)(p + l) * ld), uu[q]);
        }
      }
      aw = wave_sum(aw);
      av = wave_sum(av);
      if (lane == 0) {
        qstore(TS + ((size_t)h * 2 + 0) * nb + l, aw);
        qstore(TS + ((size_t)h * 2 + 1) * nb + l, av);
      }
    }
    return;
  }
  // tile index -> (I, J), I >= J
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const int J = t - I * (I + 1) / 2;
  const int R0 = s0 + I * TB, C0 = s0 + J * TB;

  __shared__ quat uI[TB], uJ[TB];
  __shared__ quat redr[8][TB];
  __shared__ quat redc[TB][TB + 1];
  const int tid = threadIdx.x;
  if (tid < TB) {
    const int i = R0 + tid;
    const quat v = (i < n) ? u_at(M, ld, c, i, g, u1) : qzero();
    uI[tid] = v;
    if (I == J && i < n) qstore(ucur + i, v);
  } else if (tid < 2 * TB) {
    const int i = C0 + tid - TB;
    uJ[tid - TB] = (i < n) ? u_at(M, ld, c, i, g, u1) : qzero();
  }
  __syncthreads();
  const int r = tid & (TB - 1), grp = tid >> 5;  // 8 groups of 4 columns
  const int row = R0 + r;
  quat racc = qzero();
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int cl = grp * 4 + kk;
    const int col = C0 + cl;
    quat q = qzero();
    quat cp = qzero();
    if (row < n && col < n && (I > J || row >= col)) {
      q = qload(M + row + (size_t)col * ld);
      if (I == J && row == col) q = quat{q.ar, 0.0, 0.0, 0.0};
      qfma(racc, q, uJ[cl]);
      if (I > J || row > col) qfma_cl(cp, q, uI[r]);
    }
    redc[cl][r] = cp;
  }
  redr[grp][r] = racc;
  __syncthreads();
  if (tid < TB) {
    // row sums for rows of block I
    quat y = qzero();
#pragma unroll
    for (int k = 0; k < 8; ++k) y = qadd(y, redr[k][tid]);
    quat ycol = qzero();
    for (int k = 0; k < TB; ++k) ycol = qadd(ycol, redc[tid][k]);
    if (I == J) {
      if (R0 + tid < n) qstore(P + (size_t)I * ldp + R0 + tid, qadd(y, ycol));
    } else {
      if (R0 + tid < n) qstore(P + (size_t)J * ldp + R0 + tid, y);
      if (C0 + tid < n) qstore(P + (size_t)I * ldp + C0 + tid, ycol);
    }
  }
}

// p = sum of symv partials - V (W^* u) - W (V^* u); stored in W[:, j].  Also
// stores u into M[c+1:, c], partial u^* p into KP, and (block 0) the T-factor
// column j: T[0:j, j] = -T[0:j, 0:j] (V^* u), T[j, j] = 1 (0 if u = 0) and
// zeros below, so that H_p ... H_{p+nb-1} = I - V T V^* with T stored densely.
__global__ void __launch_bounds__(RB) k_wreduce(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                                quat* __restrict__ W, int ldw, const quat* __restrict__ P, int ldp,
                                                int nt, const quat* __restrict__ TS, int nchunks, int nb,
                                                const quat* __restrict__ ucur, double* __restrict__ KP,
                                                quat* __restrict__ Tf, const double* __restrict__ scal) {
  __shared__ quat a[NBMAX], b[NBMAX];
  __shared__ double sh[RB / 64];
  for (int l = threadIdx.x; l < 2 * j; l += RB) {
    const int side = l / j, ll = l - side * j;
    quat s = qzero();
    for (int h = 0; h < nchunks; ++h) s = qadd(s, qload(TS + ((size_t)h * 2 + side) * nb + ll));
    if (side == 0) a[ll] = s; else b[ll] = s;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int r = threadIdx.x; r < nb; r += RB) {
      quat acc = qzero();
      if (r < j) {
        for (int l = r; l < j; ++l) qfma(acc, qload(Tf + r + (size_t)l * nb), b[l]);
        acc = qscale(acc, -1.0);
      } else if (r == j) {
        acc = quat{scal[0] != 0.0 ? 1.0 : 0.0, 0.0, 0.0, 0.0};
      }
      qstore(Tf + r + (size_t)j * nb, acc);
    }
  }
  const int i = c + 1 + blockIdx.x * RB + threadIdx.x;
  double kp = 0.0;
  if (i < n) {
    quat y = qzero();
    for (int k = 0; k < nt; ++k) y = qadd(y, qload(P + (size_t)k * ldp + i));
    for (int l = 0; l < j; ++l) {
      const quat vv = qload(M + i + (size_t)(p + l) * ld);
      const quat ww = qload(W + i + (size_t)l * ldw);
      y = qsub(y, qmul(vv, a[l]));
      y = qsub(y, qmul(ww, b[l]));
    }
    qstore(W + i + (size_t)j * ldw, y);
    const quat u = qload(ucur + i);
    qstore(M + i + (size_t)c * ld, u);
    kp = u.ar * y.ar + u.ai * y.ai + u.br * y.br + u.bi * y.bi;
  }
  kp = block_sum(kp, sh);
  if (threadIdx.x == 0) KP[blockIdx.x] = kp;
}

// w = p - (u^* p / 2) u
__global__ void __launch_bounds__(RB) k_wfinal(const quat* __restrict__ M, int ld, int n, int c, int j,
                                               quat* __restrict__ W, int ldw, const double* __restrict__ KP,
                                               int nkp) {
  __shared__ double sh[RB / 64];
  double t = 0.0;
  for (int i = threadIdx.x; i < nkp; i += RB) t += KP[i];
  t = block_sum(t, sh);
  const double K = 0.5 * t;
  const int i = c + 1 + blockIdx.x * RB + threadIdx.x;
  if (i < n) {
    const quat u = qload(M + i + (size_t)c * ld);
    quat y = qload(W + i + (size_t)j * ldw);
    y = qsub(y, qscale(u, K));
    qstore(W + i + (size_t)j * ldw, y);
  }
}

__global__ void k_last_diag(const quat* __restrict__ M, int ld, int n, double* __restrict__ d) {
  if (threadIdx.x == 0) d[n - 1] = M[(n - 1) + (size_t)(n - 1) * ld].ar;
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

int reduce(quat* M, int n, int ld, int nb, double* d, quat* e, const ReduceWork& w, hipStream_t s) {
  if (n <= 0) return ZQI_OK;
  if (nb < 1 || nb > NBMAX) return ZQI_ERR_ARG;
  if (n == 1) {
    hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, s, M, ld, n, d);
    return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
  }
  const int ldw = n;
  int panel = 0;
  for (int p = 0; p < n - 1; p += nb, ++panel) {
    const int wdt = (nb < n - 1 - p) ? nb : (n - 1 - p);
    quat* Tf = w.Tf + (size_t)panel * nb * nb;
    for (int j = 0; j < wdt; ++j) {
      const int c = p + j;
      const int nrow = n - c;           // rows c..n-1
      const int nblk = cdiv(nrow, RB);
      hipLaunchKernelGGL(k_col_update, dim3(nblk), dim3(RB), 0, s, M, ld, n, p, c, j, w.W, ldw, w.part);
      hipLaunchKernelGGL(k_reflector, dim3(1), dim3(RB), 0, s, M, ld, n, c, w.part, nblk, w.scal, d, e);
      const int sz = n - c - 1;         // trailing size
      const int nt = cdiv(sz, TB);
      const int ntiles = nt * (nt + 1) / 2;
      const int nch = (j > 0) ? cdiv(sz, TSR) : 0;
      {
        // algorithmic bytes: one read of the trailing lower triangle
        ProfScope ps(PROF_SYMV, s, 32.0 * 0.5 * (double)sz * (sz + 1));
        hipLaunchKernelGGL(k_symv, dim3(ntiles + nch), dim3(256), 0, s, M, ld, n, p, c, j, w.W, ldw, w.scal, w.P,
                           n, w.ucur, w.TS, nb, ntiles);
      }
      const int nb2 = cdiv(sz, RB);
      hipLaunchKernelGGL(k_wreduce, dim3(nb2), dim3(RB), 0, s, M, ld, n, p, c, j, w.W, ldw, w.P, n, nt, w.TS, nch,
                         nb, w.ucur, w.KP, Tf, w.scal);
      hipLaunchKernelGGL(k_wfinal, dim3(nb2), dim3(RB), 0, s, M, ld, n, c, j, w.W, ldw, w.KP, nb2);
    }
    // trailing update of rows/cols [p+wdt, n): M -= [V W] [W V]^*  (lower triangle)
    const int r0 = p + wdt;
    const int mt = n - r0;
    if (mt > 0) {
      QOperand A{M + r0 + (size_t)p * ld, w.W + r0, ld, ldw, wdt, false, INT_MIN / 2};
      QOperand B{w.W + r0, M + r0 + (size_t)p * ld, ldw, ld, wdt, true, INT_MIN / 2};
      // algorithmic flops: lower triangle x rank 2*wdt, 32 flops per quaternion multiply-add
      ProfScope ps(PROF_TRAIL, s, 32.0 * 0.5 * (double)mt * (mt + 1) * 2.0 * wdt);
      int rc = qgemm(mt, mt, 2 * wdt, -1.0, A, B, 1.0, M + r0 + (size_t)r0 * ld, ld, true, s);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, s, M, ld, n, d);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
