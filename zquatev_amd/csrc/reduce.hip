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
#include <algorithm>

namespace zq {

constexpr int RB = 256;      // threads per block of the row kernels
constexpr int RQ = 16;       // rows per block of the row kernels (16 threads per row split the K sums)
constexpr int SW = 32;       // matvec strip width (columns)
constexpr int SG = 64;       // matvec rows per block (one per lane)
constexpr int SYT = 512;     // matvec threads: 8 waves x 4 columns
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

// Column c of the panel (index j): apply the panel's pending update
//   M[i,c] -= sum_{l<j} V(i,l) W(c,l)^* + W(i,l) V(c,l)^*     for i in [c, n)
// and accumulate |M[i,c]|^2 over i >= c+2 (the reflector's tail norm).
// Also finishes w_{j-1} = p_{j-1} - K u_{c-1} (K = u^*p / 2 from the previous
// column's partials): each W(i, j-1), i > c, is read and rewritten by the one
// thread that uses it; W(c, j-1) is formed on the fly (no later kernel reads
// it).  Zeroes the accumulator TS that the matvec of column c fills.
// 16 threads per row, KPT panel columns each; every load is issued up front
// at a clamped address.
template <int KPT>
__global__ void __launch_bounds__(RB) k_col_update(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                                   quat* __restrict__ W, int ldw, const double* __restrict__ KP,
                                                   int nkp, double* __restrict__ part, quat* __restrict__ TS,
                                                   int nb) {
  __shared__ quat cw[NBMAX], cv[NBMAX];
  __shared__ quat red[16][RQ];
  __shared__ double sh[RB / 64];
  const int tid = threadIdx.x;
  const int r = tid & (RQ - 1), q = tid >> 4;
  const int i = c + blockIdx.x * RQ + r;
  const int ic = min(i, n - 1);
  const int jl = j > 0 ? j - 1 : 0;
  const quat mcur = qload(M + ic + (size_t)c * ld);
  const quat uprev = qload(M + ic + (size_t)(c > 0 ? c - 1 : 0) * ld);
  quat vv[KPT], ww[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int l = min(q + 16 * k, jl);
    vv[k] = qload(M + ic + (size_t)(p + l) * ld);
    ww[k] = qload(W + ic + (size_t)l * ldw);
  }
  double t = 0.0;
  for (int k = tid; k < nkp; k += RB) t += KP[k];
  quat cwl = qzero(), cvl = qzero(), ucc = qzero();
  if (tid < j) {
    cwl = qload(W + c + (size_t)tid * ldw);
    cvl = qload(M + c + (size_t)(p + tid) * ld);
    if (tid == j - 1) ucc = qload(M + c + (size_t)(c - 1) * ld);
  }
  // the matvec of this column accumulates into TS
  if (blockIdx.x == 0) {
    for (int k = tid; k < 2 * nb; k += RB) qstore(TS + k, qzero());
  }
  const double K = (j > 0) ? 0.5 * block_sum(t, sh) : 0.0;
  if (tid < j) {
    if (tid == j - 1) cwl = qsub(cwl, qscale(ucc, K));
    cw[tid] = qconj(cwl);
    cv[tid] = qconj(cvl);
  }
  __syncthreads();
  quat acc = qzero();
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int l = q + 16 * k;
    if (l < j) {
      quat wv = ww[k];
      if (l == j - 1) {
        wv = qsub(wv, qscale(uprev, K));
        if (i > c && i < n) qstore(W + i + (size_t)l * ldw, wv);
      }
      qfma(acc, vv[k], cw[l]);
      qfma(acc, wv, cv[l]);
    }
  }
  red[q][r] = acc;
  __syncthreads();
  double ss = 0.0;
  if (q == 0 && i < n) {
    quat m = mcur;
    if (j > 0) {
      quat sum = qzero();
#pragma unroll
      for (int k = 0; k < 16; ++k) sum = qadd(sum, red[k][r]);
      m = qsub(m, sum);
      qstore(M + i + (size_t)c * ld, m);
    }
    if (i >= c + 2) ss = qnorm2(m);
  }
  ss = block_sum(ss, sh);
  if (tid == 0) part[blockIdx.x] = ss;
}

// Reflector for column c: x = M[c+1:, c].  H x = alpha e_1 with
// alpha = -q |x|, q = x_1/|x_1|; u = gamma (x - alpha e_1), gamma = 1/sqrt(|x| (|x| + |x_1|)),
// so u^* u = 2.  A zero tail gives u = 0 and alpha = x_1.  Computed identically
// by every block of the matvec from the tail-norm partials.
struct Refl {
  double g;
  quat u1;
};
__device__ __forceinline__ Refl reflector(const quat& x1, double tail, quat* alpha) {
  const double ax1 = sqrt(qnorm2(x1));
  Refl R;
  if (!(tail > 0.0)) {
    R.g = 0.0;
    R.u1 = qzero();
    *alpha = x1;
  } else {
    const double nx = sqrt(ax1 * ax1 + tail);
    const quat qph = (ax1 > 0.0) ? qscale(x1, 1.0 / ax1) : quat{1.0, 0.0, 0.0, 0.0};
    const double sm = ax1 + nx;
    R.g = 1.0 / sqrt(nx * sm);
    R.u1 = qscale(qph, sm * R.g);
    *alpha = qscale(qph, -nx);
  }
  return R;
}

// u_i from the raw column entry x_i = M[i, c]
__device__ __forceinline__ quat u_from(const quat& x, int c, int i, const Refl& R) {
  return (i == c + 1) ? R.u1 : qscale(x, R.g);
}

// Hermitian matvec y = M[s0:, s0:] u (s0 = c + 1) over the lower triangle.
// Block (g, J) reads rows [s0 + SG g, +SG) x columns [s0 + SW J, +SW) (every
// lower-triangle element exactly once over the grid); wave w of 8 takes 4 of
// the 32 columns, one row per lane.  The matrix loads are issued first, so
// the reflector set-up (tail-norm reduction, x_1) runs under their latency.
// Row sums over the 32 columns (summed across the waves in LDS) go to
// Prow[J][row]; column sums over the 64 rows (a VALU-only halving butterfly,
// DPP and permlane swaps) go to Pcol[g][col]; then
//   y(i) = sum_{J <= (i-s0)/SW} Prow[J][i] + sum_{g >= (i-s0)/SG} Pcol[g][i],
// every (slot, row) written exactly once, so the sum is deterministic.
// Blocks are enumerated row-group major: group g has 2 (g + 1) strips.
// Second phase: the tall-skinny products W^* u and V^* u over the block's
// rows, groups of 8 panel columns spread over the row group's blocks, added
// into TS = [W^* u | V^* u] with FP64 atomics.
__global__ void __launch_bounds__(SYT) k_symv(const quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                              const quat* __restrict__ W, int ldw, const double* __restrict__ part,
                                              int nparts, double* __restrict__ scal, double* __restrict__ d,
                                              quat* __restrict__ e, quat* __restrict__ Prow, quat* __restrict__ Pcol,
                                              int ldp, quat* __restrict__ ucur, quat* __restrict__ TS, int nb) {
  __shared__ double sh[SYT / 64];
  __shared__ quat uJ[SW];
  __shared__ quat red[8][SG];
  __shared__ quat tsr[8][8][2];
  const int s0 = c + 1;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int b = blockIdx.x;
  // block -> (g, J): g (g + 1) <= b < (g + 1) (g + 2)
  int g = (int)((sqrt(4.0 * b + 1.0) - 1.0) * 0.5);
  while ((g + 1) * (g + 2) <= b) ++g;
  while (g * (g + 1) > b) --g;
  const int J = b - g * (g + 1);
  const int R0 = s0 + SG * g, C0 = s0 + SW * J;
  const bool active = C0 < n && R0 < n;   // uniform for the block
  const int ra = R0 + lane;
  const int rac = min(ra, n - 1);
  quat qa[4];
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    const int cgc = min(C0 + 4 * w + cc, n - 1);
    qa[cc] = qload(M + rac + (size_t)cgc * ld);
  }
  const quat xa = qload(M + rac + (size_t)c * ld);
  const quat xj = qload(M + min(C0 + (tid & (SW - 1)), n - 1) + (size_t)c * ld);
  const quat x1 = qload(M + (c + 1) + (size_t)c * ld);
  // tall-skinny task of this block (at most one): rows of group tg, panel columns 8 tl .. 8 tl + 7
  const int ngr = (j + 7) >> 3;
  const int Gt = (n - s0 + SG - 1) / SG;
  const bool has_ts = b < Gt * ngr;
  const int tg = has_ts ? b / ngr : 0, tl = has_ts ? b - (b / ngr) * ngr : 0;
  const int trow = s0 + SG * tg + (tid >> 3);
  const int tcol = 8 * tl + (tid & 7);
  const int trc = min(trow, n - 1), tcc = min(tcol, j > 0 ? j - 1 : 0);
  double tail = 0.0;
  for (int k = tid; k < nparts; k += SYT) tail += part[k];
  tail = block_sum(tail, sh);
  quat alpha;
  const Refl R = reflector(x1, tail, &alpha);
  if (b == 0 && tid == 0) {
    scal[0] = R.g;
    scal[1] = R.u1.ar; scal[2] = R.u1.ai; scal[3] = R.u1.br; scal[4] = R.u1.bi;
    d[c] = M[c + (size_t)c * ld].ar;
    qstore(e + c, alpha);
  }
  if (tid < SW) uJ[tid] = (C0 + tid < n) ? u_from(xj, c, C0 + tid, R) : qzero();
  const quat ua = (ra < n) ? u_from(xa, c, ra, R) : qzero();
  if (active && J == 0 && w == 0 && ra < n) qstore(ucur + ra, ua);
  __syncthreads();
  if (active) {
    quat acc = qzero();
    double col[16];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int cl = 4 * w + cc;
      const int cg = C0 + cl;
      quat a = (cg < n && ra < n && ra >= cg) ? qa[cc] : qzero();
      if (ra == cg) a = quat{a.ar, 0.0, 0.0, 0.0};
      qfma(acc, a, uJ[cl]);
      quat cp = qzero();
      if (ra > cg) qfma_cl(cp, a, ua);
      col[4 * cc + 0] = cp.ar; col[4 * cc + 1] = cp.ai; col[4 * cc + 2] = cp.br; col[4 * cc + 3] = cp.bi;
    }
    red[w][lane] = acc;
    // lane l ends with value (l >> 2) = component (l >> 2) & 3 of column (l >> 4), summed over the wave
    const double cs = butterfly_sum<16>(col);
    if ((lane & 3) == 0) {
      const int cg = C0 + 4 * w + (lane >> 4);
      if (cg < n) reinterpret_cast<double*>(Pcol + (size_t)g * ldp + cg)[(lane >> 2) & 3] = cs;
    }
  }
  __syncthreads();
  if (active && tid < SG) {
    const int i = R0 + tid;
    if (i < n && i >= C0) {
      quat sum = qzero();
#pragma unroll
      for (int k = 0; k < 8; ++k) sum = qadd(sum, red[k][tid]);
      qstore(Prow + (size_t)J * ldp + i, sum);
    }
  }
  // tall-skinny task: W^* u and V^* u over 64 rows for 8 panel columns
  if (has_ts) {
    const quat tw = qload(W + trc + (size_t)tcc * ldw);
    const quat tv = qload(M + trc + (size_t)(p + tcc) * ld);
    const quat tx = qload(M + trc + (size_t)c * ld);
    const bool ok = trow < n && tcol < j;
    quat aw = qzero(), av = qzero();
    if (ok) {
      const quat ur = u_from(tx, c, trow, R);
      qfma_cl(aw, tw, ur);
      qfma_cl(av, tv, ur);
    }
    // sum over the 8 rows of the wave (lanes differing in bits 3..5)
    double t8[8] = {aw.ar, aw.ai, aw.br, aw.bi, av.ar, av.ai, av.br, av.bi};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double x = t8[k] + xlane<8>(t8[k]);
      x = bf16(x, x);
      t8[k] = bf32(x, x);
    }
    if (lane < 8) {
      tsr[w][lane][0] = quat{t8[0], t8[1], t8[2], t8[3]};
      tsr[w][lane][1] = quat{t8[4], t8[5], t8[6], t8[7]};
    }
    __syncthreads();
    if (tid < 64) {
      const int side = tid >> 5, ll = (tid >> 2) & 7, comp = tid & 3;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) acc += reinterpret_cast<const double*>(&tsr[k][ll][side])[comp];
      if (8 * tl + ll < j) atomicAdd(reinterpret_cast<double*>(TS + (size_t)side * nb + 8 * tl + ll) + comp, acc);
    }
  }
}

// p = y - V (W^* u) - W (V^* u), y from the matvec partials; stored in W[:, j].  Also stores u into
// M[c+1:, c], partial u^* p into KP, and (block 0) column j of the strict
// upper part of G = V^* V (the T factor is formed from G after the panel).
template <int KPT>
__global__ void __launch_bounds__(RB) k_wreduce(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                                quat* __restrict__ W, int ldw, const quat* __restrict__ Prow,
                                                const quat* __restrict__ Pcol, int ldp, const quat* __restrict__ TS,
                                                int nb, const quat* __restrict__ ucur,
                                                double* __restrict__ KP, quat* __restrict__ Gm) {
  __shared__ quat a[NBMAX], b[NBMAX];
  __shared__ quat red[16][RQ];
  __shared__ double sh[RB / 64];
  const int tid = threadIdx.x;
  const int s0 = c + 1;
  const int r = tid & (RQ - 1), q = tid >> 4;
  const int i = s0 + blockIdx.x * RQ + r;
  const int ic = min(i, n - 1);
  const int jl = j > 0 ? j - 1 : 0;
  quat vv[KPT], ww[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int l = min(q + 16 * k, jl);
    vv[k] = qload(M + ic + (size_t)(p + l) * ld);
    ww[k] = qload(W + ic + (size_t)l * ldw);
  }
  const quat ui = qload(ucur + ic);
  quat yi = qzero();
  {
    const int li = ic - s0;
    const int G = (n - s0 + SG - 1) / SG;
#pragma unroll 4
    for (int J = q; J <= li / SW; J += 16) yi = qadd(yi, qload(Prow + (size_t)J * ldp + ic));
#pragma unroll 4
    for (int gg = li / SG + q; gg < G; gg += 16) yi = qadd(yi, qload(Pcol + (size_t)gg * ldp + ic));
  }
  if (tid < j) {
    a[tid] = qload(TS + tid);
    b[tid] = qload(TS + nb + tid);
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int rr = tid; rr < nb; rr += RB) qstore(Gm + rr + (size_t)j * nb, rr < j ? b[rr] : qzero());
  }
  quat acc = yi;   // this thread's share of y(i) minus its share of the corrections
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int l = q + 16 * k;
    if (l < j) {
      acc = qsub(acc, qmul(vv[k], a[l]));
      acc = qsub(acc, qmul(ww[k], b[l]));
    }
  }
  red[q][r] = acc;
  __syncthreads();
  double kp = 0.0;
  if (q == 0 && i < n) {
    quat pv = qzero();
#pragma unroll
    for (int k = 0; k < 16; ++k) pv = qadd(pv, red[k][r]);
    qstore(W + i + (size_t)j * ldw, pv);
    qstore(M + i + (size_t)c * ld, ui);
    kp = ui.ar * pv.ar + ui.ai * pv.ai + ui.br * pv.br + ui.bi * pv.bi;
  }
  kp = block_sum(kp, sh);
  if (tid == 0) KP[blockIdx.x] = kp;
}

// w = p - (u^* p / 2) u  (last column of a panel; the others are finished by
// the next column's k_col_update)
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

// T = (I + triu(G, 1))^{-1} for each panel: T[0:j, j] = -T[0:j, 0:j] G[0:j, j],
// T[j, j] = 1, zero below (equivalent to the usual compact-WY recurrence
// since u^* u = 2 makes every tau 1).  One workgroup per panel.
__global__ void __launch_bounds__(RB) k_tfactor(int n, int nb, const quat* __restrict__ Gall, quat* __restrict__ Tall) {
  const int panel = blockIdx.x;
  const int p = panel * nb;
  const int wdt = min(nb, n - 1 - p);
  const quat* G = Gall + (size_t)panel * nb * nb;
  quat* T = Tall + (size_t)panel * nb * nb;
  for (int jj = 0; jj < nb; ++jj) {
    for (int r = threadIdx.x; r < nb; r += RB) {
      quat acc = qzero();
      if (jj < wdt && r < jj) {
#pragma unroll 4
        for (int l = r; l < jj; ++l) qfma(acc, qload(T + r + (size_t)l * nb), qload(G + l + (size_t)jj * nb));
        acc = qscale(acc, -1.0);
      } else if (r == jj) {
        acc = quat{1.0, 0.0, 0.0, 0.0};
      }
      qstore(T + r + (size_t)jj * nb, acc);
    }
    __syncthreads();
  }
}

__global__ void k_last_diag(const quat* __restrict__ M, int ld, int n, double* __restrict__ d) {
  if (threadIdx.x == 0) d[n - 1] = M[(n - 1) + (size_t)(n - 1) * ld].ar;
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// panel columns per thread (16 threads per row) as a compile-time bound
static int kpt_for(int nb) {
  const int k = cdiv(nb, 16);
  return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8;
}

template <class... Args>
static void launch_col_update(int kpt, dim3 grid, hipStream_t s, Args... args) {
  switch (kpt) {
    case 1: hipLaunchKernelGGL(k_col_update<1>, grid, dim3(RB), 0, s, args...); break;
    case 2: hipLaunchKernelGGL(k_col_update<2>, grid, dim3(RB), 0, s, args...); break;
    case 4: hipLaunchKernelGGL(k_col_update<4>, grid, dim3(RB), 0, s, args...); break;
    default: hipLaunchKernelGGL(k_col_update<8>, grid, dim3(RB), 0, s, args...); break;
  }
}

template <class... Args>
static void launch_wreduce(int kpt, dim3 grid, hipStream_t s, Args... args) {
  switch (kpt) {
    case 1: hipLaunchKernelGGL(k_wreduce<1>, grid, dim3(RB), 0, s, args...); break;
    case 2: hipLaunchKernelGGL(k_wreduce<2>, grid, dim3(RB), 0, s, args...); break;
    case 4: hipLaunchKernelGGL(k_wreduce<4>, grid, dim3(RB), 0, s, args...); break;
    default: hipLaunchKernelGGL(k_wreduce<8>, grid, dim3(RB), 0, s, args...); break;
  }
}

int reduce(quat* M, int n, int ld, int nb, double* d, quat* e, const ReduceWork& w, hipStream_t s) {
  if (n <= 0) return ZQI_OK;
  if (nb < 1 || nb > NBMAX) return ZQI_ERR_ARG;
  if (n == 1) {
    hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, s, M, ld, n, d);
    return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
  }
  const int ldw = n;
  const int kpt = kpt_for(nb);
  int panel = 0;
  for (int p = 0; p < n - 1; p += nb, ++panel) {
    const int wdt = (nb < n - 1 - p) ? nb : (n - 1 - p);
    quat* Gp = w.Gm + (size_t)panel * nb * nb;
    for (int j = 0; j < wdt; ++j) {
      const int c = p + j;
      const int nrow = n - c;                      // rows c..n-1
      const int nblk = cdiv(nrow, RQ);
      const int nkp = cdiv(nrow, RQ);              // blocks of the previous column's k_wreduce
      launch_col_update(kpt, dim3(nblk), s, M, ld, n, p, c, j, w.W, ldw, w.KP, nkp, w.part, w.TS, nb);
      const int sz = n - c - 1;                    // trailing size
      const int G = cdiv(sz, SG);
      const int nmain = std::max(G * (G + 1), G * ((j + 7) / 8));
      {
        // algorithmic bytes: one read of the trailing lower triangle
        ProfScope ps(PROF_SYMV, s, 32.0 * 0.5 * (double)sz * (sz + 1));
        launch_k(k_symv, dim3(nmain), dim3(SYT), 0, s, M, ld, n, p, c, j, w.W, ldw, w.part, nblk,
                           w.scal, d, e, w.Prow, w.Pcol, n, w.ucur, w.TS, nb);
      }
      const int nb2 = cdiv(sz, RQ);
      launch_wreduce(kpt, dim3(nb2), s, M, ld, n, p, c, j, w.W, ldw, w.Prow, w.Pcol, n, w.TS, nb, w.ucur, w.KP, Gp);
      if (j == wdt - 1)
        hipLaunchKernelGGL(k_wfinal, dim3(cdiv(sz, RB)), dim3(RB), 0, s, M, ld, n, c, j, w.W, ldw, w.KP, nb2);
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
  hipLaunchKernelGGL(k_tfactor, dim3(panel), dim3(RB), 0, s, n, nb, w.Gm, w.Tf);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
