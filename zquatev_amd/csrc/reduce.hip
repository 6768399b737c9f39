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

constexpr int RB = 256;      // threads per block
constexpr int RR = 64;       // rows per block in the row-parallel kernels (4 waves split the K sums)
constexpr int SW = 32;       // symv strip width (columns)
constexpr int SG = 128;      // symv rows per block (2 per lane)
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
// Also finishes w_{j-1} = p_{j-1} - K u_{c-1} (K = u^*p / 2 from the previous
// column's partials): each W(i, j-1), i > c, is read and rewritten by the one
// thread that uses it; W(c, j-1) is formed on the fly (no later kernel reads it).
__global__ void __launch_bounds__(RB) k_col_update(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                                   quat* __restrict__ W, int ldw, const double* __restrict__ KP,
                                                   int nkp, double* __restrict__ part) {
  __shared__ quat cw[NBMAX], cv[NBMAX];
  __shared__ quat red[4][RR];
  __shared__ double sh[RB / 64];
  double K = 0.0;
  if (j > 0) {
    double t = 0.0;
    for (int i = threadIdx.x; i < nkp; i += RB) t += KP[i];
    K = 0.5 * block_sum(t, sh);
  }
  for (int l = threadIdx.x; l < j; l += RB) {
    quat wc = qload(W + c + (size_t)l * ldw);
    if (l == j - 1) wc = qsub(wc, qscale(qload(M + c + (size_t)(c - 1) * ld), K));
    cw[l] = qconj(wc);
    cv[l] = qconj(qload(M + c + (size_t)(p + l) * ld));
  }
  __syncthreads();
  const int r = threadIdx.x & (RR - 1), q = threadIdx.x >> 6;
  const int i = c + blockIdx.x * RR + r;
  quat acc = qzero();
  if (i < n) {
#pragma unroll 4
    for (int l = q; l < j; l += 4) {
      quat wv = qload(W + i + (size_t)l * ldw);
      if (l == j - 1) {
        wv = qsub(wv, qscale(qload(M + i + (size_t)(c - 1) * ld), K));
        if (i > c) qstore(W + i + (size_t)l * ldw, wv);
      }
      qfma(acc, qload(M + i + (size_t)(p + l) * ld), cw[l]);
      qfma(acc, wv, cv[l]);
    }
  }
  red[q][r] = acc;
  __syncthreads();
  double ss = 0.0;
  if (q == 0 && i < n) {
    quat m = qload(M + i + (size_t)c * ld);
    if (j > 0) {
      m = qsub(m, qadd(qadd(red[0][r], red[1][r]), qadd(red[2][r], red[3][r])));
      qstore(M + i + (size_t)c * ld, m);
    }
    if (i >= c + 2) ss = qnorm2(m);
  }
  ss = block_sum(ss, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = ss;
}

// Reflector for column c: x = M[c+1:, c].  H x = alpha e_1 with
// alpha = -q |x|, q = x_1/|x_1|; u = gamma (x - alpha e_1), gamma = 1/sqrt(|x| (|x| + |x_1|)),
// so u^* u = 2.  A zero tail gives u = 0 and alpha = x_1.  Computed identically
// by every block of the matvec from the tail-norm partials.
struct Refl {
  double g;
  quat u1;
};
__device__ __forceinline__ Refl reflector(const quat* M, int ld, int c, double tail, quat* alpha) {
  const quat x1 = qload(M + (c + 1) + (size_t)c * ld);
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

__device__ __forceinline__ quat u_at(const quat* M, int ld, int c, int i, const Refl& R) {
  const quat x = qscale(qload(M + i + (size_t)c * ld), R.g);
  return (i == c + 1) ? R.u1 : x;
}

// Halving butterfly over a wave: on return lane l holds v[l >> (6 - log2 N)]...
// specialised below for the two sizes used.
template <int N>
__device__ __forceinline__ double butterfly_sum(double* v) {
  // N values per lane; after the exchange with masks 32, 16, ... down to 64/N
  // each lane keeps one value, summed over the lanes that differ in those bits
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int m = 32, cnt = N / 2; cnt >= 1; m >>= 1, cnt >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int k = 0; k < cnt; ++k) {
      const double send = hi ? v[k] : v[k + cnt];
      const double keep = hi ? v[k + cnt] : v[k];
      v[k] = keep + shfl_xor_d(send, m);
    }
  }
  double r = v[0];
  // remaining lane bits (below 64/N) still hold separate partial sums
#pragma unroll
  for (int m = 32 / N; m >= 1; m >>= 1) r += shfl_xor_d(r, m);
  return r;
}

// Hermitian matvec y = M[s0:, s0:] u (s0 = c + 1) over the lower triangle.
// Block (g, J) reads rows [s0 + SG g, +SG) x columns [s0 + SW J, +SW) (every
// lower-triangle element exactly once over the grid): wave w of 8 takes 4 of
// the 32 columns, each lane two rows, and all eight loads of a lane are in
// flight together.  Row sums (over the 32 columns) go to Prow[J][row]; column
// sums (over the 128 rows, reduced across the wave with a halving butterfly)
// go to Pcol[g][col].  Then
//   y(i) = sum_{J <= (i-s0)/SW} Prow[J][i] + sum_{g >= (i-s0)/SG} Pcol[g][i].
// Blocks are enumerated row-group major: group g has 4 (g + 1) strips.
// Second phase: the tall-skinny products W^* u and V^* u over the block's rows,
// one group of 4 panel columns per block of the row group, into TS[g].
constexpr int SYT = 512;     // threads of the matvec kernel
__global__ void __launch_bounds__(SYT) k_symv(const quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                              const quat* __restrict__ W, int ldw, const double* __restrict__ part,
                                              int nparts, double* __restrict__ scal, double* __restrict__ d,
                                              quat* __restrict__ e, quat* __restrict__ Prow, quat* __restrict__ Pcol,
                                              int ldp, quat* __restrict__ ucur, quat* __restrict__ TS, int nb) {
  __shared__ double sh[SYT / 64];
  __shared__ quat uJ[SW];
  __shared__ quat red[8][SG];
  __shared__ quat tsr[8][4][2];
  double tail = 0.0;
  for (int i = threadIdx.x; i < nparts; i += SYT) tail += part[i];
  tail = block_sum(tail, sh);
  quat alpha;
  const Refl R = reflector(M, ld, c, tail, &alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[0] = R.g;
    scal[1] = R.u1.ar; scal[2] = R.u1.ai; scal[3] = R.u1.br; scal[4] = R.u1.bi;
    d[c] = M[c + (size_t)c * ld].ar;
    qstore(e + c, alpha);
  }
  const int s0 = c + 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int b = blockIdx.x;
  // block -> (g, J): 2 g (g + 1) <= b < 2 (g + 1) (g + 2)
  int g = (int)((sqrt(2.0 * b + 1.0) - 1.0) * 0.5);
  while (2 * (g + 1) * (g + 2) <= b) ++g;
  while (2 * g * (g + 1) > b) --g;
  const int J = b - 2 * g * (g + 1);
  const int R0 = s0 + SG * g, C0 = s0 + SW * J;
  const int tid = threadIdx.x;
  const bool active = C0 < n;   // uniform for the block
  if (active) {
    if (tid < SW) uJ[tid] = (C0 + tid < n) ? u_at(M, ld, c, C0 + tid, R) : qzero();
    if (J == 0 && tid < SG && R0 + tid < n) qstore(ucur + R0 + tid, u_at(M, ld, c, R0 + tid, R));
  }
  __syncthreads();
  if (active) {
    const int ra = R0 + lane, rb = R0 + 64 + lane;
    const int rac = min(ra, n - 1), rbc = min(rb, n - 1);
    quat qa[4], qb[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int cgc = min(C0 + 4 * w + cc, n - 1);
      qa[cc] = qload(M + rac + (size_t)cgc * ld);
      qb[cc] = qload(M + rbc + (size_t)cgc * ld);
    }
    const quat ua = (ra < n) ? u_at(M, ld, c, rac, R) : qzero();
    const quat ub = (rb < n) ? u_at(M, ld, c, rbc, R) : qzero();
    quat acca = qzero(), accb = qzero();
    double col[16];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int cl = 4 * w + cc;
      const int cg = C0 + cl;
      const bool okc = cg < n;
      quat a = (okc && ra < n && ra >= cg) ? qa[cc] : qzero();
      quat bq = (okc && rb < n && rb >= cg) ? qb[cc] : qzero();
      if (ra == cg) a = quat{a.ar, 0.0, 0.0, 0.0};
      if (rb == cg) bq = quat{bq.ar, 0.0, 0.0, 0.0};
      const quat uc = uJ[cl];
      qfma(acca, a, uc);
      qfma(accb, bq, uc);
      quat cp = qzero();
      if (ra > cg) qfma_cl(cp, a, ua);
      if (rb > cg) qfma_cl(cp, bq, ub);
      col[4 * cc + 0] = cp.ar; col[4 * cc + 1] = cp.ai; col[4 * cc + 2] = cp.br; col[4 * cc + 3] = cp.bi;
    }
    red[w][lane] = acca;
    red[w][64 + lane] = accb;
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
      quat y = qzero();
#pragma unroll
      for (int k = 0; k < 8; ++k) y = qadd(y, red[k][tid]);
      qstore(Prow + (size_t)J * ldp + i, y);
    }
  }
  // tall-skinny phase: block (g, J) handles panel-column groups J, J + 4(g+1), ...
  const int ngr = (j + 3) >> 2;
  const int stride = 4 * (g + 1);
  for (int lg = J; lg < ngr; lg += stride) {
    const int lo = tid & 3;
    const int row = R0 + (tid >> 2);
    const int l = 4 * lg + lo;
    quat aw = qzero(), av = qzero();
    if (row < n && l < j) {
      const quat ur = u_at(M, ld, c, row, R);
      qfma_cl(aw, qload(W + row + (size_t)l * ldw), ur);
      qfma_cl(av, qload(M + row + (size_t)(p + l) * ld), ur);
    }
#pragma unroll
    for (int m = 4; m <= 32; m <<= 1) {
      aw.ar += shfl_xor_d(aw.ar, m); aw.ai += shfl_xor_d(aw.ai, m);
      aw.br += shfl_xor_d(aw.br, m); aw.bi += shfl_xor_d(aw.bi, m);
      av.ar += shfl_xor_d(av.ar, m); av.ai += shfl_xor_d(av.ai, m);
      av.br += shfl_xor_d(av.br, m); av.bi += shfl_xor_d(av.bi, m);
    }
    if (lane < 4) {
      tsr[w][lane][0] = aw;
      tsr[w][lane][1] = av;
    }
    __syncthreads();
    if (tid < 32) {
      const int side = tid >> 4, ll = (tid >> 2) & 3, comp = tid & 3;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) acc += reinterpret_cast<const double*>(&tsr[k][ll][side])[comp];
      if (4 * lg + ll < j) reinterpret_cast<double*>(TS + ((size_t)g * 2 + side) * nb + 4 * lg + ll)[comp] = acc;
    }
    __syncthreads();
  }
}

// p = sum of matvec partials - V (W^* u) - W (V^* u); stored in W[:, j].  Also
// stores u into M[c+1:, c], partial u^* p into KP, and (block 0) the T-factor
// column j: T[0:j, j] = -T[0:j, 0:j] (V^* u), T[j, j] = 1 (0 if u = 0) and
// zeros below, so that H_p ... H_{p+nb-1} = I - V T V^* with T stored densely.
__global__ void __launch_bounds__(RB) k_wreduce(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                                quat* __restrict__ W, int ldw, const quat* __restrict__ Prow,
                                                const quat* __restrict__ Pcol, int ldp, const quat* __restrict__ TS,
                                                int nchunks, int nb, const quat* __restrict__ ucur,
                                                double* __restrict__ KP, quat* __restrict__ Gm) {
  __shared__ quat a[NBMAX], b[NBMAX];
  __shared__ quat red[4][RR];
  __shared__ double sh[RB / 64];
  for (int l = threadIdx.x; l < 2 * j; l += RB) {
    const int side = l / j, ll = l - side * j;
    quat s = qzero();
#pragma unroll 8
    for (int h = 0; h < nchunks; ++h) s = qadd(s, qload(TS + ((size_t)h * 2 + side) * nb + ll));
    if (side == 0) a[ll] = s; else b[ll] = s;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    // column j of G = V^* V (strict upper part): T = (I + triu(G, 1))^{-1} later
    for (int r = threadIdx.x; r < nb; r += RB) qstore(Gm + r + (size_t)j * nb, r < j ? b[r] : qzero());
  }
  const int s0 = c + 1;
  const int sz = n - s0;
  const int G = (sz + SG - 1) / SG;
  const int r = threadIdx.x & (RR - 1), q = threadIdx.x >> 6;
  const int i = s0 + blockIdx.x * RR + r;
  quat acc = qzero();
  if (i < n) {
    const int li = i - s0;
#pragma unroll 4
    for (int J = q; J <= li / SW; J += 4) acc = qadd(acc, qload(Prow + (size_t)J * ldp + i));
#pragma unroll 4
    for (int gg = li / SG + q; gg < G; gg += 4) acc = qadd(acc, qload(Pcol + (size_t)gg * ldp + i));
#pragma unroll 4
    for (int l = q; l < j; l += 4) {
      acc = qsub(acc, qmul(qload(M + i + (size_t)(p + l) * ld), a[l]));
      acc = qsub(acc, qmul(qload(W + i + (size_t)l * ldw), b[l]));
    }
  }
  red[q][r] = acc;
  __syncthreads();
  double kp = 0.0;
  if (q == 0 && i < n) {
    const quat y = qadd(qadd(red[0][r], red[1][r]), qadd(red[2][r], red[3][r]));
    qstore(W + i + (size_t)j * ldw, y);
    const quat u = qload(ucur + i);
    qstore(M + i + (size_t)c * ld, u);
    kp = u.ar * y.ar + u.ai * y.ai + u.br * y.br + u.bi * y.bi;
  }
  kp = block_sum(kp, sh);
  if (threadIdx.x == 0) KP[blockIdx.x] = kp;
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
    quat* Gp = w.Gm + (size_t)panel * nb * nb;
    for (int j = 0; j < wdt; ++j) {
      const int c = p + j;
      const int nrow = n - c;                      // rows c..n-1
      const int nblk = cdiv(nrow, RR);
      const int nkp = cdiv(nrow, RR);              // blocks of the previous column's k_wreduce
      hipLaunchKernelGGL(k_col_update, dim3(nblk), dim3(RB), 0, s, M, ld, n, p, c, j, w.W, ldw, w.KP, nkp, w.part);
      const int sz = n - c - 1;                    // trailing size
      const int G = cdiv(sz, SG);
      const int nmain = 2 * G * (G + 1);
      const int nch = (j > 0) ? G : 0;             // tall-skinny partials: one per row group
      {
        // algorithmic bytes: one read of the trailing lower triangle
        ProfScope ps(PROF_SYMV, s, 32.0 * 0.5 * (double)sz * (sz + 1));
        hipLaunchKernelGGL(k_symv, dim3(nmain), dim3(SYT), 0, s, M, ld, n, p, c, j, w.W, ldw, w.part, nblk,
                           w.scal, d, e, w.Prow, w.Pcol, n, w.ucur, w.TS, nb);
      }
      const int nb2 = cdiv(sz, RR);
      hipLaunchKernelGGL(k_wreduce, dim3(nb2), dim3(RB), 0, s, M, ld, n, p, c, j, w.W, ldw, w.Prow, w.Pcol, n, w.TS,
                         nch, nb, w.ucur, w.KP, Gp);
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
