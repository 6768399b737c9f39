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
#include <vector>

namespace zq {

constexpr int RB = 256;      // threads per block of the row kernels
constexpr int RQ = 16;       // rows per block of the row kernels (16 threads per row split the K sums)
constexpr int SG = 64;       // matvec rows per block (one per lane)
// Matvec tile shapes: SG rows x SW columns.  Trailing sizes up to the switch
// tsm use narrow strips (SW = 16, 256 threads: more, smaller blocks, which
// keeps the small matvecs of batches and of the end of a solve spread over
// the CUs); larger ones wide strips (SW = 32, 512 threads: fewer row
// partials for the column kernel to gather).  spg = SG / SW strips per row
// group.  The switch is a per-context tuning parameter (0 .. TSM_MAX; the
// autotuner times candidates), 768 by default.
constexpr int TSM_MAX = 2048;
__host__ __device__ constexpr int matvec_spg(int m, int tsm) { return m <= tsm ? 4 : 2; }
// matvec tiles of a trailing size with G row groups: group g holds spg (g + 1) strips
__host__ __device__ constexpr int ntiles(int G, int spg) { return spg * G * (G + 1) / 2; }
constexpr int NBMAX = 128;   // largest supported panel width
constexpr int TAILN = 64;    // trailing size from which one workgroup finishes in LDS (k_tail)

// ceil(a / b) for b > 0, and 0 for a <= 0
__device__ __forceinline__ int cdiv_d(int a, int b) { return a > 0 ? (a + b - 1) / b : 0; }

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


// Hermitian matvec over the lower triangle of the trailing matrix
// M[s0:, s0:] (s0 = c + 1), applied to the raw updated column x = M[s0:, c]
// rather than to the reflector u: u = gamma x + (u1 - gamma x1) e_1, so
//   M u = gamma (M x) + (u1 - gamma x1) M[s0:, s0],
// which k_wcol forms row by row.  The tile blocks therefore never wait for
// the reflector (no norm reduction on their path).  Each tile block also
// leaves its share of x^* M x (the lower part doubled off the diagonal) in
// QF[b], from which k_wcol forms u^* M u without another grid-wide reduction.
// Block (g, J) reads rows [s0 + SG g, +SG) x columns [s0 + SW J, +SW) (every
// lower-triangle element exactly once over the grid); wave w of NWV takes SCW of
// the SW columns, one row per lane, all loads issued up front.  Row sums over
// the SW columns (summed across the waves in LDS) go to Prow[J][row]; column
// sums over the SG rows (a VALU-only halving butterfly, DPP and permlane
// swaps) go to Pcol[g][col]; then
//   (M x)(i) = sum_{J <= (i-s0)/SW} Prow[J][i] + sum_{g >= (i-s0)/SG} Pcol[g][i],
// every (slot, row) written exactly once, so the sum is deterministic.
// Blocks are enumerated row-group major: group g has SPG (g + 1) strips.
// The max(1, ceil(8 j / 64)) blocks before the ntiles(G) tiles compute the
// reflector from the tail-norm partials (from the column itself for a
// panel's first column, j = 0) (the first of them publishes gamma,
// u1, x1 in scal and d[c], e[c]) and finish the tall-skinny products
// TS = [W^* u | V^* u]: they add the per-block partials of W^* x and V^* x
// that k_wcol left in TSP (the waves over the partial index, then the
// wave sums, in a fixed order), scale by gamma and add the row s0 term.
template <int SYT, int SW>
__global__ void __launch_bounds__(SYT) k_symv(const quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                              const quat* __restrict__ W, int ldw, const double* __restrict__ part,
                                              int nparts, double* __restrict__ scal, double* __restrict__ d,
                                              quat* __restrict__ e, quat* __restrict__ Prow, quat* __restrict__ Pcol,
                                              int ldp, quat* __restrict__ TS, int nb, const double* __restrict__ TSP,
                                              int ldtsp, double* __restrict__ QF, ZB zb) {
  constexpr int NWV = SYT / 64;    // waves per block
  constexpr int SCW = SW / NWV;     // strip columns per wave
  constexpr int SPG = SG / SW;      // strips per row group and group index
  static_assert(SW <= SG && SG % SW == 0 && SW % NWV == 0, "matvec tile shape");
  const int NT = j > 0 ? (8 * j + 63) / 64 : 0;
  const int NF = NT > 0 ? NT : 1;
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    nparts = (n - c + RQ - 1) / RQ;
    const int Gz = (n - c - 1 + SG - 1) / SG;
    if ((int)blockIdx.x >= ntiles(Gz, SPG) + NF) return;
    M = zoff(M, zb); W = zoff(W, zb); part = zoff(part, zb); scal = zoff(scal, zb); d = zoff(d, zb);
    e = zoff(e, zb); Prow = zoff(Prow, zb); Pcol = zoff(Pcol, zb); TS = zoff(TS, zb); TSP = zoff(TSP, zb);
    QF = zoff(QF, zb);
  }
  __shared__ double sh[SYT / 64];
  __shared__ double sqf[SYT / 64];
  __shared__ double red[NWV][4][SG];   // wave, component, row: conflict-free b64 accesses
  const int s0 = c + 1;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: its column loads are scalar
  // blocks [0, NF) finish the column (dispatched first, so they overlap the
  // tiles), blocks NF + b take tile b
  const int bt = blockIdx.x;
  const int b = bt - NF;
  if (b < 0) {
    // ---- finishing block: reflector, scalars, tall-skinny products
    const bool has_ts = bt < NT;
    // every load of the block is issued before the first wait: the row-s0
    // entries of W and V (for the u_1 term of TS), the diagonal entries, and
    // the partials in batches of 16 per lane (one batch up to m = 2048)
    const int e0 = 64 * bt + 4 * (tid & 15);
    const int lts = e0 >> 3, side = (e0 >> 2) & 1;
    const bool fin = has_ts && tid < 16 && lts < j;
    quat rowv = qzero();
    if (fin) rowv = side == 0 ? qload(W + s0 + (size_t)lts * ldw) : qload(M + s0 + (size_t)(p + lts) * ld);
    double mss = 0.0, mcc = 0.0;
    if (bt == 0 && tid == 0) {
      mss = M[s0 + (size_t)s0 * ld].ar;
      mcc = M[c + (size_t)c * ld].ar;
    }
    double tsacc = 0.0;
    if (has_ts) {
      const int eidx = 64 * bt + lane;
      constexpr int TB = 16;
      for (int k0 = w; k0 < nparts; k0 += NWV * TB) {
        double v[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
          const int k = k0 + NWV * u;
          v[u] = k < nparts ? TSP[(size_t)k * ldtsp + eidx] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) tsacc += v[u];
      }
    }
    const quat x1 = qload(M + s0 + (size_t)c * ld);
    double tail = 0.0;
    if (j == 0) {   // a panel's first column has no pending update: its tail norm from the column itself
      for (int i = c + 2 + tid; i < n; i += SYT) tail += qnorm2(qload(M + i + (size_t)c * ld));
    } else {        // else from the partials the column kernel left for it
      for (int k = tid; k < nparts; k += SYT) tail += part[k];
    }
    tail = block_sum(tail, sh);
    quat alpha;
    const Refl R = reflector(x1, tail, &alpha);
    if (bt == 0 && tid == 0) {
      scal[0] = R.g;
      scal[1] = R.u1.ar; scal[2] = R.u1.ai; scal[3] = R.u1.br; scal[4] = R.u1.bi;
      scal[5] = x1.ar; scal[6] = x1.ai; scal[7] = x1.br; scal[8] = x1.bi;
      // the trailing diagonal M[s0, s0] before k_wcol's update of column s0 overwrites it
      scal[9] = mss;
      d[c] = mcc;
      qstore(e + c, alpha);
    }
    if (has_ts) {
      __shared__ double tsr[NWV][64];
      tsr[w][lane] = tsacc;
      __syncthreads();
      // quaternion (side, l) = entries 64 bt + 4 tid .. + 3, entry index 8 l + 4 side + component
      if (fin) {
        double sc[4];
#pragma unroll
        for (int comp = 0; comp < 4; ++comp) {
          double t = 0.0;
#pragma unroll
          for (int k = 0; k < NWV; ++k) t += tsr[k][4 * tid + comp];
          sc[comp] = t;
        }
        quat t = qscale(quat{sc[0], sc[1], sc[2], sc[3]}, R.g);
        qfma_cl(t, rowv, R.u1);
        qstore(TS + (size_t)side * nb + lts, t);
      }
    }
    return;
  }
  // ---- tile block: block -> (g, J) with ntiles(g) <= b < ntiles(g + 1)
  int g = (int)((sqrt(8.0 * b / SPG + 1.0) - 1.0) * 0.5);
  while (ntiles(g + 1, SPG) <= b) ++g;
  while (ntiles(g, SPG) > b) --g;
  const int J = b - ntiles(g, SPG);
  const int R0 = s0 + SG * g, C0 = s0 + SW * J;
  const bool active = C0 < n && R0 < n;   // uniform for the block
  const int ra = R0 + lane;
  const int rac = min(ra, n - 1);
  // Above the diagonal (row < column: unused, masked below) a lane reads the
  // column's diagonal entry instead, a line the tile fetches anyway, so the
  // diagonal tiles' upper halves cost no memory traffic and no branch.
  quat qa[SCW];
#pragma unroll
  for (int cc = 0; cc < SCW; ++cc) {
    const int cgc = min(C0 + SCW * w + cc, n - 1);
    qa[cc] = qload(M + max(rac, cgc) + (size_t)cgc * ld);
  }
  const quat xa = qload(M + rac + (size_t)c * ld);
  // x at this wave's 4 columns (the same address across the wave: one fetch each)
  quat xc[SCW];
#pragma unroll
  for (int cc = 0; cc < SCW; ++cc) xc[cc] = qload(M + min(C0 + SCW * w + cc, n - 1) + (size_t)c * ld);
  const quat ua = (ra < n) ? xa : qzero();
  double qfl = 0.0;   // this thread's share of x^* M x over the tile
  // BF (the wide tiles): branch-free over the tile -- selects instead of
  // divergent ifs, and an inactive tile computes zeros from its clamped loads
  // and stores nothing -- so the body is one scheduling region: the
  // wave-uniform loads of x at the tile's columns issue alongside the tile
  // loads instead of after their wait, and the kernel needs 64 VGPRs instead
  // of 84 (four 512-thread blocks per CU instead of two).  2n = 8192 solve
  // -1.5 ms.  The narrow tiles keep the branches: branch-free they made the
  // 512-problem batch 1.3% slower (also when held to five blocks per CU, their
  // residency with branches), for 0.7% on single solves of 2n <= 1024.
  constexpr bool BF = SW == 32;
  if (BF || active) {
    const double nua = BF ? qnorm2(ua) : 0.0;
    quat acc = qzero();
    double col[4 * SCW];
#pragma unroll
    for (int cc = 0; cc < SCW; ++cc) {
      const int cl = SCW * w + cc;
      const int cg = C0 + cl;
      const bool in = cg < n && ra < n && ra >= cg;
      const bool dg = ra == cg;   // the diagonal: real, enters x^* M x once (below twice)
      quat a, cp = qzero();
      if constexpr (BF) {
        a = quat{in ? qa[cc].ar : 0.0, (in && !dg) ? qa[cc].ai : 0.0, (in && !dg) ? qa[cc].br : 0.0,
                 (in && !dg) ? qa[cc].bi : 0.0};
        qfl -= dg ? a.ar * nua : 0.0;
        qfma(acc, a, xc[cc]);   // a = 0 where cg >= n
        qfma_cl(cp, a, ua);
        const bool below = ra > cg;
        cp = quat{below ? cp.ar : 0.0, below ? cp.ai : 0.0, below ? cp.br : 0.0, below ? cp.bi : 0.0};
      } else {
        a = (cg < n && ra < n && ra >= cg) ? qa[cc] : qzero();
        if (ra == cg) {
          a = quat{a.ar, 0.0, 0.0, 0.0};
          qfl -= a.ar * qnorm2(ua);
        }
        qfma(acc, a, cg < n ? xc[cc] : qzero());
        if (ra > cg) qfma_cl(cp, a, ua);
      }
      col[4 * cc + 0] = cp.ar; col[4 * cc + 1] = cp.ai; col[4 * cc + 2] = cp.br; col[4 * cc + 3] = cp.bi;
    }
    red[w][0][lane] = acc.ar;
    red[w][1][lane] = acc.ai;
    red[w][2][lane] = acc.br;
    red[w][3][lane] = acc.bi;
    // x^* M x over the tile's lower part, doubled off the diagonal: 2 Re(conj(x_i) acc_i)
    qfl += 2.0 * (ua.ar * acc.ar + ua.ai * acc.ai + ua.br * acc.br + ua.bi * acc.bi);
    // lane l ends with value vi = l / LPV (component vi & 3 of column vi >> 2), summed over the wave
    constexpr int LPV = 64 / (4 * SCW);
    const double cs = butterfly_sum<4 * SCW>(col);
    if (active && lane % LPV == 0) {
      const int vi = lane / LPV;
      const int cg = C0 + SCW * w + (vi >> 2);
      if (cg < n) reinterpret_cast<double*>(Pcol + (size_t)g * ldp + cg)[vi & 3] = cs;
    }
  }
  qfl = wave_sum_d(qfl);
  if (lane == 0) sqf[w] = qfl;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < SYT / 64; ++k) t += sqf[k];
    QF[b] = t;
  }
  if (active && tid < SG) {
    const int i = R0 + tid;
    if (i < n && i >= C0) {
      quat sum = qzero();
#pragma unroll
      for (int k = 0; k < NWV; ++k) sum = qadd(sum, quat{red[k][0][tid], red[k][1][tid], red[k][2][tid], red[k][3][tid]});
      qstore(Prow + (size_t)J * ldp + i, sum);
    }
  }
}

// Column c's w_j and, inside a panel, column c+1's pending update, in one
// pass over rows s0 = c+1 .. n-1 (16 threads per row, KPT panel columns each).
//   p   = y - V (W^* u) - W (V^* u),  y = M u = gamma (M x) + M[:, s0] (u1 - gamma x1)
//         with (M x) from the matvec partials;
//   K   = u^* p / 2 = (u^* y - 2 Re(a^* b)) / 2,  a = W^* u, b = V^* u (TS), and
//         u^* y = gamma^2 x^* M x + 2 gamma Re(conj(delta) (M x)_{s0}) + M_{s0 s0} |delta|^2,
//         delta = u1 - gamma x1, x^* M x from the matvec's per-tile partials (QF) --
//         so every block has K without a grid-wide reduction of p;
//   w_j = p - K u (final), stored in W[:, j]; u stored into M[s0:, c].
// Unless column c closes its panel (j = nb - 1) or the reduction (c = n - 2),
// the block then applies column c+1's pending update: subtract the
// [V W] [W V]^*-terms of the panel's columns 0..j (row c+1's coefficients,
// including the new w_j(c+1) = p(s0) - K u1, are formed redundantly in every
// block), store the column, and leave its tail-norm partials in part and the
// per-16-row partials of W^* x, V^* x in TSP.
// PY, PQ: partial loads per thread per gather batch (quaternion row partials,
// x^* M x partials); WPE: the waves per SIMD the register budget aims at.
template <int KPT, int PY, int PQ, int WPE>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(WPE))) k_wcol(quat* __restrict__ M, int ld, int n, int p, int c, int j,
                                             quat* __restrict__ W, int ldw, const quat* __restrict__ Prow,
                                             const quat* __restrict__ Pcol, int ldp, const quat* __restrict__ TS,
                                             int nb, const double* __restrict__ scal, const double* __restrict__ QF,
                                             double* __restrict__ part, double* __restrict__ TSP, int ldtsp, int tsm,
                                             ZB zb) {
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    if (c + 1 + (int)blockIdx.x * RQ >= n) return;
    M = zoff(M, zb); W = zoff(W, zb); Prow = zoff(Prow, zb); Pcol = zoff(Pcol, zb); TS = zoff(TS, zb);
    scal = zoff(scal, zb); QF = zoff(QF, zb); part = zoff(part, zb); TSP = zoff(TSP, zb);
  }
  __shared__ quat a[NBMAX], b[NBMAX], cw[NBMAX], cv[NBMAX];
  __shared__ double red[16][8][RQ];   // k-group, (p | column update) component, row
  __shared__ double red2[8][RQ];      // the same summed over the k-groups
  __shared__ double shr[RB / 64][10];
  __shared__ quat xs[RQ], us[RQ], ws[RQ];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int s0 = c + 1;
  const bool more = (j + 1 < nb) && (c + 1 <= n - 2);   // uniform
  const int r = tid & (RQ - 1), q = tid >> 4;
  const int i = s0 + blockIdx.x * RQ + r;
  const int ic = min(i, n - 1);
  const int jl = j > 0 ? j - 1 : 0;
  // ---- loads, all issued up front
  quat vv[KPT], ww[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int l = min(q + 16 * k, jl);
    vv[k] = qload(M + ic + (size_t)(p + l) * ld);
    ww[k] = qload(W + ic + (size_t)l * ldw);
  }
  const quat xi = qload(M + ic + (size_t)c * ld);
  const quat ms0raw = qload(M + ic + (size_t)s0 * ld);   // stale column s0 = c+1 (row i)
  // The column scalars (gamma, u1, x1, M[s0, s0]) in one vector load over
  // lanes 0..9, taken apart after the partial gathers.  As scalar loads they
  // would share the scalar-load counter with the kernel arguments the gather
  // loop needs, and the loop's loads could not issue before they arrived.
  const double sv = scal[min(lane, 9)];
  quat al = qzero(), bl = qzero(), cwl = qzero(), cvl = qzero();
  if (tid < j) {
    al = qload(TS + tid);
    bl = qload(TS + nb + tid);
    cwl = qload(W + s0 + (size_t)tid * ldw);
    cvl = qload(M + s0 + (size_t)(p + tid) * ld);
  }
  const int G = (n - s0 + SG - 1) / SG;
  // The matvec partials this thread sums, gathered in fixed-size batches of
  // predicated loads so that one wait covers them all (one batch up to
  // trailing sizes of ~4k rows with the wide variant, ~1k with the light one):
  //   y(i):  slots k = q + 16 t of row i's list, Prow[J][i] for J <= li / SW
  //          (SW the strip width of this trailing size)
  //          (k < nJ), then Pcol[g][i] for g >= li / SG;
  //   x^* M x: QF[tid + RB t], t over the G (G + 1) tile partials;
  //   (M x)_{s0}: Prow[0][s0] and Pcol[g][s0] as slots tid + RB t of G + 1.
  const int li = ic - s0;
  const int spg = matvec_spg(n - s0, tsm);
  const int nJ = (spg == 4 ? li >> 4 : li >> 5) + 1, g0 = li / SG;
  const int cy = cdiv_d(nJ + G - g0 - q, 16);
  const int nq = ntiles(G, spg);
  const int cq = cdiv_d(nq - tid, RB);
  const int cw0 = cdiv_d(G + 1 - tid, RB);
  quat yi = qzero(), w0 = qzero();
  double qf = 0.0;
  for (int bt = 0; bt * PY < cy || bt * PQ < cq || bt < cw0; ++bt) {
    quat py[PY];
    double pq[PQ];
#pragma unroll
    for (int u = 0; u < PY; ++u) {
      const int t = bt * PY + u;
      const int k = q + 16 * t;
      const quat* src = k < nJ ? Prow + (size_t)k * ldp + ic : Pcol + (size_t)(g0 + k - nJ) * ldp + ic;
      py[u] = t < cy ? qload(src) : qzero();
    }
#pragma unroll
    for (int u = 0; u < PQ; ++u) {
      const int t = bt * PQ + u;
      pq[u] = t < cq ? QF[tid + RB * t] : 0.0;
    }
    const int kw = tid + RB * bt;
    const quat pw = bt < cw0 ? qload(kw == 0 ? Prow + s0 : Pcol + (size_t)(kw - 1) * ldp + s0) : qzero();
#pragma unroll
    for (int u = 0; u < PY; ++u) yi = qadd(yi, py[u]);
#pragma unroll
    for (int u = 0; u < PQ; ++u) qf += pq[u];
    w0 = qadd(w0, pw);
  }
  const double gam = lane_d(sv, 0);
  const quat u1 = quat{lane_d(sv, 1), lane_d(sv, 2), lane_d(sv, 3), lane_d(sv, 4)};
  const quat x1 = quat{lane_d(sv, 5), lane_d(sv, 6), lane_d(sv, 7), lane_d(sv, 8)};
  const double m00 = lane_d(sv, 9);   // M[s0, s0] itself is rewritten by this grid (row s0 of column s0)
  // ---- block sums: x^* M x, Re(a^* b), (M x)_{s0}, V(s0,:) a + W(s0,:) b
  {
    quat rd = qmul(cvl, al);
    qfma(rd, cwl, bl);
    double v[10] = {qf, al.ar * bl.ar + al.ai * bl.ai + al.br * bl.br + al.bi * bl.bi,
                    w0.ar, w0.ai, w0.br, w0.bi, rd.ar, rd.ai, rd.br, rd.bi};
#pragma unroll
    for (int t = 0; t < 10; ++t) v[t] = wave_sum_d(v[t]);
    if (lane == 0)
#pragma unroll
      for (int t = 0; t < 10; ++t) shr[wv][t] = v[t];
  }
  if (tid < j) {
    a[tid] = al;
    b[tid] = bl;
    cw[tid] = qconj(cwl);
    cv[tid] = qconj(cvl);
  }
  __syncthreads();
  double S[10];
#pragma unroll
  for (int t = 0; t < 10; ++t) {
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < RB / 64; ++k) x += shr[k][t];
    S[t] = x;
  }
  const quat delta = qsub(u1, qscale(x1, gam));
  const quat wrow = quat{S[2], S[3], S[4], S[5]};   // (M x)_{s0}
  const double uy = gam * gam * S[0] +
                    2.0 * gam * (delta.ar * wrow.ar + delta.ai * wrow.ai + delta.br * wrow.br + delta.bi * wrow.bi) +
                    m00 * qnorm2(delta);
  const double K = 0.5 * (uy - 2.0 * S[1]);
  // p and w_j at row s0 (= c+1), needed by every block for column c+1's update
  const quat ps0 = qsub(qadd(qscale(wrow, gam), qscale(delta, m00)), quat{S[6], S[7], S[8], S[9]});
  const quat wc1 = qsub(ps0, qscale(u1, K));
  // ---- this thread's shares: y(i) minus corrections, and column c+1's update over l < j
  // branch-free (reads at a clamped index, results selected) so that every
  // LDS read is issued before the first is used
  quat accp = qscale(yi, gam), accu = qzero();
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int l = q + 16 * k;
    const bool on = l < j;
    const int lc = on ? l : 0;
    const quat ak = a[lc], bk = b[lc], cwk = cw[lc], cvk = cv[lc];
    quat tp = qmul(vv[k], ak);
    qfma(tp, ww[k], bk);
    quat tu = qmul(vv[k], cwk);
    qfma(tu, ww[k], cvk);
    if (on) {
      accp = qsub(accp, tp);
      accu = qadd(accu, tu);
    }
  }
  red[q][0][r] = accp.ar; red[q][1][r] = accp.ai; red[q][2][r] = accp.br; red[q][3][r] = accp.bi;
  red[q][4][r] = accu.ar; red[q][5][r] = accu.ai; red[q][6][r] = accu.br; red[q][7][r] = accu.bi;
  __syncthreads();
  // sums over the 16 k-groups, one (component, row) per thread
  if (tid < 8 * RQ) {
    const int rr = tid & (RQ - 1), comp = tid >> 4;
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += red[k][comp][rr];
    red2[comp][rr] = t;
  }
  __syncthreads();
  double ss = 0.0;
  if (q == 0) {
    quat xn = qzero(), ui = qzero(), wi = qzero();
    if (i < n) {
      const quat ms0 = (i == s0) ? quat{ms0raw.ar, 0.0, 0.0, 0.0} : ms0raw;   // Hermitian diagonal: real
      const quat pv = qadd(qmul(ms0, delta), quat{red2[0][r], red2[1][r], red2[2][r], red2[3][r]});
      const quat su = quat{red2[4][r], red2[5][r], red2[6][r], red2[7][r]};
      ui = (i == s0) ? u1 : qscale(xi, gam);
      wi = qsub(pv, qscale(ui, K));
      qstore(W + i + (size_t)j * ldw, wi);
      qstore(M + i + (size_t)c * ld, ui);
      if (more) {
        // column c+1, row i: stale value minus the panel's columns 0..j
        xn = qsub(ms0raw, su);
        xn = qsub(xn, qmul(ui, qconj(wc1)));
        xn = qsub(xn, qmul(wi, qconj(u1)));
        qstore(M + i + (size_t)s0 * ld, xn);
        if (i >= s0 + 2) ss = qnorm2(xn);
        else xn = qzero();
      }
    }
    xs[r] = xn;
    us[r] = ui;
    ws[r] = wi;
  }
  if (!more) return;   // uniform
  __syncthreads();   // publishes xs, us, ws
  // the tail-norm terms live in lanes 0..15 of wave 0 (q == 0), zero elsewhere
  if (wv == 0) {
    ss = wave_sum_d(ss);
    if (lane == 0) part[blockIdx.x] = ss;
  }
  // tall-skinny partials of column c+1 over the panel columns 0..j (j + 1 of them):
  // this thread's columns l = q + 16 k against x(row r), summed over the 16 rows
  {
    constexpr int NV = (8 * KPT < 16) ? 16 : 8 * KPT;
    double v[NV];
#pragma unroll
    for (int t = 0; t < NV; ++t) v[t] = 0.0;
    const quat x = xs[r];
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int l = q + 16 * k;
      if (l <= j) {
        const quat wl = (l == j) ? ws[r] : ww[k];
        const quat vl = (l == j) ? us[r] : vv[k];
        quat ta = qzero(), tb = qzero();
        qfma_cl(ta, wl, x);
        qfma_cl(tb, vl, x);
        v[8 * k + 0] = ta.ar; v[8 * k + 1] = ta.ai; v[8 * k + 2] = ta.br; v[8 * k + 3] = ta.bi;
        v[8 * k + 4] = tb.ar; v[8 * k + 5] = tb.ai; v[8 * k + 6] = tb.br; v[8 * k + 7] = tb.bi;
      }
    }
    row16_halving<NV>(v);
#pragma unroll
    for (int t = 0; t < NV / 16; ++t) {
      const int idx = r * (NV / 16) + t;
      const int k = idx >> 3, l = q + 16 * k;
      if (k < KPT && l <= j) TSP[(size_t)blockIdx.x * ldtsp + 8 * l + (idx & 7)] = v[t];
    }
  }
}

// ---------------------------------------------------------------------------
// LDS-resident tail.  Once the trailing matrix has at most TAILN rows, one
// workgroup per problem finishes the reduction with the whole Hermitian
// trailing block in LDS (TAILN x TAILN quaternions, 128 KB, both triangles so
// that every access is a row-contiguous, conflict-free read): per column the
// reflector, y = A u, K = u^* y / 2, w = y - K u and the rank-2 update
// A -= u w^* + w u^* at once (unblocked: no panel, no W), separated by
// barriers instead of kernel launches.  Lane = row, the 4 waves split the
// columns.  It starts at a panel boundary p, after the trailing update of the
// panel before it, and writes d and e of columns p .. n-2, d[n-1], and the
// reflectors u into M below the subdiagonal (the blocked columns'
// conventions).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_tail(quat* __restrict__ M, int ld, int n, int p, const int* __restrict__ pstart,
                                              double* __restrict__ d, quat* __restrict__ e, ZB zb) {
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    p = pstart[blockIdx.z];
    M = zoff(M, zb); d = zoff(d, zb); e = zoff(e, zb);
  }
  const int ms = n - p;
  if (ms < 1 || ms > TAILN) return;   // uniform: a problem without a tail
  __shared__ quat A[TAILN * TAILN];     // A[j * TAILN + i] = element (i, j)
  __shared__ quat us[TAILN], ys[TAILN];
  __shared__ quat part[RB / 64][TAILN];
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;   // row, column group
  for (int j = ty; j < ms; j += RB / 64) {
    if (tx < ms) {
      quat a = tx >= j ? qload(M + (p + tx) + (size_t)(p + j) * ld) : qconj(qload(M + (p + j) + (size_t)(p + tx) * ld));
      if (tx == j) a = quat{a.ar, 0.0, 0.0, 0.0};
      A[j * TAILN + tx] = a;
    }
  }
  __syncthreads();
  // Every row of the block lives in wave 0 (lane = row), so the two sums over
  // rows per column (the tail norm, u^* y) are wave sums there, and the tail
  // norm of the next column is formed by wave 0 right after it updates that
  // column: four barriers per column.
  __shared__ double tnorm;
  if (ty == 0) {
    const double v = (tx > 1 && tx < ms) ? qnorm2(A[tx]) : 0.0;
    const double t = wave_sum_d(v);
    if (tx == 0) tnorm = t;
  }
  __syncthreads();
  for (int cl = 0; cl + 1 < ms; ++cl) {
    const int r0 = cl + 1;                  // first row of the trailing block
    const bool row = tx >= r0 && tx < ms;   // this lane's row is in the trailing block
    const double tail = tnorm;
    quat alpha;
    const Refl R = reflector(A[cl * TAILN + r0], tail, &alpha);
    if (tid == 0) {
      d[p + cl] = A[cl * TAILN + cl].ar;
      qstore(e + p + cl, alpha);
    }
    if (ty == 0 && row) {
      const quat ui = tx == r0 ? R.u1 : qscale(A[cl * TAILN + tx], R.g);
      us[tx] = ui;
      qstore(M + (p + tx) + (size_t)(p + cl) * ld, ui);
    }
    __syncthreads();
    quat acc = qzero();   // this lane's row of A22 u over this wave's columns
    if (row)
      for (int jj = r0 + ty; jj < ms; jj += RB / 64) qfma(acc, A[jj * TAILN + tx], us[jj]);
    part[ty][tx] = acc;
    __syncthreads();
    if (ty == 0) {
      quat y = qzero();
      double kv = 0.0;
      if (row) {
        y = part[0][tx];
#pragma unroll
        for (int k = 1; k < RB / 64; ++k) y = qadd(y, part[k][tx]);
        const quat u = us[tx];
        kv = u.ar * y.ar + u.ai * y.ai + u.br * y.br + u.bi * y.bi;   // Re(conj(u_i) y_i)
      }
      const double K = 0.5 * wave_sum_d(kv);
      if (row) ys[tx] = qsub(y, qscale(us[tx], K));   // w
    }
    __syncthreads();
    if (row) {
      const quat ui = us[tx], wi = ys[tx];
      for (int jj = r0 + ty; jj < ms; jj += RB / 64) {
        quat a = A[jj * TAILN + tx];
        a = qsub(a, qmul(ui, qconj(ys[jj])));
        a = qsub(a, qmul(wi, qconj(us[jj])));
        if (jj == tx) a = quat{a.ar, 0.0, 0.0, 0.0};   // Hermitian diagonal: real
        A[jj * TAILN + tx] = a;
      }
    }
    // tail norm of the next column, r0, whose rows this wave (jj = r0) just updated
    if (ty == 0) {
      const double v = (tx > r0 + 1 && tx < ms) ? qnorm2(A[r0 * TAILN + tx]) : 0.0;
      const double t = wave_sum_d(v);
      if (tx == 0) tnorm = t;
    }
    __syncthreads();
  }
  if (tid == 0) d[n - 1] = A[(ms - 1) * TAILN + (ms - 1)].ar;
}

int tail_start(int n, int nb) {
  if (n <= TAILN) return 0;
  const int p = (n - TAILN + nb - 1) / nb * nb;
  return p <= n - 1 ? p : n;
}

__global__ void k_last_diag(const quat* __restrict__ M, int ld, int n, double* __restrict__ d, ZB zb) {
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    M = zoff(M, zb); d = zoff(d, zb);
  }
  if (threadIdx.x == 0) d[n - 1] = M[(n - 1) + (size_t)(n - 1) * ld].ar;
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// (sized for any switch up to TSM_MAX)
int matvec_row_slots(int nmax) { return std::max(cdiv(nmax, 32), cdiv(std::min(nmax, TSM_MAX), 16)); }
int matvec_col_slots(int nmax) { return (nmax + SG - 1) / SG; }
int matvec_tile_slots(int nmax) {
  return std::max(ntiles(cdiv(nmax, SG) + 1, 2), ntiles(cdiv(std::min(nmax, TSM_MAX), SG) + 1, 4));
}

// panel columns per thread (16 threads per row) as a compile-time bound
static int kpt_for(int nb) {
  const int k = cdiv(nb, 16);
  return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8;
}

// k_wcol variants: light gathers (more blocks per CU: batches of small
// problems) up to trailing size 1536, wide ones (one gather batch) above.
template <int KPT, class... Args>
static void launch_wcol_k(int sz, dim3 grid, hipStream_t s, Args... args) {
  // register budgets that compile without spills: 3 waves per SIMD at KPT 1,
  // 2 at KPT 2 and 4, 1 at KPT 8
  constexpr int WPE = KPT == 1 ? 3 : (KPT == 8 ? 1 : 2);
  if (sz <= 1536) hipLaunchKernelGGL((k_wcol<KPT, 4, 2, WPE>), grid, dim3(RB), 0, s, args...);
  else hipLaunchKernelGGL((k_wcol<KPT, 10, 18, 1>), grid, dim3(RB), 0, s, args...);
}
template <class... Args>
static void launch_wcol(int kpt, int sz, dim3 grid, hipStream_t s, Args... args) {
  switch (kpt) {
    case 1: launch_wcol_k<1>(sz, grid, s, args...); break;
    case 2: launch_wcol_k<2>(sz, grid, s, args...); break;
    case 4: launch_wcol_k<4>(sz, grid, s, args...); break;
    default: launch_wcol_k<8>(sz, grid, s, args...); break;
  }
}

// Per panel: per column k_symv and k_wcol (which also updates the next
// column of the panel and leaves its tail-norm partials); then the trailing
// update.
int reduce(quat* M, int n, int ld, int nb, double* d, quat* e, const ReduceWork& w, hipStream_t s) {
  if (n <= 0) return ZQI_OK;
  if (nb < 1 || nb > NBMAX || w.tsm < 0 || w.tsm > TSM_MAX) return ZQI_ERR_ARG;
  if (n == 1) {
    hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, s, M, ld, n, d, ZB{nullptr, 0});
    return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
  }
  const int ldw = n;
  const int kpt = kpt_for(nb);
  const int ts = tail_start(n, nb);   // n: no tail
  for (int p = 0; p < n - 1 && p < ts; p += nb) {
    const int wdt = (nb < n - 1 - p) ? nb : (n - 1 - p);
    for (int j = 0; j < wdt; ++j) {
      const int c = p + j;
      const int nblk = cdiv(n - c, RQ);            // row blocks from c
      const int sz = n - c - 1;                    // trailing size
      const int G = cdiv(sz, SG);
      const int spg = matvec_spg(sz, w.tsm);
      const int nmain = ntiles(G, spg) + (j > 0 ? (8 * j + 63) / 64 : 1);
      {
        // algorithmic bytes: one read of the trailing lower triangle
        ProfScope ps(PROF_SYMV, s, 32.0 * 0.5 * (double)sz * (sz + 1));
        if (spg == 4)
          launch_k(k_symv<256, 16>, dim3(nmain), dim3(256), 0, s, M, ld, n, p, c, j, w.W, ldw, w.part, nblk,
                   w.scal, d, e, w.Prow, w.Pcol, n, w.TS, nb, w.TSP, w.ldtsp, w.QF, ZB{nullptr, 0});
        else
          launch_k(k_symv<512, 32>, dim3(nmain), dim3(512), 0, s, M, ld, n, p, c, j, w.W, ldw, w.part, nblk,
                   w.scal, d, e, w.Prow, w.Pcol, n, w.TS, nb, w.TSP, w.ldtsp, w.QF, ZB{nullptr, 0});
      }
      launch_wcol(kpt, sz, dim3(cdiv(sz, RQ)), s, M, ld, n, p, c, j, w.W, ldw, w.Prow, w.Pcol, n, w.TS, nb, w.scal,
                  w.QF, w.part, w.TSP, w.ldtsp, w.tsm, ZB{nullptr, 0});
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
  if (ts < n) hipLaunchKernelGGL(k_tail, dim3(1), dim3(RB), 0, s, M, ld, n, ts, nullptr, d, e, ZB{nullptr, 0});
  else hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, s, M, ld, n, d, ZB{nullptr, 0});
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

// problem z's workspace: the batch base pointer advanced by z slabs
template <class T>
static T* zshift(T* p, size_t bytes) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<B*>(p) + bytes);
}

size_t reduce_batched_tasks(int nprob, const int* nh, int nb) {
  if (nprob <= 0 || nh[0] < 2) return 1;
  return (size_t)nprob * cdiv(nh[0] - 1, nb) + 1;
}

int reduce_batched(int nprob, const int* nh, ZB zb, quat* M, int ld, int nb, double* d, quat* e, const ReduceWork& w,
                   QGemmTask* tasks, int* tstart, Staging& st, hipStream_t s) {
  if (nprob <= 0) return ZQI_OK;
  if (nb < 1 || nb > NBMAX || w.tsm < 0 || w.tsm > TSM_MAX) return ZQI_ERR_ARG;
  const int L = nh[0];
  for (int z = 1; z < nprob; ++z)
    if (nh[z] > nh[z - 1] || nh[z] < 1) return ZQI_ERR_ARG;
  const int ldw = ld;
  // problem z runs the blocked columns c < lim[z] in lockstep and its tail
  // (k_tail) from ts[z]; both are non-increasing in z (sizes descending)
  std::vector<int> ts(nprob), lim(nprob);
  for (int z = 0; z < nprob; ++z) {
    ts[z] = tail_start(nh[z], nb);
    lim[z] = std::min(ts[z], nh[z] - 1);
  }
  const int npan = L > 1 ? cdiv(L - 1, nb) : 0;
  // trailing-update tasks of every panel, problem by problem: the task of
  // problem z in panel P is what reduce() launches for z after that panel
  std::vector<QGemmTask> tt;
  std::vector<int> toff(npan + 1, 0), tmaxm(npan, 0);
  std::vector<double> tflops(npan, 0.0);   // algorithmic flops of each panel's trailing updates
  for (int P = 0; P < npan; ++P) {
    const int p = P * nb;
    toff[P] = (int)tt.size();
    for (int z = 0; z < nprob && nh[z] - 1 > p; ++z) {
      if (p >= ts[z]) continue;
      const int n = nh[z];
      const int wdt = std::min(nb, n - 1 - p);
      const int r0 = p + wdt, mt = n - r0;
      if (mt <= 0) continue;
      quat* Mz = reinterpret_cast<quat*>(reinterpret_cast<char*>(M) + (size_t)z * zb.zs);
      quat* Wz = reinterpret_cast<quat*>(reinterpret_cast<char*>(w.W) + (size_t)z * zb.zs);
      QOperand A{Mz + r0 + (size_t)p * ld, Wz + r0, ld, ldw, wdt, false, INT_MIN / 2};
      QOperand B{Wz + r0, Mz + r0 + (size_t)p * ld, ldw, ld, wdt, true, INT_MIN / 2};
      tt.push_back(QGemmTask{mt, mt, 2 * wdt, 0, 2 * wdt, ld, 1, -1.0, 1.0, A, B, Mz + r0 + (size_t)r0 * ld});
      tmaxm[P] = std::max(tmaxm[P], mt);
      tflops[P] += 32.0 * 0.5 * (double)mt * (mt + 1) * 2.0 * wdt;
    }
  }
  toff[npan] = (int)tt.size();
  if (tt.size() > reduce_batched_tasks(nprob, nh, nb)) return ZQI_ERR_ARG;
  int rc = stage_upload(st, tt.data(), tt.size() * sizeof(QGemmTask), tasks, s);
  if (rc) return rc;
  if ((rc = stage_upload(st, ts.data(), nprob * sizeof(int), tstart, s))) return rc;
  const int kpt = kpt_for(nb);
  const int cend = lim[0];
  int act = nprob;   // problems with a blocked column c: c < lim[z] (a prefix)
  for (int c = 0; c < cend; ++c) {
    while (act > 0 && lim[act - 1] <= c) --act;
    const int p = (c / nb) * nb, j = c - p;
    const int nrow = L - c;
    const int nblk = cdiv(nrow, RQ);
    const int sz = L - c - 1;
    const int NF = j > 0 ? (8 * j + 63) / 64 : 1;
    // each problem's matvec takes the tile shape of its own trailing size:
    // problems [0, k1) wide strips, [k1, act) narrow ones (sizes descending)
    int k1 = 0;
    while (k1 < act && matvec_spg(nh[k1] - c - 1, w.tsm) == 2) ++k1;
    // algorithmic bytes of a launch: one read of each member's trailing lower triangle
    auto tri_bytes = [&](int z0, int z1) {
      double b = 0.0;
      for (int z = z0; z < z1; ++z) b += 32.0 * 0.5 * (double)(nh[z] - c - 1) * (nh[z] - c);
      return b;
    };
    if (k1 > 0) {
      ProfScope ps(PROF_SYMV, s, tri_bytes(0, k1));
      launch_k(k_symv<512, 32>, dim3(ntiles(cdiv(sz, SG), 2) + NF, 1, k1), dim3(512), 0, s, M, ld, L, p, c, j, w.W,
               ldw, w.part, nblk, w.scal, d, e, w.Prow, w.Pcol, ld, w.TS, nb, w.TSP, w.ldtsp, w.QF, zb);
    }
    if (k1 < act) {
      const size_t sh = (size_t)k1 * zb.zs;
      const int szn = nh[k1] - c - 1;
      ProfScope ps(PROF_SYMV, s, tri_bytes(k1, act));
      launch_k(k_symv<256, 16>, dim3(ntiles(cdiv(szn, SG), 4) + NF, 1, act - k1), dim3(256), 0, s, zshift(M, sh), ld,
               L, p, c, j, zshift(w.W, sh), ldw, zshift(w.part, sh), nblk, zshift(w.scal, sh), zshift(d, sh),
               zshift(e, sh), zshift(w.Prow, sh), zshift(w.Pcol, sh), ld, zshift(w.TS, sh), nb, zshift(w.TSP, sh),
               w.ldtsp, zshift(w.QF, sh), ZB{zb.nvec + k1, zb.zs});
    }
    launch_wcol(kpt, sz, dim3(cdiv(sz, RQ), 1, act), s, M, ld, L, p, c, j, w.W, ldw, w.Prow, w.Pcol, ld, w.TS, nb,
                w.scal, w.QF, w.part, w.TSP, w.ldtsp, w.tsm, zb);
    if (j == nb - 1 || c == cend - 1) {
      const int P = c / nb;
      const int cnt = toff[P + 1] - toff[P];
      ProfScope ps(PROF_TRAIL, s, tflops[P]);
      if (cnt > 0 && (rc = qgemm_tasks(tasks + toff[P], cnt, tmaxm[P], tmaxm[P], false, true, true, s))) return rc;
    }
  }
  // problems without a tail take d[n-1] here; the tails overwrite theirs
  hipLaunchKernelGGL(k_last_diag, dim3(1, 1, nprob), dim3(64), 0, s, M, ld, L, d, zb);
  hipLaunchKernelGGL(k_tail, dim3(1, 1, nprob), dim3(RB), 0, s, M, ld, L, 0, tstart, d, e, zb);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
