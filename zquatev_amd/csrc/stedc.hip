// Real symmetric tridiagonal eigensolver on the device: Cuppen's divide and
// conquer with Gu-Eisenstat eigenvectors, and Sturm-count multisection for
// the eigenvalues-only path.
//
// Replaces the tridiagonal step of the reference, zhbev_("V","L",n,1,...)
// (zquatev.cc:74-84; MKL: zhbtrd phase scaling + zsteqr QL/QR).  The phase
// scaling is done on the quaternion subdiagonal (backtr.hip); this file sees
// the real tridiagonal (d, |e|) and returns the eigenvalues ascending, as
// zhbev does, with an orthonormal real eigenvector matrix.
//
// Tree: halve until a block has at most LEAF rows; each split point s couples
// rows s-1 and s with beta = e[s-1] >= 0, and T = diag(T1', T2') + beta v v^T,
// v = e_{s-1} + e_s, where T1', T2' have beta subtracted from the touching
// diagonal entries.  Leaves run implicit QL in one wavefront each; merges run
// level by level (all merges of a level in one launch):
//   prep   : stable merge of the two sorted spectra, z = [last row of Q1;
//            first row of Q2]/sqrt(2), rho = 2 beta, deflation scan (small
//            |z| or close d handled by a Givens rotation), one workgroup,
//            the scan on LDS copies up to PREP_LDS rows.
//   rot    : apply the deflation rotations to Q's columns (row-parallel).
//   gather : the kept and deflated columns into Qtmp (rows x column chunks).
//   secular: one root of 1 + rho sum z_j^2/(d_j - lambda) per wavefront,
//            safeguarded Newton on the offset from the nearer pole.
//   zhat   : Gu-Eisenstat z from the computed roots, for orthogonality.
//   buildU : U[:, i] = zhat / (d - lambda_i), normalised.
//   order  : merge the roots with the deflated values (both sorted).
//   dgemm  : Z[:, pos] = Qtmp[:, :K] U on FP64 MFMA; deflated columns copied.
#include "zq_internal.h"
#include <vector>
#include <algorithm>
#include <cfloat>

namespace zq {

constexpr int LEAF = 16;
constexpr double DEPS = DBL_EPSILON * 0.5;   // unit roundoff

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// --------------------------------------------------------------------------
// scale: s = max(max|d|, max|e|) per problem; d, e scaled by 1/s into lam, es.
// probs holds (offset, size) per problem, one workgroup each.  A tridiagonal
// holding a NaN or an infinity (an input entry of that kind, or an overflow
// in the reduction) is replaced by zeros and marks the problem's status word:
// the merges rank values by comparisons, which a NaN would make inconsistent.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_dc_scale(const int* __restrict__ probs, const double* __restrict__ d,
                                                   const double* __restrict__ e, double* __restrict__ lam,
                                                   double* __restrict__ es, double* __restrict__ scal,
                                                   int* __restrict__ info) {
  const int o = probs[2 * blockIdx.x], n = probs[2 * blockIdx.x + 1];
  d += o; e += o; lam += o; es += o;
  __shared__ double sh[16];
  double m = 0.0;
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    m = fmax(m, fabs(d[i]));
    bad |= !isfinite(d[i]);
  }
  for (int i = threadIdx.x; i < n - 1; i += 1024) {
    m = fmax(m, fabs(e[i]));
    bad |= !isfinite(e[i]);
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) m = fmax(m, __shfl_xor(m, k, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  bad = __syncthreads_or(bad);
  double s = 0.0;
  for (int k = 0; k < 16; ++k) s = fmax(s, sh[k]);
  if (!(s > 0.0) || bad) s = 1.0;
  const double inv = 1.0 / s;
  for (int i = threadIdx.x; i < n; i += 1024) lam[i] = bad ? 0.0 : d[i] * inv;
  for (int i = threadIdx.x; i < n - 1; i += 1024) es[i] = bad ? 0.0 : fabs(e[i]) * inv;
  if (threadIdx.x == 0) {
    scal[blockIdx.x] = s;
    if (bad) atomicOr(info + blockIdx.x, ZQI_INFO_NONFINITE);
  }
}

// subtract the coupling beta = es[s-1] from the diagonal entries beside each split s
__global__ void k_dc_tear(int nsplit, const int* __restrict__ splits, const double* __restrict__ es,
                          double* __restrict__ lam) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsplit) return;
  const int s = splits[t];
  const double b = es[s - 1];
  lam[s - 1] -= b;
  lam[s] -= b;
}

// --------------------------------------------------------------------------
// leaf: implicit QL with Wilkinson-type shift on a <= LEAF tridiagonal.  The
// scalar recurrence runs identically in every lane; lane k owns row k of the
// leaf's eigenvector block and applies each rotation to it.  A leaf is
// (lo, hi, problem base, problem index).
// --------------------------------------------------------------------------
// Element i (wave-uniform) of a value held one per lane.
__device__ __forceinline__ double lane_get(double v, int i) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), i);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), i);
  return __hiloint2double(hi, lo);
}

// One leaf per wavefront.  The tridiagonal lives in registers, d[k] and e[k]
// in lane k; every lane runs the (wave-uniform) QL recurrence redundantly,
// reading element i with readlane, and the owning lane keeps an update, so
// the sequential chain has no barrier and no LDS access.  Lane k owns row k
// of the leaf's eigenvectors (in LDS), to which it applies each rotation.
__global__ void __launch_bounds__(64) k_dc_leaf(const int* __restrict__ leaves, double* __restrict__ lam,
                                                const double* __restrict__ es, double* __restrict__ Z, int ldz,
                                                int* __restrict__ info, int cap) {
  __shared__ double zz[LEAF][LEAF + 1];
  const int lo = leaves[4 * blockIdx.x], hi = leaves[4 * blockIdx.x + 1];
  const int ob = leaves[4 * blockIdx.x + 2], pr = leaves[4 * blockIdx.x + 3];
  const int m = hi - lo;
  const int k = threadIdx.x;
  double dk = 0.0, ek = 0.0;
  if (k < m) {
    dk = lam[lo + k];
    ek = (k < m - 1) ? es[lo + k] : 0.0;
    for (int i = 0; i < m; ++i) zz[k][i] = (i == k) ? 1.0 : 0.0;
  }
  int fail = 0;
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    int mm;
    while (true) {
      for (mm = l; mm < m - 1; ++mm) {
        const double s = fabs(lane_get(dk, mm)) + fabs(lane_get(dk, mm + 1));
        if (fabs(lane_get(ek, mm)) <= DEPS * s) break;
      }
      if (mm == l) break;
      if (++iter > cap) { fail = 1; break; }
      const double el = lane_get(ek, l);
      double g = (lane_get(dk, l + 1) - lane_get(dk, l)) / (2.0 * el);
      double r = hypot(g, 1.0);
      g = lane_get(dk, mm) - lane_get(dk, l) + el / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      bool early = false;
      for (i = mm - 1; i >= l; --i) {
        const double ei = lane_get(ek, i);
        const double f = s * ei;
        const double b = c * ei;
        // |(f, g)|: the plain square root unless the squares could over- or
        // underflow (the chain of rotations is the leaf's critical path)
        const double rr = f * f + g * g;
        r = (rr > 1e-280 && rr < 1e280) ? sqrt(rr) : hypot(f, g);
        if (k == i + 1) ek = r;
        if (r == 0.0) {
          if (k == i + 1) dk -= p;
          if (k == mm) ek = 0.0;
          early = true;
          break;
        }
        const double ir = 1.0 / r;
        s = f * ir;
        c = g * ir;
        g = lane_get(dk, i + 1) - p;
        r = (lane_get(dk, i) - g) * s + 2.0 * c * b;
        p = s * r;
        if (k == i + 1) dk = g + p;
        g = c * r - b;
        if (k < m) {
          const double zf = zz[k][i + 1];
          const double zi = zz[k][i];
          zz[k][i + 1] = s * zi + c * zf;
          zz[k][i] = c * zi - s * zf;
        }
      }
      if (early) continue;
      if (k == l) { dk -= p; ek = g; }
      if (k == mm) ek = 0.0;
    }
    if (fail) break;
  }
  // selection sort ascending, columns follow
  for (int i = 0; i < m - 1; ++i) {
    int kmin = i;
    double vmin = lane_get(dk, i);
    for (int j = i + 1; j < m; ++j) {
      const double dj = lane_get(dk, j);
      if (dj < vmin) { vmin = dj; kmin = j; }
    }
    if (kmin != i) {
      const double di = lane_get(dk, i);
      if (k == kmin) dk = di;
      if (k == i) dk = vmin;
      if (k < m) {
        const double t = zz[k][i];
        zz[k][i] = zz[k][kmin];
        zz[k][kmin] = t;
      }
    }
  }
  if (k < m) {
    lam[lo + k] = dk;
    for (int i = 0; i < m; ++i) Z[(lo + k) + (size_t)(lo - ob + i) * ldz] = zz[k][i];
  }
  if (fail && k == 0) atomicMax(info + pr, lo - ob + 1);
}

// --------------------------------------------------------------------------
// merge kernels
// --------------------------------------------------------------------------
struct MergeBufs {
  double* dk;    // sorted d (in place modified by deflation)
  double* zk;    // sorted z
  int* src;      // sorted position -> local column
  int* keep;     // kept sorted positions (K)
  int* dfl;      // deflated sorted positions (ns - K), sorted by value
  double* rot;   // rotations: (pj, j, c, s) as doubles
  int* nrot;     // per merge
  int* kcnt;     // per merge
  double* dkk;   // kept d
  double* zkk;   // kept z
  double* dval;  // deflated values
  double* rho;   // per merge
  int* org;      // per root: pole index (into kept list)
  double* tau;   // per root: offset from the pole
  double* zhat;
  int* posK;     // final column for each root
  int* posD;     // final column for each deflated value
  double* Qtmp;
  double* U;
  int lq;        // the merge of rows [lo, hi) keeps its Qtmp and U blocks at lo * lq
  int* info;     // status word per problem (non-convergence of a secular root)
  int cap;       // secular-equation steps per root
  // The merge product Z = Qtmp U skips the zero blocks of Qtmp: a kept
  // column holds the first half's rows only, the second half's only, or (after
  // a deflation rotation across the halves) both.  Qtmp's columns and U's
  // rows are in GEMM order: first-half columns, both-halves, second-half,
  // each group in kept order.
  int* mix;      // per sorted position: the column mixes both halves
  int* gpos;     // per kept index: its GEMM position
  int* ginv;     // per GEMM position: the kept index
  int* kgrp;     // per merge: end of the first-half group, end of the both-halves group
};

__device__ __forceinline__ int count_less(const double* a, int len, double v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int count_leq(const double* a, int len, double v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Merges of up to PREP_LDS rows run the sequential deflation scan on copies
// in LDS (a dependent chain of a few accesses per entry); larger ones in
// global memory.
constexpr int PREP_LDS = 4096;

__global__ void __launch_bounds__(256) k_merge_prep(const MergeDesc* __restrict__ descs,
                                                    const double* __restrict__ lam, const double* __restrict__ es,
                                                    const double* __restrict__ Z, int ldz, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.x];
  const int lo = md.lo, mid = md.mid, hi = md.hi, ob = md.ob;
  const int ns = hi - lo, n1 = mid - lo, n2 = hi - mid;
  const double beta = es[mid - 1];
  const double rho = 2.0 * beta;
  __shared__ double s_dk[PREP_LDS], s_zk[PREP_LDS];
  __shared__ int s_keep[PREP_LDS], s_dfl[PREP_LDS], s_mix[PREP_LDS];
  const bool inl = ns <= PREP_LDS;   // uniform
  int* mix = inl ? s_mix : B.mix + lo;
  double* dk = inl ? s_dk : B.dk + lo;
  double* zk = inl ? s_zk : B.zk + lo;
  int* src = B.src + lo;
  for (int t = threadIdx.x; t < ns; t += 256) {
    double v, zr;
    int r;
    if (t < n1) {
      v = lam[lo + t];
      r = t + count_less(lam + mid, n2, v);
      zr = Z[(mid - 1) + (size_t)(lo - ob + t) * ldz];
    } else {
      v = lam[mid + t - n1];
      r = (t - n1) + count_leq(lam + lo, n1, v);
      zr = Z[mid + (size_t)(lo - ob + t) * ldz];
    }
    dk[r] = v;
    zk[r] = zr * 0.70710678118654752440;
    src[r] = t;
    mix[r] = 0;
  }
  __syncthreads();
  __shared__ double shd[4], shz[4];
  double dm = 0.0, zm = 0.0;
  for (int t = threadIdx.x; t < ns; t += 256) {
    dm = fmax(dm, fabs(dk[t]));
    zm = fmax(zm, fabs(zk[t]));
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) {
    dm = fmax(dm, __shfl_xor(dm, k, 64));
    zm = fmax(zm, __shfl_xor(zm, k, 64));
  }
  if ((threadIdx.x & 63) == 0) { shd[threadIdx.x >> 6] = dm; shz[threadIdx.x >> 6] = zm; }
  __syncthreads();
  dm = fmax(fmax(shd[0], shd[1]), fmax(shd[2], shd[3]));
  zm = fmax(fmax(shz[0], shz[1]), fmax(shz[2], shz[3]));
  const double tol = 8.0 * DEPS * fmax(dm, rho * zm);
  int* keep = inl ? s_keep : B.keep + lo;
  int* dfl = inl ? s_dfl : B.dfl + lo;
  double* rot = B.rot + 4 * (size_t)lo;
  __shared__ int sK;
  // Fast path.  The scan below changes d and z only where it rotates; if no
  // consecutive pair of non-deflated entries passes the rotation test on the
  // unrotated values, it rotates nowhere and reduces to two order-preserving
  // compactions (kept: non-negligible z; deflated: the rest, already sorted).
  // Each thread takes a contiguous chunk; scans over threads give its offsets.
  __shared__ int scnt[256], slast[256];
  const int C = (ns + 255) / 256;
  const int j0 = min(ns, (int)threadIdx.x * C), j1 = min(ns, j0 + C);
  int nk = 0, lastk = -1;
  for (int j = j0; j < j1; ++j)
    if (!(rho * fabs(zk[j]) <= tol)) { ++nk; lastk = j; }
  scnt[threadIdx.x] = nk;
  slast[threadIdx.x] = lastk;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {   // inclusive scans: sum of counts, max of last index
    const int c2 = threadIdx.x >= off ? scnt[threadIdx.x - off] : 0;
    const int l2 = threadIdx.x >= off ? slast[threadIdx.x - off] : -1;
    __syncthreads();
    scnt[threadIdx.x] += c2;
    slast[threadIdx.x] = max(slast[threadIdx.x], l2);
    __syncthreads();
  }
  const int kbase = scnt[threadIdx.x] - nk;
  int pj0 = threadIdx.x > 0 ? slast[threadIdx.x - 1] : -1;
  bool rotates = false;
  for (int j = j0; j < j1; ++j) {
    if (rho * fabs(zk[j]) <= tol) continue;
    if (pj0 >= 0) {
      double sr = zk[pj0], cr = zk[j];
      const double tt = hypot(cr, sr);
      const double t = dk[j] - dk[pj0];
      cr /= tt;
      sr = -sr / tt;
      if (fabs(t * cr * sr) <= tol) rotates = true;
    }
    pj0 = j;
  }
  if (!__syncthreads_or(rotates)) {
    int kk = kbase, dd = j0 - kbase;
    for (int j = j0; j < j1; ++j) {
      if (rho * fabs(zk[j]) <= tol) dfl[dd++] = j;
      else keep[kk++] = j;
    }
    if (threadIdx.x == 0) {
      const int K = scnt[255];
      B.nrot[blockIdx.x] = 0;
      B.kcnt[blockIdx.x] = K;
      B.rho[blockIdx.x] = rho;
      sK = K;
    }
  } else if (threadIdx.x == 0) {
    int K = 0, nd = 0, nr = 0, pj = -1;
    for (int j = 0; j < ns; ++j) {
      if (rho * fabs(zk[j]) <= tol) {
        // deflated with a negligible z: insert by value (list stays sorted)
        int q = nd++;
        while (q > 0 && dk[dfl[q - 1]] > dk[j]) { dfl[q] = dfl[q - 1]; --q; }
        dfl[q] = j;
        continue;
      }
      if (pj < 0) { pj = j; continue; }
      double s = zk[pj], c = zk[j];
      const double tt = hypot(c, s);
      const double t = dk[j] - dk[pj];
      c /= tt;
      s = -s / tt;
      if (fabs(t * c * s) <= tol) {
        zk[j] = tt;
        zk[pj] = 0.0;
        // column j now holds a combination of columns pj and j of Z
        mix[j] |= mix[pj] | ((src[pj] < n1) != (src[j] < n1));
        rot[4 * nr + 0] = (double)pj;
        rot[4 * nr + 1] = (double)j;
        rot[4 * nr + 2] = c;
        rot[4 * nr + 3] = s;
        ++nr;
        const double a = dk[pj] * c * c + dk[j] * s * s;
        dk[j] = dk[pj] * s * s + dk[j] * c * c;
        dk[pj] = a;
        int q = nd++;
        while (q > 0 && dk[dfl[q - 1]] > dk[pj]) { dfl[q] = dfl[q - 1]; --q; }
        dfl[q] = pj;
        pj = j;
      } else {
        keep[K++] = pj;
        pj = j;
      }
    }
    if (pj >= 0) keep[K++] = pj;
    B.nrot[blockIdx.x] = nr;
    B.kcnt[blockIdx.x] = K;
    B.rho[blockIdx.x] = rho;
    sK = K;
  }
  __syncthreads();
  const int K = sK;
  for (int t = threadIdx.x; t < ns; t += 256) {
    if (t < K) {
      B.dkk[lo + t] = dk[keep[t]];
      B.zkk[lo + t] = zk[keep[t]];
      if (inl) B.keep[lo + t] = keep[t];
    } else {
      B.dval[lo + t - K] = dk[dfl[t - K]];
      if (inl) B.dfl[lo + t - K] = dfl[t - K];
    }
    if (inl) B.mix[lo + t] = mix[t];
  }
}

// GEMM order of the kept columns (one workgroup per merge): first-half
// columns, then the ones mixing both halves, then second-half columns, each
// group in kept order (stable ranks from block scans over contiguous chunks).
__global__ void __launch_bounds__(256) k_merge_types(const MergeDesc* __restrict__ descs, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.x];
  const int lo = md.lo, n1 = md.mid - md.lo;
  const int K = B.kcnt[blockIdx.x];
  const int* keep = B.keep + lo;
  auto type_of = [&](int k) {
    const int sp = keep[k];
    return B.mix[lo + sp] ? 1 : (B.src[lo + sp] < n1 ? 0 : 2);
  };
  const int C = (K + 255) / 256;
  const int k0 = min(K, (int)threadIdx.x * C), k1 = min(K, k0 + C);
  int c[3] = {0, 0, 0};
  for (int k = k0; k < k1; ++k) ++c[type_of(k)];
  __shared__ int sc[3][256];
#pragma unroll
  for (int g = 0; g < 3; ++g) sc[g][threadIdx.x] = c[g];
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {   // inclusive scans of the three counts
    int v[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) v[g] = threadIdx.x >= off ? sc[g][threadIdx.x - off] : 0;
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 3; ++g) sc[g][threadIdx.x] += v[g];
    __syncthreads();
  }
  const int n0 = sc[0][255], nm = sc[1][255];
  int pos[3] = {sc[0][threadIdx.x] - c[0], n0 + sc[1][threadIdx.x] - c[1], n0 + nm + sc[2][threadIdx.x] - c[2]};
  for (int k = k0; k < k1; ++k) {
    const int p = pos[type_of(k)]++;
    B.gpos[lo + k] = p;
    B.ginv[lo + p] = k;
  }
  if (threadIdx.x == 0) {
    B.kgrp[2 * blockIdx.x] = n0;
    B.kgrp[2 * blockIdx.x + 1] = n0 + nm;
  }
}

// The deflation rotations, applied in order to the merge's columns of Z
// (one thread per row).
__global__ void __launch_bounds__(256) k_merge_rot(const MergeDesc* __restrict__ descs, double* __restrict__ Z,
                                                   int ldz, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.y];
  const int lo = md.lo, ns = md.hi - md.lo, cb = md.lo - md.ob;
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int nr = B.nrot[blockIdx.y];
  if (r >= ns || nr == 0) return;
  const int* src = B.src + lo;
  const double* rot = B.rot + 4 * (size_t)lo;
  double* zr = Z + (lo + r);
  for (int t = 0; t < nr; ++t) {
    const int pj = (int)rot[4 * t], j = (int)rot[4 * t + 1];
    const double c = rot[4 * t + 2], s = rot[4 * t + 3];
    const size_t cp = (size_t)(cb + src[pj]) * ldz, cj = (size_t)(cb + src[j]) * ldz;
    const double x = zr[cp], y = zr[cj];
    zr[cp] = c * x + s * y;
    zr[cj] = c * y - s * x;
  }
}

// Qtmp = the merge's block of Z with its columns in the order kept (in GEMM
// order), then deflated.  Grid: x = 256-row chunks, y = GATHER_COLS-column chunks, z = merge.
constexpr int GATHER_COLS = 8;
__global__ void __launch_bounds__(256) k_merge_gather(const MergeDesc* __restrict__ descs,
                                                      const double* __restrict__ Z, int ldz, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.z];
  const int lo = md.lo, ns = md.hi - md.lo, cb = md.lo - md.ob;
  const int k0 = blockIdx.y * GATHER_COLS;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (k0 >= ns || r >= ns) return;
  const int K = B.kcnt[blockIdx.z];
  const int* src = B.src + lo;
  double* Q = B.Qtmp + (size_t)lo * B.lq;
  const double* zr = Z + (lo + r);
  int col[GATHER_COLS];
#pragma unroll
  for (int u = 0; u < GATHER_COLS; ++u) {
    const int k = min(k0 + u, ns - 1);
    const int sp = (k < K) ? B.keep[lo + B.ginv[lo + k]] : B.dfl[lo + k - K];
    col[u] = cb + src[sp];
  }
  double v[GATHER_COLS];
#pragma unroll
  for (int u = 0; u < GATHER_COLS; ++u) v[u] = zr[(size_t)col[u] * ldz];
#pragma unroll
  for (int u = 0; u < GATHER_COLS; ++u)
    if (k0 + u < ns) Q[r + (size_t)(k0 + u) * ns] = v[u];
}

// One root of f(t) = 1 + rho sum_j z_j^2 / (delta_j - t) per wavefront, the 64
// lanes splitting the sum.  Origin: the nearer pole (chosen by the sign of f
// at the midpoint of the interval), so t is small and the differences
// delta_j - t are formed without cancellation.  Iteration: Newton on the
// pole-free G(t) = (t - p_lo)(p_hi - t) f(t), safeguarded by a bisection
// bracket, until the step falls below a few ulps of t, starting from the root
// of the two-pole model fitted at the midpoint.
struct SecSums {
  double r, rp;   // rest = rho sum over the non-pole terms, and its derivative
};

__device__ __forceinline__ SecSums sec_rest(const double* d, const double* z, int K, double dorg, double rho, double t,
                                            int plo, int phi) {
  const int lane = threadIdx.x & 63;
  double r = 0.0, rp = 0.0;
  for (int jj = lane; jj < K; jj += 64) {
    if (jj == plo || jj == phi) continue;
    const double zj = z[jj];
    const double inv = 1.0 / ((d[jj] - dorg) - t);
    const double q = zj * zj * inv;
    r += q;
    rp += q * inv;
  }
  r = wave_sum_d(r);
  rp = wave_sum_d(rp);
  return SecSums{rho * r, rho * rp};
}

// A block serves ROOTS roots of one merge.  Every Newton step re-reads the
// merge's poles d and weights z, so merges of up to SEC_LDS kept values
// first stage both in LDS (one read from L2 per block instead of one per
// root and step; the arithmetic and its order are unchanged).  The LDS is
// sized for the level's largest merge, and levels of small merges (the many
// deep merges of a batch) use blocks of 4 roots, so that they keep their
// occupancy.
constexpr int SEC_LDS = 4096;

__device__ __forceinline__ void secular_root(const MergeDesc& md, const MergeBufs& B, int K, int i, const double* d,
                                             const double* z) {
  const int lo = md.lo;
  const double rho = B.rho[blockIdx.y];
  const int lane = threadIdx.x & 63;
  int org, plo, phi;
  double a, b, pl, ph;   // bracket (a, b) and pole positions relative to the origin
  double t0;             // first iterate, relative to the origin
  if (i < K - 1) {
    const double gap = d[i + 1] - d[i];
    const double half = 0.5 * gap;
    // f at the midpoint with origin d_i, and the part of it from the other poles
    double s = 0.0, sr = 0.0;
    for (int jj = lane; jj < K; jj += 64) {
      const double zj = z[jj];
      const double q = zj * zj / ((d[jj] - d[i]) - half);
      s += q;
      if (jj != i && jj != i + 1) sr += q;
    }
    s = wave_sum_d(s);
    sr = wave_sum_d(sr);
    plo = i;
    phi = i + 1;
    if (1.0 + rho * s >= 0.0) { org = i; a = 0.0; b = half; pl = 0.0; ph = gap; }
    else { org = i + 1; a = -half; b = 0.0; pl = -gap; ph = 0.0; }
    // First iterate: the root in the bracket of the two-pole model
    //   c + rho z_i^2 / (pl - t) + rho z_{i+1}^2 / (ph - t),  c = 1 + rho (other poles at the midpoint),
    // i.e. of -c t^2 + (c (pl + ph) + zl + zh) t - zl ph - zh pl (pl ph = 0), written
    // about the origin so that a root a few ulps from its pole keeps its
    // relative accuracy (the rational initial guess of LAPACK's dlaed4, in its
    // simplest form).  The quadratic has one root in (pl, ph); it is the one of
    // the two stable-formula roots that lies in (a, b), else the midpoint.
    const double zl = rho * z[i] * z[i], zh = rho * z[i + 1] * z[i + 1];
    const double c = 1.0 + rho * sr;
    const double qa = -c, qb = c * (pl + ph) + zl + zh, qc = -zl * ph - zh * pl;
    const double disc = fmax(qb * qb - 4.0 * qa * qc, 0.0);
    const double qq = -0.5 * (qb + copysign(sqrt(disc), qb));
    const double r1 = (qa != 0.0) ? qq / qa : 0.5 * (a + b);
    const double r2 = (qq != 0.0) ? qc / qq : 0.5 * (a + b);
    t0 = (r1 > a && r1 < b) ? r1 : r2;
  } else {
    double zz = 0.0;
    for (int jj = lane; jj < K; jj += 64) zz += z[jj] * z[jj];
    zz = wave_sum_d(zz);
    org = i; plo = i; phi = -1;
    a = 0.0; b = rho * zz; pl = 0.0; ph = 0.0;
    // First iterate: the root zl / c of the one-pole model c + rho z_i^2 / (0 - t),
    // c = 1 + rho (the other poles at the midpoint of the bracket)
    const double tm = 0.5 * (a + b);
    double sr = 0.0;
    for (int jj = lane; jj < K - 1; jj += 64) {
      const double zj = z[jj];
      sr += zj * zj / ((d[jj] - d[i]) - tm);
    }
    sr = wave_sum_d(sr);
    const double c = 1.0 + rho * sr;
    const double r = rho * z[i] * z[i] / c;
    t0 = (c > 0.0 && r > a && r < b) ? r : tm;
  }
  const double dorg = d[org];
  const double zl2 = rho * z[plo] * z[plo];
  const double zh2 = (phi >= 0) ? rho * z[phi] * z[phi] : 0.0;
  double t = (t0 > a && t0 < b) ? t0 : 0.5 * (a + b);
  bool conv = false;
  for (int it = 0; it < B.cap; ++it) {
    const SecSums S = sec_rest(d, z, K, dorg, rho, t, plo, phi);
    const double one = 1.0 + S.r;
    double G, Gp;
    if (phi >= 0) {
      const double tl = t - pl, th = ph - t;
      G = tl * th * one - zl2 * th + zh2 * tl;
      Gp = (th - tl) * one + tl * th * S.rp + zl2 + zh2;
    } else {
      const double tl = t - pl;
      G = tl * one - zl2;
      Gp = one + tl * S.rp;
    }
    if (G == 0.0) { a = b = t; conv = true; break; }
    if (G < 0.0) a = t; else b = t;
    // A Newton correction below the stopping size ends the iteration before
    // the bracket test: t is an end of the bracket now, and t - dt can round
    // onto it, which must not be taken for a step out of the bracket.
    const double dt = G / Gp;
    if (fabs(dt) <= 4.0 * DEPS * fabs(t)) {
      const double tc = t - dt;
      if (tc >= a && tc <= b) t = tc;
      conv = true;
      break;
    }
    double tn = t - dt;
    if (!(tn > a && tn < b)) tn = 0.5 * (a + b);
    const double step = fabs(tn - t);
    t = tn;
    if (!(b > a) || step <= 4.0 * DEPS * fabs(t) || !(0.5 * (a + b) > a && 0.5 * (a + b) < b)) {
      conv = true;
      break;
    }
  }
  // never return the pole itself
  if (org == i && !(t > 0.0)) t = b > 0.0 ? b : DBL_MIN;
  if (org != i && !(t < 0.0)) t = a < 0.0 ? a : -DBL_MIN;
  if (lane == 0) {
    B.org[lo + i] = org;
    B.tau[lo + i] = t;
    // no convergence within the step cap: reported as zhbev reports a QL failure (info > 0)
    if (!conv) atomicMax(B.info + md.pr, lo - md.ob + i + 1);
  }
}

// kst: doubles of dynamic LDS per staged array (the level's largest merge,
// 0 when it exceeds SEC_LDS and the roots read d and z from global memory)
template <int ROOTS>
__global__ void __launch_bounds__(64 * ROOTS) k_merge_secular(const MergeDesc* __restrict__ descs, MergeBufs B,
                                                             int kst) {
  const MergeDesc md = descs[blockIdx.y];
  const int K = B.kcnt[blockIdx.y];
  if ((int)blockIdx.x * ROOTS >= K) return;   // block-uniform: no root of this merge here
  const int i = blockIdx.x * ROOTS + (threadIdx.x >> 6);
  const double* d = B.dkk + md.lo;
  const double* z = B.zkk + md.lo;
  if (K <= kst) {   // block-uniform
    extern __shared__ double sdz[];
    double* sd = sdz;
    double* sz = sdz + kst;
    for (int q = threadIdx.x; q < K; q += 64 * ROOTS) {
      sd[q] = d[q];
      sz[q] = z[q];
    }
    __syncthreads();
    if (i < K) secular_root(md, B, K, i, sd, sz);   // wave-uniform
  } else if (i < K) {
    secular_root(md, B, K, i, d, z);
  }
}

// Gu-Eisenstat: zhat_j^2 = prod_i (lambda_i - d_j) / (rho prod_{i != j} (d_i - d_j)),
// one wavefront per j, the lanes splitting the product.
__global__ void __launch_bounds__(256) k_merge_zhat(const MergeDesc* __restrict__ descs, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.y];
  const int lo = md.lo;
  const int K = B.kcnt[blockIdx.y];
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= K) return;   // wave-uniform
  const int lane = threadIdx.x & 63;
  const double rho = B.rho[blockIdx.y];
  const double* d = B.dkk + lo;
  const int* org = B.org + lo;
  const double* tau = B.tau + lo;
  const double dj = d[j];
  double prod = 1.0;
  for (int i = lane; i < K; i += 64) {
    const double num = (d[org[i]] - dj) + tau[i];
    prod *= (i == j) ? num / rho : num / (d[i] - dj);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) prod *= __shfl_xor(prod, m, 64);
  if (lane == 0) B.zhat[lo + j] = copysign(sqrt(fmax(prod, 0.0)), B.zkk[lo + j]);
}

__global__ void __launch_bounds__(256) k_merge_buildU(const MergeDesc* __restrict__ descs, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.y];
  const int lo = md.lo, ns = md.hi - md.lo;
  const int K = B.kcnt[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= K) return;
  const double* d = B.dkk + lo;
  const double* zh = B.zhat + lo;
  const int o = B.org[lo + i];
  const double t = B.tau[lo + i];
  const double dorg = d[o];
  double* Ucol = B.U + (size_t)lo * B.lq + (size_t)i * ns;
  const int* gpos = B.gpos + lo;
  double ss = 0.0;
  for (int j = threadIdx.x; j < K; j += 256) {   // row j of the kept order at its GEMM position
    const double v = zh[j] / ((d[j] - dorg) - t);
    Ucol[gpos[j]] = v;
    ss += v * v;
  }
  __shared__ double sh[4];
  ss = wave_sum_d(ss);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = ss;
  __syncthreads();
  const double inv = 1.0 / sqrt(sh[0] + sh[1] + sh[2] + sh[3]);
  for (int j = threadIdx.x; j < K; j += 256) Ucol[j] *= inv;
}

__global__ void __launch_bounds__(256) k_merge_order(const MergeDesc* __restrict__ descs, double* __restrict__ lam,
                                                     MergeBufs B) {
  const MergeDesc md = descs[blockIdx.x];
  const int lo = md.lo, ns = md.hi - md.lo;
  const int K = B.kcnt[blockIdx.x];
  const int nd = ns - K;
  // roots are written to lam only after every thread has read the deflated list
  double* dv = B.dval + lo;
  const double* d = B.dkk + lo;
  for (int t = threadIdx.x; t < ns; t += 256) {
    if (t < K) {
      const double v = d[B.org[lo + t]] + B.tau[lo + t];
      B.posK[lo + t] = t + count_less(dv, nd, v);
      B.zhat[lo + t] = v;  // zhat no longer needed: keep the root value
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ns; t += 256) {
    if (t < nd) {
      const double v = dv[t];
      // number of roots <= v: roots are sorted ascending
      int a = 0, b = K;
      while (a < b) {
        const int m = (a + b) >> 1;
        if (B.zhat[lo + m] <= v) a = m + 1; else b = m;
      }
      B.posD[lo + t] = t + a;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ns; t += 256) {
    if (t < K) lam[lo + B.posK[lo + t]] = B.zhat[lo + t];
    else lam[lo + B.posD[lo + t - K]] = dv[t - K];
  }
}

__global__ void __launch_bounds__(256) k_merge_copydefl(const MergeDesc* __restrict__ descs, double* __restrict__ Z,
                                                        int ldz, MergeBufs B) {
  const MergeDesc md = descs[blockIdx.z];
  const int lo = md.lo, ns = md.hi - md.lo;
  const int K = B.kcnt[blockIdx.z];
  const int t = blockIdx.y;
  if (t >= ns - K) return;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= ns) return;
  const double* Q = B.Qtmp + (size_t)lo * B.lq;
  Z[(lo + r) + (size_t)(lo - md.ob + B.posD[lo + t]) * ldz] = Q[r + (size_t)(K + t) * ns];
}

// Eigenvalues of a problem whose input held a NaN or an infinity are NaN.
__global__ void k_dc_unscale(const int* __restrict__ probs, const double* __restrict__ lam,
                             const double* __restrict__ scal, const int* __restrict__ info, double* __restrict__ w) {
  const int o = probs[2 * blockIdx.y], n = probs[2 * blockIdx.y + 1];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool bad = (info[blockIdx.y] & ZQI_INFO_NONFINITE) != 0;
  if (i < n) w[o + i] = bad ? __builtin_nan("") : lam[o + i] * scal[blockIdx.y];
}

// --------------------------------------------------------------------------
// eigenvalues only: Sturm-count multisection, one eigenvalue per wavefront, on the
// scaled tridiagonal.
// --------------------------------------------------------------------------
// The recurrence is a chain of divisions; its operands are loaded eight at a
// time ahead of the chain so that the loads do not sit on it.
__device__ __forceinline__ int sturm_count(int n, const double* d, const double* e2, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++cnt;
  int i = 1;
  for (; i + 8 <= n; i += 8) {
    double dd[8], ee[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dd[u] = d[i + u];
      ee[u] = e2[i + u - 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      q = (dd[u] - x) - ee[u] / q;
      if (fabs(q) < pivmin) q = -pivmin;
      if (q < 0.0) ++cnt;
    }
  }
  for (; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Eigenvalues of problem blockIdx.y by multisection, MS_LANES lanes per
// eigenvalue (64 / MS_LANES per wavefront): each step the group's lanes
// evaluate Sturm counts at the points a + (b - a) (l + 1) / (MS_LANES + 1), and
// the first lane whose count exceeds k brackets the eigenvalue with its left
// neighbour, so that count(a) <= k < count(b) holds after every step
// (whatever the rounding of the counts) and the bracket shrinks
// (MS_LANES + 1)x per step.  More lanes mean fewer steps of the chain of
// divisions each lane runs, but more count work in total: 64 lanes are
// faster while the eigenvalues do not fill the device (n up to ~2.5k), 16
// above.
template <int MS_LANES>
__global__ void __launch_bounds__(256) k_multisect(const int* __restrict__ probs, const double* __restrict__ d,
                                                   const double* __restrict__ e2, double gl, double gu, double pivmin,
                                                   const double* __restrict__ scal, int* __restrict__ info,
                                                   double* __restrict__ w, int cap) {
  const int o = probs[2 * blockIdx.y], n = probs[2 * blockIdx.y + 1];
  const int lane = threadIdx.x & 63;
  const int grp = lane / MS_LANES, gl0 = grp * MS_LANES, li = lane - gl0;
  const int k = blockIdx.x * (256 / MS_LANES) + threadIdx.x / MS_LANES;
  const int wave_k0 = blockIdx.x * (256 / MS_LANES) + (threadIdx.x >> 6) * (64 / MS_LANES);
  if (wave_k0 >= n) return;   // whole wavefront
  d += o; e2 += o; w += o;
  double a = gl, b = gu;
  bool live = k < n;          // this group still narrowing (uniform within the group)
  // bracket width down to a few ulps of the eigenvalue, or of the matrix
  // norm (1 after scaling) for eigenvalues near zero: an absolute error of
  // eps ||T|| is what any backward-stable eigensolver commits
  for (int it = 0; it < cap; ++it) {
    if (live && b - a <= 2.0 * DEPS * fmax(fmax(fabs(a), fabs(b)), 1.0)) live = false;
    if (!__any(live)) break;
    const double x = a + (b - a) * ((double)(li + 1) * (1.0 / (MS_LANES + 1)));
    const bool up = live && sturm_count(n, d, e2, x, pivmin) > k;
    const unsigned long long all = __ballot(up);
    const unsigned long long grpmask = MS_LANES == 64 ? all : ((all >> gl0) & ((1ull << (MS_LANES % 64)) - 1));
    const int f = grpmask ? __ffsll((long long)grpmask) - 1 : MS_LANES;
    const double xb = __shfl(x, gl0 + (f < MS_LANES ? f : MS_LANES - 1));
    const double xl = __shfl(x, gl0 + (f > 0 ? f - 1 : 0));
    if (live) {
      const double na = f == 0 ? a : xl;
      const double nbd = f < MS_LANES ? xb : b;
      if (na == a && nbd == b) live = false;   // the points no longer split the bracket
      a = na;
      b = nbd;
    }
  }
  if (live && b - a <= 2.0 * DEPS * fmax(fmax(fabs(a), fabs(b)), 1.0)) live = false;   // narrowed by the last step
  if (li == 0 && k < n) {
    const bool bad = (info[blockIdx.y] & ZQI_INFO_NONFINITE) != 0;
    w[k] = bad ? __builtin_nan("") : 0.5 * (a + b) * scal[blockIdx.y];
    if (live) atomicMax(info + blockIdx.y, k + 1);   // bracket still open after the step cap
  }
}

__global__ void k_square(int n, const double* __restrict__ es, double* __restrict__ e2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n - 1) e2[i] = es[i] * es[i];
}

// Eigenvalues only, after k_dc_scale: Sturm-count multisection on every
// scaled tridiagonal (their Gershgorin intervals lie in [-3, 3]).
static int values_kernels(int nprob, int nmax, int ntot, const int* iprob, double* w, const StedcWork& sw,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_square, dim3(cdiv(ntot, 256)), dim3(256), 0, s, ntot, sw.es, sw.e2);
  const double pivmin = DBL_MIN * 4.0;
  if (nmax <= 2560)
    hipLaunchKernelGGL(k_multisect<64>, dim3(cdiv(nmax, 4), nprob), dim3(256), 0, s, iprob, sw.lam, sw.e2,
                       -3.0 - 1e-12, 3.0 + 1e-12, pivmin, sw.scal, sw.info, w, sw.cap_bisect);
  else
    hipLaunchKernelGGL(k_multisect<16>, dim3(cdiv(nmax, 16), nprob), dim3(256), 0, s, iprob, sw.lam, sw.e2,
                       -3.0 - 1e-12, 3.0 + 1e-12, pivmin, sw.scal, sw.info, w, sw.cap_bisect);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

__global__ void k_set_prob(int* __restrict__ probs, int o, int n) {
  probs[0] = o;
  probs[1] = n;
}

int eigvals_one(int n, const double* d, const double* e, double* w, const StedcWork& sw, hipStream_t s) {
  if (n <= 0) return ZQI_OK;
  int* iprob = sw.ints;   // (offset, size) written by a kernel: nothing is uploaded from the host
  hipLaunchKernelGGL(k_set_prob, dim3(1), dim3(1), 0, s, iprob, 0, n);
  hipLaunchKernelGGL(k_dc_scale, dim3(1), dim3(1024), 0, s, iprob, d, e, sw.lam, sw.es, sw.scal, sw.info);
  return values_kernels(1, n, n, iprob, w, sw, s);
}

__global__ void k_status_out(const int* __restrict__ word, int* __restrict__ out) { out[0] = info_code(word[0]); }

int status_out(const int* word, int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_status_out, dim3(1), dim3(1), 0, s, word, out);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

// --------------------------------------------------------------------------
static void build_tree(int lo, int hi, int ob, int pr, int depth, std::vector<int>& leaves,
                       std::vector<std::vector<MergeDesc>>& lv, std::vector<int>& splits) {
  if (hi - lo <= LEAF) {
    leaves.insert(leaves.end(), {lo, hi, ob, pr});
    return;
  }
  const int mid = lo + (hi - lo) / 2;
  if ((int)lv.size() <= depth) lv.resize(depth + 1);
  lv[depth].push_back(MergeDesc{lo, mid, hi, ob, pr});
  splits.push_back(mid);
  build_tree(lo, mid, ob, pr, depth + 1, leaves, lv, splits);
  build_tree(mid, hi, ob, pr, depth + 1, leaves, lv, splits);
}

size_t stedc_int_slots(int ntot, int nprob) {
  // per problem of size n at most L = 2 ceil(n / LEAF) leaves (4 ints each),
  // L - 1 splits (1 int) and L - 1 merges (5 ints), and (offset, size)
  return (size_t)22 * (ntot / LEAF + nprob) + 2 * (size_t)nprob + 64;
}

int stedc(int nprob, const int* offs, const int* ns, int ntot, int lq, const double* d, const double* e, double* w,
          const StedcWork& sw, bool vectors, hipStream_t s) {
  if (nprob <= 0 || ntot <= 0) return ZQI_OK;
  int nmax = 0;
  std::vector<int> probs(2 * (size_t)nprob);
  for (int z = 0; z < nprob; ++z) {
    probs[2 * z] = offs[z];
    probs[2 * z + 1] = ns[z];
    nmax = std::max(nmax, ns[z]);
  }
  if (nmax <= 0) return ZQI_OK;
  std::vector<int> leaves, splits;
  std::vector<std::vector<MergeDesc>> levels;
  if (vectors)
    for (int z = 0; z < nprob; ++z)
      if (ns[z] > 0) build_tree(offs[z], offs[z] + ns[z], offs[z], z, 0, leaves, levels, splits);
  // device table: merge descriptors (deepest level first), leaves, splits, problems
  std::vector<MergeDesc> flat;
  std::vector<size_t> level_off;
  for (int dpt = (int)levels.size() - 1; dpt >= 0; --dpt) {
    level_off.push_back(flat.size());
    for (auto& md : levels[dpt]) flat.push_back(md);
  }
  static_assert(sizeof(MergeDesc) == 5 * sizeof(int), "MergeDesc is uploaded as 5 ints");
  std::vector<int> host(5 * flat.size() + leaves.size() + splits.size() + probs.size());
  size_t o = 0;
  for (auto& md : flat) {
    host[o++] = md.lo;
    host[o++] = md.mid;
    host[o++] = md.hi;
    host[o++] = md.ob;
    host[o++] = md.pr;
  }
  const size_t o_leaf = o;
  for (int v : leaves) host[o++] = v;
  const size_t o_split = o;
  for (int v : splits) host[o++] = v;
  const size_t o_prob = o;
  for (int v : probs) host[o++] = v;
  if (host.size() > sw.int_slots) return ZQI_ERR_ARG;
  int rc = stage_upload(*sw.stage, host.data(), host.size() * sizeof(int), sw.ints, s);
  if (rc) return rc;
  const MergeDesc* idesc = reinterpret_cast<const MergeDesc*>(sw.ints);
  const int* ileaf = sw.ints + o_leaf;
  const int* isplit = sw.ints + o_split;
  const int* iprob = sw.ints + o_prob;

  hipLaunchKernelGGL(k_dc_scale, dim3(nprob), dim3(1024), 0, s, iprob, d, e, sw.lam, sw.es, sw.scal, sw.info);
  if (!vectors) return values_kernels(nprob, nmax, ntot, iprob, w, sw, s);
  const int nleaf = (int)leaves.size() / 4;
  const int ldz = ntot;
  if ((rc = fill_zero(sw.Z, (size_t)ntot * lq * sizeof(double), s))) return rc;
  if (!splits.empty())
    hipLaunchKernelGGL(k_dc_tear, dim3(cdiv((int)splits.size(), 256)), dim3(256), 0, s, (int)splits.size(), isplit,
                       sw.es, sw.lam);
  hipLaunchKernelGGL(k_dc_leaf, dim3(nleaf), dim3(64), 0, s, ileaf, sw.lam, sw.es, sw.Z, ldz, sw.info, sw.cap_leaf);
  MergeBufs B{sw.dk, sw.zk, sw.src, sw.keep, sw.dfl, sw.rot, sw.nrot, sw.kcnt, sw.dkk, sw.zkk, sw.dval, sw.rho,
              sw.org, sw.tau, sw.zhat, sw.posK, sw.posD, sw.Qtmp, sw.U, lq, sw.info, sw.cap_secular,
              sw.mix, sw.gpos, sw.ginv, sw.kgrp};
  for (size_t li = 0; li < level_off.size(); ++li) {
    const size_t off = level_off[li];
    const size_t cnt = (li + 1 < level_off.size() ? level_off[li + 1] : flat.size()) - off;
    const MergeDesc* dd = idesc + off;
    int nsmax = 0;
    for (size_t q = 0; q < cnt; ++q) nsmax = std::max(nsmax, flat[off + q].hi - flat[off + q].lo);
    const int nm = (int)cnt;
    hipLaunchKernelGGL(k_merge_prep, dim3(nm), dim3(256), 0, s, dd, sw.lam, sw.es, sw.Z, ldz, B);
    hipLaunchKernelGGL(k_merge_rot, dim3(cdiv(nsmax, 256), nm), dim3(256), 0, s, dd, sw.Z, ldz, B);
    hipLaunchKernelGGL(k_merge_types, dim3(nm), dim3(256), 0, s, dd, B);
    hipLaunchKernelGGL(k_merge_gather, dim3(cdiv(nsmax, 256), cdiv(nsmax, GATHER_COLS), nm), dim3(256), 0, s, dd,
                       sw.Z, ldz, B);
    {
      const int kst = nsmax <= SEC_LDS ? nsmax : 0;
      const unsigned shm = (unsigned)(2 * kst * sizeof(double));
      if (nsmax <= 512)
        hipLaunchKernelGGL(k_merge_secular<4>, dim3(cdiv(nsmax, 4), nm), dim3(256), shm, s, dd, B, kst);
      else
        hipLaunchKernelGGL(k_merge_secular<16>, dim3(cdiv(nsmax, 16), nm), dim3(1024), shm, s, dd, B, kst);
    }
    hipLaunchKernelGGL(k_merge_zhat, dim3(cdiv(nsmax, 4), nm), dim3(256), 0, s, dd, B);
    hipLaunchKernelGGL(k_merge_buildU, dim3(nsmax, nm), dim3(256), 0, s, dd, B);
    hipLaunchKernelGGL(k_merge_order, dim3(nm), dim3(256), 0, s, dd, sw.lam, B);
    {
      ProfScope ps(PROF_MERGE, s, 0.0);
      rc = dgemm_merges(dd, nm, nsmax, sw.kcnt, sw.kgrp, sw.Qtmp, sw.U, lq, sw.Z, ldz, sw.posK, s);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_merge_copydefl, dim3(cdiv(nsmax, 256), nsmax, nm), dim3(256), 0, s, dd, sw.Z, ldz, B);
  }
  hipLaunchKernelGGL(k_dc_unscale, dim3(cdiv(nmax, 256), nprob), dim3(256), 0, s, iprob, sw.lam, sw.scal, sw.info, w);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
