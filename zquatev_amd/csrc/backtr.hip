// Eigenvector back-transform X = Q Phi Z.
//
// Replaces the reference's explicit Q accumulation (blocked.cc:477-544,
// unblocked.cc:75-80, :121-126) followed by U = Q0 Z, V = Q1 Z
// (zquatev.cc:86-90).  Q is never formed.  The reflectors u_c, kept in place
// below the subdiagonal of M, are grouped in blocks of wbt consecutive ones
// (independent of the reduction's panel width) with compact-WY factors,
// H_p ... H_{p+w-1} = I - V T V^*, T = (I + triu(V^* V, 1))^{-1} (every tau is
// 1 because u^* u = 2).  For all blocks at once: G = V^* V (split-K FP64
// MFMA products and their sums), T, and VT = V T^*.  Then per block, last
// block first, two MFMA products on X's rows r0 = p + 1 and below:
//   Y = VT^* X[r0:, :]   (= T V^* X, split over the long dimension),
//   X[r0:, :] -= V Y.
// Phi is the diagonal of unit quaternions with Phi^* T_q Phi real (zhbtrd's
// phase scaling for kd = 1, applied here to the quaternion subdiagonal); Z is
// the real eigenvector matrix of the tridiagonal.
#include "zq_internal.h"
#include <climits>
#include <algorithm>
#include <type_traits>
#include <vector>

namespace zq {

// phi_0 = 1, phi_{c+1} = r_c phi_c with r_c = e_c / |e_c| (1 where e_c = 0):
// a prefix product of unit quaternions, chunked over one workgroup.
constexpr int PH_T = 256;
__global__ void __launch_bounds__(PH_T) k_phases(int n, const quat* __restrict__ e, quat* __restrict__ phi, ZB zb) {
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    e = zoff(e, zb); phi = zoff(phi, zb);
  }
  __shared__ quat tot[PH_T];
  __shared__ quat off[PH_T];
  const int t = threadIdx.x;
  const int chunk = (n + PH_T - 1) / PH_T;
  const int a = t * chunk, b = min(n, a + chunk);
  // local prefix: L_k = r_{k-1} ... r_a for k in (a, b], L_a = 1
  quat L = quat{1.0, 0.0, 0.0, 0.0};
  for (int k = a; k < b; ++k) {
    if (k > a) {
      const quat ek = qload(e + k - 1);
      const double ae = sqrt(qnorm2(ek));
      const quat r = (ae > 0.0) ? qscale(ek, 1.0 / ae) : quat{1.0, 0.0, 0.0, 0.0};
      L = qmul(r, L);
    }
    qstore(phi + k, L);
  }
  // chunk total maps phi_a to phi_b: T = r_{b-1} L_{b-1}
  quat T = L;
  if (b > a && b < n) {
    const quat ek = qload(e + b - 1);
    const double ae = sqrt(qnorm2(ek));
    const quat r = (ae > 0.0) ? qscale(ek, 1.0 / ae) : quat{1.0, 0.0, 0.0, 0.0};
    T = qmul(r, L);
  }
  tot[t] = T;
  __syncthreads();
  if (t == 0) {
    quat cur = quat{1.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < PH_T; ++k) {
      off[k] = cur;
      if (k * chunk < n) cur = qmul(tot[k], cur);
    }
  }
  __syncthreads();
  const quat o = off[t];
  for (int k = a; k < b; ++k) qstore(phi + k, qmul(qload(phi + k), o));
}

// X[i, j] = phi_i Z[i, j]  (in a batch, problem z's Z block starts at row zrow[z])
__global__ void __launch_bounds__(256) k_phase_apply(int n, const quat* __restrict__ phi, const double* __restrict__ Z,
                                                     int ldz, quat* __restrict__ X, int ldx, ZB zb,
                                                     const int* __restrict__ zrow) {
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    phi = zoff(phi, zb); X = zoff(X, zb);
    Z += zrow[blockIdx.z];
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || j >= n) return;
  const double z = Z[i + (size_t)j * ldz];
  qstore(X + i + (size_t)j * ldx, qscale(qload(phi + i), z));
}

// out = sum_z parts[z] (each part `total` quaternions) for a list of tasks (task blockIdx.z)
__global__ void __launch_bounds__(256) k_sum_parts_b(const SumTask* __restrict__ tasks) {
  const SumTask t = tasks[blockIdx.z];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.total) return;
  quat s = qzero();
  for (int z = 0; z < t.nz; ++z) s = qadd(s, qload_g(t.parts + (size_t)z * t.total + i));
  qstore_g(t.out + i, s);
}

// T = (I + triu(G, 1))^{-1} for each block: T[0:j, j] = -T[0:j, 0:j] G[0:j, j],
// T[j, j] = 1, zero below (the compact-WY recurrence with every tau = 1).
// G and T are stored per block of sub * WB reflectors (leading dimension
// sub * WB); workgroup x factors the diagonal WB x WB sub-block h = x % sub
// of stored block x / sub (problem blockIdx.z) and clears the part of T left
// of it (rows h WB .. + WB, columns 0 .. h WB).  For sub > 1 the blocks above
// the diagonal are formed afterwards by GEMMs, pairs of neighbouring
// sub-blocks of width w = WB, 2 WB, ... merged level by level:
// T12 = -T1 (G12 T2).
constexpr int TF_WB = 64;
__global__ void __launch_bounds__(256) k_tfactor(int n, const quat* __restrict__ Gall, quat* __restrict__ Tall,
                                                 int sub, ZB zb) {
  constexpr int WB = TF_WB;
  if (zb.nvec) {
    n = zb.nvec[blockIdx.z];
    Gall = zoff(Gall, zb); Tall = zoff(Tall, zb);
  }
  const int ldt = sub * WB;
  const int sb = blockIdx.x / sub, h = blockIdx.x - sb * sub;
  if (sb >= (n - 1 + ldt - 1) / ldt) return;
  const int p = sb * ldt + h * WB;
  const int wdt = min(WB, n - 1 - p);   // <= 0: no reflectors here, T is the identity
  const quat* G = Gall + (size_t)sb * ldt * ldt + (size_t)h * WB * (ldt + 1);
  quat* T = Tall + (size_t)sb * ldt * ldt + (size_t)h * WB * (ldt + 1);
  // Columns in groups of TB = 256 / WB, one per wave set: thread (r, c) (its
  // row r of T, column J + c of the group) first applies the columns l < J
  // (T rows re-read from LDS, the group's G entries staged in LDS); then the
  // threads of c = 0 run the recurrence inside the group from those
  // partials.  Every entry sums l = r .. jj-1 in increasing order, as the
  // column-by-column recurrence does.  T stays in LDS (its rows are re-read
  // for every later group) and is written out at the end.
  constexpr int TB = 256 / WB;
  __shared__ quat gs[TB][WB];
  __shared__ quat ax[TB][WB];
  __shared__ quat Ts[WB][WB];   // Ts[col][row]
  const int r = threadIdx.x % WB, cq = threadIdx.x / WB;
  for (int J = 0; J < WB; J += TB) {
    const int jq = J + cq;
    if (J >= wdt) {   // columns past the block's reflectors: identity (uniform branch)
      Ts[jq][r] = r == jq ? quat{1.0, 0.0, 0.0, 0.0} : qzero();
      continue;
    }
    gs[cq][r] = (jq < wdt && r < jq) ? qload(G + r + (size_t)jq * ldt) : qzero();
    __syncthreads();
    quat acc = qzero();
#pragma unroll 4
    for (int l = r; l < J; ++l) qfma(acc, Ts[l][r], gs[cq][l]);
    ax[cq][r] = acc;
    __syncthreads();
    if (cq == 0) {
      quat tb[TB];
#pragma unroll
      for (int c = 0; c < TB; ++c) {
        const int jj = J + c;
        quat a = ax[c][r];
#pragma unroll
        for (int cc = 0; cc < c; ++cc)
          if (r <= J + cc) qfma(a, tb[cc], gs[c][J + cc]);
        if (jj < wdt && r < jj) tb[c] = qscale(a, -1.0);
        else tb[c] = (r == jj) ? quat{1.0, 0.0, 0.0, 0.0} : qzero();
        Ts[jj][r] = tb[c];
      }
    }
    __syncthreads();
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < WB * WB; idx += 256) {
    const int col = idx / WB, row = idx - col * WB;
    qstore(T + row + (size_t)col * ldt, Ts[col][row]);
  }
  for (int idx = threadIdx.x; idx < WB * h * WB; idx += 256) {   // below the diagonal, left of this sub-block
    const int col = idx / WB, row = idx - col * WB;
    qstore(T - (size_t)h * WB * ldt + row + (size_t)col * ldt, qzero());
  }
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

size_t backtr_vt_slots(int n, int wbt) {
  size_t t = 0;
  for (int p = 0; p < n - 1; p += wbt) t += (size_t)(n - p - 1) * wbt;
  return t + 1;
}

size_t backtransform_batched_tasks(int nprob, const int* nh, int wbt, int nsplit) {
  if (nprob <= 0 || nh[0] < 2) return 1;
  // per block: G and Y split products (<= nsplit each), VT and the X update,
  // and two products per merge of T sub-blocks (wbt / 64 - 1 merges)
  return (size_t)nprob * cdiv(nh[0] - 1, wbt) * (2 * std::max(nsplit, 1) + 2 * (wbt / 64) + 2) + 1;
}

size_t backtransform_batched_sums(int nprob, const int* nh, int wbt) {
  if (nprob <= 0 || nh[0] < 2) return 1;
  return (size_t)nprob * cdiv(nh[0] - 1, wbt) * 2 + 1;
}

int backtransform_batched(int nprob, const int* nh, ZB zb, const int* zrow_d, const quat* M, int ld, const quat* e,
                          const double* Z, int ldz, quat* X, int ldx, const BacktrWork& w, QGemmTask* tasks,
                          SumTask* sums, Staging& st, hipStream_t s) {
  if (nprob <= 0) return ZQI_OK;
  const int L = nh[0];
  const int wbt = w.wbt;
  hipLaunchKernelGGL(k_phases, dim3(1, 1, nprob), dim3(PH_T), 0, s, L, e, w.phi, zb);
  hipLaunchKernelGGL(k_phase_apply, dim3(cdiv(L, 256), L, nprob), dim3(256), 0, s, L, w.phi, Z, ldz, X, ldx, zb,
                     zrow_d);
  if (hipGetLastError() != hipSuccess) return ZQI_ERR_HIP;
  if (L < 2) return ZQI_OK;
  auto at = [&](auto* p, int z) {
    using T = std::remove_pointer_t<decltype(p)>;
    return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(p) + (size_t)z * zb.zs);
  };
  const int nblk = cdiv(L - 1, wbt);
  // block b of problem z: reflectors p = b wbt .. p + wdt - 1 acting on rows r0 = p + 1 .. n - 1
  struct Blk { int p, wdt, r0, mr; size_t vt; };
  auto block = [&](int z, int b) {
    const int n = nh[z];
    Blk k;
    k.p = b * wbt;
    k.wdt = std::min(wbt, n - 1 - k.p);
    k.r0 = k.p + 1;
    k.mr = n - k.r0;
    size_t v = 0;
    for (int q = 0; q < b; ++q) v += (size_t)(n - q * wbt - 1) * wbt;
    k.vt = v;
    return k;
  };
  std::vector<QGemmTask> tt;
  std::vector<SumTask> ss;
  // ---- phase A (every block of every problem): G = V^* V, T, VT = V T^*
  const int gt0 = 0;
  int gmaxk = 0;
  for (int z = 0; z < nprob; ++z)
    for (int b = 0; b < nblk && b * wbt < nh[z] - 1; ++b) {
      const Blk k = block(z, b);
      const quat* V = at(M, z) + k.r0 + (size_t)k.p * ld;
      QOperand Vc{V, V, ld, ld, INT_MAX, true, -1};
      QOperand Vn{V, V, ld, ld, INT_MAX, false, -1};
      quat* part = at(w.part, z) + (size_t)b * w.nsplit * wbt * wbt;
      int kchunk = 0;
      const int nz = splitk_count(k.mr, w.nsplit, &kchunk);
      if ((size_t)(b + 1) * w.nsplit * wbt * wbt > w.part_slots) return ZQI_ERR_ARG;
      for (int q = 0; q < nz; ++q) {
        const int kb = q * kchunk, ke = std::min(k.mr, kb + kchunk);
        tt.push_back(QGemmTask{k.wdt, k.wdt, k.mr, kb, ke, wbt, 0, 1.0, 0.0, Vc, Vn, part + (size_t)q * wbt * wbt});
      }
      ss.push_back(SumTask{part, at(w.G, z) + (size_t)b * wbt * wbt, wbt * wbt, nz});
      gmaxk = std::max(gmaxk, k.wdt);
    }
  const int ngt = (int)tt.size() - gt0;
  const int ngs = (int)ss.size();
  // T sub-blocks merged level by level: at width w, the pair (q, q + w) of
  // each block gets T12 = -T1 (G12 T2), P = G12 T2 in the (then free)
  // split-K partial space; w1 / w2 reflectors in the two halves
  const int tsub = wbt / TF_WB;
  struct Level { int p0, np, t0, nt; };
  std::vector<Level> levels;
  for (int wl = TF_WB; wl < wbt; wl *= 2) {
    Level lv{};
    auto pairs = [&](bool second) {
      for (int z = 0; z < nprob; ++z)
        for (int b = 0; b < nblk && b * wbt < nh[z] - 1; ++b) {
          const Blk k = block(z, b);
          for (int q = 0; q + wl < wbt; q += 2 * wl) {
            const int w1 = std::min(wl, std::max(0, k.wdt - q)), w2 = std::min(wl, std::max(0, k.wdt - q - wl));
            if (w2 <= 0) continue;
            quat* T = at(w.T, z) + (size_t)b * wbt * wbt;
            quat* P = at(w.part, z) + (size_t)b * wbt * wbt + (size_t)(q / (2 * wl)) * wl * wl;
            if (!second) {
              const quat* G12 = at(w.G, z) + (size_t)b * wbt * wbt + q + (size_t)(q + wl) * wbt;
              const quat* T2 = T + (q + wl) + (size_t)(q + wl) * wbt;
              QOperand Ga{G12, G12, wbt, wbt, INT_MAX, false, INT_MIN / 2};
              QOperand Tb{T2, T2, wbt, wbt, INT_MAX, false, INT_MIN / 2};
              tt.push_back(QGemmTask{w1, w2, w2, 0, w2, wl, 0, 1.0, 0.0, Ga, Tb, P});
            } else {
              const quat* T1 = T + q + (size_t)q * wbt;
              QOperand Ta{T1, T1, wbt, wbt, INT_MAX, false, INT_MIN / 2};
              QOperand Pb{P, P, wl, wl, INT_MAX, false, INT_MIN / 2};
              tt.push_back(QGemmTask{w1, w2, w1, 0, w1, wbt, 0, -1.0, 0.0, Ta, Pb, T + q + (size_t)(q + wl) * wbt});
            }
          }
        }
    };
    lv.p0 = (int)tt.size();
    pairs(false);
    lv.np = (int)tt.size() - lv.p0;
    lv.t0 = (int)tt.size();
    pairs(true);
    lv.nt = (int)tt.size() - lv.t0;
    levels.push_back(lv);
  }
  const int vt0 = (int)tt.size();
  int vtmaxm = 0;
  for (int z = 0; z < nprob; ++z)
    for (int b = 0; b < nblk && b * wbt < nh[z] - 1; ++b) {
      const Blk k = block(z, b);
      if (k.vt + (size_t)k.mr * wbt > w.vt_slots) return ZQI_ERR_ARG;
      const quat* V = at(M, z) + k.r0 + (size_t)k.p * ld;
      const quat* T = at(w.T, z) + (size_t)b * wbt * wbt;
      QOperand Vn{V, V, ld, ld, INT_MAX, false, -1};
      QOperand Tc{T, T, wbt, wbt, INT_MAX, true, INT_MIN / 2};
      tt.push_back(QGemmTask{k.mr, k.wdt, k.wdt, 0, k.wdt, k.mr, 0, 1.0, 0.0, Vn, Tc, at(w.VT, z) + k.vt});
      vtmaxm = std::max(vtmaxm, k.mr);
    }
  const int nvt = (int)tt.size() - vt0;
  // ---- phase B, last block first: Y = VT^* X[r0:, :] (split), sums, X[r0:, :] -= V Y
  struct Step { int y0, ny, s0, ns, x0, nx, maxw, maxn, maxtot, maxmr; double fy, fx; };
  std::vector<Step> steps;
  for (int b = nblk - 1; b >= 0; --b) {
    Step q{};
    q.y0 = (int)tt.size();
    q.s0 = (int)ss.size();
    std::vector<QGemmTask> tx;
    for (int z = 0; z < nprob && b * wbt < nh[z] - 1; ++z) {
      const int n = nh[z];
      const Blk k = block(z, b);
      const quat* V = at(M, z) + k.r0 + (size_t)k.p * ld;
      const quat* VT = at(w.VT, z) + k.vt;
      quat* Xz = at(X, z);
      quat* part = at(w.part, z);
      quat* Y = at(w.Y, z);
      QOperand VTc{VT, VT, k.mr, k.mr, INT_MAX, true, INT_MIN / 2};
      QOperand Xo{Xz + k.r0, Xz + k.r0, ldx, ldx, INT_MAX, false, INT_MIN / 2};
      int kchunk = 0;
      const int nz = splitk_count(k.mr, w.nsplit, &kchunk);
      if ((size_t)nz * k.wdt * n > w.part_slots) return ZQI_ERR_ARG;
      for (int c = 0; c < nz; ++c) {
        const int kb = c * kchunk, ke = std::min(k.mr, kb + kchunk);
        tt.push_back(QGemmTask{k.wdt, n, k.mr, kb, ke, k.wdt, 0, 1.0, 0.0, VTc, Xo, part + (size_t)c * k.wdt * n});
      }
      ss.push_back(SumTask{part, Y, k.wdt * n, nz});
      QOperand Vn{V, V, ld, ld, INT_MAX, false, -1};
      QOperand Yo{Y, Y, k.wdt, k.wdt, INT_MAX, false, INT_MIN / 2};
      tx.push_back(QGemmTask{k.mr, n, k.wdt, 0, k.wdt, ldx, 0, -1.0, 1.0, Vn, Yo, Xz + k.r0});
      q.maxw = std::max(q.maxw, k.wdt);
      q.maxn = std::max(q.maxn, n);
      q.maxtot = std::max(q.maxtot, k.wdt * n);
      q.maxmr = std::max(q.maxmr, k.mr);
      q.fy += 32.0 * k.wdt * (double)n * k.mr;
      q.fx += 32.0 * k.mr * (double)n * k.wdt;
    }
    q.ny = (int)tt.size() - q.y0;
    q.ns = (int)ss.size() - q.s0;
    q.x0 = (int)tt.size();
    tt.insert(tt.end(), tx.begin(), tx.end());
    q.nx = (int)tx.size();
    steps.push_back(q);
  }
  if (tt.size() > backtransform_batched_tasks(nprob, nh, wbt, w.nsplit) ||
      ss.size() > backtransform_batched_sums(nprob, nh, wbt))
    return ZQI_ERR_ARG;
  int rc = stage_upload(st, tt.data(), tt.size() * sizeof(QGemmTask), tasks, s);
  if (!rc) rc = stage_upload(st, ss.data(), ss.size() * sizeof(SumTask), sums, s);
  if (rc) return rc;
  // phase A launches
  if ((rc = qgemm_tasks(tasks + gt0, ngt, gmaxk, gmaxk, true, false, false, s))) return rc;
  hipLaunchKernelGGL(k_sum_parts_b, dim3(cdiv(wbt * wbt, 256), 1, ngs), dim3(256), 0, s, sums);
  if (wbt < TF_WB || wbt % TF_WB || (tsub & (tsub - 1))) return ZQI_ERR_ARG;
  hipLaunchKernelGGL(k_tfactor, dim3(nblk * tsub, 1, nprob), dim3(256), 0, s, L, w.G, w.T, tsub, zb);
  for (size_t li = 0; li < levels.size(); ++li) {
    const int wl = TF_WB << li;
    if ((rc = qgemm_tasks(tasks + levels[li].p0, levels[li].np, wl, wl, false, false, false, s))) return rc;
    if ((rc = qgemm_tasks(tasks + levels[li].t0, levels[li].nt, wl, wl, false, false, false, s))) return rc;
  }
  if ((rc = qgemm_tasks(tasks + vt0, nvt, vtmaxm, wbt, false, true, false, s))) return rc;
  // phase B launches
  for (const Step& q : steps) {
    {
      ProfScope ps(PROF_BACKTR, s, q.fy);
      if ((rc = qgemm_tasks(tasks + q.y0, q.ny, q.maxw, q.maxn, true, false, false, s))) return rc;
    }
    hipLaunchKernelGGL(k_sum_parts_b, dim3(cdiv(q.maxtot, 256), 1, q.ns), dim3(256), 0, s, sums + q.s0);
    {
      ProfScope ps(PROF_BACKTR, s, q.fx);
      if ((rc = qgemm_tasks(tasks + q.x0, q.nx, q.maxmr, q.maxn, false, false, false, s))) return rc;
    }
  }
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
