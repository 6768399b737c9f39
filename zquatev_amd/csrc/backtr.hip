// Eigenvector back-transform X = Q Phi Z.
//
// Replaces the reference's explicit Q accumulation (blocked.cc:477-544,
// unblocked.cc:75-80, :121-126) followed by U = Q0 Z, V = Q1 Z
// (zquatev.cc:86-90).  Q is never formed: the panel reflectors, kept in place
// below the subdiagonal of M with their compact-WY factors T
// (H_p ... H_{p+nb-1} = I - V T V^*), are applied to Phi Z panel by panel in
// reverse order, each as three FP64 MFMA products:
//   Y = V^* X[r0:, :]   (split over the long K dimension),
//   Y2 = T Y,
//   X[r0:, :] -= V Y2.
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

// Y[b, col] = sum_z Ypart[z][b, col]   (w x nn, ld w)
__global__ void __launch_bounds__(256) k_sum_parts(int total, int nz, const quat* __restrict__ Yp, quat* __restrict__ Y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  quat s = qzero();
  for (int z = 0; z < nz; ++z) s = qadd(s, qload(Yp + (size_t)z * total + i));
  qstore(Y + i, s);
}

// the same sum for a list of tasks (task blockIdx.z)
__global__ void __launch_bounds__(256) k_sum_parts_b(const SumTask* __restrict__ tasks) {
  const SumTask t = tasks[blockIdx.z];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.total) return;
  quat s = qzero();
  for (int z = 0; z < t.nz; ++z) s = qadd(s, qload(t.parts + (size_t)z * t.total + i));
  qstore(t.out + i, s);
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

int backtransform(int n, int nb, const quat* M, int ld, const quat* Tf, const quat* e, const double* Z, int ldz,
                  quat* X, int ldx, const BacktrWork& w, hipStream_t s) {
  if (n <= 0) return ZQI_OK;
  hipLaunchKernelGGL(k_phases, dim3(1), dim3(PH_T), 0, s, n, e, w.phi, ZB{nullptr, 0});
  hipLaunchKernelGGL(k_phase_apply, dim3(cdiv(n, 256), n), dim3(256), 0, s, n, w.phi, Z, ldz, X, ldx,
                     ZB{nullptr, 0}, (const int*)nullptr);
  if (hipGetLastError() != hipSuccess) return ZQI_ERR_HIP;
  if (n == 1) return ZQI_OK;
  const int npanel = cdiv(n - 1, nb);
  for (int panel = npanel - 1; panel >= 0; --panel) {
    const int p = panel * nb;
    const int wdt = (nb < n - 1 - p) ? nb : (n - 1 - p);
    const int r0 = p + 1;
    const int mr = n - r0;
    // V: rows r0.., cols p..p+wdt-1 of M; u_{p+t} lives in rows >= p+t+1, i.e. local row s >= t
    QOperand Vc{M + r0 + (size_t)p * ld, M + r0 + (size_t)p * ld, ld, ld, INT_MAX, true, -1};
    const int nz = splitk_count(mr, w.nsplit, nullptr);
    if ((size_t)nz * wdt * n > w.ypart_slots) return ZQI_ERR_ARG;
    int rc;
    {
      ProfScope ps(PROF_BACKTR, s, 32.0 * (double)mr * wdt * n);
      rc = qgemm_cn_splitk(wdt, n, mr, Vc, X + r0, ldx, w.Ypart, w.nsplit, s);
    }
    if (rc) return rc;
    const int total = wdt * n;
    hipLaunchKernelGGL(k_sum_parts, dim3(cdiv(total, 256)), dim3(256), 0, s, total, nz, w.Ypart, w.Y);
    const quat* T = Tf + (size_t)panel * nb * nb;
    QOperand To{T, T, nb, nb, INT_MAX, false, INT_MIN / 2};
    QOperand Yo{w.Y, w.Y, wdt, wdt, INT_MAX, false, INT_MIN / 2};
    rc = qgemm(wdt, n, wdt, 1.0, To, Yo, 0.0, w.Y2, wdt, false, s);
    if (rc) return rc;
    QOperand Vn{M + r0 + (size_t)p * ld, M + r0 + (size_t)p * ld, ld, ld, INT_MAX, false, -1};
    QOperand Y2o{w.Y2, w.Y2, wdt, wdt, INT_MAX, false, INT_MIN / 2};
    {
      ProfScope ps(PROF_BACKTR, s, 32.0 * (double)mr * wdt * n);
      rc = qgemm(mr, n, wdt, -1.0, Vn, Y2o, 1.0, X + r0, ldx, false, s);
    }
    if (rc) return rc;
  }
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

size_t backtransform_batched_tasks(int nprob, const int* nh, int nb, int nsplit) {
  if (nprob <= 0 || nh[0] < 2) return 1;
  return (size_t)nprob * cdiv(nh[0] - 1, nb) * (std::max(nsplit, 1) + 2) + 1;
}

int backtransform_batched(int nprob, const int* nh, ZB zb, const int* zrow_h, const int* zrow_d, int nb,
                          const quat* M, int ld, const quat* Tf, const quat* e, const double* Z, int ldz, quat* X,
                          int ldx, const BacktrWork& w, QGemmTask* tasks, SumTask* sums, Staging& st, hipStream_t s) {
  (void)zrow_h;
  if (nprob <= 0) return ZQI_OK;
  const int L = nh[0];
  hipLaunchKernelGGL(k_phases, dim3(1, 1, nprob), dim3(PH_T), 0, s, L, e, w.phi, zb);
  hipLaunchKernelGGL(k_phase_apply, dim3(cdiv(L, 256), L, nprob), dim3(256), 0, s, L, w.phi, Z, ldz, X, ldx, zb,
                     zrow_d);
  if (hipGetLastError() != hipSuccess) return ZQI_ERR_HIP;
  if (L < 2) return ZQI_OK;
  auto at = [&](auto* p, int z) {
    using T = std::remove_pointer_t<decltype(p)>;
    return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(p) + (size_t)z * zb.zs);
  };
  // per panel (last first): the split products Y = V^* X, their sums, Y2 = T Y, X -= V Y2,
  // each problem's tasks exactly those backtransform() launches for it
  const int npan = cdiv(L - 1, nb);
  struct Step { int y0, ny, s0, ns, t0, nt, x0, nx, maxm_y, maxn, maxtot, maxmr; };
  std::vector<QGemmTask> tt;
  std::vector<SumTask> ss;
  std::vector<Step> steps;
  for (int panel = npan - 1; panel >= 0; --panel) {
    const int p = panel * nb, r0 = p + 1;
    Step st_{};
    st_.y0 = (int)tt.size();
    st_.s0 = (int)ss.size();
    std::vector<QGemmTask> ty, tx;
    for (int z = 0; z < nprob && nh[z] - 1 > p; ++z) {
      const int n = nh[z];
      const int wdt = std::min(nb, n - 1 - p);
      const int mr = n - r0;
      const quat* Mz = at(M, z);
      quat* Xz = at(X, z);
      quat* Yp = at(w.Ypart, z);
      quat* Yz = at(w.Y, z);
      quat* Y2z = at(w.Y2, z);
      const quat* Tz = at(Tf, z) + (size_t)panel * nb * nb;
      QOperand Vc{Mz + r0 + (size_t)p * ld, Mz + r0 + (size_t)p * ld, ld, ld, INT_MAX, true, -1};
      QOperand Bo{Xz + r0, Xz + r0, ldx, ldx, INT_MAX, false, INT_MIN / 2};
      int kchunk = 0;
      const int nz = splitk_count(mr, w.nsplit, &kchunk);
      if ((size_t)nz * wdt * n > w.ypart_slots) return ZQI_ERR_ARG;
      for (int q = 0; q < nz; ++q) {
        const int kb = q * kchunk, ke = std::min(mr, kb + kchunk);
        tt.push_back(QGemmTask{wdt, n, mr, kb, ke, wdt, 0, 1.0, 0.0, Vc, Bo, Yp + (size_t)q * wdt * n});
      }
      ss.push_back(SumTask{Yp, Yz, wdt * n, nz});
      QOperand To{Tz, Tz, nb, nb, INT_MAX, false, INT_MIN / 2};
      QOperand Yo{Yz, Yz, wdt, wdt, INT_MAX, false, INT_MIN / 2};
      ty.push_back(QGemmTask{wdt, n, wdt, 0, wdt, wdt, 0, 1.0, 0.0, To, Yo, Y2z});
      QOperand Vn{Mz + r0 + (size_t)p * ld, Mz + r0 + (size_t)p * ld, ld, ld, INT_MAX, false, -1};
      QOperand Y2o{Y2z, Y2z, wdt, wdt, INT_MAX, false, INT_MIN / 2};
      tx.push_back(QGemmTask{mr, n, wdt, 0, wdt, ldx, 0, -1.0, 1.0, Vn, Y2o, Xz + r0});
      st_.maxm_y = std::max(st_.maxm_y, wdt);
      st_.maxn = std::max(st_.maxn, n);
      st_.maxtot = std::max(st_.maxtot, wdt * n);
      st_.maxmr = std::max(st_.maxmr, mr);
    }
    st_.ny = (int)tt.size() - st_.y0;
    st_.ns = (int)ss.size() - st_.s0;
    st_.t0 = (int)tt.size();
    tt.insert(tt.end(), ty.begin(), ty.end());
    st_.nt = (int)ty.size();
    st_.x0 = (int)tt.size();
    tt.insert(tt.end(), tx.begin(), tx.end());
    st_.nx = (int)tx.size();
    steps.push_back(st_);
  }
  if (tt.size() > backtransform_batched_tasks(nprob, nh, nb, w.nsplit)) return ZQI_ERR_ARG;
  int rc = stage_upload(st, tt.data(), tt.size() * sizeof(QGemmTask), tasks, s);
  if (!rc) rc = stage_upload(st, ss.data(), ss.size() * sizeof(SumTask), sums, s);
  if (rc) return rc;
  for (const Step& q : steps) {
    {
      ProfScope ps(PROF_BACKTR, s, 0.0);
      if ((rc = qgemm_tasks(tasks + q.y0, q.ny, q.maxm_y, q.maxn, s))) return rc;
    }
    hipLaunchKernelGGL(k_sum_parts_b, dim3(cdiv(q.maxtot, 256), 1, q.ns), dim3(256), 0, s, sums + q.s0);
    if ((rc = qgemm_tasks(tasks + q.t0, q.nt, q.maxm_y, q.maxn, s))) return rc;
    {
      ProfScope ps(PROF_BACKTR, s, 0.0);
      if ((rc = qgemm_tasks(tasks + q.x0, q.nx, q.maxmr, q.maxn, s))) return rc;
    }
  }
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
