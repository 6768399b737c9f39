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

namespace zq {

// phi_0 = 1, phi_{c+1} = r_c phi_c with r_c = e_c / |e_c| (1 where e_c = 0):
// a prefix product of unit quaternions, chunked over one workgroup.
constexpr int PH_T = 256;
__global__ void __launch_bounds__(PH_T) k_phases(int n, const quat* __restrict__ e, quat* __restrict__ phi) {
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

// X[i, j] = phi_i Z[i, j]
__global__ void __launch_bounds__(256) k_phase_apply(int n, const quat* __restrict__ phi, const double* __restrict__ Z,
                                                     int ldz, quat* __restrict__ X, int ldx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
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

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

int backtransform(int n, int nb, const quat* M, int ld, const quat* Tf, const quat* e, const double* Z, int ldz,
                  quat* X, int ldx, const BacktrWork& w, hipStream_t s) {
  if (n <= 0) return ZQI_OK;
  hipLaunchKernelGGL(k_phases, dim3(1), dim3(PH_T), 0, s, n, e, w.phi);
  hipLaunchKernelGGL(k_phase_apply, dim3(cdiv(n, 256), n), dim3(256), 0, s, n, w.phi, Z, ldz, X, ldx);
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

}  // namespace zq
