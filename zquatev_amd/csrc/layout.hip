// Data movement at the two ends of the solve.
//
// In:  the caller's D (nld2 x n2 complex128, column-major) holds [A; B] in its
//      left half (zquatev.h:46, zquatev.cc:47-54).  It becomes the quaternion
//      matrix M_ij = (A_ij, B_ij) (lower triangle used).
// Out: X_ij = (U_ij, V_ij) is written back as [[U, -conj V], [V, conj U]]
//      (zquatev.cc:86-98), columns i and i + n forming the Kramers pair.
#include "zq_internal.h"

namespace zq {

__global__ void __launch_bounds__(256) k_load_left(int n, const double* __restrict__ D, int ld2, quat* __restrict__ M,
                                                   int ld) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || i < j) return;
  const double2* col = reinterpret_cast<const double2*>(D) + (size_t)j * ld2;
  const double2 a = col[i];
  const double2 b = col[n + i];
  qstore(M + i + (size_t)j * ld, quat{a.x, a.y, b.x, b.y});
}

__global__ void __launch_bounds__(256) k_store_vec(int n, const quat* __restrict__ X, int ldx, double* __restrict__ D,
                                                   int ld2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const quat x = qload(X + i + (size_t)j * ldx);
  double2* left = reinterpret_cast<double2*>(D) + (size_t)j * ld2;
  double2* right = reinterpret_cast<double2*>(D) + (size_t)(j + n) * ld2;
  left[i] = make_double2(x.ar, x.ai);          // U
  left[n + i] = make_double2(x.br, x.bi);      // V
  right[i] = make_double2(-x.br, x.bi);        // -conj V
  right[n + i] = make_double2(x.ar, -x.ai);    // conj U
}

__global__ void k_offdiag_abs(int n, const quat* __restrict__ e, double* __restrict__ ea) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n - 1) ea[i] = sqrt(qnorm2(qload(e + i)));
}

int load_left_half(int n, const double* D, int ld2, quat* M, int ld, hipStream_t s) {
  hipLaunchKernelGGL(k_load_left, dim3((n + 255) / 256, n), dim3(256), 0, s, n, D, ld2, M, ld);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

int store_vectors(int n, const quat* X, int ldx, double* D, int ld2, hipStream_t s) {
  hipLaunchKernelGGL(k_store_vec, dim3((n + 255) / 256, n), dim3(256), 0, s, n, X, ldx, D, ld2);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

int offdiag_abs(int n, const quat* e, double* ea, hipStream_t s) {
  if (n < 2) return ZQI_OK;
  hipLaunchKernelGGL(k_offdiag_abs, dim3((n + 254) / 256), dim3(256), 0, s, n, e, ea);
  return hipGetLastError() == hipSuccess ? ZQI_OK : ZQI_ERR_HIP;
}

}  // namespace zq
