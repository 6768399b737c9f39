// zquatev.h -- drop-in C++ entry point, the same declaration as the
// reference's public API (zquatev.h:54):
//
//   namespace ts { int zquatev(const int n2, std::complex<double>* const D,
//                              const int nld2, double* const eig); }
//
// Diagonalises H = [[A, -conj B], [B, conj A]] (A Hermitian, B complex
// antisymmetric) given the left half [A; B] of D (nld2 x n2, column-major).
// On return eig[0:n2/2] holds the distinct eigenvalues in ascending order and D
// holds [[U, -conj V], [V, conj U]]: columns i and i + n2/2 are the Kramers pair
// of eig[i].  Returns 0 on success, -1 for an odd or negative n2, -2 for a null
// D, -3 for nld2 < n2, -4 for a null eig, and the codes of zq.h for device
// errors.  Linking libzquatev_amd provides this symbol; the work runs on HIP
// device 0 through a process-wide context.
#ifndef ZQUATEV_AMD_ZQUATEV_H
#define ZQUATEV_AMD_ZQUATEV_H

#include <complex>

namespace ts {
int zquatev(const int n2, std::complex<double>* const D, const int nld2, double* const eig);
}

#endif
