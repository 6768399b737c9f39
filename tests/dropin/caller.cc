// A caller of the drop-in entry point, written the way the reference's own
// harness calls it (test.cc:97-101): it includes only the drop-in header
// (the declaration of zquatev.h:54), links libzquatev_amd, and calls
// ts::zquatev on a host matrix.  The test suite runs it as a child process.
//
// Usage: zquatev_caller IN OUT
//   IN:  int32 n2, int32 ld2, then ld2 * n2 complex128, column-major (the
//        left half [A; B] is what the solver reads)
//   OUT: int32 info, then n2/2 doubles (eigenvalues), then the ld2 * n2
//        complex128 of D as ts::zquatev left it
#include "zquatev.h"

#include <complex>
#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int hdr[2] = {0, 0};
  if (std::fread(hdr, sizeof(int), 2, in) != 2) return 2;
  const int n2 = hdr[0], ld2 = hdr[1];
  if (n2 < 0 || ld2 < 1) return 2;
  std::vector<std::complex<double>> D((size_t)ld2 * (n2 > 0 ? n2 : 1));
  if (std::fread(D.data(), sizeof(std::complex<double>), (size_t)ld2 * n2, in) != (size_t)ld2 * n2) return 2;
  std::fclose(in);
  std::vector<double> eig(n2 / 2 > 0 ? n2 / 2 : 1);
  const int info = ts::zquatev(n2, D.data(), ld2, eig.data());
  std::FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  std::fwrite(&info, sizeof(int), 1, out);
  std::fwrite(eig.data(), sizeof(double), (size_t)(n2 / 2), out);
  std::fwrite(D.data(), sizeof(std::complex<double>), (size_t)ld2 * n2, out);
  std::fclose(out);
  return 0;
}
