// Host-code checks run under AddressSanitizer and UndefinedBehaviorSanitizer
// (host code only: the library is built with -fsanitize after -Xarch_host,
// see zquatev_amd/csrc/Makefile, target `sanitize`).  Everything here runs
// without a GPU: argument validation of every entry point, error strings,
// the batch workspace planner (group building, batch layout and the split of
// heavy-tailed batches into runs) on many random size lists, and the
// failure paths of context and server creation when no device is present.
#include "zq.h"
#include "zquatev.h"

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

static int fails = 0;
#define CHECK(cond)                                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      ++fails;                                                           \
    }                                                                    \
  } while (0)

int main() {
  // argument validation (nothing reaches a device)
  std::vector<std::complex<double>> D(16);
  std::vector<double> eig(2);
  CHECK(zq_zquatev(3, reinterpret_cast<double*>(D.data()), 4, eig.data()) == -1);
  CHECK(zq_zquatev(4, nullptr, 4, eig.data()) == -2);
  CHECK(zq_zquatev(4, reinterpret_cast<double*>(D.data()), 2, eig.data()) == -3);
  CHECK(zq_zquatev(4, reinterpret_cast<double*>(D.data()), 4, nullptr) == -4);
  CHECK(zq_zquatev(0, nullptr, 1, nullptr) == 0);
  CHECK(ts::zquatev(5, D.data(), 4, eig.data()) == -1);
  CHECK(zq_zquatev_device(nullptr, 4, nullptr, 4, nullptr, 'V') == ZQ_ERR_ARG);
  CHECK(zq_eigvals_noalloc(2050, reinterpret_cast<double*>(D.data()), 2050, eig.data(), nullptr, nullptr) == -5);
  CHECK(zq_eigvals_noalloc(7, reinterpret_cast<double*>(D.data()), 8, eig.data(), nullptr, nullptr) == -1);
  const int one = 4;
  CHECK(zq_zquatev_batched_device(nullptr, 1, &one, nullptr, &one, nullptr, 'V', nullptr) == ZQ_ERR_ARG);
  CHECK(zq_last_info(nullptr) == ZQ_ERR_ARG);
  CHECK(zq_set_tile_shape(nullptr, 0, 0) == ZQ_ERR_ARG);
  CHECK(zq_server_set_budget(nullptr, 1.0) == ZQ_ERR_ARG);
  for (int code : {0, -1, -2, -3, -4, -5, 3, ZQ_ERR_HIP, ZQ_ERR_NOMEM, ZQ_ERR_ARG, ZQ_ERR_GRAPH, ZQ_ERR_NONFINITE,
                   ZQ_ERR_BUDGET})
    CHECK(zq_error_string(code) != nullptr && zq_error_string(code)[0] != 0);
  // no device in this process: creation fails cleanly
  zq_context* ctx = reinterpret_cast<zq_context*>(1);
  const int rc = zq_create(&ctx, 0);
  CHECK(rc != 0 ? ctx == nullptr : true);
  if (rc == 0) zq_destroy(ctx);
  zq_server* srv = nullptr;
  const int rs = zq_server_create(&srv, 0, 2, 8, 0.0);
  if (rs == 0) zq_server_destroy(srv);
  else CHECK(srv == nullptr);
  // the batch workspace planner on random batches, with and without a budget
  std::mt19937 rng(12345);
  for (int trial = 0; trial < 300; ++trial) {
    const int count = 1 + (int)(rng() % 200);
    std::vector<int> n2(count);
    for (int& v : n2) {
      const int kind = (int)(rng() % 10);
      v = kind == 0 ? 0 : 2 * (1 + (int)(rng() % (kind == 1 ? 4096 : 600)));
    }
    const int jobz = (trial & 1) ? 'V' : 'N';
    size_t whole = 0;
    std::vector<int> run(count);
    const int k = zq_batch_workspace(nullptr, count, n2.data(), jobz, 0, &whole, run.data());
    CHECK(k >= 0);
    const size_t budget = whole / (2 + trial % 5) + 1;
    size_t need = 0;
    const int runs = zq_batch_workspace(nullptr, count, n2.data(), jobz, budget, &need, run.data());
    CHECK(runs >= k);
    std::vector<int> members;
    for (int r = 0; r < runs; ++r) {
      members.clear();
      for (int i = 0; i < count; ++i)
        if (run[i] == r) members.push_back(n2[i]);
      CHECK(!members.empty());
      size_t got = 0;
      CHECK(zq_batch_workspace(nullptr, (int)members.size(), members.data(), jobz, 0, &got, nullptr) >= 1);
      if (members.size() > 1) CHECK(got <= budget);
    }
    for (int i = 0; i < count; ++i) CHECK((n2[i] == 0) == (run[i] == -1));
  }
  const int bad = 63;
  CHECK(zq_batch_workspace(nullptr, 1, &bad, 'V', 0, nullptr, nullptr) == -3);
  CHECK(zq_batch_workspace(nullptr, -1, nullptr, 'V', 0, nullptr, nullptr) == -2);
  std::printf("host checks: %d failure(s)\n", fails);
  return fails ? 1 : 0;
}
