// hj_pmc -- profiling driver without Python: a device-to-device copy of known
// size (the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters),
// then K level-1 steps of an N x N SVD (ALG 9, cycwor/cycwor) on a
// deterministic pseudo-random matrix in HBM, launched step by step; K = 0:
// one whole sweep as the library runs it (staged launches).
//
//   hj_pmc N K [COPY_MIB]
//   rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc -- gpujachx_amd/lib/hj_pmc 8192 16
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/hypjac.h"

#define CHECK(x)                                                                  \
  do {                                                                            \
    const hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                       \
      (void)fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

int main(int argc, char **argv)
{
  if (argc < 3) {
    (void)fprintf(stderr, "%s N K [COPY_MIB]\n", argv[0]);
    return 2;
  }
  const unsigned n = (unsigned)atoi(argv[1]);
  const long steps = atol(argv[2]);
  const size_t copy_mib = (argc > 3) ? (size_t)atol(argv[3]) : 1024u;
  if (!n || (n % 64u) || steps < 0) {
    (void)fprintf(stderr, "N must be a positive multiple of 64, K >= 0\n");
    return 2;
  }
  if (configureGPU(0) < 0)
    return 1;
  hj_init_strats_builtin("cycwor", 32u, "cycwor", n / 16u);

  // calibration: copy_mib MiB read + written by one copy kernel
  const size_t cbytes = copy_mib << 20;
  void *ca = nullptr, *cb = nullptr;
  CHECK(hipMalloc(&ca, cbytes));
  CHECK(hipMalloc(&cb, cbytes));
  CHECK(hipMemset(ca, 0, cbytes));
  CHECK(hipMemcpy(cb, ca, cbytes, hipMemcpyDeviceToDevice));
  CHECK(hipDeviceSynchronize());
  CHECK(hipFree(ca));
  CHECK(hipFree(cb));

  // G: columns with graded norms, entries from a fixed linear congruential
  // sequence (any full-rank matrix does for counting bytes)
  std::vector<double> h((size_t)n * n);
  unsigned long long st = 0x9e3779b97f4a7c15ull;
  for (size_t j = 0; j < n; ++j) {
    const double sc = std::pow(10.0, -6.0 * (double)j / n);
    for (size_t i = 0; i < n; ++i) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      h[j * n + i] = sc * ((double)(st >> 11) * 0x1p-53 - 0.5);
    }
  }
  double *dG = nullptr, *dD = nullptr;
  CHECK(hipMalloc(&dG, sizeof(double) * h.size()));
  CHECK(hipMalloc(&dD, sizeof(double) * n));
  CHECK(hipMemcpy(dG, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  unsigned sw = 0u;
  unsigned long long gs = 0ull, gb = 0ull;
  const unsigned cap = steps ? 0u : hj_set_max_sweeps(1u);
  hj_set_step_limit(steps);
  const int rc = hypJacL2_dev(9u, n, n, n, dG, n, nullptr, 0u, dD, &sw, &gs, &gb, nullptr, nullptr);
  hj_set_step_limit(0);
  if (!steps) hj_set_max_sweeps(cap);
  CHECK(hipDeviceSynchronize());
  (void)printf("hj_pmc: n=%u steps=%ld (0: one sweep, staged %d) rc=%d copy=%zu MiB\n", n, steps,
               hj_last_staged(), rc, copy_mib);
  CHECK(hipFree(dG));
  CHECK(hipFree(dD));
  return rc ? 1 : 0;
}
