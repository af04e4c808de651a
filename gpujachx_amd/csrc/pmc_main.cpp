// hj_pmc -- profiling driver without Python: a device-to-device copy of known
// size (the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters),
// then K level-1 steps of an N x N SVD (ALG 9, cycwor/cycwor) on a
// deterministic pseudo-random matrix in HBM, launched step by step; K = 0:
// one whole sweep as the library runs it (staged launches).  Batch mode: the
// same calibration copy, then one persistent launch of COUNT independent
// (H)SVDs (ALG 11): N uniform in {64, ..., 1024}, M in {N, 2N}, half with
// NPLUS = N, half NPLUS = N/2, entries of a fixed pseudo-random sequence
// times column scales 10^U(-3,3); prints the batch's algorithmic flops and
// bytes (from the sweeps each problem ran) and its time.
//
// Learned mode: the calibration copy, then a whole SVD (ALG 9, cycwor,
// learned skip + split Gram, at most 80 sweeps) of an N x N column-major
// float64 matrix read from a file, with the skip model's float32
// parameters read from another (gpujachx_amd/pmc_learned.py writes both);
// prints the run's sweeps and executed / skipped visits as JSON.
//
//   hj_pmc N K [COPY_MIB]
//   hj_pmc batch COUNT WAVES WG_PER_CU STRATEGY LONGEST_FIRST [COPY_MIB]
//   hj_pmc learned N G_FILE MODEL_FILE THRESHOLD MAX_SKIPS [COPY_MIB]
//   rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc -- gpujachx_amd/lib/hj_pmc 8192 16
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// copy_mib MiB read + written by one copy kernel (the counters' calibration)
static int calibration_copy(size_t copy_mib)
{
  const size_t cbytes = copy_mib << 20;
  void *ca = nullptr, *cb = nullptr;
  CHECK(hipMalloc(&ca, cbytes));
  CHECK(hipMalloc(&cb, cbytes));
  CHECK(hipMemset(ca, 0, cbytes));
  CHECK(hipMemcpy(cb, ca, cbytes, hipMemcpyDeviceToDevice));
  CHECK(hipDeviceSynchronize());
  CHECK(hipFree(ca));
  CHECK(hipFree(cb));
  return 0;
}

static unsigned long long g_lcg = 0x9e3779b97f4a7c15ull;
static double next_unit()  // uniform in [0, 1)
{
  g_lcg = g_lcg * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(g_lcg >> 11) * 0x1p-53;
}

// block-pair visits of one sweep of order n1 under the batch's strategy
// choice: the preferred family where the library can build it, else BrentL,
// else mmstep (steps x n1/2)
static unsigned long long visits_per_sweep(unsigned strategy, unsigned n1)
{
  const char *pref = (strategy == HJ_BATCH_STRAT_MMSTEP) ? "mmstep"
                     : (strategy == HJ_BATCH_STRAT_BRENTL) ? "BrentL" : "cycwor";
  const char *fams[3] = {pref, "BrentL", "mmstep"};
  for (const char *f : fams) {
    unsigned steps = 0u;
    if (!hj_strat_generate(f, n1, nullptr, &steps))
      return (unsigned long long)steps * (n1 / 2u);
  }
  return 0ull;
}

static int batch_main(int argc, char **argv)
{
  if (argc < 7) {
    (void)fprintf(stderr, "%s batch COUNT WAVES WG_PER_CU STRATEGY LONGEST_FIRST [COPY_MIB]\n", argv[0]);
    return 2;
  }
  const unsigned count = (unsigned)atoi(argv[2]);
  hj_batch_cfg cfg;
  hj_batch_default_cfg(&cfg);
  cfg.waves = (unsigned)atoi(argv[3]);
  cfg.wg_per_cu = (unsigned)atoi(argv[4]);
  cfg.strategy = (unsigned)atoi(argv[5]);
  cfg.longest_first = (unsigned)atoi(argv[6]);
  const size_t copy_mib = (argc > 7) ? (size_t)atol(argv[7]) : 1024u;
  if (configureGPU(0) < 0)
    return 1;
  if (calibration_copy(copy_mib))
    return 1;
  std::vector<hj_batch_desc> d(count);
  std::vector<void *> bufs;
  for (unsigned k = 0u; k < count; ++k) {
    const unsigned n = 64u * (1u + (unsigned)(next_unit() * 16.0));
    const unsigned m = n * (1u + (unsigned)(next_unit() * 2.0));
    const bool hyp = next_unit() < 0.5;
    std::vector<double> h((size_t)m * n);
    for (unsigned j = 0u; j < n; ++j) {
      const double sc = std::pow(10.0, 6.0 * next_unit() - 3.0);
      for (unsigned i = 0u; i < m; ++i)
        h[(size_t)j * m + i] = sc * (next_unit() - 0.5);
    }
    double *G = nullptr, *V = nullptr, *D = nullptr;
    CHECK(hipMalloc(&G, sizeof(double) * h.size()));
    CHECK(hipMalloc(&V, sizeof(double) * n * n));
    CHECK(hipMalloc(&D, sizeof(double) * n));
    CHECK(hipMemcpy(G, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    bufs.push_back(G);
    bufs.push_back(V);
    bufs.push_back(D);
    d[k] = {11u, m, n, hyp ? n / 2u : n, m, n, G, V, D};
  }
  int err = 0;
  void *plan = hj_batch_plan(d.data(), count, &cfg, &err);
  if (!plan) {
    (void)fprintf(stderr, "hj_batch_plan failed: %d\n", err);
    return 1;
  }
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  CHECK(hipEventRecord(e0, nullptr));
  const int rc = hj_batch_run(plan, nullptr);
  CHECK(hipEventRecord(e1, nullptr));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  std::vector<unsigned> sw(count);
  hj_batch_results(plan, sw.data(), nullptr, nullptr, nullptr, nullptr);
  double flops = 0.0, bytes = 0.0, visits = 0.0, sweeps = 0.0;
  for (unsigned k = 0u; k < count; ++k) {
    const double v = (double)visits_per_sweep(cfg.strategy, d[k].ncol / 16u) * sw[k];
    visits += v;
    sweeps += sw[k];
    flops += v * 2048.0 * (2.0 * d[k].nrow + d[k].ncol);
    bytes += v * 256.0 * (3.0 * d[k].nrow + 2.0 * d[k].ncol);
  }
  (void)printf("hj_pmc batch: count=%u waves=%u wg_per_cu=%u strategy=%u longest_first=%u rc=%d "
               "ms=%.3f mean_sweeps=%.3f visits=%.6e alg_flops=%.6e alg_bytes=%.6e "
               "tflops=%.3f alg_GBs=%.1f copy=%zu MiB\n",
               count, cfg.waves, cfg.wg_per_cu, cfg.strategy, cfg.longest_first, rc, ms,
               sweeps / count, visits, flops, bytes, flops / (ms * 1e-3) / 1e12,
               bytes / (ms * 1e-3) / 1e9, copy_mib);
  hj_batch_free(plan);
  for (void *b : bufs)
    CHECK(hipFree(b));
  return rc ? 1 : 0;
}

static int learned_main(int argc, char **argv)
{
  if (argc < 7) {
    (void)fprintf(stderr, "%s learned N G_FILE MODEL_FILE THRESHOLD MAX_SKIPS [COPY_MIB]\n", argv[0]);
    return 2;
  }
  const unsigned n = (unsigned)atoi(argv[2]);
  const float thr = (float)atof(argv[5]);
  const unsigned max_skips = (unsigned)atoi(argv[6]);
  const size_t copy_mib = (argc > 7) ? (size_t)atol(argv[7]) : 1024u;
  if (!n || (n % 64u)) {
    (void)fprintf(stderr, "N must be a positive multiple of 64\n");
    return 2;
  }
  std::vector<double> h((size_t)n * n);
  std::vector<float> model;
  {
    FILE *f = fopen(argv[3], "rb");
    if (!f || fread(h.data(), sizeof(double), h.size(), f) != h.size()) {
      (void)fprintf(stderr, "cannot read %s\n", argv[3]);
      return 1;
    }
    (void)fclose(f);
    f = fopen(argv[4], "rb");
    if (!f) {
      (void)fprintf(stderr, "cannot read %s\n", argv[4]);
      return 1;
    }
    float v;
    while (fread(&v, sizeof(float), 1u, f) == 1u)
      model.push_back(v);
    (void)fclose(f);
  }
  if (configureGPU(0) < 0)
    return 1;
  if (calibration_copy(copy_mib))
    return 1;
  hj_init_strats_builtin("cycwor", 32u, "cycwor", n / 16u);
  if (hj_set_skip_model(model.data(), (unsigned)model.size(), thr, max_skips)) {
    (void)fprintf(stderr, "skip model: wrong parameter count %zu\n", model.size());
    return 1;
  }
  hj_set_form(HJ_FORM_LEARNED_SKIP | HJ_FORM_SPLIT_GRAM);
  const unsigned cap = hj_set_max_sweeps(80u);
  double *dG = nullptr, *dD = nullptr;
  CHECK(hipMalloc(&dG, sizeof(double) * h.size()));
  CHECK(hipMalloc(&dD, sizeof(double) * n));
  CHECK(hipMemcpy(dG, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  unsigned sw = 0u;
  unsigned long long gs = 0ull, gb = 0ull;
  const int rc = hypJacL2_dev(9u, n, n, n, dG, n, nullptr, 0u, dD, &sw, &gs, &gb, nullptr, nullptr);
  CHECK(hipDeviceSynchronize());
  hj_set_form(0u);
  (void)hj_set_max_sweeps(cap);
  unsigned per[256];
  unsigned long long skipped = 0ull, visits = 0ull, splits = 0ull;
  hj_last_skips(per, &skipped, &visits);
  hj_last_splits(per, &splits);
  unsigned long long launches = 0ull;
  (void)hj_kernel_ms(5, &launches);
  (void)printf("{\"n\": %u, \"form\": \"learned+split\", \"ret\": %d, \"sweeps\": %u, "
               "\"visits_executed\": %llu, \"visits_skipped\": %llu, \"split_grams\": %llu, "
               "\"staged\": %d, \"staged_launches\": %llu, \"copy_mib\": %zu}\n",
               n, rc, sw, visits, skipped, splits, hj_last_staged(), launches, copy_mib);
  CHECK(hipFree(dG));
  CHECK(hipFree(dD));
  return rc ? 1 : 0;
}

int main(int argc, char **argv)
{
  if (argc >= 2 && !strcmp(argv[1], "batch"))
    return batch_main(argc, argv);
  if (argc >= 2 && !strcmp(argv[1], "learned"))
    return learned_main(argc, argv);
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

  if (calibration_copy(copy_mib))
    return 1;

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
