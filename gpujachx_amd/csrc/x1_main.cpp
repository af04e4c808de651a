// x1_main.cpp -- command-line test driver with the reference's surface
// (1/main.cpp is 1/main.cu:10-160):
//
//   x1 DEV SDY SNP0 SNP1 ALG H5F H5G [H5R]
//
// reads /H5G/IDADIM = {LDA, M, N, NPLUS} (int32[4]) and /H5G/G (float64,
// Fortran LDA x N) from H5F, runs hypJacL2 on device DEV with the strategy
// tables SNP0 (order 32) and SNP1 (order ceil(H5G/16)) from the plugin SDY,
// prints the per-sweep lines and the totals, and, if H5R is given, writes a
// new group H5G there with IDADIM, ISTATS, TIMING, G, V (if ALG & 2) and D.
// The reference's build options USE_QR / USE_DRMAC are chosen at run time by
// the environment variable HJ_FACT = qr | drmac (default: Cholesky).
#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hypjac.h"

#define H5CALL(call)                                                                 \
  do {                                                                               \
    const herr_t e_ = static_cast<herr_t>(call);                                     \
    if (e_ < 0) {                                                                    \
      (void)fprintf(stderr, "HDF5 error %d @ %s(%d)!\n", (int)e_, __FILE__, __LINE__); \
      exit(EXIT_FAILURE);                                                            \
    }                                                                                \
  } while (0)

#define X1_DIE(msg)                                                              \
  do {                                                                           \
    (void)fprintf(stderr, "[ERROR] %s(%d): %s\n", __FILE__, __LINE__, (msg));    \
    exit(EXIT_FAILURE);                                                          \
  } while (0)

int main(int argc, char *argv[])
{
  if ((8 > argc) || (9 < argc)) {
    (void)fprintf(stderr, "%s DEV SDY SNP0 SNP1 ALG H5F H5G [H5R]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const char *const ca_exe = argv[0];
  const char *const ca_dev = argv[1];
  const char *const ca_sdy = argv[2];
  const char *const ca_snp0 = argv[3];
  const char *const ca_snp1 = argv[4];
  const char *const ca_alg = argv[5];
  const char *const ca_h5f = argv[6];
  const char *const ca_h5g = argv[7];
  const char *const ca_h5r = (9 == argc) ? argv[8] : nullptr;

  const unsigned n0 = HYPJACL1_NCOLB << 1u;
  const unsigned n1 = (static_cast<unsigned>(atoi(ca_h5g)) + HYPJACL1_NCOLB - 1u) / HYPJACL1_NCOLB;
  init_strats(ca_sdy, ca_snp0, n0, ca_snp1, n1);

  if (const char *f = getenv("HJ_FACT")) {
    if (!strcmp(f, "qr"))
      (void)hj_set_factorization(HJ_FACT_QR);
    else if (!strcmp(f, "drmac"))
      (void)hj_set_factorization(HJ_FACT_QR_DRMAC);
    else if (strcmp(f, "cholesky"))
      X1_DIE("HJ_FACT must be qr, drmac or cholesky");
  }

  const int dev = atoi(ca_dev);
  const int dcc = configureGPU(dev);
  if (dcc < 30) {
    char msg[128];
    (void)snprintf(msg, sizeof(msg), "Device %d has CC %d < 30", dev, dcc);
    X1_DIE(msg);
  }
  const unsigned routine = static_cast<unsigned>(atoi(ca_alg));
  if (!fexist(ca_h5f)) {
    char msg[1024];
    (void)snprintf(msg, sizeof(msg), "non-existent H5F(%s)", ca_h5f);
    X1_DIE(msg);
  }

  H5CALL(H5open());
  hid_t fid = -1, gid = -1;
  H5CALL(fid = H5Fopen(ca_h5f, H5F_ACC_RDONLY, H5P_DEFAULT));
  H5CALL(gid = H5Gopen2(fid, ca_h5g, H5P_DEFAULT));
  int idadim[HYPJAC_IDADIM_SIZE] = {0, 0, 0, 0};
  H5CALL(H5LTread_dataset_int(gid, "IDADIM", idadim));
  unsigned ldhG = static_cast<unsigned>(idadim[0]);
  const unsigned nrow = static_cast<unsigned>(idadim[1]);
  const unsigned ncol = static_cast<unsigned>(idadim[2]);
  const unsigned nplus = static_cast<unsigned>(idadim[3]);
  if (ldhG < nrow) X1_DIE("IDADIM: LDA < M");
  std::vector<double> hG((size_t)ldhG * ncol, 0.0);
  H5CALL(H5LTread_dataset_double(gid, "G", hG.data()));
  H5CALL(H5Gclose(gid));
  H5CALL(H5Fclose(fid));

  const bool full = (routine & HYPJAC_FULL_SVD) != 0u;
  std::vector<double> hV;
  unsigned ldhV = 0u;
  if (full) {
    hV.assign((size_t)ldhG * ncol, 0.0);
    ldhV = ldhG;
  }
  std::vector<double> hD(ncol, 0.0);
  unsigned glbSwp = 0u;
  unsigned long long glb_s = 0ull, glb_b = 0ull;
  double timing[4] = {-0.0, -0.0, -0.0, -0.0};
  const int ret = hypJacL2(routine, nrow, ncol, nplus, hG.data(), ldhG, full ? hV.data() : nullptr,
                           ldhV, hD.data(), &glbSwp, &glb_s, &glb_b, timing);
  if (ret)
    (void)fprintf(stderr, "%s: error %d\n", ca_exe, ret);
  else {
    (void)fprintf(stdout, "GLB_ROT_S(%20llu), GLB_ROT_B(%20llu)\n", glb_s, glb_b);
    (void)fflush(stdout);
    (void)fprintf(stdout, "%#12.6f s %2u sweeps\n", *timing, glbSwp);
    (void)fflush(stdout);
  }

  if (ca_h5r) {
    H5CALL(fid = fexist(ca_h5r) ? H5Fopen(ca_h5r, H5F_ACC_RDWR, H5P_DEFAULT)
                                : H5Fcreate(ca_h5r, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT));
    H5CALL(gid = H5Gcreate2(fid, ca_h5g, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    idadim[0] = static_cast<int>(ldhG);
    hsize_t dims[2] = {HYPJAC_IDADIM_SIZE, 1};
    H5CALL(H5LTmake_dataset_int(gid, "IDADIM", 1, dims, idadim));
    int istats[4] = {(ret < 0) ? ret : static_cast<int>(glbSwp), static_cast<int>(STRAT0),
                     static_cast<int>(STRAT1), ret};
    dims[0] = 4;
    H5CALL(H5LTmake_dataset_int(gid, "ISTATS", 1, dims, istats));
    H5CALL(H5LTmake_dataset_double(gid, "TIMING", 1, dims, timing));
    dims[0] = ncol;
    dims[1] = ldhG;
    H5CALL(H5LTmake_dataset_double(gid, "G", 2, dims, hG.data()));
    if (full) {
      dims[1] = ldhV;
      H5CALL(H5LTmake_dataset_double(gid, "V", 2, dims, hV.data()));
    }
    dims[0] = ncol;
    dims[1] = 1;
    H5CALL(H5LTmake_dataset_double(gid, "D", 1, dims, hD.data()));
    H5CALL(H5Gclose(gid));
    H5CALL(H5Fclose(fid));
  }
  H5CALL(H5close());
  return ret;
}
