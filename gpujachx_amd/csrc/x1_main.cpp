// x1_main.cpp -- the x1 driver: one (hyperbolic) SVD of an HDF5 input on one GPU.
//
//   x1 DEV SDY SNP0 SNP1 ALG H5F H5G [H5R]
//
// What it keeps of the reference's driver is its surface, not its code:
//  * the usage line and the argument meaning (1/main.cu:12-15): device,
//    strategy plugin, level-0 and level-1 family names, ALG, input file,
//    group, optional result file; the level-1 order is ceil(H5G / 16), so
//    the group must be named by the column count (1/main.cu:32);
//  * the input group (1/main.cu:54-60): IDADIM int32[4] = {LDA, M, N, NPLUS}
//    and G float64 with C dimensions [N][LDA] (Fortran LDA x N);
//  * stdout: the library's per-sweep BLK_SWP lines, then the GLB_ROT_S /
//    GLB_ROT_B totals and the wall time with the sweep count (1/main.cu:98-100);
//  * the result group, created anew in H5R (1/main.cu:105-140): IDADIM,
//    ISTATS int32[4] = {sweeps or the error code, STRAT0, STRAT1, return
//    code}, TIMING float64[4], G [N][LDA], V [N][LDA] when ALG & 2, D [N];
//  * the exit status is hypJacL2's return code.
// The reference's compile-time factorization variants (USE_QR, USE_DRMAC)
// are chosen at run time by HJ_FACT = cholesky | qr | drmac.
#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hypjac.h"

namespace {

[[noreturn]] void fail(const std::string &what)
{
  (void)fprintf(stderr, "x1: %s\n", what.c_str());
  exit(EXIT_FAILURE);
}

// An HDF5 status or identifier; negative values are failures.
template <typename T>
T h5(T v, const char *what)
{
  if (v < 0)
    fail(std::string("HDF5 failed: ") + what);
  return v;
}

// Owns one open HDF5 file or group.
class H5Object {
 public:
  H5Object(hid_t id, herr_t (*closer)(hid_t)) : id_(id), closer_(closer) {}
  H5Object(const H5Object &) = delete;
  H5Object &operator=(const H5Object &) = delete;
  ~H5Object()
  {
    if (id_ >= 0) (void)closer_(id_);
  }
  hid_t id() const { return id_; }

 private:
  hid_t id_;
  herr_t (*closer_)(hid_t);
};

struct Options {
  int device = 0;
  const char *plugin = nullptr;  // strat.so
  const char *fam0 = nullptr;    // level-0 family (order 32)
  const char *fam1 = nullptr;    // level-1 family (order ceil(N / 16))
  unsigned alg = 0u;
  const char *input = nullptr;
  const char *group = nullptr;
  const char *result = nullptr;  // optional
};

bool parse(int argc, char *argv[], Options &o)
{
  if (argc != 8 && argc != 9)
    return false;
  o.device = atoi(argv[1]);
  o.plugin = argv[2];
  o.fam0 = argv[3];
  o.fam1 = argv[4];
  o.alg = static_cast<unsigned>(atoi(argv[5]));
  o.input = argv[6];
  o.group = argv[7];
  o.result = (argc == 9) ? argv[8] : nullptr;
  return true;
}

void choose_factorization()
{
  const char *f = getenv("HJ_FACT");
  if (!f || !strcmp(f, "cholesky"))
    return;
  if (!strcmp(f, "qr"))
    (void)hj_set_factorization(HJ_FACT_QR);
  else if (!strcmp(f, "drmac"))
    (void)hj_set_factorization(HJ_FACT_QR_DRMAC);
  else
    fail("HJ_FACT must be cholesky, qr or drmac");
}

// The input problem: G with leading dimension lda (>= m), n columns, the
// first nplus of them with J = +1.
struct Problem {
  unsigned lda = 0u, m = 0u, n = 0u, nplus = 0u;
  std::vector<double> G;
};

Problem load(const char *path, const char *group)
{
  if (!fexist(path))
    fail(std::string("no input file ") + path);
  H5Object file(h5(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), "open input"), H5Fclose);
  H5Object grp(h5(H5Gopen2(file.id(), group, H5P_DEFAULT), "open input group"), H5Gclose);
  int dim[HYPJAC_IDADIM_SIZE] = {0, 0, 0, 0};
  h5(H5LTread_dataset_int(grp.id(), "IDADIM", dim), "read IDADIM");
  for (int d : dim)
    if (d < 0)
      fail("IDADIM holds a negative size");
  Problem p;
  p.lda = static_cast<unsigned>(dim[0]);
  p.m = static_cast<unsigned>(dim[1]);
  p.n = static_cast<unsigned>(dim[2]);
  p.nplus = static_cast<unsigned>(dim[3]);
  if (p.lda < p.m)
    fail("IDADIM: LDA is below M");
  p.G.assign(static_cast<size_t>(p.lda) * p.n, 0.0);
  h5(H5LTread_dataset_double(grp.id(), "G", p.G.data()), "read G");
  return p;
}

struct Result {
  int ret = 0;
  unsigned sweeps = 0u;
  unsigned long long rot_s = 0ull, rot_b = 0ull;
  double timing[4] = {-0.0, -0.0, -0.0, -0.0};
  std::vector<double> V, D;
};

// Runs hypJacL2 in place on p.G (which becomes U); V gets the input's
// leading dimension, as the reference's driver allocates it.
Result solve(unsigned alg, Problem &p)
{
  Result r;
  const bool withV = (alg & HYPJAC_FULL_SVD) != 0u;
  if (withV)
    r.V.assign(static_cast<size_t>(p.lda) * p.n, 0.0);
  r.D.assign(p.n, 0.0);
  r.ret = hypJacL2(alg, p.m, p.n, p.nplus, p.G.data(), p.lda, withV ? r.V.data() : nullptr,
                   withV ? p.lda : 0u, r.D.data(), &r.sweeps, &r.rot_s, &r.rot_b, r.timing);
  return r;
}

void store(const char *path, const char *group, unsigned alg, const Problem &p, const Result &r)
{
  H5Object file(h5(fexist(path) ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                                : H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT),
                   "open result file"),
                H5Fclose);
  H5Object grp(h5(H5Gcreate2(file.id(), group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                  "create result group"),
               H5Gclose);
  const int dim[HYPJAC_IDADIM_SIZE] = {static_cast<int>(p.lda), static_cast<int>(p.m),
                                       static_cast<int>(p.n), static_cast<int>(p.nplus)};
  const int stats[4] = {r.ret < 0 ? r.ret : static_cast<int>(r.sweeps), static_cast<int>(STRAT0),
                        static_cast<int>(STRAT1), r.ret};
  const hsize_t four[1] = {4u};
  h5(H5LTmake_dataset_int(grp.id(), "IDADIM", 1, four, dim), "write IDADIM");
  h5(H5LTmake_dataset_int(grp.id(), "ISTATS", 1, four, stats), "write ISTATS");
  h5(H5LTmake_dataset_double(grp.id(), "TIMING", 1, four, r.timing), "write TIMING");
  const hsize_t mat[2] = {p.n, p.lda};
  h5(H5LTmake_dataset_double(grp.id(), "G", 2, mat, p.G.data()), "write G");
  if (alg & HYPJAC_FULL_SVD)
    h5(H5LTmake_dataset_double(grp.id(), "V", 2, mat, r.V.data()), "write V");
  const hsize_t vec[1] = {p.n};
  h5(H5LTmake_dataset_double(grp.id(), "D", 1, vec, r.D.data()), "write D");
}

}  // namespace

int main(int argc, char *argv[])
{
  Options o;
  if (!parse(argc, argv, o)) {
    (void)fprintf(stderr, "%s DEV SDY SNP0 SNP1 ALG H5F H5G [H5R]\n", argv[0]);
    return EXIT_FAILURE;
  }
  choose_factorization();
  const unsigned order0 = 2u * HYPJACL1_NCOLB;
  const unsigned order1 = (static_cast<unsigned>(atoi(o.group)) + HYPJACL1_NCOLB - 1u) / HYPJACL1_NCOLB;
  init_strats(o.plugin, o.fam0, order0, o.fam1, order1);
  (void)configureGPU(o.device);  // ends the process on a device it cannot use

  h5(H5open(), "initialise");
  Problem p = load(o.input, o.group);
  const Result r = solve(o.alg, p);
  if (r.ret) {
    (void)fprintf(stderr, "%s: error %d\n", argv[0], r.ret);
  }
  else {
    (void)printf("GLB_ROT_S(%20llu), GLB_ROT_B(%20llu)\n", r.rot_s, r.rot_b);
    (void)printf("%#12.6f s %2u sweeps\n", r.timing[0], r.sweeps);
    (void)fflush(stdout);
  }
  if (o.result)
    store(o.result, o.group, o.alg, p, r);
  h5(H5close(), "close");
  return r.ret;
}
