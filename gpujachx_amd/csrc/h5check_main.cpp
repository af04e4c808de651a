// h5check_main.cpp -- accuracy verifier of an x1 result, restating the
// metrics of the reference's offline post-processor var/pp80.f90 (80-bit
// REAL(10) with OpenMP) in C++ long double:
//
//   h5check FH5 GH5 RH5
//
// FH5/GH5: the input group (IDADIM, G, optional LAMBDA); RH5/GH5: the result
// group written by x1 (IDADIM, G = U, V if ALG & 2, D).  Prints and writes
// into the result group
//   NRMS[11]: ||U^T U - I||_M,F (pp80.f90:151-206), ||U U^T - I||_M,F
//             (:207-268, when M == N), ||V^T J V - J||_M,F and
//             ||V J V^T - J||_M,F (:414-650), and, since synthetic inputs carry
//             no (A, IPIV), the residual of the factorisation itself,
//             ||G V - U sqrt|D| ||_M,F and its largest column-relative value
//             in place of pp80's ||A U - U D||_M,F,R (:663-752);
//   ABSERR[N+1], RELERR[N+1]: LAMBDA - D after both are sorted descending,
//             the last entry the one of largest magnitude (:759-841), when
//             the input has LAMBDA.
#include <hdf5.h>
#include <hdf5_hl.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

typedef long double LD;

#define H5CALL(call)                                                         \
  do {                                                                       \
    if ((call) < 0) {                                                        \
      fprintf(stderr, "HDF5 error @ %s(%d)\n", __FILE__, __LINE__);          \
      exit(EXIT_FAILURE);                                                    \
    }                                                                        \
  } while (0)

// max |X - S| over the lower triangle and Frobenius norm (symmetric X)
static void sym_err(size_t n, const std::function<LD(size_t, size_t)> &X,
                    const std::function<LD(size_t)> &S, double &mx, double &fro)
{
  LD m = 0, offsq = 0, dsq = 0;
#pragma omp parallel for reduction(max : m) reduction(+ : offsq, dsq) schedule(dynamic, 8)
  for (long jj = 0; jj < (long)n; ++jj) {
    const size_t j = (size_t)jj;
    const LD dj = std::fabs(X(j, j) - S(j));
    m = std::max(m, dj);
    dsq += dj * dj;
    for (size_t i = j + 1; i < n; ++i) {
      const LD v = std::fabs(X(i, j));
      m = std::max(m, v);
      offsq += v * v;
    }
  }
  mx = (double)m;
  fro = (double)std::sqrt(2 * offsq + dsq);
}

int main(int argc, char **argv)
{
  if (argc != 4) {
    fprintf(stderr, "%s FH5 GH5 RH5\n", argv[0]);
    return EXIT_FAILURE;
  }
  const char *fh5 = argv[1], *gh5 = argv[2], *rh5 = argv[3];
  H5CALL(H5open());
  int idadim[4];
  hid_t fid, gid;
  H5CALL(fid = H5Fopen(fh5, H5F_ACC_RDONLY, H5P_DEFAULT));
  H5CALL(gid = H5Gopen2(fid, gh5, H5P_DEFAULT));
  H5CALL(H5LTread_dataset_int(gid, "IDADIM", idadim));
  const size_t lda = idadim[0], m = idadim[1], n = idadim[2], p = idadim[3];
  std::vector<double> G(lda * n), lambda;
  H5CALL(H5LTread_dataset_double(gid, "G", G.data()));
  if (H5LTfind_dataset(gid, "LAMBDA")) {
    lambda.resize(n);
    H5CALL(H5LTread_dataset_double(gid, "LAMBDA", lambda.data()));
  }
  H5CALL(H5Gclose(gid));
  H5CALL(H5Fclose(fid));

  H5CALL(fid = H5Fopen(rh5, H5F_ACC_RDWR, H5P_DEFAULT));
  H5CALL(gid = H5Gopen2(fid, gh5, H5P_DEFAULT));
  int rdim[4];
  H5CALL(H5LTread_dataset_int(gid, "IDADIM", rdim));
  const size_t ldr = rdim[0];
  std::vector<double> U(ldr * n), D(n), V;
  H5CALL(H5LTread_dataset_double(gid, "G", U.data()));
  H5CALL(H5LTread_dataset_double(gid, "D", D.data()));
  const bool haveV = H5LTfind_dataset(gid, "V") > 0;
  size_t ldv = 0;
  if (haveV) {
    hsize_t dims[2];
    H5CALL(H5LTget_dataset_info(gid, "V", dims, nullptr, nullptr));
    ldv = dims[1];
    V.resize(ldv * n);
    H5CALL(H5LTread_dataset_double(gid, "V", V.data()));
  }
  auto u = [&](size_t i, size_t j) { return (LD)U[j * ldr + i]; };
  auto v = [&](size_t i, size_t j) { return (LD)V[j * ldv + i]; };
  auto jsg = [&](size_t i) { return (i < p) ? (LD)1 : (LD)-1; };

  double nrms[11];
  for (double &x : nrms) x = -1.0;
  sym_err(n, [&](size_t i, size_t j) { LD s = 0; for (size_t k = 0; k < m; ++k) s += u(k, i) * u(k, j); return s; },
          [](size_t) { return (LD)1; }, nrms[0], nrms[1]);
  printf("|| U^T U - I ||_M %.16e\n|| U^T U - I ||_F %.16e\n", nrms[0], nrms[1]);
  if (m == n) {
    sym_err(m, [&](size_t i, size_t j) { LD s = 0; for (size_t k = 0; k < n; ++k) s += u(i, k) * u(j, k); return s; },
            [](size_t) { return (LD)1; }, nrms[2], nrms[3]);
    printf("|| U U^T - I ||_M %.16e\n|| U U^T - I ||_F %.16e\n", nrms[2], nrms[3]);
  }
  if (haveV) {
    sym_err(n, [&](size_t i, size_t j) { LD s = 0; for (size_t k = 0; k < n; ++k) s += v(k, i) * jsg(k) * v(k, j); return s; },
            jsg, nrms[4], nrms[5]);
    sym_err(n, [&](size_t i, size_t j) { LD s = 0; for (size_t k = 0; k < n; ++k) s += v(i, k) * jsg(k) * v(j, k); return s; },
            jsg, nrms[6], nrms[7]);
    printf("|| V^T J V - J ||_M %.16e\n|| V^T J V - J ||_F %.16e\n", nrms[4], nrms[5]);
    printf("|| V J V^T - J ||_M %.16e\n|| V J V^T - J ||_F %.16e\n", nrms[6], nrms[7]);
    // residual of G V = U sqrt|D|
    LD mx = 0, fsq = 0, rel = 0;
#pragma omp parallel for reduction(max : mx, rel) reduction(+ : fsq) schedule(dynamic, 4)
    for (long jj = 0; jj < (long)n; ++jj) {
      const size_t j = (size_t)jj;
      const LD sj = std::sqrt(std::fabs((LD)D[j]));
      LD cn = 0, cm = 0;
      for (size_t i = 0; i < m; ++i) {
        LD s = 0;
        for (size_t k = 0; k < n; ++k) s += (LD)G[k * lda + i] * v(k, j);
        const LD r = std::fabs(s - u(i, j) * sj);
        cm = std::max(cm, r);
        cn += r * r;
      }
      mx = std::max(mx, cm);
      fsq += cn;
      if (sj > 0) rel = std::max(rel, std::sqrt(cn) / sj);
    }
    nrms[8] = (double)mx;
    nrms[9] = (double)std::sqrt(fsq);
    nrms[10] = (double)rel;
    printf("|| G V - U S ||_M %.16e\n|| G V - U S ||_F %.16e\n|| G V - U S ||_R %.16e\n", nrms[8],
           nrms[9], nrms[10]);
  }
  hsize_t d1[1] = {11};
  if (H5LTfind_dataset(gid, "NRMS")) H5CALL(H5Ldelete(gid, "NRMS", H5P_DEFAULT));
  H5CALL(H5LTmake_dataset_double(gid, "NRMS", 1, d1, nrms));

  if (!lambda.empty()) {
    std::vector<double> dd(D), ll(lambda);
    std::sort(dd.begin(), dd.end(), std::greater<double>());
    std::sort(ll.begin(), ll.end(), std::greater<double>());
    std::vector<double> ae(n + 1), re(n + 1);
    double amax = 0.0, rmax = 0.0;
    for (size_t i = 0; i < n; ++i) {
      ae[i] = ll[i] - dd[i];
      re[i] = (ll[i] != 0.0) ? ae[i] / ll[i] : ae[i];
      if (std::fabs(ae[i]) > std::fabs(amax)) amax = ae[i];
      if (std::fabs(re[i]) > std::fabs(rmax)) rmax = re[i];
    }
    ae[n] = amax;
    re[n] = rmax;
    printf("max abs err %.16e\nmax rel err %.16e\n", amax, rmax);
    d1[0] = n + 1;
    for (const char *nm : {"ABSERR", "RELERR"})
      if (H5LTfind_dataset(gid, nm)) H5CALL(H5Ldelete(gid, nm, H5P_DEFAULT));
    H5CALL(H5LTmake_dataset_double(gid, "ABSERR", 1, d1, ae.data()));
    H5CALL(H5LTmake_dataset_double(gid, "RELERR", 1, d1, re.data()));
  }
  H5CALL(H5Gclose(gid));
  H5CALL(H5Fclose(fid));
  H5CALL(H5close());
  return 0;
}
