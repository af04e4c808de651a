// h5gen_main.cpp -- writes an x1 input group in the reference's HDF5 layout
// (README.md:53-83, 1/main.cu:49-73): /GROUP/IDADIM = int32 {LDA, M, N,
// NPLUS} and /GROUP/G = float64 [N][LDA] (Fortran LDA x N).  Optional
// datasets: LAMBDA (float64[N], the exact eigenvalues j_i sigma_i^2, for the
// verifier) and, to stage a result group for tests, V and D.
//
//   h5gen OUT.h5 GROUP LDA M N NPLUS G.bin [LAMBDA.bin | V.bin D.bin]
//
// *.bin are raw little-endian float64 arrays (G: LDA*N, V: N*N column-major).
#include <hdf5.h>
#include <hdf5_hl.h>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static std::vector<double> read_bin(const char *fn, size_t n)
{
  std::vector<double> v(n);
  FILE *f = fopen(fn, "rb");
  if (!f || fread(v.data(), sizeof(double), n, f) != n) {
    fprintf(stderr, "h5gen: cannot read %zu doubles from %s\n", n, fn);
    exit(EXIT_FAILURE);
  }
  fclose(f);
  return v;
}

#define H5CALL(call)                                                         \
  do {                                                                       \
    if ((call) < 0) {                                                        \
      fprintf(stderr, "HDF5 error @ %s(%d)\n", __FILE__, __LINE__);          \
      exit(EXIT_FAILURE);                                                    \
    }                                                                        \
  } while (0)

int main(int argc, char **argv)
{
  if (argc != 8 && argc != 9 && argc != 10) {
    fprintf(stderr, "%s OUT.h5 GROUP LDA M N NPLUS G.bin [LAMBDA.bin | V.bin D.bin]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const char *out = argv[1], *grp = argv[2];
  int idadim[4] = {atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6])};
  const size_t lda = (size_t)idadim[0], n = (size_t)idadim[2];
  std::vector<double> G = read_bin(argv[7], lda * n);
  struct stat st;
  hid_t fid;
  if (stat(out, &st) == 0)
    fid = H5Fopen(out, H5F_ACC_RDWR, H5P_DEFAULT);
  else
    fid = H5Fcreate(out, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  H5CALL(fid);
  const hid_t gid = H5Gcreate2(fid, grp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5CALL(gid);
  hsize_t d1[1] = {4};
  H5CALL(H5LTmake_dataset_int(gid, "IDADIM", 1, d1, idadim));
  hsize_t d2[2] = {n, lda};
  H5CALL(H5LTmake_dataset_double(gid, "G", 2, d2, G.data()));
  if (argc == 9) {
    std::vector<double> L = read_bin(argv[8], n);
    d1[0] = n;
    H5CALL(H5LTmake_dataset_double(gid, "LAMBDA", 1, d1, L.data()));
  }
  if (argc == 10) {
    std::vector<double> V = read_bin(argv[8], n * n), D = read_bin(argv[9], n);
    d2[0] = n;
    d2[1] = n;
    H5CALL(H5LTmake_dataset_double(gid, "V", 2, d2, V.data()));
    d1[0] = n;
    H5CALL(H5LTmake_dataset_double(gid, "D", 1, d1, D.data()));
  }
  H5CALL(H5Gclose(gid));
  H5CALL(H5Fclose(fid));
  return 0;
}
