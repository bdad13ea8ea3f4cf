// mf_driver.cpp -- `mf-solve`: Matrix Market in, solve, report.
//
// Replaces the reference's four executables (main-cg, main-bicg,
// main-cg-mixed, main-bicg-mixed; src/main-cg.cu:1630-1673,
// src/main-bicg.cu:1696-1726, src/main-cg-mixed.cu:3006-3042,
// src/main-bicg-mixed.cu:2168-2193):
//   mf-solve [--cg|--bicgstab] [--mixed] [--multikernel] A.mtx
// CG: x* = ones, b = A x* over the leading n x n block (n truncated to a
// multiple of 16 as the reference does); BiCGSTAB: the solver sets b to the
// row sums of A.  Non-square input prints "unequal" and exits 0, as there.
// The policy of --mixed runs follows MF_COMPAT_POLICY (default: rule).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "millefeuille.h"

int main(int argc, char **argv) {
  bool bicg = false, mixed = false, multi = false;
  const char *path = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--bicgstab")) bicg = true;
    else if (!std::strcmp(argv[i], "--cg")) bicg = false;
    else if (!std::strcmp(argv[i], "--mixed")) mixed = true;
    else if (!std::strcmp(argv[i], "--multikernel")) multi = true;
    else path = argv[i];
  }
  if (!path) {
    std::fprintf(stderr, "usage: %s [--cg|--bicgstab] [--mixed] [--multikernel] A.mtx\n", argv[0]);
    return 2;
  }
  int m, n, nnzR, isSym;
  int *RowPtr, *ColIdx;
  double *Val;
  read_Dmatrix_32(&m, &n, &nnzR, &RowPtr, &ColIdx, &Val, &isSym, (char *)path);
  if (m < 0) {
    std::fprintf(stderr, "mf-solve: %s\n", mf_last_error());
    return 1;
  }
  if (m != n) {
    std::printf("unequal\n");
    return 0;
  }
  if (bicg) {
    std::vector<double> rhs(n, 1.0), x(n, 0.0);
    int rc = mixed ? bicgstab_mixed(RowPtr, ColIdx, Val, rhs.data(), x.data(), nnzR, (char *)path, n)
           : multi ? bicgstab(RowPtr, ColIdx, Val, rhs.data(), x.data(), nnzR, (char *)path, n)
                   : bicgstab_sync(RowPtr, ColIdx, Val, rhs.data(), x.data(), nnzR, (char *)path, n);
    mf_csr a = {m, n, nnzR, RowPtr, ColIdx, Val, isSym};
    mf_csr_free(&a);
    return rc < 0 ? 1 : 0;
  }
  std::vector<float> low(nnzR);
  for (int i = 0; i < nnzR; ++i) low[i] = (float)Val[i];
  const int ori = n;
  n = (n / 16) * 16;
  std::vector<double> X(n, 1.0), B(n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = RowPtr[i]; k < RowPtr[i + 1]; ++k)
      if (ColIdx[k] < n) B[i] += Val[k] * X[ColIdx[k]];
  int iter = 0;
  if (mixed)
    cg_solve_mixed(RowPtr, ColIdx, Val, low.data(), X.data(), B.data(), n, &iter, 10, 1e-5,
                   (char *)path, nnzR, ori);
  else if (multi)
    cg_solve_reduce(RowPtr, ColIdx, Val, low.data(), X.data(), B.data(), n, &iter, 10, 1e-5,
                    (char *)path, nnzR, ori);
  else
    cg_solve_sync(RowPtr, ColIdx, Val, low.data(), X.data(), B.data(), n, &iter, 10, 1e-5,
                  (char *)path, nnzR, ori);
  double err = 0.0;
  for (int i = 0; i < n; ++i) err = std::max(err, std::abs(X[i] - 1.0));
  std::printf("max |x - 1| = %e\n", err);
  mf_csr a = {m, ori, nnzR, RowPtr, ColIdx, Val, isSym};
  mf_csr_free(&a);
  return 0;
}
