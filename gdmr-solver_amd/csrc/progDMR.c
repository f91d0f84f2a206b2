/*
 * progDMR -- control program for libgDMR.so (the role of progDMR.c in seanMolesky/gDMR-solver).
 *
 * Like the reference driver (progDMR.c:135-182) it talks to the library only through gDMRC.h:
 * initLS, operator import, then a sequence of solves where the first starts fresh (mode 0) and
 * the later ones recycle the deflation subspace (mode = i, progDMR.c:169).  The operator is the
 * 2-D convection-diffusion stencil (default 64 x 64, m = 30, k = 10, tol 1e-8, 4 N(0,1) RHS from
 * seed 0) instead of a time()-seeded dense Laplacian, so runs are reproducible.
 *
 * usage: progDMR [N] [m] [k] [tol] [numRHS]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "gDMRC.h"

int main(int argc, char **argv) {
  int N = argc > 1 ? atoi(argv[1]) : 64;
  int m = argc > 2 ? atoi(argv[2]) : 30;
  int k = argc > 3 ? atoi(argv[3]) : 10;
  double tol = argc > 4 ? atof(argv[4]) : 1e-8;
  int numRHS = argc > 5 ? atoi(argv[5]) : 4;
  int n = N * N;
  int devList[1] = {0};
  long long nnz = dmrGenConvDiff2D(N, 0.1, 1.0, 0.5, 0.0, 0, 0, NULL, NULL, NULL);
  int *rp = (int *)malloc(sizeof(int) * (n + 1));
  int *ci = (int *)malloc(sizeof(int) * nnz);
  double *va = (double *)malloc(sizeof(double) * nnz);
  double *b = (double *)malloc(sizeof(double) * n * numRHS);
  double *x = (double *)calloc(n, sizeof(double));
  if (!rp || !ci || !va || !b || !x) return 1;
  dmrGenConvDiff2D(N, 0.1, 1.0, 0.5, 0.0, 0, 0, rp, ci, va);
  dmrRandNormal(0, 1, (long long)n * numRHS, b);

  initLS(32, 512, 0, n, devList, 1, m, k);
  if (impCsrLS(rp, ci, va, (int)nnz) != 0) {
    fprintf(stderr, "progDMR: operator import failed\n");
    return 1;
  }
  for (int i = 0; i < numRHS; ++i) {
    int numIts = 0;
    for (int j = 0; j < n; ++j) x[j] = 0.0;
    impImgLS(b + (size_t)i * n);
    impSrcLS(x);
    double res = solveLS(i, tol, n, &numIts);
    /* True residual ||b - A x|| / ||b|| on the host. */
    expSrcLS(x);
    double rr = 0.0, bb = 0.0;
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int e = rp[r]; e < rp[r + 1]; ++e) s += va[e] * x[ci[e]];
      double d = b[(size_t)i * n + r] - s;
      rr += d * d;
      bb += b[(size_t)i * n + r] * b[(size_t)i * n + r];
    }
    printf("gDMR solve %d: residual %.3e (true %.3e) after %d iterations\n", i, res,
           sqrt(rr / bb), numIts);
  }
  freeLS();
  free(rp);
  free(ci);
  free(va);
  free(b);
  free(x);
  return 0;
}
