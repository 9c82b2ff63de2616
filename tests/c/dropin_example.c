/* A C caller of the drop-in drivers through include/propack.h (the
 * Fortran ABI: arguments by reference, hidden CHARACTER lengths last):
 * DLANSVD with a C host APROD, then SLANSVD_IRL with the library's device
 * operator.  Built by the test build step; run by
 * tests/test_dropin_programs.py. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "propack.h"
#include "propack_amd.h"

enum { M = 250, N = 180, PER = 5, KMAX = 180, NSV = 6 };
static int rowptr[M + 1], colind[M * PER];
static double val[M * PER];

static void build(void) {
  int q = 0;
  rowptr[0] = 0;
  for (int i = 0; i < M; ++i) {
    for (int t = 1; t <= PER; ++t, ++q) {
      colind[q] = (i * 11 + t * 29 + t * t) % N;
      val[q] = cos(0.23 * i + 0.7 * t) + 0.05 * t;
    }
    rowptr[i + 1] = q;
  }
}

static void my_aprod(const char* tr, const int* m, const int* n, const double* x, double* y,
                     double* dparm, int* iparm, size_t len) {
  (void)dparm; (void)iparm; (void)len;
  if (*tr == 'n' || *tr == 'N') {
    for (int i = 0; i < *m; ++i) {
      double s = 0;
      for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) s += val[q] * x[colind[q]];
      y[i] = s;
    }
  } else {
    for (int j = 0; j < *n; ++j) y[j] = 0;
    for (int i = 0; i < *m; ++i)
      for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) y[colind[q]] += val[q] * x[i];
  }
}

int main(void) {
  build();
  static double U[M * (KMAX + 1)], V[N * KMAX], sigma[KMAX], bnd[KMAX];
  double dopt[4] = {1.4901161193847656e-08, 1.8189894035458565e-12, 0.0, 0.002}, dparm[1];
  double work[1], tol = 1e-12;
  int iwork[1], iopt[2] = {1, 1}, iparm[2] = {0, 0}, info = 0, k = NSV, m = M, n = N,
      kmax = KMAX, one = 1;
  for (int i = 0; i < M; ++i) U[i] = sin(0.3 * i + 0.1);
  dlansvd_("y", "y", &m, &n, &k, &kmax, my_aprod, U, &m, sigma, bnd, V, &n, &tol, work, &one,
           iwork, &one, dopt, iopt, &info, dparm, iparm, 1, 1);
  printf("HOST_INFO_K %d %d\nHOST_SIGMA", info, k);
  for (int i = 0; i < NSV; ++i) printf(" %.17e", sigma[i]);
  printf("\n");

  /* single precision through the device operator */
  static float Us[M * (KMAX + 1)], Vs[N * KMAX], ss[KMAX], bs[KMAX], vs[M * PER];
  for (int q = 0; q < M * PER; ++q) vs[q] = (float)val[q];
  pk_csr_t op;
  if (pk_csr_create(M, N, rowptr[M], rowptr, colind, vs, PK_F32, PK_HOST, &op) != 0) return 2;
  pk_csr_to_iparm(op, iparm);
  for (int i = 0; i < M; ++i) Us[i] = (float)sin(0.3 * i + 0.1);
  float sopt[4] = {3.4526698e-04f, 6.4155380e-06f, 0.0f, 0.002f}, sparm[1], swork[1],
        stol = 1e-5f;
  int dim = 24, p = 24 - NSV, neig = NSV, maxit = 1000;
  slansvd_irl_("L", "y", "y", &m, &n, &dim, &p, &neig, &maxit, pk_scsr_aprod_, Us, &m, ss, bs,
               Vs, &n, &stol, swork, &one, iwork, &one, sopt, iopt, &info, sparm, iparm, 1, 1,
               1);
  printf("DEV_INFO_K %d %d %d\nDEV_SIGMA", info, neig, dim);
  for (int i = 0; i < NSV; ++i) printf(" %.9e", ss[i]);
  printf("\n");
  pk_csr_destroy(op);
  return 0;
}
