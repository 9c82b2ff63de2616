/*
 * Fortran-ABI declarations of the drop-in drivers exported by
 * libpropack_amd.so, for C/C++ callers.  The argument lists are the
 * reference's (PROPACK 2.1): every argument by reference, hidden CHARACTER
 * lengths appended as size_t in argument order, arrays column-major.
 *
 *   dlansvd_      replaces DLANSVD      (double/dlansvd.F:1-3)
 *   dlansvd_irl_  replaces DLANSVD_IRL  (double/dlansvd_irl.F:1-3)
 *   slansvd_      replaces SLANSVD      (single/slansvd.F:1-3)
 *   slansvd_irl_  replaces SLANSVD_IRL  (single/slansvd_irl.F:1-3)
 *   zlansvd_      replaces ZLANSVD      (complex16/zlansvd.F:1-3)
 *   zlansvd_irl_  replaces ZLANSVD_IRL  (complex16/zlansvd_irl.F:1-3)
 *   clansvd_      replaces CLANSVD      (complex8/clansvd.F:1-3)
 *   clansvd_irl_  replaces CLANSVD_IRL  (complex8/clansvd_irl.F:1-3)
 *   clearstat_ / printstat_ over COMMON /timing/ (double/printstat.F:4-75)
 *
 * APROD has the reference callback signature (double/dlanbpro.F:28-41).
 * Complex arrays (COMPLEX*16 / COMPLEX) are (re, im) pairs, passed as void*.
 * Passing pk_dcsr_aprod_ / pk_scsr_aprod_ / pk_zcsr_aprod_ / pk_ccsr_aprod_
 * (propack_amd.h) with IPARM filled
 * by pk_csr_to_iparm keeps the operator on the GPU; any other APROD is
 * called on the host.  WORK/IWORK are accepted and unused; INFO <= -100
 * reports a device-layer error (PK_ERR_* of propack_amd.h).
 */
#ifndef PROPACK_H
#define PROPACK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*pk_daprod_t)(const char* transa, const int* m, const int* n, const double* x,
                            double* y, double* dparm, int* iparm, size_t len_transa);
typedef void (*pk_saprod_t)(const char* transa, const int* m, const int* n, const float* x,
                            float* y, float* sparm, int* iparm, size_t len_transa);

void dlansvd_(const char* jobu, const char* jobv, const int* m, const int* n, int* k,
              const int* kmax, pk_daprod_t aprod, double* U, const int* ldu, double* sigma,
              double* bnd, double* V, const int* ldv, const double* tolin, double* work,
              const int* lwork, int* iwork, const int* liwork, double* doption, int* ioption,
              int* info, double* dparm, int* iparm, size_t len_jobu, size_t len_jobv);

void dlansvd_irl_(const char* which, const char* jobu, const char* jobv, const int* m,
                  const int* n, int* dim, const int* p, int* neig, const int* maxiter,
                  pk_daprod_t aprod, double* U, const int* ldu, double* sigma, double* bnd,
                  double* V, const int* ldv, const double* tolin, double* work,
                  const int* lwork, int* iwork, const int* liwork, double* doption,
                  int* ioption, int* info, double* dparm, int* iparm, size_t len_which,
                  size_t len_jobu, size_t len_jobv);

void slansvd_(const char* jobu, const char* jobv, const int* m, const int* n, int* k,
              const int* kmax, pk_saprod_t aprod, float* U, const int* ldu, float* sigma,
              float* bnd, float* V, const int* ldv, const float* tolin, float* work,
              const int* lwork, int* iwork, const int* liwork, float* soption, int* ioption,
              int* info, float* sparm, int* iparm, size_t len_jobu, size_t len_jobv);

void slansvd_irl_(const char* which, const char* jobu, const char* jobv, const int* m,
                  const int* n, int* dim, const int* p, int* neig, const int* maxiter,
                  pk_saprod_t aprod, float* U, const int* ldu, float* sigma, float* bnd,
                  float* V, const int* ldv, const float* tolin, float* work, const int* lwork,
                  int* iwork, const int* liwork, float* soption, int* ioption, int* info,
                  float* sparm, int* iparm, size_t len_which, size_t len_jobu,
                  size_t len_jobv);

/* APROD of the complex trees (complex16/zlanbpro.F:28-41): TRANSA 'n' or
 * 'c' (A^H). */
typedef void (*pk_zaprod_t)(const char* transa, const int* m, const int* n, const void* x,
                            void* y, void* zparm, int* iparm, size_t len_transa);

void zlansvd_(const char* jobu, const char* jobv, const int* m, const int* n, int* k,
              const int* kmax, pk_zaprod_t aprod, void* U, const int* ldu, double* sigma,
              double* bnd, void* V, const int* ldv, const double* tolin, double* work,
              const int* lwork, void* zwork, const int* lzwrk, int* iwork, const int* liwork,
              double* doption, int* ioption, int* info, void* zparm, int* iparm,
              size_t len_jobu, size_t len_jobv);

void zlansvd_irl_(const char* which, const char* jobu, const char* jobv, const int* m,
                  const int* n, int* dim, const int* p, int* neig, const int* maxiter,
                  pk_zaprod_t aprod, void* U, const int* ldu, double* sigma, double* bnd,
                  void* V, const int* ldv, const double* tolin, double* work, const int* lwork,
                  void* zwork, const int* lzwrk, int* iwork, const int* liwork,
                  double* doption, int* ioption, int* info, void* zparm, int* iparm,
                  size_t len_which, size_t len_jobu, size_t len_jobv);

void clansvd_(const char* jobu, const char* jobv, const int* m, const int* n, int* k,
              const int* kmax, pk_zaprod_t aprod, void* U, const int* ldu, float* sigma,
              float* bnd, void* V, const int* ldv, const float* tolin, float* work,
              const int* lwork, void* cwork, const int* lcwrk, int* iwork, const int* liwork,
              float* soption, int* ioption, int* info, void* cparm, int* iparm,
              size_t len_jobu, size_t len_jobv);

void clansvd_irl_(const char* which, const char* jobu, const char* jobv, const int* m,
                  const int* n, int* dim, const int* p, int* neig, const int* maxiter,
                  pk_zaprod_t aprod, void* U, const int* ldu, float* sigma, float* bnd,
                  void* V, const int* ldv, const float* tolin, float* work, const int* lwork,
                  void* cwork, const int* lcwrk, int* iwork, const int* liwork,
                  float* soption, int* ioption, int* info, void* cparm, int* iparm,
                  size_t len_which, size_t len_jobu, size_t len_jobv);

void clearstat_(void);
void printstat_(void);

#ifdef __cplusplus
}
#endif
#endif
