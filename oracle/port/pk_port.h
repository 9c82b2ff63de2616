/*
 * Test infrastructure only.  Plain-C restatement of PROPACK's Lanczos
 * bidiagonalization SVD drivers (reference: double/ and single/ trees of
 * PROPACK 2.1).  Used by tests/ as an independent checker and by bench.py's
 * cpu_baseline leg when the compiled reference is absent.  Never linked into
 * the product library.
 *
 * Symbols: pkp_d_* (double) and pkp_s_* (single).  Arrays are column-major,
 * 0-based in C, leading dimensions as in the reference.
 */
#ifndef PK_PORT_H
#define PK_PORT_H
#include <stdint.h>

/* Counters with the meaning of COMMON /timing/ (double/stat.h:12-15). */
typedef struct {
    int nopx, nreorth, ndot, nbsvd, nrestart, nlandim, nsing;
} pkp_stats;

#define PKP_DECL(R, P)                                                        \
    typedef void (*P##aprod_fn)(char trans, const R *x, R *y, void *ctx);     \
    void P##set_csr(int m, int n, const int *rp, const int *ci, const R *va,  \
                    const int *rpt, const int *cit, const R *vat);            \
    void P##csr_aprod(char trans, const R *x, R *y, void *ctx);               \
    void P##larnv_pm1(int n, R *x);                                           \
    R P##lapy2(R x, R y);                                                     \
    void P##lartg(R f, R g, R *c, R *s, R *r);                                \
    void P##set_mu(int k, R *mu, const int *index, R val);                    \
    void P##compute_int(const R *mu, int j, R delta, R eta, int *index);      \
    void P##update_mu(R *mumax, R *mu, const R *nu, int j, const R *alpha,    \
                      const R *beta, R anorm, R eps1);                        \
    void P##update_nu(R *numax, const R *mu, R *nu, int j, const R *alpha,    \
                      const R *beta, R anorm, R eps1);                        \
    void P##refinebounds(int n, int k, const R *theta, R *bound, R tol,       \
                         R eps34);                                            \
    void P##bdqr(int ignorelast, char jobq, int n, R *D, R *E, R *c1, R *c2,  \
                 R *Qt, int ldq);                                             \
    void P##bsvdstep(char jobu, char jobv, int m, int n, int k, R sigma,      \
                     R *D, R *E, R *U, int ldu, R *V, int ldv);               \
    void P##reorth(int n, int k, const R *V, int ldv, R *vnew, R *normvnew,   \
                   const int *index, R alpha, int iflag);                     \
    void P##gemm_ovwr_left(char transb, int m, int n, int k, R *A, int lda,   \
                           const R *B, int ldb);                              \
    int P##lanbpro(int m, int n, int k0, int *k, P##aprod_fn aprod,           \
                   void *ctx, R *U, int ldu, R *V, int ldv, R *alpha,         \
                   R *beta, R *rnorm, R *doption, const int *ioption);        \
    int P##lansvd(char jobu, char jobv, int m, int n, int *k, int kmax,       \
                  P##aprod_fn aprod, void *ctx, R *U, int ldu, R *sigma,      \
                  R *bnd, R *V, int ldv, R tolin, R *doption,                 \
                  const int *ioption);                                        \
    int P##lansvd_irl(char which, char jobu, char jobv, int m, int n,         \
                      int *dim, int p, int *neig, int maxiter,                \
                      P##aprod_fn aprod, void *ctx, R *U, int ldu, R *sigma,  \
                      R *bnd, R *V, int ldv, R tolin, R *doption,             \
                      const int *ioption);                                    \
    void P##get_stats(pkp_stats *out);                                        \
    void P##clear_stats(void);

PKP_DECL(double, pkp_d_)
PKP_DECL(float, pkp_s_)

#endif
