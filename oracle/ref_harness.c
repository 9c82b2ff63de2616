/*
 * Test infrastructure only (oracle side).  A small C harness linked with the
 * reference PROPACK 2.1 objects compiled from /root/reference by
 * oracle/Makefile.  It supplies what the reference leaves to its caller:
 *
 *   - an APROD operator over a CSR matrix and its explicit transpose
 *     (the callback interface of double/dlansvd.F:20-33: Y = A*X for 'n',
 *     Y = A^T*X for 't'; Y fully overwritten, DPARM/IPARM unused);
 *   - workspace sizing per double/dlansvd.F:54-65 and double/dlansvd_irl.F:64-75;
 *   - access to COMMON /timing/ (double/stat.h:12-15).
 *
 * Built twice: -DPREC_D links against the double tree, -DPREC_S against the
 * single tree.  Exported names carry the precision letter (ref_d_*, ref_s_*).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#if defined(PREC_D)
typedef double real_t;
#define P(name) ref_d_##name
#define LANSVD dlansvd_
#define LANSVD_IRL dlansvd_irl_
#elif defined(PREC_S)
typedef float real_t;
#define P(name) ref_s_##name
#define LANSVD slansvd_
#define LANSVD_IRL slansvd_irl_
#else
#error "define PREC_D or PREC_S"
#endif

typedef void (*aprod_fn)(const char *, const int *, const int *, const real_t *,
                         real_t *, real_t *, int *, size_t);

/* Fortran entry points of the reference (flang: lower case + '_', hidden
 * CHARACTER lengths appended as size_t).  Argument order: double/dlansvd.F:1-3
 * and double/dlansvd_irl.F:1-3. */
void LANSVD(const char *jobu, const char *jobv, const int *m, const int *n,
            int *k, const int *kmax, aprod_fn aprod, real_t *U, const int *ldu,
            real_t *sigma, real_t *bnd, real_t *V, const int *ldv,
            const real_t *tolin, real_t *work, const int *lwork, int *iwork,
            const int *liwork, real_t *doption, int *ioption, int *info,
            real_t *dparm, int *iparm, size_t ljobu, size_t ljobv);
void LANSVD_IRL(const char *which, const char *jobu, const char *jobv,
                const int *m, const int *n, int *dim, const int *p, int *neig,
                const int *maxiter, aprod_fn aprod, real_t *U, const int *ldu,
                real_t *sigma, real_t *bnd, real_t *V, const int *ldv,
                const real_t *tolin, real_t *work, const int *lwork, int *iwork,
                const int *liwork, real_t *doption, int *ioption, int *info,
                real_t *dparm, int *iparm, size_t lwhich, size_t ljobu,
                size_t ljobv);

/* COMMON /timing/ as declared in double/stat.h:12-15 (8 integers, 14 reals,
 * nlandim, 3 reals, nsing). */
extern struct {
    int nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd;
    float t1[14];
    int nlandim;
    float t2[3];
    int nsing;
} timing_;

/* ---- CSR operator (A and explicit A^T), shared by both transposes ---- */
static struct {
    int m, n;
    const int *rp, *ci;
    const real_t *va;
    const int *rpt, *cit;
    const real_t *vat;
    int64_t napply;
} op;

static void csr_apply(int rows, const int *rp, const int *ci, const real_t *va,
                      const real_t *x, real_t *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rows; ++i) {
        double s = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) s += (double)va[p] * (double)x[ci[p]];
        y[i] = (real_t)s;
    }
}

static void ref_aprod(const char *transa, const int *m, const int *n,
                      const real_t *x, real_t *y, real_t *dparm, int *iparm,
                      size_t ltransa)
{
    (void)m; (void)n; (void)dparm; (void)iparm; (void)ltransa;
    op.napply++;
    if (transa[0] == 'n' || transa[0] == 'N')
        csr_apply(op.m, op.rp, op.ci, op.va, x, y);
    else
        csr_apply(op.n, op.rpt, op.cit, op.vat, x, y);
}

void P(set_operator)(int m, int n, const int *rp, const int *ci,
                     const real_t *va, const int *rpt, const int *cit,
                     const real_t *vat)
{
    op.m = m; op.n = n;
    op.rp = rp; op.ci = ci; op.va = va;
    op.rpt = rpt; op.cit = cit; op.vat = vat;
    op.napply = 0;
}

void P(aprod)(char trans, const real_t *x, real_t *y)
{
    ref_aprod(&trans, &op.m, &op.n, x, y, NULL, NULL, 1);
}

void P(set_threads)(int nt) { omp_set_num_threads(nt); }
int P(max_threads)(void) { return omp_get_max_threads(); }

extern int MKL_Set_Threading_Layer(int);
/* Sequential MKL: the reference calls BLAS from inside its own OpenMP
 * regions (dreorth.F:174,199; dlansvd_irl.F:387-395), one BLAS per thread. */
void P(init)(void) { MKL_Set_Threading_Layer(1 /* MKL_THREADING_SEQUENTIAL */); }

void P(clear_counters)(void) { memset(&timing_, 0, sizeof(timing_)); }

/* out[0..7] = nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart,
 * nbsvd; out[8] = nlandim; out[9] = nsing. */
void P(get_counters)(int *out)
{
    out[0] = timing_.nopx; out[1] = timing_.nreorth; out[2] = timing_.ndot;
    out[3] = timing_.nreorthu; out[4] = timing_.nreorthv; out[5] = timing_.nitref;
    out[6] = timing_.nrestart; out[7] = timing_.nbsvd;
    out[8] = timing_.nlandim; out[9] = timing_.nsing;
}

static int imax(int a, int b) { return a > b ? a : b; }

/* Non-restarted driver with workspace sized per double/dlansvd.F:54-65
 * (vector case, NB = 16). */
int P(lansvd)(char jobu, char jobv, int m, int n, int *k, int kmax,
              real_t *U, int ldu, real_t *sigma, real_t *bnd, real_t *V,
              int ldv, real_t tolin, real_t *doption, int *ioption)
{
    int mn = imax(m, n);
    int lwork = m + n + 9 * kmax + 5 * kmax * kmax + 4 +
                imax(3 * kmax * kmax + 4 * kmax + 4, 16 * mn);
    int liwork = 8 * kmax;
    real_t *work = calloc((size_t)lwork, sizeof(real_t));
    int *iwork = calloc((size_t)liwork, sizeof(int));
    int info = 0;
    real_t dparm[1] = {0};
    int iparm[1] = {0};
    LANSVD(&jobu, &jobv, &m, &n, k, &kmax, ref_aprod, U, &ldu, sigma, bnd, V,
           &ldv, &tolin, work, &lwork, iwork, &liwork, doption, ioption, &info,
           dparm, iparm, 1, 1);
    free(work);
    free(iwork);
    return info;
}

/* Restarted driver with workspace sized per double/dlansvd_irl.F:64-75. */
int P(lansvd_irl)(char which, char jobu, char jobv, int m, int n, int *dim,
                  int p, int *neig, int maxiter, real_t *U, int ldu,
                  real_t *sigma, real_t *bnd, real_t *V, int ldv, real_t tolin,
                  real_t *doption, int *ioption)
{
    int kmax = *dim;
    int mn = imax(m, n);
    int lwork = m + n + 10 * kmax + 5 * kmax * kmax + 5 +
                imax(imax(3 * kmax * kmax + 4 * kmax + 4, 16 * mn),
                     imax(mn, 4 * kmax + 4));
    int liwork = 8 * kmax;
    real_t *work = calloc((size_t)lwork, sizeof(real_t));
    int *iwork = calloc((size_t)liwork, sizeof(int));
    int info = 0;
    real_t dparm[1] = {0};
    int iparm[1] = {0};
    LANSVD_IRL(&which, &jobu, &jobv, &m, &n, dim, &p, neig, &maxiter, ref_aprod,
               U, &ldu, sigma, bnd, V, &ldv, &tolin, work, &lwork, iwork,
               &liwork, doption, ioption, &info, dparm, iparm, 1, 1, 1);
    free(work);
    free(iwork);
    return info;
}
