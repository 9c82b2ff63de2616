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
 * Built four times: -DPREC_D links against the double tree, -DPREC_S against
 * the single tree, -DPREC_Z against complex16/, -DPREC_C against complex8/.
 * Exported names carry the precision letter (ref_d_*, ref_s_*, ref_z_*,
 * ref_c_*).  The complex trees take APROD with TRANSA = 'c' for Y = A^H*X
 * (complex16/zlanbpro.F:297) and a complex workspace ZWORK(LZWRK) beside
 * the real one (complex16/zlansvd.F:1-3).
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
#elif defined(PREC_Z)
typedef double real_t;
#define CPLX 1
#define P(name) ref_z_##name
#define LANSVD zlansvd_
#define LANSVD_IRL zlansvd_irl_
#elif defined(PREC_C)
typedef float real_t;
#define CPLX 1
#define P(name) ref_c_##name
#define LANSVD clansvd_
#define LANSVD_IRL clansvd_irl_
#else
#error "define PREC_D, PREC_S, PREC_Z or PREC_C"
#endif

/* Element of the bases and of the operator: real, or (re, im) pairs laid
 * out as Fortran COMPLEX. */
#ifdef CPLX
typedef struct { real_t re, im; } elem_t;
#else
typedef real_t elem_t;
#endif

typedef void (*aprod_fn)(const char *, const int *, const int *, const elem_t *,
                         elem_t *, elem_t *, int *, size_t);

/* Fortran entry points of the reference (flang: lower case + '_', hidden
 * CHARACTER lengths appended as size_t).  Argument order: double/dlansvd.F:1-3
 * and double/dlansvd_irl.F:1-3. */
#ifdef CPLX
/* complex16/zlansvd.F:1-3, complex16/zlansvd_irl.F:1-3 (and the complex8
 * twins): ZWORK(LZWRK) follows WORK(LWORK). */
#define ZWORK_ARGS elem_t *zwork, const int *lzwrk,
#define ZWORK_PASS zwork, &lzwrk,
#else
#define ZWORK_ARGS
#define ZWORK_PASS
#endif
void LANSVD(const char *jobu, const char *jobv, const int *m, const int *n,
            int *k, const int *kmax, aprod_fn aprod, elem_t *U, const int *ldu,
            real_t *sigma, real_t *bnd, elem_t *V, const int *ldv,
            const real_t *tolin, real_t *work, const int *lwork, ZWORK_ARGS
            int *iwork, const int *liwork, real_t *doption, int *ioption,
            int *info, elem_t *dparm, int *iparm, size_t ljobu, size_t ljobv);
void LANSVD_IRL(const char *which, const char *jobu, const char *jobv,
                const int *m, const int *n, int *dim, const int *p, int *neig,
                const int *maxiter, aprod_fn aprod, elem_t *U, const int *ldu,
                real_t *sigma, real_t *bnd, elem_t *V, const int *ldv,
                const real_t *tolin, real_t *work, const int *lwork, ZWORK_ARGS
                int *iwork, const int *liwork, real_t *doption, int *ioption,
                int *info, elem_t *dparm, int *iparm, size_t lwhich,
                size_t ljobu, size_t ljobv);

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
    const elem_t *va;
    const int *rpt, *cit;
    const elem_t *vat;
    int64_t napply;
} op;

/* y = M x for CSR M; conj != 0 conjugates M's values (A^H from the stored
 * A^T).  Sums in double. */
static void csr_apply(int rows, const int *rp, const int *ci, const elem_t *va,
                      const elem_t *x, elem_t *y, int conj)
{
    (void)conj;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rows; ++i) {
#ifdef CPLX
        double sr = 0.0, si = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            const double ar = va[p].re, ai = conj ? -(double)va[p].im : (double)va[p].im;
            const double xr = x[ci[p]].re, xi = x[ci[p]].im;
            sr += ar * xr - ai * xi;
            si += ar * xi + ai * xr;
        }
        y[i].re = (real_t)sr;
        y[i].im = (real_t)si;
#else
        double s = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) s += (double)va[p] * (double)x[ci[p]];
        y[i] = (real_t)s;
#endif
    }
}

static void ref_aprod(const char *transa, const int *m, const int *n,
                      const elem_t *x, elem_t *y, elem_t *dparm, int *iparm,
                      size_t ltransa)
{
    (void)m; (void)n; (void)dparm; (void)iparm; (void)ltransa;
    op.napply++;
    if (transa[0] == 'n' || transa[0] == 'N')
        csr_apply(op.m, op.rp, op.ci, op.va, x, y, 0);
    else  /* 't' (real trees) or 'c' (complex trees) */
        csr_apply(op.n, op.rpt, op.cit, op.vat, x, y, transa[0] == 'c' || transa[0] == 'C');
}

void P(set_operator)(int m, int n, const int *rp, const int *ci,
                     const elem_t *va, const int *rpt, const int *cit,
                     const elem_t *vat)
{
    op.m = m; op.n = n;
    op.rp = rp; op.ci = ci; op.va = va;
    op.rpt = rpt; op.cit = cit; op.vat = vat;
    op.napply = 0;
}

void P(aprod)(char trans, const elem_t *x, elem_t *y)
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
              elem_t *U, int ldu, real_t *sigma, real_t *bnd, elem_t *V,
              int ldv, real_t tolin, real_t *doption, int *ioption)
{
    int mn = imax(m, n);
    int lwork = m + n + 9 * kmax + 5 * kmax * kmax + 4 +
                imax(3 * kmax * kmax + 4 * kmax + 4, 16 * mn);
    int liwork = 8 * kmax;
    real_t *work = calloc((size_t)lwork, sizeof(real_t));
    int *iwork = calloc((size_t)liwork, sizeof(int));
    int info = 0;
    elem_t dparm[1];
    int iparm[1] = {0};
    memset(dparm, 0, sizeof dparm);
#ifdef CPLX
    /* ZWORK: MAX(M,N) for DLANBPRO's start vector (complex16/zlanbpro.F:
     * 83-84) and at least one row block per thread for ZDGEMM_OVWR_LEFT
     * (complex16/zgemm_ovwr.F:17-19). */
    int lzwrk = 16 * mn + 64 * (kmax + 2);
    elem_t *zwork = calloc((size_t)lzwrk, sizeof(elem_t));
#endif
    LANSVD(&jobu, &jobv, &m, &n, k, &kmax, ref_aprod, U, &ldu, sigma, bnd, V,
           &ldv, &tolin, work, &lwork, ZWORK_PASS iwork, &liwork, doption,
           ioption, &info, dparm, iparm, 1, 1);
#ifdef CPLX
    free(zwork);
#endif
    free(work);
    free(iwork);
    return info;
}

/* Restarted driver with workspace sized per double/dlansvd_irl.F:64-75. */
int P(lansvd_irl)(char which, char jobu, char jobv, int m, int n, int *dim,
                  int p, int *neig, int maxiter, elem_t *U, int ldu,
                  real_t *sigma, real_t *bnd, elem_t *V, int ldv, real_t tolin,
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
    elem_t dparm[1];
    int iparm[1] = {0};
    memset(dparm, 0, sizeof dparm);
#ifdef CPLX
    int lzwrk = 16 * mn + 64 * (kmax + 2);
    elem_t *zwork = calloc((size_t)lzwrk, sizeof(elem_t));
#endif
    LANSVD_IRL(&which, &jobu, &jobv, &m, &n, dim, &p, neig, &maxiter, ref_aprod,
               U, &ldu, sigma, bnd, V, &ldv, &tolin, work, &lwork, ZWORK_PASS
               iwork, &liwork, doption, ioption, &info, dparm, iparm, 1, 1, 1);
#ifdef CPLX
    free(zwork);
#endif
    free(work);
    free(iwork);
    return info;
}
