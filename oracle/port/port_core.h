/*
 * Test infrastructure only: the body of the plain-C restatement, compiled
 * once per precision (port_d.c, port_s.c).  Macros expected:
 *   R        real type (double / float)
 *   P(x)     symbol prefix (pkp_d_##x / pkp_s_##x)
 *   EPS_R    unit roundoff as LAPACK xLAMCH('E') returns it (2^-53 / 2^-24)
 *   SFMIN_R  xLAMCH('S')
 *   GEMV_ GEMM_ BDSQR_ BDSDC_   Fortran BLAS/LAPACK names
 *   SQRT ABS  math functions of the precision
 * Each routine names the reference routine (file:line) it restates.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "pk_port.h"

void GEMV_(const char *t, const int *m, const int *n, const R *alpha,
           const R *A, const int *lda, const R *x, const int *incx,
           const R *beta, R *y, const int *incy, size_t lt);
void GEMM_(const char *ta, const char *tb, const int *m, const int *n,
           const int *k, const R *alpha, const R *A, const int *lda,
           const R *B, const int *ldb, const R *beta, R *Cm, const int *ldc,
           size_t l1, size_t l2);
void BDSQR_(const char *uplo, const int *n, const int *ncvt, const int *nru,
            const int *ncc, R *d, R *e, R *vt, const int *ldvt, R *u,
            const int *ldu, R *c, const int *ldc, R *work, int *info,
            size_t l);
void BDSDC_(const char *uplo, const char *compq, const int *n, R *d, R *e,
            R *u, const int *ldu, R *vt, const int *ldvt, R *q, int *iq,
            R *work, int *iwork, int *info, size_t l1, size_t l2);

static pkp_stats st;

void P(get_stats)(pkp_stats *out) { *out = st; }
void P(clear_stats)(void) { memset(&st, 0, sizeof st); }

#define COL(A, ld, j) ((A) + (size_t)((j) - 1) * (size_t)(ld)) /* 1-based column */
#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

/* ---------------- level-1 helpers (double/dblasext.F:6-196) ------------- */
static R nrm2(int n, const R *x)
{
    /* dblasext.F:6-31 (OpenMP build): sqrt of the plain sum of squares. */
    R s = 0;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    return n > 0 ? SQRT(s) : (R)0;
}
static R dot(int n, const R *x, const R *y)
{
    R s = 0;
    for (int i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}
static void axpy(int n, R a, const R *x, R *y)
{
    for (int i = 0; i < n; ++i) y[i] = a * x[i] + y[i];
}
static void scal(int n, R a, R *x)
{
    for (int i = 0; i < n; ++i) x[i] = a * x[i];
}

/* ---------------- CSR operator ---------------- */
static struct {
    int m, n;
    const int *rp, *ci, *rpt, *cit;
    const R *va, *vat;
} csr;

void P(set_csr)(int m, int n, const int *rp, const int *ci, const R *va,
                const int *rpt, const int *cit, const R *vat)
{
    csr.m = m; csr.n = n;
    csr.rp = rp; csr.ci = ci; csr.va = va;
    csr.rpt = rpt; csr.cit = cit; csr.vat = vat;
}

void P(csr_aprod)(char trans, const R *x, R *y, void *ctx)
{
    (void)ctx;
    int rows = (trans == 'n' || trans == 'N') ? csr.m : csr.n;
    const int *rp = (trans == 'n' || trans == 'N') ? csr.rp : csr.rpt;
    const int *ci = (trans == 'n' || trans == 'N') ? csr.ci : csr.cit;
    const R *va = (trans == 'n' || trans == 'N') ? csr.va : csr.vat;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rows; ++i) {
        double s = 0;
        for (int p = rp[i]; p < rp[i + 1]; ++p) s += (double)va[p] * (double)x[ci[p]];
        y[i] = (R)s;
    }
}

/* ---------------- DLARNV(2, (1,3,5,7), n, x) ----------------
 * Restates double/Lapack_Util/dlarnv.f + dlaruv.f for IDIST=2 from the
 * fixed seed used by double/dgetu0.F:47-50: a 48-bit multiplicative
 * congruential generator with multiplier a = 33952834046453; the g-th
 * output (0-based) is 2*((s0 * a^(g+1) mod 2^48) / 2^48) - 1 with
 * s0 = 1*2^36 + 3*2^24 + 5*2^12 + 7.  The 48-bit value is converted the
 * way dlaruv.f does, from four 12-bit limbs in Horner form. */
static void larnv_pm1_from(long long first, int n, R *x)
{
    const uint64_t mask = (1ULL << 48) - 1, a = 33952834046453ULL;
    uint64_t s = (1ULL << 36) + (3ULL << 24) + (5ULL << 12) + 7ULL;
    const R r = (R)1 / (R)4096;
    /* skip the first `first` elements: xLARNV advances its seed by the
     * count it generates (dlaruv.f sets ISEED to the last state) */
    for (long long g = 0; g < first; ++g) s = (uint64_t)(((unsigned __int128)s * a) & mask);
    for (int g = 0; g < n; ++g) {
        s = (uint64_t)(((unsigned __int128)s * a) & mask);
        R i1 = (R)(s >> 36), i2 = (R)((s >> 24) & 4095);
        R i3 = (R)((s >> 12) & 4095), i4 = (R)(s & 4095);
        R u = r * (i1 + r * (i2 + r * (i3 + r * i4)));
        x[g] = (R)2 * u - (R)1;
    }
}

void P(larnv_pm1)(int n, R *x) { larnv_pm1_from(0, n, x); }

/* DLAPY2 (double/Lapack_Util/dlapy2.f): sqrt(x^2+y^2) without overflow. */
R P(lapy2)(R x, R y)
{
    R xa = ABS(x), ya = ABS(y);
    R w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0) return w;
    return w * SQRT((R)1 + (z / w) * (z / w));
}

/* DLARTG (double/Lapack_Util/dlartg.f, LAPACK 3.0): plane rotation with
 * r = sqrt(f^2+g^2), scaled by powers of the base when f,g are extreme;
 * sign fixed so that c > 0 when |f| > |g|. */
void P(lartg)(R f, R g, R *c, R *s, R *r)
{
    if (g == 0) { *c = 1; *s = 0; *r = f; return; }
    if (f == 0) { *c = 0; *s = 1; *r = g; return; }
    const R safmin = SFMIN_R, epsl = EPS_R;
    /* safmn2 = base^int(log(safmin/eps)/log(base)/2) */
    const R safmn2 = (R)pow(2.0, (double)(int)(log((double)(safmin / epsl)) / log(2.0) / 2.0));
    const R safmx2 = (R)1 / safmn2;
    R f1 = f, g1 = g, rr, cs, sn;
    R scale = MAXI(ABS(f1), ABS(g1));
    int count = 0;
    if (scale >= safmx2) {
        do { count++; f1 *= safmn2; g1 *= safmn2; scale = MAXI(ABS(f1), ABS(g1)); }
        while (scale >= safmx2);
        rr = SQRT(f1 * f1 + g1 * g1); cs = f1 / rr; sn = g1 / rr;
        for (int i = 0; i < count; ++i) rr *= safmx2;
    } else if (scale <= safmn2) {
        do { count++; f1 *= safmx2; g1 *= safmx2; scale = MAXI(ABS(f1), ABS(g1)); }
        while (scale <= safmn2);
        rr = SQRT(f1 * f1 + g1 * g1); cs = f1 / rr; sn = g1 / rr;
        for (int i = 0; i < count; ++i) rr *= safmn2;
    } else {
        rr = SQRT(f1 * f1 + g1 * g1); cs = f1 / rr; sn = g1 / rr;
    }
    if (ABS(f) > ABS(g) && cs < 0) { cs = -cs; sn = -sn; rr = -rr; }
    *c = cs; *s = sn; *r = rr;
}

static void rot(int n, R *x, R *y, R c, R s)
{
    for (int i = 0; i < n; ++i) {
        R t = c * x[i] + s * y[i];
        y[i] = c * y[i] - s * x[i];
        x[i] = t;
    }
}

/* ---------------- omega recurrence (double/dlanbpro.F:555-723) ----------
 * Arrays mu, nu, alpha, beta and index are 1-based here (element 0 unused)
 * to keep the recurrences legible against the reference. */

/* dset_mu (dlanbpro.F:555-577): mu(p:q) = val over every interval. */
void P(set_mu)(int k, R *mu, const int *index, R val)
{
    int i = 1;
    while (index[i] <= k && index[i] > 0) {
        for (int j = index[i]; j <= index[i + 1]; ++j) mu[j] = val;
        i += 2;
    }
}

/* dcompute_int (dlanbpro.F:581-624): intervals around |mu| > delta,
 * widened while |mu| >= eta, terminated by j+1. */
void P(compute_int)(const R *mu, int j, R delta, R eta, int *index)
{
    if (delta < eta) return;
    int ip = 0, i = 0, k, s;
    index[1] = 0;
    while (i < j) {
        for (k = i + 1; k <= j; ++k)
            if (ABS(mu[k]) > delta) break;
        if (k > j) break;
        int lo = MAXI(i, 1);
        for (s = k; s >= lo; --s)
            if (ABS(mu[s]) < eta) break;
        /* loop exhausted => s = lo - 1 (Fortran DO variable past the end) */
        index[++ip] = s + 1;
        for (i = s + 1; i <= j; ++i)
            if (ABS(mu[i]) < eta) break;
        index[++ip] = i - 1;
    }
    index[++ip] = j + 1;
}

static R fsign(R a, R b) { return b >= 0 ? ABS(a) : -ABS(a); }

/* dupdate_mu (dlanbpro.F:628-679). */
void P(update_mu)(R *mumax, R *mu, const R *nu, int j, const R *alpha,
                  const R *beta, R anorm, R eps1)
{
    R d;
    if (j == 1) {
        d = eps1 * (P(lapy2)(alpha[j], beta[j]) + alpha[1]) + eps1 * anorm;
        mu[1] = eps1 / beta[1];
        *mumax = ABS(mu[1]);
    } else {
        mu[1] = alpha[1] * nu[1] - alpha[j] * mu[1];
        d = eps1 * (P(lapy2)(alpha[j], beta[j]) + alpha[1]) + eps1 * anorm;
        mu[1] = (mu[1] + fsign(d, mu[1])) / beta[j];
        *mumax = ABS(mu[1]);
        for (int k = 2; k <= j - 1; ++k) {
            mu[k] = alpha[k] * nu[k] + beta[k - 1] * nu[k - 1] - alpha[j] * mu[k];
            d = eps1 * (P(lapy2)(alpha[j], beta[j]) + P(lapy2)(alpha[k], beta[k - 1])) +
                eps1 * anorm;
            mu[k] = (mu[k] + fsign(d, mu[k])) / beta[j];
            *mumax = MAXI(*mumax, ABS(mu[k]));
        }
        mu[j] = beta[j - 1] * nu[j - 1];
        d = eps1 * (P(lapy2)(alpha[j], beta[j]) + P(lapy2)(alpha[j], beta[j - 1])) +
            eps1 * anorm;
        mu[j] = (mu[j] + fsign(d, mu[j])) / beta[j];
        *mumax = MAXI(*mumax, ABS(mu[j]));
    }
    mu[j + 1] = 1;
}

/* dupdate_nu (dlanbpro.F:684-723). */
void P(update_nu)(R *numax, const R *mu, R *nu, int j, const R *alpha,
                  const R *beta, R anorm, R eps1)
{
    if (j > 1) {
        *numax = 0;
        for (int k = 1; k <= j - 1; ++k) {
            nu[k] = beta[k] * mu[k + 1] + alpha[k] * mu[k] - beta[j - 1] * nu[k];
            R d = eps1 * (P(lapy2)(alpha[k], beta[k]) + P(lapy2)(alpha[j], beta[j - 1])) +
                  eps1 * anorm;
            nu[k] = (nu[k] + fsign(d, nu[k])) / alpha[j];
            *numax = MAXI(*numax, ABS(nu[k]));
        }
        nu[j] = 1;
    }
}

/* ---------------- bidiagonal helpers (double/dbsvd.F) ---------------- */

/* drefinebounds (dbsvd.F:162-231): merge bounds of clustered Ritz values,
 * then sharpen by the gap theorem.  theta, bound 0-based. */
void P(refinebounds)(int n, int k, const R *theta, R *bound, R tol, R eps34)
{
    if (k <= 1) return;
    for (int i = 1; i <= k; ++i) {
        for (int l = -1; l <= 1; l += 2) {
            if ((l == 1 && i < k) || (l == -1 && i > 1)) {
                if (ABS(theta[i - 1] - theta[i + l - 1]) < eps34 * theta[i - 1]) {
                    if (bound[i - 1] > tol && bound[i + l - 1] > tol) {
                        bound[i + l - 1] = P(lapy2)(bound[i - 1], bound[i + l - 1]);
                        bound[i - 1] = 0;
                    }
                }
            }
        }
    }
    for (int i = 1; i <= k; ++i) {
        if (i < k || k == n) {
            R gap;
            if (i == 1) {
                gap = ABS(theta[0] - theta[1]) - MAXI(bound[0], bound[1]);
            } else if (i == n) {
                gap = ABS(theta[i - 2] - theta[i - 1]) - MAXI(bound[i - 2], bound[i - 1]);
            } else {
                gap = ABS(theta[i - 1] - theta[i]) - MAXI(bound[i - 1], bound[i]);
                R g2 = ABS(theta[i - 2] - theta[i - 1]) - MAXI(bound[i - 2], bound[i - 1]);
                gap = gap < g2 ? gap : g2;
            }
            if (gap > bound[i - 1]) bound[i - 1] = bound[i - 1] * (bound[i - 1] / gap);
        }
    }
}

/* dbdqr (dbsvd.F:87-157): QR of the (n+1) x n lower bidiagonal matrix by
 * Givens rotations; [.. c1 c2]' = Q'[0..0 1]'; Qt = Q^T if jobq = 'Y'.
 * D, E 0-based; Qt column-major with leading dimension ldq. */
void P(bdqr)(int ignorelast, char jobq, int n, R *D, R *E, R *c1, R *c2,
             R *Qt, int ldq)
{
    if (n < 1) return;
    int wantq = (jobq == 'Y' || jobq == 'y');
#define QT(i, j) Qt[(size_t)((j) - 1) * ldq + (i) - 1]
    if (wantq) {
        for (int j = 1; j <= n + 1; ++j) {
            for (int i = 1; i <= n + 1; ++i) QT(i, j) = 0;
            QT(j, j) = 1;
        }
    }
    R cs, sn, r;
    int i;
    for (i = 1; i <= n - 1; ++i) {
        P(lartg)(D[i - 1], E[i - 1], &cs, &sn, &r);
        D[i - 1] = r;
        E[i - 1] = sn * D[i];
        D[i] = cs * D[i];
        if (wantq) {
            for (int j = 1; j <= i; ++j) {
                QT(i + 1, j) = -sn * QT(i, j);
                QT(i, j) = cs * QT(i, j);
            }
            QT(i, i + 1) = sn;
            QT(i + 1, i + 1) = cs;
        }
    }
    /* i == n here, as the Fortran DO variable after the loop */
    if (!ignorelast) {
        P(lartg)(D[n - 1], E[n - 1], &cs, &sn, &r);
        D[n - 1] = r;
        E[n - 1] = 0;
        *c1 = sn;
        *c2 = cs;
        if (wantq) {
            for (int j = 1; j <= i; ++j) {
                QT(i + 1, j) = -sn * QT(i, j);
                QT(i, j) = cs * QT(i, j);
            }
            QT(i, i + 1) = sn;
            QT(i + 1, i + 1) = cs;
        }
    }
#undef QT
}

/* dbsvdstep (dbsvd.F:5-82): one implicit LQ sweep with shift sigma on the
 * lower bidiagonal (D,E), accumulating the rotations into columns of U, V. */
void P(bsvdstep)(char jobu, char jobv, int m, int n, int k, R sigma, R *D,
                 R *E, R *U, int ldu, R *V, int ldv)
{
    if (k <= 1) return;
    int dou = (jobu == 'y' || jobu == 'Y'), dov = (jobv == 'y' || jobv == 'Y');
    R c, s, r;
    R x = D[0] * D[0] - sigma * sigma;
    R y = E[0] * D[0];
    for (int i = 1; i <= k - 1; ++i) {
        if (i > 1) P(lartg)(x, y, &c, &s, &E[i - 2]);
        else P(lartg)(x, y, &c, &s, &r);
        x = c * D[i - 1] + s * E[i - 1];
        E[i - 1] = -s * D[i - 1] + c * E[i - 1];
        D[i - 1] = x;
        y = s * D[i];
        D[i] = c * D[i];
        if (dou && m > 0) rot(m, COL(U, ldu, i), COL(U, ldu, i + 1), c, s);
        P(lartg)(x, y, &c, &s, &D[i - 1]);
        x = c * E[i - 1] + s * D[i];
        D[i] = -s * E[i - 1] + c * D[i];
        E[i - 1] = x;
        y = s * E[i];
        E[i] = c * E[i];
        if (dov && n > 0) rot(n, COL(V, ldv, i), COL(V, ldv, i + 1), c, s);
    }
    P(lartg)(x, y, &c, &s, &E[k - 2]);
    x = c * D[k - 1] + s * E[k - 1];
    E[k - 1] = -s * D[k - 1] + c * E[k - 1];
    D[k - 1] = x;
    if (dou && m > 0) rot(m, COL(U, ldu, k), COL(U, ldu, k + 1), c, s);
}

/* ---------------- Gram-Schmidt (double/dreorth.F, double/dmgs.risc.F) ---- */

/* dcgs (dreorth.F:106-210): for each interval [p,q] of INDEX (1-based),
 * vnew -= V(:,p:q) * (V(:,p:q)^T vnew). */
static void cgs(int n, int k, const R *V, int ldv, R *vnew, const int *index,
                R *work)
{
    const R one = 1, zero = 0, mone = -1;
    const int inc = 1;
    int i = 1;
    while (index[i] <= k && index[i] > 0) {
        int p = index[i], q = index[i + 1], l = q - p + 1;
        st.ndot += l;
        if (l > 0) {
            GEMV_("T", &n, &l, &one, COL(V, ldv, p), &ldv, vnew, &inc, &zero, work, &inc, 1);
            GEMV_("N", &n, &l, &mone, COL(V, ldv, p), &ldv, work, &inc, &one, vnew, &inc, 1);
        }
        i += 2;
    }
}

/* dmgs (dmgs.risc.F:11-83): column-sequential modified Gram-Schmidt with the
 * reference's software-pipelined update order. */
static void mgs(int n, int k, const R *V, int ldv, R *vnew, const int *index)
{
    if (k <= 0 || n <= 0) return;
    int ib = 1, p = index[ib], q = index[ib + 1];
    while (p <= k && p > 0 && p <= q) {
        st.ndot += q - p + 1;
        R coef = dot(n, COL(V, ldv, p), vnew);
        for (int i = p + 1; i <= q; ++i) {
            const R *vp = COL(V, ldv, i - 1), *vi = COL(V, ldv, i);
            R nc = 0;
            for (int j = 0; j < n; ++j) {
                R vn0 = vnew[j] - coef * vp[j];
                nc = nc + vn0 * vi[j];
                vnew[j] = vn0;
            }
            coef = nc;
        }
        axpy(n, -coef, COL(V, ldv, q), vnew);
        ib += 2;
        p = index[ib];
        q = index[ib + 1];
    }
}

/* dreorth (dreorth.F:5-101): iterate CGS/MGS up to NTRY=5 times until
 * ||vnew'|| > alpha ||vnew||; otherwise vnew = 0. */
void P(reorth)(int n, int k, const R *V, int ldv, R *vnew, R *normvnew,
               const int *index, R alpha, int iflag)
{
    if (k <= 0 || n <= 0) return;
    R *work = malloc(sizeof(R) * (size_t)(k + 1));
    int ok = 0;
    for (int itry = 1; itry <= 5; ++itry) {
        R n0 = *normvnew;
        if (iflag == 1) cgs(n, k, V, ldv, vnew, index, work);
        else mgs(n, k, V, ldv, vnew, index);
        st.ndot += k;
        *normvnew = nrm2(n, vnew);
        if (*normvnew > alpha * n0) { ok = 1; break; }
    }
    if (!ok) {
        *normvnew = 0;
        memset(vnew, 0, sizeof(R) * (size_t)n);
    }
    st.nreorth++;
    free(work);
}

/* dsafescal (dsafescal.F:4-55): x /= alpha, through a scaled path when
 * |alpha| < sfmin (dlascl's stepwise scaling by powers of the base). */
static void safescal(int n, R alpha, R *x)
{
    if (ABS(alpha) >= SFMIN_R) {
        scal(n, (R)1 / alpha, x);
    } else {
        /* dlascl('G', cfrom=alpha, cto=1): multiply by cto/cfrom in safe
         * steps (double/Lapack_Util/dlascl.f). */
        R smlnum = SFMIN_R, bignum = (R)1 / smlnum;
        R cfromc = alpha, ctoc = 1;
        int done = 0;
        while (!done) {
            R cfrom1 = cfromc * smlnum, mul;
            R cto1 = ctoc / bignum;
            if (ABS(cfrom1) > ABS(ctoc) && ctoc != 0) { mul = smlnum; cfromc = cfrom1; }
            else if (ABS(cto1) > ABS(cfromc)) { mul = bignum; ctoc = cto1; }
            else { mul = ctoc / cfromc; done = 1; }
            scal(n, mul, x);
        }
    }
}

/* dgetu0 (dgetu0.F:11-89): random vector (fixed seed) mapped through op(A),
 * then orthogonalized against U(:,1:j); ||A|| estimate = ||op(A)x||/||x||. */
static int getu0(char transa, int m, int n, int j, int ntry, R *u0, R *u0norm,
                 const R *U, int ldu, P(aprod_fn) aprod, void *ctx, int icgs,
                 R *anormest)
{
    int tr = (transa == 'n' || transa == 'N');
    int rsize = tr ? n : m, usize = tr ? m : n;
    R *work = malloc(sizeof(R) * (size_t)rsize);
    int ierr = -1;
    /* kappa here is the single-precision literal 0.717 (dgetu0.F:28-29). */
    const R kap = (R)0.717f;
    /* the seed is set once per call (dgetu0.F:47-50); each try's DLARNV
     * continues the sequence (dgetu0.F:67) */
    for (int itry = 1; itry <= ntry; ++itry) {
        larnv_pm1_from((long long)(itry - 1) * rsize, rsize, work);
        R nrm = nrm2(rsize, work);
        aprod(tr ? 'n' : 't', work, u0, ctx);
        st.nopx++;
        *u0norm = nrm2(usize, u0);
        *anormest = *u0norm / nrm;
        if (j >= 1) {
            int index[4] = {0, 1, j, j + 1};
            P(reorth)(usize, j, U, ldu, u0, u0norm, index, kap, icgs);
        }
        if (*u0norm > 0) { ierr = 0; break; }
    }
    free(work);
    return ierr;
}

/* ---------------- dlanbpro (double/dlanbpro.F:1-549) ----------------
 * alpha, beta are B(:,1), B(:,2) (0-based: alpha[j-1] = B(j,1)).
 * Returns ierr; *k may shrink on an invariant subspace. */
int P(lanbpro)(int m, int n, int k0, int *kp, P(aprod_fn) aprod, void *ctx,
               R *U, int ldu, R *V, int ldv, R *alpha_, R *beta_, R *rnorm,
               R *doption, const int *ioption)
{
    const R one = 1, zero = 0, FUDGE = (R)1.01, kappa = (R)0.717;
    int k = *kp, ierr = 0;
    R eps = EPS_R;
    R eps34 = (R)pow((double)eps, 0.75);
    (void)eps34;
    R epsn = (R)MAXI(m, n) * eps;
    R epsn2 = SQRT((R)MAXI(m, n)) * eps;
    R delta, eta, anorm, anormest, s, a1, b1, amax = 0, alpha, beta;
    int full_reorth, force_reorth, j0, i, j;
    /* 1-based views of B(:,1), B(:,2) */
    R *B1 = alpha_ - 1, *B2 = beta_ - 1;

    delta = doption[0] < zero ? SQRT(eps / (R)k) : doption[0];
    eta = doption[1] < zero ? (R)pow((double)eps, 0.75) / SQRT((R)k) : doption[1];
    full_reorth = (delta <= eta || delta == zero);
    if (doption[2] > zero) anorm = doption[2];
    else if (k0 > 0) {
        anorm = P(lapy2)(B1[1], B2[1]);
        if (anorm <= zero) { ierr = -1; goto done; }
    } else anorm = zero;
    ierr = 0;

    if (*rnorm == zero) {
        getu0('n', m, n, k0, 3, COL(U, ldu, k0 + 1), rnorm, U, ldu, aprod, ctx,
              ioption[0], &anormest);
        anorm = MAXI(anorm, anormest);
    }

    R *mu = calloc((size_t)k + 3, sizeof(R));   /* 1-based */
    R *nu = calloc((size_t)k + 3, sizeof(R));
    int *index = calloc((size_t)2 * k + 4, sizeof(int));
    R *scratch = calloc((size_t)MAXI(m, n) + 1, sizeof(R));

    if (k0 == 0) {
        amax = zero;
        alpha = zero;
        beta = *rnorm;
        force_reorth = 0;
        if (n > m)
            getu0('n', m, n, 0, 1, scratch, &s, U, ldu, aprod, ctx, ioption[0], &anormest);
        else
            getu0('t', m, n, 0, 1, scratch, &s, V, ldv, aprod, ctx, ioption[0], &anormest);
        anorm = MAXI(anorm, FUDGE * anormest);
        j0 = 1;
        if (beta != zero) safescal(m, beta, COL(U, ldu, 1));
        mu[1] = one;
        nu[1] = one;
    } else {
        force_reorth = 1;
        alpha = B1[k0];
        beta = *rnorm;
        if (k0 < k && beta * delta < anorm * eps) {
            full_reorth = 1;
            ierr = k0;
        }
        index[1] = 1; index[2] = k0; index[3] = k0 + 1;
        scal(m, *rnorm, COL(U, ldu, k0 + 1));
        P(reorth)(m, k0, U, ldu, COL(U, ldu, k0 + 1), rnorm, index, kappa, ioption[0]);
        safescal(m, *rnorm, COL(U, ldu, k0 + 1));
        P(set_mu)(k0, mu, index, epsn2);
        P(set_mu)(k0, nu, index, epsn2);
        beta = *rnorm;
        B2[k0] = beta;
        amax = zero;
        for (j = 1; j <= k0; ++j) {
            amax = MAXI(amax, MAXI(B1[j], B2[j]));
            if (j == 1) anorm = MAXI(anorm, FUDGE * alpha);
            else if (j == 2) {
                a1 = B2[1] / amax;
                a1 = FUDGE * amax * SQRT((B1[1] / amax) * (B1[1] / amax) + a1 * a1 + B1[2] / amax * a1);
                anorm = MAXI(anorm, a1);
            } else {
                a1 = B1[j - 1] / amax;
                b1 = B2[j - 1] / amax;
                a1 = FUDGE * amax * SQRT(a1 * a1 + b1 * b1 + a1 * B2[j - 2] / amax + B1[j] / amax * b1);
                anorm = MAXI(anorm, a1);
            }
        }
        j0 = k0 + 1;
    }
    R numax = zero, mumax = zero;

    for (j = j0; j <= k; ++j) {
        /* alpha_j v_j = A^T u_j - beta_j v_{j-1}  (dlanbpro.F:288-316) */
        aprod('t', COL(U, ldu, j), COL(V, ldv, j), ctx);
        st.nopx++;
        if (j == 1) {
            alpha = nrm2(n, COL(V, ldv, j));
            anorm = MAXI(anorm, FUDGE * alpha);
        } else {
            axpy(n, -beta, COL(V, ldv, j - 1), COL(V, ldv, j));
            alpha = nrm2(n, COL(V, ldv, j));
            if (j > 1 && ioption[1] > 0 && alpha < kappa * beta) {
                for (i = 1; i <= ioption[1]; ++i) {
                    s = dot(n, COL(V, ldv, j - 1), COL(V, ldv, j));
                    axpy(n, -s, COL(V, ldv, j - 1), COL(V, ldv, j));
                    if (beta != zero) {
                        beta = beta + s;
                        B2[j - 1] = beta;
                    }
                    s = nrm2(n, COL(V, ldv, j));
                    if (s >= kappa * alpha) break;
                    alpha = s;
                }
                nu[j - 1] = eps;
                alpha = s;
            }
            B1[j] = alpha;
            amax = MAXI(amax, alpha);
            if (j == 2) {
                a1 = B2[1] / amax;
                a1 = FUDGE * amax * SQRT((B1[1] / amax) * (B1[1] / amax) + a1 * a1 + B1[2] / amax * a1);
            } else {
                a1 = B1[j - 1] / amax;
                b1 = B2[j - 1] / amax;
                a1 = FUDGE * amax * SQRT(a1 * a1 + b1 * b1 + a1 * B2[j - 2] / amax + B1[j] / amax * b1);
            }
            anorm = MAXI(anorm, a1);
        }
        /* nu recurrence and reorthogonalization of v_j (dlanbpro.F:340-367) */
        if (!full_reorth && alpha != zero)
            P(update_nu)(&numax, mu, nu, j, B1, B2, anorm, epsn2);
        if ((full_reorth || numax > delta || force_reorth) && alpha != zero) {
            if (full_reorth || eta == zero) {
                index[1] = 1; index[2] = j - 1; index[3] = j;
            } else if (!force_reorth) {
                P(compute_int)(nu, j - 1, delta, eta, index);
            }
            P(reorth)(n, j - 1, V, ldv, COL(V, ldv, j), &alpha, index, kappa, ioption[0]);
            P(set_mu)(j - 1, nu, index, eps);
            numax = eta;
            force_reorth = !force_reorth;
        }
        /* invariant subspace for alpha (dlanbpro.F:372-409) */
        if (alpha < anorm * epsn && j < k) {
            *rnorm = alpha;
            alpha = zero;
            int e2 = getu0('t', m, n, j - 1, 3, COL(V, ldv, j), &alpha, V, ldv, aprod,
                           ctx, ioption[0], &anormest);
            (void)e2;
            if (alpha == zero) {
                k = j - 1;
                ierr = -j;
                goto cleanup;
            } else {
                safescal(n, alpha, COL(V, ldv, j));
                alpha = zero;
                force_reorth = 1;
                if (delta > zero) full_reorth = 0;
            }
        } else if (j > 1 && !full_reorth && j < k && (delta * alpha < anorm * eps)) {
            ierr = j;
        }
        B1[j] = alpha;
        if (alpha != zero) safescal(n, alpha, COL(V, ldv, j));

        /* beta_{j+1} u_{j+1} = A v_j - alpha_j u_j  (dlanbpro.F:420-443) */
        aprod('n', COL(V, ldv, j), COL(U, ldu, j + 1), ctx);
        st.nopx++;
        axpy(m, -alpha, COL(U, ldu, j), COL(U, ldu, j + 1));
        beta = nrm2(m, COL(U, ldu, j + 1));
        if (ioption[1] > 0 && beta < kappa * alpha) {
            for (i = 1; i <= ioption[1]; ++i) {
                s = dot(m, COL(U, ldu, j), COL(U, ldu, j + 1));
                axpy(m, -s, COL(U, ldu, j), COL(U, ldu, j + 1));
                if (alpha != zero) {
                    alpha = alpha + s;
                    B1[j] = alpha;
                }
                s = nrm2(m, COL(U, ldu, j + 1));
                if (s >= kappa * beta) break;
                beta = s;
            }
            mu[j] = eps;
            beta = s;
        }
        B2[j] = beta;
        amax = MAXI(amax, beta);
        if (j <= 1) {
            a1 = P(lapy2)(B1[1], B2[1]);
        } else {
            a1 = B1[j] / amax;
            a1 = amax * SQRT(a1 * a1 + (B2[j] / amax) * (B2[j] / amax) + a1 * B2[j - 1] / amax);
        }
        anorm = MAXI(anorm, a1);

        /* mu recurrence and reorthogonalization of u_{j+1} (dlanbpro.F:451-498) */
        if (!full_reorth && beta != zero)
            P(update_mu)(&mumax, mu, nu, j, B1, B2, anorm, epsn2);
        if ((full_reorth || mumax > delta || force_reorth) && beta != zero) {
            if (full_reorth || eta == zero) {
                index[1] = 1; index[2] = j; index[3] = j + 1;
            } else if (!force_reorth) {
                P(compute_int)(mu, j, delta, eta, index);
            } else {
                for (i = 1; i <= 2 * j + 1; ++i)
                    if (index[i] == j) { index[i] = j + 1; break; }
            }
            P(reorth)(m, j, U, ldu, COL(U, ldu, j + 1), &beta, index, kappa, ioption[0]);
            P(set_mu)(j, mu, index, eps);
            mumax = eta;
            force_reorth = !force_reorth;
        }
        /* invariant subspace for beta (dlanbpro.F:503-539) */
        if (beta < anorm * epsn && j < k) {
            *rnorm = beta;
            beta = zero;
            getu0('n', m, n, j, 3, COL(U, ldu, j + 1), &beta, U, ldu, aprod, ctx,
                  ioption[0], &anormest);
            if (beta == zero) {
                k = j;
                ierr = -j;
                goto cleanup;
            } else {
                safescal(m, beta, COL(U, ldu, j + 1));
                beta = zero;
                force_reorth = 1;
                if (delta > zero) full_reorth = 0;
            }
        } else if (!full_reorth && j < k && (delta * beta < anorm * eps)) {
            ierr = j;
        }
        B2[j] = beta;
        if (beta != zero && beta != one) safescal(m, beta, COL(U, ldu, j + 1));
        *rnorm = beta;
    }
cleanup:
    free(mu); free(nu); free(index); free(scratch);
done:
    doption[2] = anorm;
    *kp = k;
    return ierr;
}

/* ---------------- dgemm_ovwr_left (double/dgemm_ovwr.F:56-89) ----------
 * A(m x n) <- A(:,1:k) * op(B), op(B) k x n, row-blocked through a buffer. */
void P(gemm_ovwr_left)(char transb, int m, int n, int k, R *A, int lda,
                       const R *B, int ldb)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int bs = 512;
    const R one = 1, zero = 0;
    R *buf = malloc(sizeof(R) * (size_t)bs * n);
    char tb[2] = {transb, 0};
    for (int i = 0; i < m; i += bs) {
        int rows = MINI(bs, m - i);
        GEMM_("N", tb, &rows, &n, &k, &one, A + i, &lda, B, &ldb, &zero, buf, &rows, 1, 1);
        for (int j = 0; j < n; ++j)
            memcpy(A + i + (size_t)j * lda, buf + (size_t)j * rows, sizeof(R) * rows);
    }
    free(buf);
}

/* ---------------- dritzvec (double/dritzvec.F:1-200) ---------------- */
static void ritzvec(char which, char jobu, char jobv, int m, int n, int k,
                    int dim, R *D, R *E, R *U, int ldu, R *V, int ldv)
{
    int d1 = dim + 1;
    R *Mt = calloc((size_t)d1 * d1, sizeof(R));
    R *Qt = calloc((size_t)dim * dim, sizeof(R));
    R *Pm = calloc((size_t)dim * dim, sizeof(R));
    R *X = calloc((size_t)d1 * d1, sizeof(R));
    R c1 = 0, c2 = 0, dd[1];
    int id[1], info;
    P(bdqr)(dim == MINI(m, n), jobu, dim, D, E, &c1, &c2, Mt, d1);
    {
        size_t lw = (size_t)3 * dim * dim + 4 * (size_t)dim + 8;
        R *w = calloc(lw, sizeof(R));
        int *iw = calloc((size_t)8 * dim + 8, sizeof(int));
        BDSDC_("U", "I", &dim, D, E, Pm, &dim, Qt, &dim, dd, id, w, iw, &info, 1, 1);
        free(w); free(iw);
    }
    /* X(1:dim, 1:dim+1) = Pm^T * Mt(1:dim, 1:dim+1)  (dritzvec.F:130-131) */
    for (int c = 0; c < d1; ++c)
        for (int r = 0; r < dim; ++r) {
            R s = 0;
            for (int l = 0; l < dim; ++l) s += Pm[(size_t)r * dim + l] * Mt[(size_t)c * d1 + l];
            X[(size_t)c * d1 + r] = s;
        }
    int mstart = (which == 's' || which == 'S') ? dim - k + 1 : 1;
    if (jobu == 'y' || jobu == 'Y')
        P(gemm_ovwr_left)('T', m, k, d1, U, ldu, X + (mstart - 1), d1);
    if (jobv == 'y' || jobv == 'Y')
        P(gemm_ovwr_left)('T', n, k, dim, V, ldv, Qt + (mstart - 1), dim);
    free(Mt); free(Qt); free(Pm); free(X);
}

/* ---------------- dlansvd (double/dlansvd.F:1-291) ---------------- */
int P(lansvd)(char jobu, char jobv, int m, int n, int *kp, int kmax,
              P(aprod_fn) aprod, void *ctx, R *U, int ldu, R *sigma, R *bnd,
              R *V, int ldv, R tolin, R *doption, const int *ioption)
{
    const R one = 1, zero = 0;
    int k = *kp, info = 0, ierr = 0, lapinfo;
    R eps = EPS_R, eps34 = (R)pow((double)eps, 0.75);
    R epsn = (R)MAXI(m, n) * eps / (R)2;
    int lanmax = MINI(MINI(n + 1, m + 1), kmax);
    R tol = MINI(one, MAXI((R)16 * eps, tolin));
    R anorm = zero, rnorm;
    R *B = calloc((size_t)2 * lanmax, sizeof(R));      /* alpha | beta */
    R *B1 = calloc((size_t)2 * lanmax, sizeof(R));
    R *wb = calloc((size_t)lanmax + 2, sizeof(R));
    R *work = calloc((size_t)4 * lanmax + 8, sizeof(R));
    R dummy[2] = {0, 0};
    const int ione = 1, izero = 0;

    rnorm = nrm2(m, U);
    if (rnorm == zero)
        getu0('n', m, n, 0, 1, U, &rnorm, U, ldu, aprod, ctx, ioption[0], &anorm);

    st.nsing = k;
    int neig = 0, jold = 0, j = MINI(k + MAXI(8, k) + 1, lanmax), dj;
    while (neig < k) {
        int jj = j;
        ierr = P(lanbpro)(m, n, jold, &jj, aprod, ctx, U, ldu, V, ldv, B, B + lanmax,
                          &rnorm, doption, ioption);
        j = jj;
        jold = j;
        memcpy(B1, B, sizeof(R) * 2 * lanmax);
        memset(wb, 0, sizeof(R) * (size_t)(j + 1));
        P(bdqr)(j == MINI(m, n), 'N', j, B1, B1 + lanmax, &wb[j - 1], &wb[j], NULL, lanmax + 1);
        BDSQR_("U", &j, &izero, &ione, &izero, B1, B1 + lanmax, dummy, &ione, wb, &ione,
               dummy, &ione, work, &lapinfo, 1);
        st.nbsvd++;
        if (j > 5) anorm = B1[0];
        else anorm = MAXI(anorm, B1[0]);
        for (int i = 0; i < j; ++i) wb[i] = ABS(rnorm * wb[i]);
        P(refinebounds)(MINI(m, n), j, B1, wb, epsn * anorm, eps34);
        for (int i = 0; i < MINI(j, k); ++i) bnd[i] = wb[i];
        int i = 0;
        neig = 0;
        while (i < MINI(j, k)) {
            if (wb[i] <= tol * B1[i]) {
                sigma[neig++] = B1[i];
                i++;
            } else i = k;
        }
        if (ierr < 0) {
            if (j < k) info = j;
            break;
        }
        if (j >= lanmax) {
            if (neig < k) info = -1;
            break;
        }
        if (neig > 1) {
            dj = MINI(j / 2, ((k - neig) * (j - 6)) / (2 * neig + 1));
            dj = MINI(100, MAXI(2, dj));
        } else {
            dj = j / 2;
            dj = MINI(100, MAXI(10, dj));
        }
        j = MINI(j + dj, lanmax);
    }
    if ((neig >= k || info > 0) && (jobu == 'y' || jobu == 'Y' || jobv == 'y' || jobv == 'Y'))
        ritzvec('L', jobu, jobv, m, n, neig, jold, B, B + lanmax, U, ldu, V, ldv);
    *kp = neig;
    st.nlandim = j;
    free(B); free(B1); free(wb); free(work);
    return info;
}

/* ---------------- dlansvd_irl (double/dlansvd_irl.F:1-419) ---------------- */
int P(lansvd_irl)(char which, char jobu, char jobv, int m, int n, int *dimp,
                  int p, int *neigp, int maxiter, P(aprod_fn) aprod, void *ctx,
                  R *U, int ldu, R *sigma, R *bnd, R *V, int ldv, R tolin,
                  R *doption, const int *ioption)
{
    const R one = 1, zero = 0;
    int neig = *neigp, info = 0, ierr = 0, lapinfo;
    int small = (which == 's' || which == 'S');
    R eps = EPS_R, eps34 = (R)pow((double)eps, 0.75);
    R epsn = (R)MAXI(m, n) * eps / (R)2;
    int dim = MINI(MINI(*dimp, n + 1), m + 1);
    *dimp = dim;
    int k = dim - p;
    R tol = MINI(one, MAXI((R)16 * eps, tolin));
    R anorm = zero, rnorm;
    R *al = calloc((size_t)dim, sizeof(R)), *be = calloc((size_t)dim, sizeof(R));
    R *al1 = calloc((size_t)dim, sizeof(R)), *be1 = calloc((size_t)dim, sizeof(R));
    R *wb = calloc((size_t)dim + 2, sizeof(R)), *shift = calloc((size_t)dim, sizeof(R));
    R *Pm = calloc((size_t)(dim + 1) * (dim + 1), sizeof(R));
    R *Qm = calloc((size_t)dim * dim, sizeof(R));
    R *work = calloc((size_t)4 * dim + 8, sizeof(R));
    R dummy[2] = {0, 0};
    const int ione = 1, izero = 0;

    rnorm = nrm2(m, U);
    if (rnorm == zero)
        getu0('n', m, n, 0, 1, U, &rnorm, U, ldu, aprod, ctx, ioption[0], &anorm);

    int iter = 0, nconv = 0, kold = 0;
    while (nconv < neig && iter < maxiter) {
        int kk = dim;
        ierr = P(lanbpro)(m, n, kold, &kk, aprod, ctx, U, ldu, V, ldv, al, be, &rnorm,
                          doption, ioption);
        dim = kk; /* DLANBPRO lowers K (= DIM) at an invariant subspace (dlansvd_irl.F:213) */
        kold = k;
        memcpy(al1, al, sizeof(R) * dim);
        memcpy(be1, be, sizeof(R) * dim);
        memset(wb, 0, sizeof(R) * (size_t)(dim + 1));
        P(bdqr)(dim == MINI(m, n), 'N', dim, al1, be1, &wb[dim - 1], &wb[dim], NULL, dim + 1);
        BDSQR_("U", &dim, &izero, &ione, &izero, al1, be1, dummy, &ione, wb, &ione, dummy,
               &ione, work, &lapinfo, 1);
        st.nbsvd++;
        if (dim > 5) anorm = al1[0];
        else anorm = MAXI(anorm, al1[0]);
        for (int i = 0; i < dim; ++i) wb[i] = ABS(rnorm * wb[i]);
        if (small) P(refinebounds)(MINI(m, n), dim, al1, wb, epsn * anorm, eps34);
        else P(refinebounds)(MINI(m, n), MINI(dim, neig), al1, wb, epsn * anorm, eps34);
        nconv = 0;
        if (small) {
            for (int i = dim - neig + 1; i <= dim; ++i) {
                if (wb[i - 1] <= tol * al1[0]) {
                    sigma[nconv] = al1[i - 1];
                    bnd[nconv] = wb[i - 1];
                    nconv++;
                }
            }
        } else {
            int i = 1;
            while (i <= MINI(dim, neig)) {
                if (wb[i - 1] <= tol * al1[i - 1]) {
                    sigma[nconv] = al1[i - 1];
                    bnd[nconv] = wb[i - 1];
                    nconv++;
                    i++;
                } else i = k + 1;
            }
        }
        if (ierr < 0) {
            if (dim < k) info = dim;
            break;
        }
        if (nconv < neig) {
            /* exact shifts with the relative-gap filter (dlansvd_irl.F:318-344) */
            memset(shift, 0, sizeof(R) * (size_t)(dim - k));
            int nshft = 0;
            if (small) {
                for (int i = 1; i <= k; ++i) {
                    R relgap = al1[i - 1] - wb[i - 1] - al1[dim - neig - 1];
                    shift[nshft++] = relgap > doption[3] * al1[dim - neig - 1] ? al1[i - 1] : al1[0];
                }
            } else {
                for (int i = dim; i >= k + 1; --i) {
                    R relgap = al1[k - 1] - (al1[i - 1] + wb[i - 1]);
                    shift[nshft++] = relgap > doption[3] * al1[k - 1] ? al1[i - 1] : (R)0;
                }
            }
            memset(Pm, 0, sizeof(R) * (size_t)(dim + 1) * (dim + 1));
            memset(Qm, 0, sizeof(R) * (size_t)dim * dim);
            for (int i = 0; i < dim + 1; ++i) Pm[(size_t)i * (dim + 2)] = one;
            for (int i = 0; i < dim; ++i) Qm[(size_t)i * (dim + 1)] = one;
            for (int i = dim; i >= k + 1; --i)
                P(bsvdstep)('y', 'y', dim + 1, dim, i, shift[dim - i], al, be, Pm, dim + 1,
                            Qm, dim);
            P(gemm_ovwr_left)('N', m, k + 1, dim + 1, U, ldu, Pm, dim + 1);
            P(gemm_ovwr_left)('N', n, k, dim, V, ldv, Qm, dim);
            rnorm = be[k - 1];
            st.nrestart++;
        }
        iter++;
    }
    if ((nconv >= neig || info > 0) && (jobu == 'y' || jobu == 'Y' || jobv == 'y' || jobv == 'Y')) {
        memcpy(al1, al, sizeof(R) * dim);
        memcpy(be1, be, sizeof(R) * dim);
        ritzvec(which, jobu, jobv, m, n, nconv, dim, al1, be1, U, ldu, V, ldv);
    }
    *neigp = nconv;
    *dimp = dim;
    st.nlandim = dim;
    free(al); free(be); free(al1); free(be1); free(wb); free(shift);
    free(Pm); free(Qm); free(work);
    return info;
}
