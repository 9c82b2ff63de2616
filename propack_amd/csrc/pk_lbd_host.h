// Host-side pieces of a Lanczos bidiagonalization call shared by the real
// engine (Impl<T>, DLANBPRO / SLANBPRO) and the complex one (ZImpl<R>,
// ZLANBPRO / CLANBPRO): the set-up before the first step, the host
// execution of a step's phases, the outputs, DSAFESCAL and the restart /
// Ritz GEMM.  The real and complex trees of the reference share these
// lines (complex16/zlanbpro.F differs from double/dlanbpro.F only in its
// vector types, in APROD's 'c' and in the extended local
// reorthogonalization), so one copy serves both.
//
// I supplies: e, strm, Real (the real type of the storage precision), the
// basis accessors (U, V, ucol, vcol, scratch), eff_length, resolve_timers,
// and the vector operations scale, zero, getu0, reorth, reo_pass, gemm; and
// the two stages whose arithmetic differs between the trees:
//   spmv_step(c, vside)  the operator application of one side of step c.j
//                        with its control stage (ctl_v_spmv / ctl_u_spmv)
//   elr_step(c, vside)   one extended local reorthogonalization of a side
// Included by pk_engine.hip after its helpers (PK_TRY, PK_HIP, PK_LAUNCH,
// Mach, lapy2, set_mu, stream_wait, set_error).
#pragma once

namespace pk {

// dsafescal (double/dsafescal.F:4-55, complex16/zsafescal.F:4-56): x /=
// alpha; below the safe minimum the factor is applied in the stepwise way
// of dlascl / zlascl (a complex vector scales its real and imaginary parts
// by the same real factors).
template <class I, class P>
int lbd_safescal(I& im, long long L, double alpha, P* x) {
  using R = typename I::Real;
  if (std::fabs(alpha) >= Mach<R>::sfmin) return im.scale(x, L, 1.0 / alpha);
  const double smlnum = Mach<R>::sfmin, bignum = 1.0 / smlnum;
  double cfromc = alpha, ctoc = 1.0;
  bool done = false;
  while (!done) {
    const double cfrom1 = cfromc * smlnum, cto1 = ctoc / bignum;
    double mul;
    if (std::fabs(cfrom1) > std::fabs(ctoc) && ctoc != 0.0) {
      mul = smlnum;
      cfromc = cfrom1;
    } else if (std::fabs(cto1) > std::fabs(cfromc)) {
      mul = bignum;
      ctoc = cto1;
    } else {
      mul = ctoc / cfromc;
      done = true;
    }
    PK_TRY(im.scale(x, L, mul));
  }
  return PK_OK;
}

// Execute the current phase of the step on the host (same control
// functions as the device chain, plus the rare paths).
template <class I>
int lbd_host_phase(I& im, LbdCtlRef& c) {
  Engine* e = im.e;
  const int j = c.j, m = e->m, n = e->n;
  double anormest = 0;
  int e2 = 0;
  if (c.phase >= PH_V_INVARIANT && c.phase <= PH_U_ZERO) e->cnt.rare[c.phase - PH_V_INVARIANT]++;
  switch (c.phase) {
    case PH_V_SPMV:
      PK_TRY(im.spmv_step(c, true));
      break;
    case PH_V_ELR:
      PK_TRY(im.elr_step(c, true));
      break;
    case PH_V_REO: {
      double s2 = 0;
      PK_TRY(im.reo_pass(n, im.V(), e->ldv, im.vcol(j), c, &s2));
      ctl_v_reo(c, s2);
      break;
    }
    case PH_V_INVARIANT:  // dlanbpro.F:372-409
      c.rnorm = c.alpha;
      c.alpha = 0.0;
      PK_TRY(im.getu0('t', j - 1, 3, im.vcol(j), &c.alpha, im.V(), e->ldv, &anormest, &e2));
      c.ierr = e2;
      if (c.alpha == 0.0) {
        c.k = j - 1;
        c.ierr = -j;
        c.phase = PH_DONE;
      } else {
        PK_TRY(lbd_safescal(im, n, c.alpha, im.vcol(j)));
        c.alpha = 0.0;
        c.force_reorth = 1;
        if (c.delta > 0.0) c.full_reorth = 0;
        c.B1[j] = 0.0;
        c.vscale = 0;
        c.phase = PH_U_SPMV;
      }
      break;
    case PH_V_TINY:
      PK_TRY(lbd_safescal(im, n, c.alpha, im.vcol(j)));
      c.vscale = 0;
      c.phase = PH_U_SPMV;
      break;
    case PH_V_ZERO:  // dreorth.F:96-99
      PK_TRY(im.zero(im.vcol(j), n));
      c.alpha = 0.0;
      c.nreorth++;
      ctl_v_after_reorth(c);
      break;
    case PH_U_SPMV:
      PK_TRY(im.spmv_step(c, false));
      break;
    case PH_U_ELR:
      PK_TRY(im.elr_step(c, false));
      break;
    case PH_U_REO: {
      double s2 = 0;
      PK_TRY(im.reo_pass(m, im.U(), e->ldu, im.ucol(j + 1), c, &s2));
      ctl_u_reo(c, s2);
      break;
    }
    case PH_U_INVARIANT:  // dlanbpro.F:503-539
      c.rnorm = c.beta;
      c.beta = 0.0;
      PK_TRY(im.getu0('n', j, 3, im.ucol(j + 1), &c.beta, im.U(), e->ldu, &anormest, &e2));
      c.ierr = e2;
      if (c.beta == 0.0) {
        c.k = j;
        c.ierr = -j;
        c.phase = PH_DONE;
      } else {
        PK_TRY(lbd_safescal(im, m, c.beta, im.ucol(j + 1)));
        c.beta = 0.0;
        c.force_reorth = 1;
        if (c.delta > 0.0) c.full_reorth = 0;
        c.B2[j] = 0.0;
        c.uscale = 0;
        c.rnorm = 0.0;
        c.j = j + 1;
        c.phase = c.j > c.k ? PH_DONE : PH_V_SPMV;
      }
      break;
    case PH_U_TINY:
      PK_TRY(lbd_safescal(im, m, c.beta, im.ucol(j + 1)));
      c.uscale = 0;
      c.rnorm = c.beta;
      c.j = j + 1;
      c.phase = c.j > c.k ? PH_DONE : PH_V_SPMV;
      break;
    case PH_U_ZERO:
      PK_TRY(im.zero(im.ucol(j + 1), m));
      c.beta = 0.0;
      c.nreorth++;
      ctl_u_after_reorth(c);
      break;
    default:
      set_error("lanbpro: unknown phase %d", c.phase);
      return PK_ERR_ARG;
  }
  return PK_OK;
}

// Setup of a DLANBPRO call (double/dlanbpro.F:134-282,
// complex16/zlanbpro.F:138-290): the state of the first step goes to
// e->hctl.  *finished: the call already returned (as the reference does
// for a zero B at k0 > 0).
template <class I>
int lbd_begin(I& im, int k0, int k_in, double* alpha_, double* beta_, double* rnorm,
              double* doption, const int* ioption, int* ierr_out, bool* finished) {
  using R = typename I::Real;
  Engine* e = im.e;
  *finished = false;
  const double one = 1.0, zero = 0.0, FUDGE = 1.01, kappa = 0.717;
  const int m = e->m, n = e->n;
  int k = k_in, ierr = 0;
  const double eps = Mach<R>::eps;
  const double leff = im.eff_length();
  const double epsn = leff * eps;
  const double epsn2 = std::sqrt(leff) * eps;
  double delta, eta, anorm, anormest = 0, s = 0, a1, b1, amax = 0, alpha, beta;
  bool full_reorth, force_reorth;
  int j0;
  if (k > e->kmax || k0 < 0 || k0 > k) {
    set_error("lanbpro: k=%d k0=%d outside the engine's kmax=%d", k, k0, e->kmax);
    return PK_ERR_ARG;
  }
  LbdCtlRef& c = *e->hctl;
  std::memset(static_cast<LbdCtl*>(&c), 0, sizeof(LbdCtl));
  c.cap = e->kmax;
  std::memset(e->harr, 0, static_cast<size_t>(CtlArrays::bytes(e->kmax)));
  double* mu = c.mu;
  double* nu = c.nu;
  int* index = c.index;
  double* B1 = c.B1;  // B(j,1) = B1[j]
  double* B2 = c.B2;  // B(j,2) = B2[j]
  for (int i = 1; i <= k; ++i) {
    B1[i] = alpha_[i - 1];
    B2[i] = beta_[i - 1];
  }

  delta = doption[0] < zero ? std::sqrt(eps / k) : doption[0];
  eta = doption[1] < zero ? std::pow(eps, 0.75) / std::sqrt(static_cast<double>(k)) : doption[1];
  full_reorth = (delta <= eta || delta == zero);
  if (doption[2] > zero) {
    anorm = doption[2];
  } else if (k0 > 0) {
    anorm = lapy2(B1[1], B2[1]);
    if (anorm <= zero) {
      *ierr_out = -1;
      doption[2] = anorm;
      *finished = true;
      return PK_OK;
    }
  } else {
    anorm = zero;
  }
  ierr = 0;

  if (*rnorm == zero) {
    int e2;
    PK_TRY(im.getu0('n', k0, 3, im.ucol(k0 + 1), rnorm, im.U(), e->ldu, &anormest, &e2));
    ierr = e2;
    anorm = std::max(anorm, anormest);
  }

  if (k0 == 0) {
    amax = zero;
    alpha = zero;
    beta = *rnorm;
    force_reorth = false;
    int e2;
    if (n > m)
      PK_TRY(im.getu0('n', 0, 1, im.scratch(1), &s, im.U(), e->ldu, &anormest, &e2));
    else
      PK_TRY(im.getu0('t', 0, 1, im.scratch(1), &s, im.V(), e->ldv, &anormest, &e2));
    ierr = e2;
    anorm = std::max(anorm, FUDGE * anormest);
    j0 = 1;
    if (beta != zero) PK_TRY(lbd_safescal(im, m, beta, im.ucol(1)));
    mu[1] = one;
    nu[1] = one;
  } else {
    force_reorth = true;
    alpha = B1[k0];
    beta = *rnorm;
    if (k0 < k && beta * delta < anorm * eps) {
      full_reorth = true;
      ierr = k0;
    }
    index[1] = 1;
    index[2] = k0;
    index[3] = k0 + 1;
    PK_TRY(im.scale(im.ucol(k0 + 1), m, *rnorm));
    PK_TRY(im.reorth(m, k0, im.U(), e->ldu, im.ucol(k0 + 1), rnorm, index, kappa));
    PK_TRY(lbd_safescal(im, m, *rnorm, im.ucol(k0 + 1)));
    set_mu(k0, mu, index, epsn2);
    set_mu(k0, nu, index, epsn2);
    beta = *rnorm;
    B2[k0] = beta;
    amax = zero;
    for (int j = 1; j <= k0; ++j) {
      amax = std::max(amax, std::max(B1[j], B2[j]));
      if (j == 1) {
        anorm = std::max(anorm, FUDGE * alpha);
      } else if (j == 2) {
        a1 = B2[1] / amax;
        a1 = FUDGE * amax * std::sqrt((B1[1] / amax) * (B1[1] / amax) + a1 * a1 + B1[2] / amax * a1);
        anorm = std::max(anorm, a1);
      } else {
        a1 = B1[j - 1] / amax;
        b1 = B2[j - 1] / amax;
        a1 = FUDGE * amax * std::sqrt(a1 * a1 + b1 * b1 + a1 * B2[j - 2] / amax + B1[j] / amax * b1);
        anorm = std::max(anorm, a1);
      }
    }
    j0 = k0 + 1;
  }

  c.m = m;
  c.n = n;
  c.k = k;
  c.elr = ioption[1];
  c.delta = delta;
  c.eta = eta;
  c.eps = eps;
  c.epsn = epsn;
  c.epsn2 = epsn2;
  c.sfmin = Mach<R>::sfmin;
  c.kappa = kappa;
  c.fudge = FUDGE;
  c.j = j0;
  c.phase = j0 <= k ? PH_V_SPMV : PH_DONE;
  c.full_reorth = full_reorth;
  c.force_reorth = force_reorth;
  c.ierr = ierr;
  c.alpha = alpha;
  c.beta = beta;
  c.anorm = anorm;
  c.amax = amax;
  c.numax = zero;
  c.mumax = zero;
  c.rnorm = *rnorm;
  return PK_OK;
}

// End of the call: outputs in the reference's places (dlanbpro.F:540-548).
template <class I>
int lbd_end(I& im, int k0, int* kp, double* alpha_, double* beta_, double* rnorm,
            double* doption, int* ierr_out) {
  Engine* e = im.e;
  LbdCtlRef& c = *e->hctl;
  const double* B1 = c.B1;
  const double* B2 = c.B2;
  // The 1/beta scaling of u_{k+1} is deferred to the start of the next
  // step (see the chains); the last step of the call applies it here.
  if (c.uscale) {
    PK_TRY(im.scale(im.ucol(c.j), e->m, c.uscale_f));
    c.uscale = 0;
  }

  for (int i = 1; i <= c.k; ++i) {
    alpha_[i - 1] = B1[i];
    beta_[i - 1] = B2[i];
  }
  if (k0 > 0 && k0 <= c.k) beta_[k0 - 1] = B2[k0];
  doption[2] = c.anorm;
  *kp = c.k;
  *ierr_out = c.ierr;
  *rnorm = c.rnorm;
  e->cnt.nopx += c.nopx;
  e->cnt.nreorth += c.nreorth;
  e->cnt.ndot += c.ndot;
  e->cnt.ncols_touched += c.ncols_touched;
  e->cnt.nretry += c.nretry;
  e->cnt.nwarn += c.nwarn;
  PK_HIP(stream_wait(im.strm));
  return im.resolve_timers();
}

// X(:, 1:ncols) = X(:, 1:kdim) * op(B) into the spare buffer, then swap
// (DGEMM_OVWR_LEFT, double/dgemm_ovwr.F:56-89; ZDGEMM_OVWR_LEFT,
// complex16/zgemm_ovwr.F:6-39, has a real B as well).
template <class I>
int lbd_update(I& im, char which, char transb, int kdim, int ncols, const double* B, int ldb) {
  Engine* e = im.e;
  const bool isu = (which == 'u' || which == 'U');
  const bool tb = (transb == 't' || transb == 'T');
  const int maxk = isu ? e->kmax + 1 : e->kmax;
  if (kdim < 1 || ncols < 1 || kdim > maxk || ncols > maxk) {
    set_error("update: kdim=%d ncols=%d outside [1, %d]", kdim, ncols, maxk);
    return PK_ERR_ARG;
  }
  if (ldb < (tb ? ncols : kdim)) {
    set_error("update: ldb=%d too small", ldb);
    return PK_ERR_ARG;
  }
  void** buf = isu ? e->U : e->V;
  int& cur = isu ? e->ucur : e->vcur;
  const long long ld = isu ? e->ldu : e->ldv;
  const long long L = isu ? e->m : e->n;
  if (!buf[1 - cur]) {
    PK_HIP(hipMalloc(&buf[1 - cur], e->tbytes * ld * maxk));
  }
  // The call returns with the GEMM in flight: the host refills hB only
  // after the stream has read its previous contents, and dB is rewritten by
  // a copy that follows the previous GEMM on the same stream.
  if (e->evB) PK_HIP(hipEventSynchronize(e->evB));
  for (int j = 0; j < ncols; ++j)
    for (int l = 0; l < kdim; ++l)
      e->hB[static_cast<size_t>(j) * kdim + l] =
          tb ? B[static_cast<size_t>(l) * ldb + j] : B[static_cast<size_t>(j) * ldb + l];
  PK_HIP(hipMemcpyAsync(e->dB, e->hB, sizeof(double) * kdim * ncols, hipMemcpyHostToDevice,
                        im.strm));
  if (!e->evB) PK_HIP(hipEventCreateWithFlags(&e->evB, hipEventDisableTiming));
  PK_HIP(hipEventRecord(e->evB, im.strm));
  PK_TRY(im.gemm(L, buf[cur], ld, kdim, ncols, buf[1 - cur]));
  cur = 1 - cur;
  return PK_OK;
}

}  // namespace pk
