// Scalar control of one Lanczos bidiagonalization step (DLANBPRO,
// double/dlanbpro.F:283-546) as a phase machine over a plain state struct.
//
// The same __host__ __device__ functions run either in single-workgroup
// control kernels on the GPU (device-resident steps, no host round trip) or
// on the host (rare paths and the host executor); both see identical,
// deterministic reductions, so both make identical decisions.  Arrays are
// 1-based as in the reference (element 0 unused).
#pragma once

#include <cmath>

#ifdef __HIPCC__
#include "pk_common.h"
#else
// Host-only compilation (the control logic is plain C++).
#define __host__
#define __device__
#define __forceinline__ inline
namespace pk {
constexpr int kMaxCols = 512;
}
#endif

namespace pk {

// Largest step j whose arrays the control kernels stage in LDS; later steps
// (Lanczos dimensions above it) run the same control code on the arrays in
// device memory.
constexpr int kCtlLdsJ = 500;

// An omega-recurrence update a control stage left to run before its
// remainder (the control kernels spread it over the workgroup).
enum Pending : int { PEND_NONE = 0, PEND_NU = 1, PEND_MU = 2 };

enum Phase : int {
  // phases a device chain executes
  PH_V_SPMV = 1,  // alpha_j v_j = A^T u_j - beta_j v_{j-1}   (dlanbpro.F:288-296)
  PH_V_ELR = 2,   // extended local reorthogonalization of v_j (:301-316)
  PH_V_REO = 3,   // iterated CGS of v_j against V(:, intervals) (:348-367)
  PH_U_SPMV = 4,  // beta_{j+1} u_{j+1} = A v_j - alpha_j u_j (:420-424)
  PH_U_ELR = 5,   // (:429-443)
  PH_U_REO = 6,   // (:471-498)
  // rare paths, executed on the host only
  PH_V_INVARIANT = 20,  // alpha below anorm*epsn: new random direction (:372-409)
  PH_V_TINY = 21,       // |alpha| below the safe minimum: stepwise scaling
  PH_V_ZERO = 22,       // five CGS passes failed: v_j = 0 (dreorth.F:96-98)
  PH_U_INVARIANT = 23,  // (:503-539)
  PH_U_TINY = 24,
  PH_U_ZERO = 25,
  PH_DONE = 100
};

// Scalar state of a DLANBPRO call.  16-byte aligned, size a multiple of 16:
// control kernels stage it in LDS with 16-byte loads.  The arrays of the
// call (mu, nu, B, INDEX, the CGS column list) live in a separate buffer
// sized for the engine's Lanczos dimension (CtlArrays).
struct alignas(16) LbdCtl {
  // parameters of the call
  int m, n, k, elr;
  double delta, eta, eps, epsn, epsn2, sfmin, kappa, fudge;
  // step state
  int j, phase, full_reorth, force_reorth, ierr;
  double alpha, beta, anorm, amax, numax, mumax, rnorm;
  int elr_i;
  double elr_s;       // coefficient of the pending ELR update
  int reo_try, reo_k, reo_ncols, reo_nd;
  double reo_norm0;   // ||w|| before the current CGS pass
  int vscale, uscale; // pending 1/alpha, 1/beta scalings of v_j, u_{j+1}
  double vscale_f, uscale_f;
  long long nopx, nreorth, ndot, ncols_touched, nretry;
  long long npass_v, npass_u, ncols_v, ncols_u;  // CGS passes and columns streamed
  long long nwarn;    // steps that set IERR = j > 0 (near invariance, :406-409,536-539)
  int pending;        // an omega update the step still has to run (Pending)
  int cap;            // capacity of the arrays: the largest k of the engine
};

// Layout of the array buffer of capacity cap (1-based arrays, element 0
// unused): mu, nu, B1 = B(:,1), B2 = B(:,2) as doubles, then INDEX and the
// 0-based CGS column list as ints; every segment starts on 16 bytes.
struct CtlArrays {
  __host__ __device__ static int dlen(int cap) { return (cap + 4) & ~1; }      // >= cap + 3
  __host__ __device__ static int ilen(int cap) { return (2 * cap + 8) & ~3; }  // >= 2 cap + 5
  __host__ __device__ static int clen(int cap) { return (cap + 4) & ~3; }      // >= cap + 1
  __host__ __device__ static long long bytes(int cap) {
    return 32LL * dlen(cap) + 4LL * (ilen(cap) + clen(cap));
  }
  __host__ __device__ static double* mu(void* b, int cap) { return static_cast<double*>(b); }
  __host__ __device__ static double* nu(void* b, int cap) {
    return static_cast<double*>(b) + dlen(cap);
  }
  __host__ __device__ static double* B1(void* b, int cap) {
    return static_cast<double*>(b) + 2 * dlen(cap);
  }
  __host__ __device__ static double* B2(void* b, int cap) {
    return static_cast<double*>(b) + 3 * dlen(cap);
  }
  __host__ __device__ static int* index(void* b, int cap) {
    return reinterpret_cast<int*>(static_cast<double*>(b) + 4 * dlen(cap));
  }
  __host__ __device__ static int* cols(void* b, int cap) { return index(b, cap) + ilen(cap); }
};

// The state the control functions work on: the scalars plus pointers to
// the arrays (host memory for the host executor, device memory for the
// control kernels past the LDS window).
struct LbdCtlRef : LbdCtl {
  double *mu, *nu, *B1, *B2;
  int *index, *cols;
  __host__ __device__ void bind(void* arrays) {
    mu = CtlArrays::mu(arrays, cap);
    nu = CtlArrays::nu(arrays, cap);
    B1 = CtlArrays::B1(arrays, cap);
    B2 = CtlArrays::B2(arrays, cap);
    index = CtlArrays::index(arrays, cap);
    cols = CtlArrays::cols(arrays, cap);
  }
};

// ---- scalar helpers (double/dlanbpro.F:555-723, LAPACK DLAPY2) ----
// Contraction of a*b+c into FMA is disabled so host and device round alike.

__host__ __device__ inline double hd_lapy2(double x, double y) {
#pragma clang fp contract(off)
  const double xa = fabs(x), ya = fabs(y);
  const double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
  if (z == 0.0) return w;
  return w * sqrt(1.0 + (z / w) * (z / w));
}

__host__ __device__ inline double hd_sign(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }
__host__ __device__ inline double hd_max(double a, double b) { return a > b ? a : b; }

// dset_mu (dlanbpro.F:555-577)
__host__ __device__ inline void hd_set_mu(int k, double* mu, const int* index, double val) {
  for (int i = 1; index[i] <= k && index[i] > 0; i += 2)
    for (int j = index[i]; j <= index[i + 1]; ++j) mu[j] = val;
}

// dcompute_int (dlanbpro.F:581-624)
__host__ __device__ inline void hd_compute_int(const double* mu, int j, double delta, double eta,
                                               int* index) {
  if (delta < eta) return;
  int ip = 0, i = 0, k, s;
  index[1] = 0;
  while (i < j) {
    for (k = i + 1; k <= j; ++k)
      if (fabs(mu[k]) > delta) break;
    if (k > j) break;
    const int lo = i > 1 ? i : 1;
    for (s = k; s >= lo; --s)
      if (fabs(mu[s]) < eta) break;
    index[++ip] = s + 1;
    for (i = s + 1; i <= j; ++i)
      if (fabs(mu[i]) < eta) break;
    index[++ip] = i - 1;
  }
  index[++ip] = j + 1;
}

// dupdate_mu (dlanbpro.F:628-679)
__host__ __device__ inline void hd_update_mu(double* mumax, double* mu, const double* nu, int j,
                                             const double* alpha, const double* beta,
                                             double anorm, double eps1) {
#pragma clang fp contract(off)
  double d;
  if (j == 1) {
    d = eps1 * (hd_lapy2(alpha[j], beta[j]) + alpha[1]) + eps1 * anorm;
    mu[1] = eps1 / beta[1];
    *mumax = fabs(mu[1]);
  } else {
    mu[1] = alpha[1] * nu[1] - alpha[j] * mu[1];
    d = eps1 * (hd_lapy2(alpha[j], beta[j]) + alpha[1]) + eps1 * anorm;
    mu[1] = (mu[1] + hd_sign(d, mu[1])) / beta[j];
    *mumax = fabs(mu[1]);
    for (int k = 2; k <= j - 1; ++k) {
      mu[k] = alpha[k] * nu[k] + beta[k - 1] * nu[k - 1] - alpha[j] * mu[k];
      d = eps1 * (hd_lapy2(alpha[j], beta[j]) + hd_lapy2(alpha[k], beta[k - 1])) + eps1 * anorm;
      mu[k] = (mu[k] + hd_sign(d, mu[k])) / beta[j];
      *mumax = hd_max(*mumax, fabs(mu[k]));
    }
    mu[j] = beta[j - 1] * nu[j - 1];
    d = eps1 * (hd_lapy2(alpha[j], beta[j]) + hd_lapy2(alpha[j], beta[j - 1])) + eps1 * anorm;
    mu[j] = (mu[j] + hd_sign(d, mu[j])) / beta[j];
    *mumax = hd_max(*mumax, fabs(mu[j]));
  }
  mu[j + 1] = 1.0;
}

// dupdate_nu (dlanbpro.F:684-723)
__host__ __device__ inline void hd_update_nu(double* numax, const double* mu, double* nu, int j,
                                             const double* alpha, const double* beta,
                                             double anorm, double eps1) {
#pragma clang fp contract(off)
  if (j > 1) {
    *numax = 0.0;
    for (int k = 1; k <= j - 1; ++k) {
      nu[k] = beta[k] * mu[k + 1] + alpha[k] * mu[k] - beta[j - 1] * nu[k];
      const double d =
          eps1 * (hd_lapy2(alpha[k], beta[k]) + hd_lapy2(alpha[j], beta[j - 1])) + eps1 * anorm;
      nu[k] = (nu[k] + hd_sign(d, nu[k])) / alpha[j];
      *numax = hd_max(*numax, fabs(nu[k]));
    }
    nu[j] = 1.0;
  }
}

// Column list of INDEX's intervals for a CGS against columns 1..k
// (dcgs, dreorth.F:158-206); nd = sum of interval lengths (ndot).  The list
// holds at most k <= cap columns.
template <class C>
__host__ __device__ inline void hd_build_cols(C& c, int k) {
  int nc = 0, nd = 0;
  for (int i = 1; c.index[i] <= k && c.index[i] > 0; i += 2) {
    const int p = c.index[i], q = c.index[i + 1];
    nd += q - p + 1;
    for (int col = p; col <= q; ++col) c.cols[nc++] = col - 1;
  }
  c.reo_k = k;
  c.reo_ncols = nc;
  c.reo_nd = nd;
  c.reo_try = 0;
}

// ---------------------------------------------------------------- V side --
template <class C>
__host__ __device__ inline void ctl_v_post(C& c) {
#pragma clang fp contract(off)
  const int j = c.j;
  if (c.alpha < c.anorm * c.epsn && j < c.k) {
    c.phase = PH_V_INVARIANT;
    return;
  } else if (j > 1 && !c.full_reorth && j < c.k && (c.delta * c.alpha < c.anorm * c.eps)) {
    c.ierr = j;
    c.nwarn++;
  }
  c.B1[j] = c.alpha;
  c.vscale = 0;
  if (c.alpha != 0.0) {
    if (fabs(c.alpha) >= c.sfmin) {
      c.vscale = 1;
      c.vscale_f = 1.0 / c.alpha;
    } else {
      c.phase = PH_V_TINY;
      return;
    }
  }
  c.phase = PH_U_SPMV;
}

template <class C>
__host__ __device__ inline void ctl_v_after_reorth(C& c) {
  hd_set_mu(c.j - 1, c.nu, c.index, c.eps);
  c.numax = c.eta;
  c.force_reorth = !c.force_reorth;
  ctl_v_post(c);
}

// Remainder of the nu step after the omega update (dlanbpro.F:321-367).
template <class C>
__host__ __device__ inline void ctl_v_nu_post(C& c) {
  const int j = c.j;
  if ((c.full_reorth || c.numax > c.delta || c.force_reorth) && c.alpha != 0.0) {
    if (c.full_reorth || c.eta == 0.0) {
      c.index[1] = 1;
      c.index[2] = j - 1;
      c.index[3] = j;
    } else if (!c.force_reorth) {
      hd_compute_int(c.nu, j - 1, c.delta, c.eta, c.index);
    }
    if (j - 1 <= 0) {  // dreorth returns at once for k <= 0 (dreorth.F:79)
      ctl_v_after_reorth(c);
      return;
    }
    hd_build_cols(c, j - 1);
    c.reo_norm0 = c.alpha;
    c.phase = PH_V_REO;
    return;
  }
  ctl_v_post(c);
}

// nu step (dlanbpro.F:318-320): the omega update is left pending.
template <class C>
__host__ __device__ inline void ctl_v_nu(C& c) {
  if (!c.full_reorth && c.alpha != 0.0) {
    c.pending = PEND_NU;
    return;
  }
  ctl_v_nu_post(c);
}

template <class C>
__host__ __device__ inline void ctl_v_fin(C& c) {
#pragma clang fp contract(off)
  const int j = c.j;
  const double* B1 = c.B1;
  const double* B2 = c.B2;
  c.B1[j] = c.alpha;
  c.amax = hd_max(c.amax, c.alpha);
  double a1, b1;
  if (j == 2) {
    a1 = B2[1] / c.amax;
    a1 = c.fudge * c.amax * sqrt((B1[1] / c.amax) * (B1[1] / c.amax) + a1 * a1 + B1[2] / c.amax * a1);
  } else {
    a1 = B1[j - 1] / c.amax;
    b1 = B2[j - 1] / c.amax;
    a1 = c.fudge * c.amax * sqrt(a1 * a1 + b1 * b1 + a1 * B2[j - 2] / c.amax + B1[j] / c.amax * b1);
  }
  c.anorm = hd_max(c.anorm, a1);
  ctl_v_nu(c);
}

// After the fused SpMV A^T u_j - beta v_{j-1}: s2 = ||v_j||^2, sd = v_{j-1}.v_j.
template <class C>
__host__ __device__ inline void ctl_v_spmv_s(C& c, double s2, double sd) {
#pragma clang fp contract(off)
  c.nopx++;
  c.uscale = 0;
  c.alpha = sqrt(s2);
  if (c.j == 1) {
    c.anorm = hd_max(c.anorm, c.fudge * c.alpha);
    ctl_v_nu(c);
    return;
  }
  if (c.elr > 0 && c.alpha < c.kappa * c.beta) {
    c.elr_i = 0;
    c.elr_s = sd;
    c.phase = PH_V_ELR;
    return;
  }
  ctl_v_fin(c);
}

// After v_j -= elr_s v_{j-1}: ns2 = ||v_j||^2, nsd = v_{j-1}.v_j.
template <class C>
__host__ __device__ inline void ctl_v_elr_s(C& c, double ns2, double nsd) {
#pragma clang fp contract(off)
  const double s = c.elr_s;
  if (c.beta != 0.0) {
    c.beta = c.beta + s;
    c.B2[c.j - 1] = c.beta;
  }
  const double sn = sqrt(ns2);
  c.elr_i++;
  if (sn >= c.kappa * c.alpha || c.elr_i >= c.elr) {
    c.nu[c.j - 1] = c.eps;
    c.alpha = sn;
    ctl_v_fin(c);
  } else {
    c.alpha = sn;
    c.elr_s = nsd;
  }
}

// After one CGS pass of v_j: s2 = ||v_j||^2 (dreorth.F:85-99).
template <class C>
__host__ __device__ inline void ctl_v_reo_s(C& c, double s2) {
#pragma clang fp contract(off)
  c.ndot += c.reo_nd + c.reo_k;
  c.ncols_touched += c.reo_ncols;
  c.npass_v++;
  c.ncols_v += c.reo_ncols;
  if (c.reo_try > 0) c.nretry++;
  c.reo_try++;
  const double nrm = sqrt(s2);
  if (nrm > c.kappa * c.reo_norm0) {
    c.alpha = nrm;
    c.nreorth++;
    ctl_v_after_reorth(c);
  } else if (c.reo_try >= 5) {
    c.phase = PH_V_ZERO;
  } else {
    c.reo_norm0 = nrm;
  }
}

// ---------------------------------------------------------------- U side --
template <class C>
__host__ __device__ inline void ctl_u_post(C& c) {
#pragma clang fp contract(off)
  const int j = c.j;
  if (c.beta < c.anorm * c.epsn && j < c.k) {
    c.phase = PH_U_INVARIANT;
    return;
  } else if (!c.full_reorth && j < c.k && (c.delta * c.beta < c.anorm * c.eps)) {
    c.ierr = j;
    c.nwarn++;
  }
  c.B2[j] = c.beta;
  c.uscale = 0;
  if (c.beta != 0.0 && c.beta != 1.0) {
    if (fabs(c.beta) >= c.sfmin) {
      c.uscale = 1;
      c.uscale_f = 1.0 / c.beta;
    } else {
      c.phase = PH_U_TINY;
      return;
    }
  }
  c.rnorm = c.beta;
  c.j = j + 1;
  c.phase = c.j > c.k ? PH_DONE : PH_V_SPMV;
}

template <class C>
__host__ __device__ inline void ctl_u_after_reorth(C& c) {
  hd_set_mu(c.j, c.mu, c.index, c.eps);
  c.mumax = c.eta;
  c.force_reorth = !c.force_reorth;
  ctl_u_post(c);
}

// Remainder of the mu step after the omega update (dlanbpro.F:449-498).
template <class C>
__host__ __device__ inline void ctl_u_mu_post(C& c) {
  const int j = c.j;
  if ((c.full_reorth || c.mumax > c.delta || c.force_reorth) && c.beta != 0.0) {
    if (c.full_reorth || c.eta == 0.0) {
      c.index[1] = 1;
      c.index[2] = j;
      c.index[3] = j + 1;
    } else if (!c.force_reorth) {
      hd_compute_int(c.mu, j, c.delta, c.eta, c.index);
    } else {
      for (int i = 1; i <= 2 * j + 1; ++i)
        if (c.index[i] == j) {
          c.index[i] = j + 1;
          break;
        }
    }
    hd_build_cols(c, j);
    c.reo_norm0 = c.beta;
    c.phase = PH_U_REO;
    return;
  }
  ctl_u_post(c);
}

// mu step (dlanbpro.F:446-448): the omega update is left pending.
template <class C>
__host__ __device__ inline void ctl_u_mu(C& c) {
  if (!c.full_reorth && c.beta != 0.0) {
    c.pending = PEND_MU;
    return;
  }
  ctl_u_mu_post(c);
}

template <class C>
__host__ __device__ inline void ctl_u_fin(C& c) {
#pragma clang fp contract(off)
  const int j = c.j;
  c.B2[j] = c.beta;
  c.amax = hd_max(c.amax, c.beta);
  double a1;
  if (j <= 1) {
    a1 = hd_lapy2(c.B1[1], c.B2[1]);
  } else {
    a1 = c.B1[j] / c.amax;
    a1 = c.amax * sqrt(a1 * a1 + (c.B2[j] / c.amax) * (c.B2[j] / c.amax) + a1 * c.B2[j - 1] / c.amax);
  }
  c.anorm = hd_max(c.anorm, a1);
  ctl_u_mu(c);
}

template <class C>
__host__ __device__ inline void ctl_u_spmv_s(C& c, double s2, double sd) {
#pragma clang fp contract(off)
  c.nopx++;
  c.vscale = 0;
  c.beta = sqrt(s2);
  if (c.elr > 0 && c.beta < c.kappa * c.alpha) {
    c.elr_i = 0;
    c.elr_s = sd;
    c.phase = PH_U_ELR;
    return;
  }
  ctl_u_fin(c);
}

template <class C>
__host__ __device__ inline void ctl_u_elr_s(C& c, double ns2, double nsd) {
#pragma clang fp contract(off)
  const double s = c.elr_s;
  if (c.alpha != 0.0) {
    c.alpha = c.alpha + s;
    c.B1[c.j] = c.alpha;
  }
  const double sn = sqrt(ns2);
  c.elr_i++;
  if (sn >= c.kappa * c.beta || c.elr_i >= c.elr) {
    c.mu[c.j] = c.eps;
    c.beta = sn;
    ctl_u_fin(c);
  } else {
    c.beta = sn;
    c.elr_s = nsd;
  }
}

template <class C>
__host__ __device__ inline void ctl_u_reo_s(C& c, double s2) {
#pragma clang fp contract(off)
  c.ndot += c.reo_nd + c.reo_k;
  c.ncols_touched += c.reo_ncols;
  c.npass_u++;
  c.ncols_u += c.reo_ncols;
  if (c.reo_try > 0) c.nretry++;
  c.reo_try++;
  const double nrm = sqrt(s2);
  if (nrm > c.kappa * c.reo_norm0) {
    c.beta = nrm;
    c.nreorth++;
    ctl_u_after_reorth(c);
  } else if (c.reo_try >= 5) {
    c.phase = PH_U_ZERO;
  } else {
    c.reo_norm0 = nrm;
  }
}

// Runs a pending omega update serially, then the rest of its step.
template <class C>
__host__ __device__ inline void ctl_resume(C& c) {
  if (c.pending == PEND_NU) {
    c.pending = PEND_NONE;
    hd_update_nu(&c.numax, c.mu, c.nu, c.j, c.B1, c.B2, c.anorm, c.epsn2);
    ctl_v_nu_post(c);
  } else if (c.pending == PEND_MU) {
    c.pending = PEND_NONE;
    hd_update_mu(&c.mumax, c.mu, c.nu, c.j, c.B1, c.B2, c.anorm, c.epsn2);
    ctl_u_mu_post(c);
  }
}

// Control stages of one step.  The *_s forms stop at a pending omega update
// (the control kernels run it in parallel); these complete the stage.
// After the fused SpMV A^T u_j - beta v_{j-1}: s2 = ||v_j||^2, sd = v_{j-1}.v_j.
template <class C>
__host__ __device__ inline void ctl_v_spmv(C& c, double s2, double sd) {
  ctl_v_spmv_s(c, s2, sd);
  ctl_resume(c);
}
// After v_j -= elr_s v_{j-1}: ns2 = ||v_j||^2, nsd = v_{j-1}.v_j.
template <class C>
__host__ __device__ inline void ctl_v_elr(C& c, double ns2, double nsd) {
  ctl_v_elr_s(c, ns2, nsd);
  ctl_resume(c);
}
// After one CGS pass of v_j: s2 = ||v_j||^2.
template <class C>
__host__ __device__ inline void ctl_v_reo(C& c, double s2) {
  ctl_v_reo_s(c, s2);
  ctl_resume(c);
}
template <class C>
__host__ __device__ inline void ctl_u_spmv(C& c, double s2, double sd) {
  ctl_u_spmv_s(c, s2, sd);
  ctl_resume(c);
}
template <class C>
__host__ __device__ inline void ctl_u_elr(C& c, double ns2, double nsd) {
  ctl_u_elr_s(c, ns2, nsd);
  ctl_resume(c);
}
template <class C>
__host__ __device__ inline void ctl_u_reo(C& c, double s2) {
  ctl_u_reo_s(c, s2);
  ctl_resume(c);
}

// Gate of a device-chain kernel: runs only while the control state is in
// phase `phase` of step `j` (and, if flag is set, while *flag != 0).
struct Gate {
  const LbdCtl* c = nullptr;
  int phase = 0, j = 0;
  int flag = 0;  // 0 none, 1 vscale, 2 uscale
  __device__ __forceinline__ bool off() const {
    if (!c) return false;
    if (c->phase != phase || c->j != j) return true;
    if (flag == 1) return c->vscale == 0;
    if (flag == 2) return c->uscale == 0;
    return false;
  }
};

}  // namespace pk
