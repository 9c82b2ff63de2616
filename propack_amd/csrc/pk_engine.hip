// Lanczos bidiagonalization engine (the reference's L4 routine DLANBPRO with
// its L3/L2 helpers DREORTH/DCGS, DGETU0, DSAFESCAL and the pd* level-1
// wrappers) on the GPU.
//
// The bases U, V, the operator and every vector operation live on the
// device.  The scalar control of the step (alpha/beta, the extended local
// reorthogonalization test, the mu/nu omega-recurrence and the choice of
// reorthogonalization intervals) is evaluated in double on the host from
// deterministic device reductions, in exactly the order of the reference.
// Each device reduction that feeds a decision costs one stream
// synchronization; counters record them.
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <algorithm>

#include "pk_internal.h"

namespace pk {

// One message buffer for the process: batched drivers raise errors on
// their host worker threads and the caller reads them on its own thread.
// Writers and readers take the lock; a reader gets a copy in its thread's
// own buffer, which stays valid until that thread's next call.
static char g_err[512] = "";
static std::mutex g_err_lock;

void set_error(const char* fmt, ...) {
  char msg[sizeof g_err];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_err_lock);
  std::memcpy(g_err, msg, sizeof g_err);
}

const char* last_error() {
  static thread_local char copy[sizeof g_err];
  std::lock_guard<std::mutex> lk(g_err_lock);
  std::memcpy(copy, g_err, sizeof copy);
  return copy;
}

#define PK_HIP(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));       \
      return e_ == hipErrorOutOfMemory ? PK_ERR_OOM : PK_ERR_HIP;                        \
    }                                                                                    \
  } while (0)
#define PK_TRY(expr)        \
  do {                      \
    int r_ = (expr);        \
    if (r_ < 0) return r_;  \
  } while (0)
#define PK_LAUNCH(expr)                                                     \
  do {                                                                      \
    if ((expr) < 0) {                                                       \
      set_error("%s:%d launch failed: %s", __FILE__, __LINE__, #expr);      \
      return PK_ERR_HIP;                                                    \
    }                                                                       \
  } while (0)

// Unit roundoff and safe minimum as LAPACK xLAMCH('E') / xLAMCH('S') return
// them for the storage precision (double/Lapack_Util/dlamch.f).
template <typename T>
struct Mach;
template <>
struct Mach<double> {
  static constexpr double eps = 0x1p-53;
  static constexpr double sfmin = 0x1p-1022;
};
template <>
struct Mach<float> {
  static constexpr double eps = 0x1p-24;
  static constexpr double sfmin = 0x1p-126;
};

// DLARNV(IDIST=2) from the fixed seed ISEED=(1,3,5,7) used by DGETU0
// (double/dgetu0.F:47-50): a 48-bit multiplicative congruential sequence
// with multiplier 33952834046453; output g (0-based) is
// 2*((s0*a^(g+1) mod 2^48)/2^48) - 1, the 48-bit value converted from four
// 12-bit limbs in Horner form in the target precision, as xLARUV does.  The
// engine generates it on the device (launch_larnv_pm1); this host loop over
// the same helpers is the C-ABI's pk_larnv_pm1.
template <typename T>
static void larnv_pm1(long long n, T* x) {
  unsigned long long s = kLcgSeed;
  for (long long g = 0; g < n; ++g) {
    s = lcg_mulmod(s, kLcgA);
    x[g] = lcg_pm1<T>(s);
  }
}

// The scalar routines of the step are the __host__ __device__ ones of
// pk_lbd.h, shared with the device control kernels.
static double lapy2(double x, double y) { return hd_lapy2(x, y); }
static void set_mu(int k, double* mu, const int* index, double val) { hd_set_mu(k, mu, index, val); }

// Waits for stream s by polling.  The engine's waits sit between dependent
// host steps of a solve (control state back, small reductions read); a
// blocking wait adds the host thread's wake-up latency to each of them.
static hipError_t stream_wait(hipStream_t s) {
  hipError_t r;
  while ((r = hipStreamQuery(s)) == hipErrorNotReady) __builtin_ia32_pause();
  return r;
}

// Accumulates the pending timed regions of e (the stream has passed them).
void profile_resolve(Engine* e) {
  for (const KTimer& t : e->prof.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      e->prof.ms[t.cls] += ms;
      e->prof.bytes[t.cls] += t.bytes;
      e->prof.flops[t.cls] += t.flops;
      e->prof.count[t.cls] += 1;
    }
    e->prof.pool.push_back(t.a);
    e->prof.pool.push_back(t.b);
  }
  e->prof.pending.clear();
}

}  // namespace pk

#include "pk_lbd_host.h"

namespace pk {

template <typename T>
struct Impl {
  using Real = T;
  Engine* e;
  hipStream_t strm;
  explicit Impl(Engine* eng) : e(eng), strm(eng->stream) {}

  T* U() const { return static_cast<T*>(e->u()); }
  T* V() const { return static_cast<T*>(e->v()); }
  T* ucol(int j) const { return U() + static_cast<long long>(j - 1) * e->ldu; }
  T* vcol(int j) const { return V() + static_cast<long long>(j - 1) * e->ldv; }
  T* scratch(int i) const {
    return static_cast<T*>(e->vec) + static_cast<long long>(i) * std::max(e->m, e->n);
  }

  int read(int offset, int nvals) {
    PK_HIP(hipMemcpyAsync(e->hred + offset, e->red + offset, sizeof(double) * nvals,
                          hipMemcpyDeviceToHost, strm));
    PK_HIP(stream_wait(strm));
    e->cnt.nsync++;
    return resolve_timers();
  }

  // ---- live kernel timing (HIP events on the engine stream) ----
  hipEvent_t ev_get() {
    hipEvent_t ev = nullptr;
    if (!e->prof.pool.empty()) {
      ev = e->prof.pool.back();
      e->prof.pool.pop_back();
    } else if (hipEventCreate(&ev) != hipSuccess) {
      ev = nullptr;
    }
    return ev;
  }
  // Every timed region costs two marker packets (an idle gap of a few
  // microseconds each) on the stream, so only the selected classes are timed.
  hipEvent_t tstart(int cls) {
    if (!e->prof.on || !((e->prof.mask >> cls) & 1u)) return nullptr;
    hipEvent_t a = ev_get();
    if (a) (void)hipEventRecord(a, strm);
    return a;
  }
  void tstop(hipEvent_t a, int cls, double bytes, double flops = 0.0) {
    if (!a) return;
    hipEvent_t b = ev_get();
    if (!b) return;
    (void)hipEventRecord(b, strm);
    e->prof.pending.push_back(KTimer{cls, a, b, bytes, flops});
  }
  // Called after a stream synchronization: every pending pair has completed.
  int resolve_timers() {
    profile_resolve(e);
    return PK_OK;
  }
  // Algorithmic bytes of y = op(A) x [- c yprev] (compulsory traffic: every
  // nonzero and index once, x and y once, yprev once).
  double spmv_bytes(bool tr, bool with_prev) const {
    const double sv = e->op ? static_cast<double>(e->op->val_bytes) : 0.0;
    const double sb = static_cast<double>(e->tbytes);
    const double nnz = e->op ? static_cast<double>(e->op->nnz) : 0.0;
    const double rows = tr ? e->n : e->m, cols = tr ? e->m : e->n;
    return nnz * (sv + 4.0) + (rows + 1) * 4.0 + cols * sb + rows * sb +
           (with_prev ? rows * sb : 0.0);
  }

  // Effective vector length of the rounding model (see Engine::rounding).
  double eff_length() const {
    return e->rounding ? 4.0 : static_cast<double>(std::max(e->m, e->n));
  }

  // ||x|| of a device vector (pdnrm2).
  int nrm2(const T* x, long long L, double* out) {
    if (L <= 0) {
      *out = 0.0;
      return PK_OK;
    }
    const int nb = launch_norm<T>(strm, L, x, e->part);
    PK_LAUNCH(nb);
    PK_LAUNCH(launch_reduce(strm, 1, nb, e->part, e->red, Gate{}));
    PK_TRY(read(0, 1));
    *out = std::sqrt(e->hred[0]);
    return PK_OK;
  }

  // y = op(A) x - c * yprev; *s2 = ||y||^2, *sd = yprev.y (0 without yprev).
  int apply(char trans, const T* x, T* y, const T* yprev, double c, double* s2, double* sd) {
    const bool tr = (trans == 't' || trans == 'T');
    const long long rows_out = tr ? e->n : e->m, rows_in = tr ? e->m : e->n;
    int nb;
    if (e->op) {
      hipEvent_t t0 = tstart(tr ? KC_SPMV_T : KC_SPMV_N);
      nb = csr_apply<T>(e->op, strm, trans, x, y, yprev, sval(c), e->part, Gate{}, e->sw_y,
                        e->sw_l, e->sw_y_len);
      PK_LAUNCH(nb);
      tstop(t0, tr ? KC_SPMV_T : KC_SPMV_N, spmv_bytes(tr, yprev != nullptr),
            2.0 * static_cast<double>(e->op->nnz));
    } else if (e->aprod) {
      PK_TRY(host_apply(trans, x, y, rows_in, rows_out));
      if (yprev)
        nb = launch_axpy_norm_dot<T>(strm, rows_out, y, yprev, sval(c), e->part, Gate{});
      else
        nb = launch_norm<T>(strm, rows_out, y, e->part);
      PK_LAUNCH(nb);
      if (!yprev) {
        PK_LAUNCH(launch_reduce(strm, 1, nb, e->part, e->red, Gate{}));
        PK_TRY(read(0, 1));
        *s2 = e->hred[0];
        *sd = 0.0;
        return PK_OK;
      }
    } else {
      set_error("engine has no operator");
      return PK_ERR_ARG;
    }
    PK_LAUNCH(launch_reduce(strm, 2, nb, e->part, e->red, Gate{}));
    PK_TRY(read(0, 2));
    *s2 = e->hred[0];
    *sd = yprev ? e->hred[1] : 0.0;
    return PK_OK;
  }

  // Host callback: X staged device -> pinned host, Y back to the device.
  int host_apply(char trans, const T* x, T* y, long long rows_in, long long rows_out) {
    const char t = (trans == 't' || trans == 'T') ? 't' : 'n';
    const int m = e->m, n = e->n;
    const bool same = (e->aprod_prec == PK_F64) == (sizeof(T) == 8);
    if (same) {
      PK_HIP(hipMemcpyAsync(e->hx, x, sizeof(T) * rows_in, hipMemcpyDeviceToHost, strm));
      PK_HIP(hipStreamSynchronize(strm));
    } else {
      PK_HIP(hipMemcpyAsync(e->hy, x, sizeof(T) * rows_in, hipMemcpyDeviceToHost, strm));
      PK_HIP(hipStreamSynchronize(strm));
      const T* src = static_cast<const T*>(e->hy);
      if (e->aprod_prec == PK_F64)
        for (long long i = 0; i < rows_in; ++i) static_cast<double*>(e->hx)[i] = src[i];
      else
        for (long long i = 0; i < rows_in; ++i) static_cast<float*>(e->hx)[i] = src[i];
    }
    e->cnt.nsync++;
    if (e->aprod_prec == PK_F64)
      reinterpret_cast<aprod_d_fn>(e->aprod)(&t, &m, &n, static_cast<const double*>(e->hx),
                                             static_cast<double*>(e->hy),
                                             static_cast<double*>(e->dparm), e->iparm, 1);
    else
      reinterpret_cast<aprod_s_fn>(e->aprod)(&t, &m, &n, static_cast<const float*>(e->hx),
                                             static_cast<float*>(e->hy),
                                             static_cast<float*>(e->dparm), e->iparm, 1);
    if (!same) {
      T* dst = static_cast<T*>(e->hx);
      if (e->aprod_prec == PK_F64)
        for (long long i = 0; i < rows_out; ++i)
          dst[i] = static_cast<T>(static_cast<double*>(e->hy)[i]);
      else
        for (long long i = 0; i < rows_out; ++i)
          dst[i] = static_cast<T>(static_cast<float*>(e->hy)[i]);
      PK_HIP(hipMemcpyAsync(y, e->hx, sizeof(T) * rows_out, hipMemcpyHostToDevice, strm));
    } else {
      PK_HIP(hipMemcpyAsync(y, e->hy, sizeof(T) * rows_out, hipMemcpyHostToDevice, strm));
    }
    return PK_OK;
  }

  int scale(T* x, long long L, double a) {
    hipEvent_t t0 = tstart(KC_VEC);
    PK_LAUNCH(launch_scale<T>(strm, L, x, sval(a), Gate{}));
    tstop(t0, KC_VEC, 2.0 * sizeof(T) * L);
    return PK_OK;
  }

  // dsafescal (double/dsafescal.F:4-55)
  int safescal(long long L, double alpha, T* x) { return lbd_safescal(*this, L, alpha, x); }

  // dreorth + dcgs (double/dreorth.F:5-210): iterated block classical
  // Gram-Schmidt of vnew against basis(:, INDEX intervals), up to 5 passes
  // until ||vnew'|| > alpha ||vnew||; otherwise vnew = 0.  INDEX is 1-based.
  int reorth(long long L, int k, T* basis, long long ld, T* vnew, double* normvnew,
             const int* index, double alpha) {
    if (k <= 0 || L <= 0) return PK_OK;
    int ncols = 0;
    long long nd = 0;
    for (int i = 1; index[i] <= k && index[i] > 0; i += 2) {
      const int p = index[i], q = index[i + 1];
      nd += q - p + 1;
      for (int c = p; c <= q; ++c) {
        if (ncols >= e->ccap) {
          set_error("reorthogonalization touches more than %d columns", e->ccap);
          return PK_ERR_NOTSUP;
        }
        e->hcols[ncols++] = c - 1;
      }
    }
    if (ncols > 0)
      PK_HIP(hipMemcpyAsync(e->dcols, e->hcols, sizeof(int) * ncols, hipMemcpyHostToDevice, strm));
    bool ok = false;
    for (int itry = 1; itry <= 5; ++itry) {
      const double n0 = *normvnew;
      e->cnt.ndot += nd + k;
      e->cnt.ncols_touched += ncols;
      if (ncols > 0) {
        const double sb = static_cast<double>(sizeof(T));
        hipEvent_t t0 = tstart(KC_GEMVT);
        const int nb = launch_gemvt<T>(strm, L, basis, ld, e->dcols, ncols, nullptr, vnew, e->part,
                                       Gate{});
        PK_LAUNCH(nb);
        tstop(t0, KC_GEMVT, (ncols + 1.0) * L * sb, 2.0 * ncols * L);
        PK_LAUNCH(launch_reduce(strm, ncols, nb, e->part, e->red, Gate{}));
        hipEvent_t t1 = tstart(KC_GEMV);
        const int nb2 =
            launch_gemv_update<T>(strm, L, basis, ld, e->dcols, ncols, nullptr, e->red, vnew,
                                  e->part, Gate{});
        PK_LAUNCH(nb2);
        tstop(t1, KC_GEMV, (ncols + 2.0) * L * sb, 2.0 * ncols * L);
        PK_LAUNCH(launch_reduce(strm, 1, nb2, e->part, e->red + e->ccap, Gate{}));
        PK_TRY(read(e->ccap, 1));
        *normvnew = std::sqrt(e->hred[e->ccap]);
      } else {
        PK_TRY(nrm2(vnew, L, normvnew));
      }
      if (*normvnew > alpha * n0) {
        ok = true;
        break;
      }
    }
    if (!ok) {
      *normvnew = 0.0;
      PK_LAUNCH(launch_fill_zero_bytes(strm, vnew, sizeof(T) * L));
    }
    e->cnt.nreorth++;
    return PK_OK;
  }

  // dgetu0 (double/dgetu0.F:11-89) writing op(A) x into u0 (device), then
  // orthogonalizing against basis(:, 1:j).  Returns ierr (0 or -1).
  int getu0(char transa, int j, int ntry, T* u0, double* u0norm, T* basis, long long ld,
            double* anormest, int* ierr) {
    const bool tn = (transa == 'n' || transa == 'N');
    const long long rsize = tn ? e->n : e->m, usize = tn ? e->m : e->n;
    T* x = scratch(0);
    // kappa is the single-precision literal 0.717 in dgetu0.F:28-29.
    const double kap = static_cast<double>(0.717f);
    *ierr = -1;
    for (int itry = 1; itry <= ntry; ++itry) {
      // DLARNV advances the seed DGETU0 set once per call (double/dgetu0.F:
      // 47-50,67): try t continues the sequence at element (t-1)*rsize
      PK_LAUNCH(launch_larnv_pm1<T>(strm, rsize, x, (itry - 1) * rsize));
      double nrm;
      PK_TRY(nrm2(x, rsize, &nrm));
      double s2, sd;
      PK_TRY(apply(tn ? 'n' : 't', x, u0, nullptr, 0.0, &s2, &sd));
      e->cnt.nopx++;
      *u0norm = std::sqrt(s2);
      *anormest = *u0norm / nrm;
      if (j >= 1) {
        int index[4] = {0, 1, j, j + 1};
        PK_TRY(reorth(usize, j, basis, ld, u0, u0norm, index, kap));
      }
      if (*u0norm > 0) {
        *ierr = 0;
        break;
      }
    }
    return PK_OK;
  }

  // One CGS pass of vnew against basis(:, c.cols) on the host executor:
  // returns ||vnew'||^2.
  int reo_pass(long long L, const T* basis, long long ld, T* vnew, const LbdCtlRef& c,
               double* s2) {
    const int ncols = c.reo_ncols;
    if (ncols > 0) {
      std::memcpy(e->hcols, c.cols, sizeof(int) * ncols);
      PK_HIP(hipMemcpyAsync(e->dcols, e->hcols, sizeof(int) * ncols, hipMemcpyHostToDevice, strm));
      const double sb = static_cast<double>(sizeof(T));
      hipEvent_t t0 = tstart(KC_GEMVT);
      const int nb = launch_gemvt<T>(strm, L, basis, ld, e->dcols, ncols, nullptr, vnew, e->part,
                                     Gate{});
      PK_LAUNCH(nb);
      tstop(t0, KC_GEMVT, (ncols + 1.0) * L * sb, 2.0 * ncols * L);
      PK_LAUNCH(launch_reduce(strm, ncols, nb, e->part, e->red, Gate{}));
      hipEvent_t t1 = tstart(KC_GEMV);
      const int nb2 = launch_gemv_update<T>(strm, L, basis, ld, e->dcols, ncols, nullptr, e->red,
                                            vnew, e->part, Gate{});
      PK_LAUNCH(nb2);
      tstop(t1, KC_GEMV, (ncols + 2.0) * L * sb, 2.0 * ncols * L);
      PK_LAUNCH(launch_reduce(strm, 1, nb2, e->part, e->red + e->ccap, Gate{}));
      PK_TRY(read(e->ccap, 1));
      *s2 = e->hred[e->ccap];
    } else {
      double nrm;
      PK_TRY(nrm2(vnew, L, &nrm));
      *s2 = nrm * nrm;
    }
    return PK_OK;
  }

  int axpy_read(long long L, T* w, const T* v, double coef, double* ns2, double* nsd) {
    const int nb = launch_axpy_norm_dot<T>(strm, L, w, v, sval(coef), e->part, Gate{});
    PK_LAUNCH(nb);
    PK_LAUNCH(launch_reduce(strm, 2, nb, e->part, e->red, Gate{}));
    PK_TRY(read(0, 2));
    *ns2 = e->hred[0];
    *nsd = e->hred[1];
    return PK_OK;
  }

  // Execute the current phase of the step on the host (lbd_host_phase).
  int host_phase(LbdCtlRef& c) { return lbd_host_phase(*this, c); }

  int zero(T* x, long long L) {
    PK_LAUNCH(launch_fill_zero_bytes(strm, x, sizeof(T) * L));
    return PK_OK;
  }

  // The operator application of one side of step c.j with its control stage
  // (double/dlanbpro.F:288-296 for v, :420-424 for u).
  int spmv_step(LbdCtlRef& c, bool vside) {
    const int j = c.j;
    double s2 = 0, sd = 0;
    if (vside) {
      if (c.uscale) {
        PK_TRY(scale(ucol(j), e->m, c.uscale_f));
        c.uscale = 0;
      }
      PK_TRY(apply('t', ucol(j), vcol(j), j > 1 ? vcol(j - 1) : nullptr, j > 1 ? c.beta : 0.0,
                   &s2, &sd));
      ctl_v_spmv(c, s2, sd);
    } else {
      if (c.vscale) {
        PK_TRY(scale(vcol(j), e->n, c.vscale_f));
        c.vscale = 0;
      }
      PK_TRY(apply('n', vcol(j), ucol(j + 1), ucol(j), c.alpha, &s2, &sd));
      ctl_u_spmv(c, s2, sd);
    }
    return PK_OK;
  }

  // One extended local reorthogonalization of a side (dlanbpro.F:301-316,
  // :429-443).
  int elr_step(LbdCtlRef& c, bool vside) {
    const int j = c.j;
    double s2 = 0, sd = 0;
    if (vside) {
      PK_TRY(axpy_read(e->n, vcol(j), vcol(j - 1), c.elr_s, &s2, &sd));
      ctl_v_elr(c, s2, sd);
    } else {
      PK_TRY(axpy_read(e->m, ucol(j + 1), ucol(j), c.elr_s, &s2, &sd));
      ctl_u_elr(c, s2, sd);
    }
    return PK_OK;
  }

  // Enqueue the fixed kernel chain of step J.  Every kernel is gated on the
  // control state being in its phase of step J; control kernels advance the
  // state.  Nothing here waits for the device.
  int enqueue_step(int J, int elr, int tries) {
    LbdCtl* dc = e->dctl;
    const int* dcl = CtlArrays::cols(e->darr, e->kmax);  // CGS column list (device)
    const int m = e->m, n = e->n;
    const double sb = static_cast<double>(sizeof(T));
    // ---- v_J = A^T u_J - beta v_{J-1} ----
    PK_LAUNCH(launch_scale<T>(strm, m, ucol(J), sptr(&dc->uscale_f), Gate{dc, PH_V_SPMV, J, 2}));
    hipEvent_t t0 = tstart(KC_SPMV_T);
    int nb = csr_apply<T>(e->op, strm, 't', ucol(J), vcol(J), J > 1 ? vcol(J - 1) : nullptr,
                          sptr(&dc->beta), e->part, Gate{dc, PH_V_SPMV, J, 0}, e->sw_y,
                          e->sw_l, e->sw_y_len);
    PK_LAUNCH(nb);
    tstop(t0, KC_SPMV_T, spmv_bytes(true, J > 1), 2.0 * static_cast<double>(e->op->nnz));
    PK_LAUNCH(launch_ctl(strm, dc, e->darr, e->kmax, CTL_V_SPMV, PH_V_SPMV, J, e->part, nb));
    if (J > 1) {
      for (int i = 0; i < elr; ++i) {
        nb = launch_axpy_norm_dot<T>(strm, n, vcol(J), vcol(J - 1), sptr(&dc->elr_s), e->part,
                                     Gate{dc, PH_V_ELR, J, 0});
        PK_LAUNCH(nb);
        PK_LAUNCH(launch_ctl(strm, dc, e->darr, e->kmax, CTL_V_ELR, PH_V_ELR, J, e->part, nb));
      }
      for (int t = 0; t < tries; ++t) {
        const Gate g{dc, PH_V_REO, J, 0};
        hipEvent_t t1 = tstart(KC_GEMVT);
        nb = launch_gemvt<T>(strm, n, V(), e->ldv, dcl, J - 1, &dc->reo_ncols, vcol(J),
                             e->part, g);
        PK_LAUNCH(nb);
        tstop(t1, KC_GEMVT, 0.0);
        PK_LAUNCH(launch_reduce(strm, J - 1, nb, e->part, e->red, g, &dc->reo_ncols));
        hipEvent_t t2 = tstart(KC_GEMV);
        const int nb2 = launch_gemv_update<T>(strm, n, V(), e->ldv, dcl, J - 1,
                                              &dc->reo_ncols, e->red, vcol(J), e->part, g);
        PK_LAUNCH(nb2);
        tstop(t2, KC_GEMV, 0.0);
        PK_LAUNCH(launch_ctl(strm, dc, e->darr, e->kmax, CTL_V_REO, PH_V_REO, J, e->part, nb2));
      }
    }
    // ---- u_{J+1} = A v_J - alpha u_J ----
    PK_LAUNCH(launch_scale<T>(strm, n, vcol(J), sptr(&dc->vscale_f), Gate{dc, PH_U_SPMV, J, 1}));
    hipEvent_t t3 = tstart(KC_SPMV_N);
    nb = csr_apply<T>(e->op, strm, 'n', vcol(J), ucol(J + 1), ucol(J), sptr(&dc->alpha), e->part,
                      Gate{dc, PH_U_SPMV, J, 0}, e->sw_y, e->sw_l, e->sw_y_len);
    PK_LAUNCH(nb);
    tstop(t3, KC_SPMV_N, spmv_bytes(false, true), 2.0 * static_cast<double>(e->op->nnz));
    PK_LAUNCH(launch_ctl(strm, dc, e->darr, e->kmax, CTL_U_SPMV, PH_U_SPMV, J, e->part, nb));
    for (int i = 0; i < elr; ++i) {
      nb = launch_axpy_norm_dot<T>(strm, m, ucol(J + 1), ucol(J), sptr(&dc->elr_s), e->part,
                                   Gate{dc, PH_U_ELR, J, 0});
      PK_LAUNCH(nb);
      PK_LAUNCH(launch_ctl(strm, dc, e->darr, e->kmax, CTL_U_ELR, PH_U_ELR, J, e->part, nb));
    }
    for (int t = 0; t < tries; ++t) {
      const Gate g{dc, PH_U_REO, J, 0};
      hipEvent_t t1 = tstart(KC_GEMVT);
      nb = launch_gemvt<T>(strm, m, U(), e->ldu, dcl, J, &dc->reo_ncols, ucol(J + 1),
                           e->part, g);
      PK_LAUNCH(nb);
      tstop(t1, KC_GEMVT, 0.0);
      PK_LAUNCH(launch_reduce(strm, J, nb, e->part, e->red, g, &dc->reo_ncols));
      hipEvent_t t2 = tstart(KC_GEMV);
      const int nb2 = launch_gemv_update<T>(strm, m, U(), e->ldu, dcl, J, &dc->reo_ncols,
                                            e->red, ucol(J + 1), e->part, g);
      PK_LAUNCH(nb2);
      tstop(t2, KC_GEMV, 0.0);
      PK_LAUNCH(launch_ctl(strm, dc, e->darr, e->kmax, CTL_U_REO, PH_U_REO, J, e->part, nb2));
    }
    (void)sb;
    return PK_OK;
  }

  // Run steps c.j..c.k on the device until done or until the state needs
  // the host (a rare path, or more CGS passes than the chain holds).
  int run_device(LbdCtlRef& c, int elr) {
    const long long pv0 = c.npass_v, pu0 = c.npass_u, cv0 = c.ncols_v, cu0 = c.ncols_u;
    const size_t ab = static_cast<size_t>(CtlArrays::bytes(e->kmax));
    PK_HIP(hipMemcpyAsync(e->dctl, static_cast<LbdCtl*>(&c), sizeof(LbdCtl),
                          hipMemcpyHostToDevice, strm));
    PK_HIP(hipMemcpyAsync(e->darr, e->harr, ab, hipMemcpyHostToDevice, strm));
    for (int J = c.j; J <= c.k; ++J) PK_TRY(enqueue_step(J, elr, e->dev_tries));
    PK_HIP(hipMemcpyAsync(static_cast<LbdCtl*>(&c), e->dctl, sizeof(LbdCtl),
                          hipMemcpyDeviceToHost, strm));
    PK_HIP(hipMemcpyAsync(e->harr, e->darr, ab, hipMemcpyDeviceToHost, strm));
    PK_HIP(stream_wait(strm));
    e->cnt.nsync++;
    // algorithmic bytes of the CGS passes the chain actually ran
    if (e->prof.on) {
      const double sb = static_cast<double>(sizeof(T));
      const double pv = static_cast<double>(c.npass_v - pv0), pu = static_cast<double>(c.npass_u - pu0);
      const double cv = static_cast<double>(c.ncols_v - cv0), cu = static_cast<double>(c.ncols_u - cu0);
      if ((e->prof.mask >> KC_GEMVT) & 1u) {
        e->prof.bytes[KC_GEMVT] += sb * (e->n * (cv + pv) + e->m * (cu + pu));
        e->prof.flops[KC_GEMVT] += 2.0 * (e->n * cv + e->m * cu);
      }
      if ((e->prof.mask >> KC_GEMV) & 1u) {
        e->prof.bytes[KC_GEMV] += sb * (e->n * (cv + 2 * pv) + e->m * (cu + 2 * pu));
        e->prof.flops[KC_GEMV] += 2.0 * (e->n * cv + e->m * cu);
      }
    }
    return resolve_timers();
  }

  // dlanbpro (double/dlanbpro.F:1-549).  alpha_/beta_ are B(:,1), B(:,2).
  int lanbpro(int k0, int* kp, double* alpha_, double* beta_, double* rnorm, double* doption,
              const int* ioption, int* ierr_out) {
    bool finished = false;
    PK_TRY(lanbpro_begin(k0, *kp, alpha_, beta_, rnorm, doption, ioption, ierr_out, &finished));
    if (finished) return PK_OK;
    PK_TRY(lanbpro_loop());
    return lanbpro_end(k0, kp, alpha_, beta_, rnorm, doption, ierr_out);
  }

  // Setup of a DLANBPRO call (lbd_begin).
  int lanbpro_begin(int k0, int k_in, double* alpha_, double* beta_, double* rnorm,
                    double* doption, const int* ioption, int* ierr_out, bool* finished) {
    return lbd_begin(*this, k0, k_in, alpha_, beta_, rnorm, doption, ioption, ierr_out, finished);
  }

  // Steps of the call from the state in e->hctl until it is done: device
  // chains from PH_V_SPMV, host execution of every other phase.
  int lanbpro_loop() {
    LbdCtlRef& c = *e->hctl;
    const bool device = e->mode == 0 && e->op != nullptr;
    while (c.phase != PH_DONE) {
      if (device && c.phase == PH_V_SPMV) {
        const int j_before = c.j;
        PK_TRY(run_device(c, c.elr));
        // A chain always advances at least one phase; if it did not, finish
        // the step on the host rather than enqueue the same chain again.
        if (c.phase != PH_V_SPMV || c.j != j_before) continue;
      }
      // a chain that stopped in a CGS phase ran out of passes: hold two
      // passes in the engine's later chains
      if (device && e->tries_auto && e->dev_tries < 2 &&
          (c.phase == PH_V_REO || c.phase == PH_U_REO))
        e->dev_tries = 2;
      PK_TRY(host_phase(c));
    }
    return PK_OK;
  }

  // End of the call (lbd_end).
  int lanbpro_end(int k0, int* kp, double* alpha_, double* beta_, double* rnorm, double* doption,
                  int* ierr_out) {
    return lbd_end(*this, k0, kp, alpha_, beta_, rnorm, doption, ierr_out);
  }

  // Residuals of the triplets (sigma[i], u_{i+1}, v_{i+1}) held in src's
  // bases (src has the engine's shape, any precision), with this engine's
  // operator: res[2i] = ||A v - s u||, res[2i+1] = ||A^T u - s v||.  The
  // vectors are converted to T, and each residual is the fused epilogue of
  // one SpMV (y = op(A) x - s yprev, ||y||^2 in FP64).
  int residuals(const Engine* src, int k, const double* sigma, double* res) {
    const long long mx = std::max(e->m, e->n);
    if (!e->kres) PK_HIP(hipMalloc(&e->kres, e->tbytes * mx * 3));
    T* x = static_cast<T*>(e->kres);
    T* yp = x + mx;
    T* y = yp + mx;
    PK_HIP(hipStreamSynchronize(src->stream));
    auto load = [&](char which, int col, T* dst) -> int {
      const bool isu = which == 'u';
      const long long L = isu ? e->m : e->n;
      const char* base = static_cast<const char*>(isu ? src->u() : src->v()) +
                         src->tbytes * static_cast<size_t>(col - 1) * (isu ? src->ldu : src->ldv);
      const int r = src->prec == PK_F64
                        ? launch_convert<double, T>(strm, L, reinterpret_cast<const double*>(base), dst)
                        : launch_convert<float, T>(strm, L, reinterpret_cast<const float*>(base), dst);
      return r < 0 ? PK_ERR_HIP : PK_OK;
    };
    for (int i = 0; i < k; ++i) {
      double s2 = 0, sd = 0;
      PK_TRY(load('v', i + 1, x));
      PK_TRY(load('u', i + 1, yp));
      PK_TRY(apply('n', x, y, yp, sigma[i], &s2, &sd));
      res[2 * i] = std::sqrt(s2);
      PK_TRY(apply('t', yp, y, x, sigma[i], &s2, &sd));
      res[2 * i + 1] = std::sqrt(s2);
    }
    return PK_OK;
  }

  // X(:, 1:ncols) = X(:, 1:kdim) * op(B) into the spare buffer, then swap.
  int update(char which, char transb, int kdim, int ncols, const double* B, int ldb) {
    return lbd_update(*this, which, transb, kdim, ncols, B, ldb);
  }

  // C = A(:, 1:kdim) * dB (kdim x ncols) on FP64 matrix cores, L rows.
  int gemm(long long L, const void* A, long long ld, int kdim, int ncols, void* C) {
    hipEvent_t t0 = tstart(KC_GEMM);
    PK_LAUNCH(launch_gemm_mfma<T>(strm, L, static_cast<const T*>(A), ld, kdim, e->dB, kdim,
                                  ncols, static_cast<T*>(C), ld));
    tstop(t0, KC_GEMM, static_cast<double>(sizeof(T)) * L * (kdim + ncols),
          2.0 * static_cast<double>(L) * kdim * ncols);
    return PK_OK;
  }
};

// ---- complex bases: ZLANBPRO / CLANBPRO ----
// Extended local reorthogonalization of the complex tree
// (complex16/zlanbpro.F:316-335 for v, :449-466 for u): the coefficient
// s = v_{j-1}^H v_j is complex, so B keeps its entry (the real tree adds s
// to it); otherwise the stage is the real one.  nsd is the real part of the
// next coefficient.
static void zctl_v_elr(LbdCtlRef& c, double ns2, double nsd) {
  const double sn = std::sqrt(ns2);
  c.elr_i++;
  if (sn >= c.kappa * c.alpha || c.elr_i >= c.elr) {
    c.nu[c.j - 1] = c.eps;
    c.alpha = sn;
    ctl_v_fin(c);
  } else {
    c.alpha = sn;
    c.elr_s = nsd;
  }
  ctl_resume(c);
}

static void zctl_u_elr(LbdCtlRef& c, double ns2, double nsd) {
  const double sn = std::sqrt(ns2);
  c.elr_i++;
  if (sn >= c.kappa * c.beta || c.elr_i >= c.elr) {
    c.mu[c.j] = c.eps;
    c.beta = sn;
    ctl_u_fin(c);
  } else {
    c.beta = sn;
    c.elr_s = nsd;
  }
  ctl_resume(c);
}

template <typename R>
struct CxOf;
template <>
struct CxOf<double> {
  using type = double2;
};
template <>
struct CxOf<float> {
  using type = float2;
};

// The engine over complex bases (PK_C128: R = double, PK_C64: R = float).
// Every step runs on the host executor (the phase machine of pk_lbd.h), one
// synchronization per reduction; norms, real scalings, the random start
// vector and the GEMMs run the real kernels on the 2L-long real view of a
// column, the operator, dot products and Gram-Schmidt the complex kernels
// of pk_cplx.hip.
template <typename R>
struct ZImpl {
  using Real = R;
  using T = typename CxOf<R>::type;
  Engine* e;
  hipStream_t strm;
  explicit ZImpl(Engine* eng) : e(eng), strm(eng->stream) {}

  T* U() const { return static_cast<T*>(e->u()); }
  T* V() const { return static_cast<T*>(e->v()); }
  T* ucol(int j) const { return U() + static_cast<long long>(j - 1) * e->ldu; }
  T* vcol(int j) const { return V() + static_cast<long long>(j - 1) * e->ldv; }
  T* scratch(int i) const {
    return static_cast<T*>(e->vec) + static_cast<long long>(i) * std::max(e->m, e->n);
  }
  static R* re(T* p) { return reinterpret_cast<R*>(p); }
  static const R* re(const T* p) { return reinterpret_cast<const R*>(p); }
  double* hsum() const { return e->red + 2 * e->ccap; }  // slot of the CGS norm

  int read(int offset, int nvals) {
    PK_HIP(hipMemcpyAsync(e->hred + offset, e->red + offset, sizeof(double) * nvals,
                          hipMemcpyDeviceToHost, strm));
    PK_HIP(stream_wait(strm));
    e->cnt.nsync++;
    return resolve_timers();
  }
  int resolve_timers() {
    profile_resolve(e);
    return PK_OK;
  }
  double eff_length() const {
    return e->rounding ? 4.0 : static_cast<double>(std::max(e->m, e->n));
  }
  // live kernel timing: the real engine's event helpers (they touch only the
  // engine's stream and event pool)
  hipEvent_t tstart(int cls) { return Impl<R>(e).tstart(cls); }
  void tstop(hipEvent_t a, int cls, double bytes, double flops) {
    Impl<R>(e).tstop(a, cls, bytes, flops);
  }

  // ||x|| (pdznrm2, complex16/zblasext.F:6-30)
  int nrm2(const T* x, long long L, double* out) {
    if (L <= 0) {
      *out = 0.0;
      return PK_OK;
    }
    const int nb = launch_norm<R>(strm, 2 * L, re(x), e->part);
    PK_LAUNCH(nb);
    PK_LAUNCH(launch_reduce(strm, 1, nb, e->part, e->red, Gate{}));
    PK_TRY(read(0, 1));
    *out = std::sqrt(e->hred[0]);
    return PK_OK;
  }

  int scale(T* x, long long L, double a) {
    PK_LAUNCH(launch_scale<R>(strm, 2 * L, re(x), sval(a), Gate{}));
    return PK_OK;
  }
  int zero(T* x, long long L) {
    PK_LAUNCH(launch_fill_zero_bytes(strm, x, sizeof(T) * L));
    return PK_OK;
  }

  // y = op(A) x - c * yprev (trans 'n' or 'c' = A^H); *s2 = ||y||^2,
  // (*sdr, *sdi) = yprev^H y (0 without yprev).
  int apply(char trans, const T* x, T* y, const T* yprev, double c, double* s2, double* sdr,
            double* sdi) {
    const bool tr = trans != 'n' && trans != 'N';
    const long long rows_out = tr ? e->n : e->m, rows_in = tr ? e->m : e->n;
    int nb;
    if (e->op) {
      // algorithmic bytes: every nonzero and index once, row pointers, x and
      // y once, yprev once
      const double sb = static_cast<double>(sizeof(T)), nnz = static_cast<double>(e->op->nnz);
      const double bytes = nnz * (static_cast<double>(e->op->val_bytes) + 4.0) +
                           (rows_out + 1) * 4.0 + rows_in * sb + rows_out * sb +
                           (yprev ? rows_out * sb : 0.0);
      hipEvent_t t0 = tstart(tr ? KC_SPMV_T : KC_SPMV_N);
      nb = zcsr_apply<R>(e->op, strm, tr ? 'c' : 'n', x, y, yprev, c, e->part, e->sw_y);
      PK_LAUNCH(nb);
      tstop(t0, tr ? KC_SPMV_T : KC_SPMV_N, bytes, 8.0 * nnz);
    } else if (e->aprod) {
      PK_TRY(host_apply(tr ? 'c' : 'n', x, y, rows_in, rows_out));
      if (!yprev) {
        double nrm;
        PK_TRY(nrm2(y, rows_out, &nrm));
        *s2 = nrm * nrm;
        *sdr = *sdi = 0.0;
        return PK_OK;
      }
      nb = launch_zaxpy_norm_dot<R>(strm, rows_out, y, yprev, c, 0.0, e->part);
      PK_LAUNCH(nb);
    } else {
      set_error("engine has no operator");
      return PK_ERR_ARG;
    }
    PK_LAUNCH(launch_reduce(strm, 3, nb, e->part, e->red, Gate{}));
    PK_TRY(read(0, 3));
    *s2 = e->hred[0];
    *sdr = yprev ? e->hred[1] : 0.0;
    *sdi = yprev ? e->hred[2] : 0.0;
    return PK_OK;
  }

  // Host callback (complex16/zlanbpro.F:28-41 with 'n' / 'c'): X staged
  // device -> pinned host, Y back; the callback's precision is the engine's.
  int host_apply(char trans, const T* x, T* y, long long rows_in, long long rows_out) {
    const int m = e->m, n = e->n;
    PK_HIP(hipMemcpyAsync(e->hx, x, sizeof(T) * rows_in, hipMemcpyDeviceToHost, strm));
    PK_HIP(hipStreamSynchronize(strm));
    e->cnt.nsync++;
    typedef void (*zaprod_fn)(const char*, const int*, const int*, const void*, void*, void*,
                              int*, size_t);
    reinterpret_cast<zaprod_fn>(e->aprod)(&trans, &m, &n, e->hx, e->hy, e->dparm, e->iparm, 1);
    PK_HIP(hipMemcpyAsync(y, e->hy, sizeof(T) * rows_out, hipMemcpyHostToDevice, strm));
    return PK_OK;
  }

  // One block classical Gram-Schmidt pass of vnew against basis(:, cols)
  // (zcgs, complex16/zreorth.F:120-205): h = basis^H vnew, vnew -= basis h;
  // returns ||vnew'||^2.
  int cgs_pass(long long L, const T* basis, long long ld, T* vnew, int ncols, double* s2) {
    const double sb = static_cast<double>(sizeof(T));
    hipEvent_t t0 = tstart(KC_GEMVT);
    const int nb = launch_zgemvt<R>(strm, L, basis, ld, e->dcols, ncols, vnew, e->part);
    PK_LAUNCH(nb);
    tstop(t0, KC_GEMVT, (ncols + 1.0) * L * sb, 8.0 * ncols * L);
    PK_LAUNCH(launch_reduce(strm, 2 * ncols, nb, e->part, e->red, Gate{}));
    hipEvent_t t1 = tstart(KC_GEMV);
    const int nb2 = launch_zgemv_update<R>(strm, L, basis, ld, e->dcols, ncols, e->red, vnew,
                                           e->part);
    PK_LAUNCH(nb2);
    tstop(t1, KC_GEMV, (ncols + 2.0) * L * sb, 8.0 * ncols * L);
    PK_LAUNCH(launch_reduce(strm, 1, nb2, e->part, hsum(), Gate{}));
    PK_TRY(read(2 * e->ccap, 1));
    *s2 = e->hred[2 * e->ccap];
    return PK_OK;
  }

  // zreorth (complex16/zreorth.F:5-110): up to 5 passes until
  // ||vnew'|| > alpha ||vnew||; otherwise vnew = 0.  INDEX is 1-based.
  int reorth(long long L, int k, T* basis, long long ld, T* vnew, double* normvnew,
             const int* index, double alpha) {
    if (k <= 0 || L <= 0) return PK_OK;
    int ncols = 0;
    long long nd = 0;
    for (int i = 1; index[i] <= k && index[i] > 0; i += 2) {
      const int p = index[i], q = index[i + 1];
      nd += q - p + 1;
      for (int c = p; c <= q; ++c) {
        if (ncols >= e->ccap) {
          set_error("reorthogonalization touches more than %d columns", e->ccap);
          return PK_ERR_NOTSUP;
        }
        e->hcols[ncols++] = c - 1;
      }
    }
    if (ncols > 0)
      PK_HIP(hipMemcpyAsync(e->dcols, e->hcols, sizeof(int) * ncols, hipMemcpyHostToDevice, strm));
    bool ok = false;
    for (int itry = 1; itry <= 5; ++itry) {
      const double n0 = *normvnew;
      e->cnt.ndot += nd + k;
      e->cnt.ncols_touched += ncols;
      if (ncols > 0) {
        double s2;
        PK_TRY(cgs_pass(L, basis, ld, vnew, ncols, &s2));
        *normvnew = std::sqrt(s2);
      } else {
        PK_TRY(nrm2(vnew, L, normvnew));
      }
      if (*normvnew > alpha * n0) {
        ok = true;
        break;
      }
    }
    if (!ok) {
      *normvnew = 0.0;
      PK_TRY(zero(vnew, L));
    }
    e->cnt.nreorth++;
    return PK_OK;
  }

  // zgetu0 (complex16/zgetu0.F:11-90): a ZLARNV(IDIST=2) vector -- real and
  // imaginary parts drawn in turn from the xLARUV sequence, so the complex
  // vector of length r is the real sequence of length 2r -- through op(A),
  // then orthogonalized against basis(:, 1:j).
  int getu0(char transa, int j, int ntry, T* u0, double* u0norm, T* basis, long long ld,
            double* anormest, int* ierr) {
    const bool tn = (transa == 'n' || transa == 'N');
    const long long rsize = tn ? e->n : e->m, usize = tn ? e->m : e->n;
    T* x = scratch(0);
    const double kap = static_cast<double>(0.717f);  // single-precision literal, zgetu0.F:28-29
    *ierr = -1;
    for (int itry = 1; itry <= ntry; ++itry) {
      PK_LAUNCH(launch_larnv_pm1<R>(strm, 2 * rsize, re(x), (itry - 1) * 2 * rsize));
      double nrm;
      PK_TRY(nrm2(x, rsize, &nrm));
      double s2, sr, si;
      PK_TRY(apply(tn ? 'n' : 'c', x, u0, nullptr, 0.0, &s2, &sr, &si));
      e->cnt.nopx++;
      *u0norm = std::sqrt(s2);
      *anormest = *u0norm / nrm;
      if (j >= 1) {
        int index[4] = {0, 1, j, j + 1};
        PK_TRY(reorth(usize, j, basis, ld, u0, u0norm, index, kap));
      }
      if (*u0norm > 0) {
        *ierr = 0;
        break;
      }
    }
    return PK_OK;
  }

  // One CGS pass of the step (host executor): returns ||vnew'||^2.
  int reo_pass(long long L, const T* basis, long long ld, T* vnew, const LbdCtlRef& c,
               double* s2) {
    const int ncols = c.reo_ncols;
    if (ncols > 0) {
      std::memcpy(e->hcols, c.cols, sizeof(int) * ncols);
      PK_HIP(hipMemcpyAsync(e->dcols, e->hcols, sizeof(int) * ncols, hipMemcpyHostToDevice, strm));
      return cgs_pass(L, basis, ld, vnew, ncols, s2);
    }
    double nrm;
    PK_TRY(nrm2(vnew, L, &nrm));
    *s2 = nrm * nrm;
    return PK_OK;
  }

  // The operator application of one side of step c.j with its control stage
  // (complex16/zlanbpro.F:297-311 for v, :444-452 for u).
  int spmv_step(LbdCtlRef& c, bool vside) {
    const int j = c.j;
    double s2 = 0, sr = 0, si = 0;
    if (vside) {
      if (c.uscale) {
        PK_TRY(scale(ucol(j), e->m, c.uscale_f));
        c.uscale = 0;
      }
      PK_TRY(apply('c', ucol(j), vcol(j), j > 1 ? vcol(j - 1) : nullptr, j > 1 ? c.beta : 0.0,
                   &s2, &sr, &si));
      e->elr_im = si;
      ctl_v_spmv(c, s2, sr);
    } else {
      if (c.vscale) {
        PK_TRY(scale(vcol(j), e->n, c.vscale_f));
        c.vscale = 0;
      }
      PK_TRY(apply('n', vcol(j), ucol(j + 1), ucol(j), c.alpha, &s2, &sr, &si));
      e->elr_im = si;
      ctl_u_spmv(c, s2, sr);
    }
    return PK_OK;
  }

  // v_j -= s v_{j-1} with s = v_{j-1}^H v_j (pzaxpy with -s,
  // complex16/zlanbpro.F:321-323; u side :454-456).
  int elr_step(LbdCtlRef& c, bool vside) {
    const int j = c.j;
    const long long L = vside ? e->n : e->m;
    T* w = vside ? vcol(j) : ucol(j + 1);
    const T* v = vside ? vcol(j - 1) : ucol(j);
    const int nb = launch_zaxpy_norm_dot<R>(strm, L, w, v, c.elr_s, e->elr_im, e->part);
    PK_LAUNCH(nb);
    PK_LAUNCH(launch_reduce(strm, 3, nb, e->part, e->red, Gate{}));
    PK_TRY(read(0, 3));
    e->elr_im = e->hred[2];
    if (vside)
      zctl_v_elr(c, e->hred[0], e->hred[1]);
    else
      zctl_u_elr(c, e->hred[0], e->hred[1]);
    return PK_OK;
  }

  // zlanbpro (complex16/zlanbpro.F:1-553): every phase on the host executor.
  int lanbpro(int k0, int* kp, double* alpha_, double* beta_, double* rnorm, double* doption,
              const int* ioption, int* ierr_out) {
    bool finished = false;
    PK_TRY(lbd_begin(*this, k0, *kp, alpha_, beta_, rnorm, doption, ioption, ierr_out,
                     &finished));
    if (finished) return PK_OK;
    LbdCtlRef& c = *e->hctl;
    while (c.phase != PH_DONE) PK_TRY(lbd_host_phase(*this, c));
    return lbd_end(*this, k0, kp, alpha_, beta_, rnorm, doption, ierr_out);
  }

  // Restart / Ritz GEMM on the real view: a complex column of length L is a
  // real column of length 2L and B is real (complex16/zgemm_ovwr.F:6-39).
  int gemm(long long L, const void* A, long long ld, int kdim, int ncols, void* C) {
    hipEvent_t t0 = tstart(KC_GEMM);
    PK_LAUNCH(launch_gemm_mfma<R>(strm, 2 * L, static_cast<const R*>(A), 2 * ld, kdim, e->dB,
                                  kdim, ncols, static_cast<R*>(C), 2 * ld));
    tstop(t0, KC_GEMM, static_cast<double>(sizeof(T)) * L * (kdim + ncols),
          4.0 * static_cast<double>(L) * kdim * ncols);
    return PK_OK;
  }
  int update(char which, char transb, int kdim, int ncols, const double* B, int ldb) {
    return lbd_update(*this, which, transb, kdim, ncols, B, ldb);
  }
};

// Pieces of a DLANBPRO call for the batched step engine (pk_batch.hip).
int engine_lanbpro_begin(Engine* e, int k0, int k, double* alpha, double* beta, double* rnorm,
                         double* doption, const int* ioption, int* ierr, bool* finished) {
  if (k > e->kmax || k0 < 0 || k0 > k) {
    set_error("lanbpro: k=%d k0=%d outside the engine's kmax=%d", k, k0, e->kmax);
    return PK_ERR_ARG;
  }
  if (prec_is_complex(e->prec)) {
    set_error("batched steps take real engines only");
    return PK_ERR_NOTSUP;
  }
  if (e->prec == PK_F64)
    return Impl<double>(e).lanbpro_begin(k0, k, alpha, beta, rnorm, doption, ioption, ierr,
                                         finished);
  return Impl<float>(e).lanbpro_begin(k0, k, alpha, beta, rnorm, doption, ioption, ierr,
                                      finished);
}

int engine_lanbpro_loop(Engine* e) {
  if (e->prec == PK_F64) return Impl<double>(e).lanbpro_loop();
  return Impl<float>(e).lanbpro_loop();
}

int engine_lanbpro_end(Engine* e, int k0, int* kp, double* alpha, double* beta, double* rnorm,
                       double* doption, int* ierr) {
  if (e->prec == PK_F64)
    return Impl<double>(e).lanbpro_end(k0, kp, alpha, beta, rnorm, doption, ierr);
  return Impl<float>(e).lanbpro_end(k0, kp, alpha, beta, rnorm, doption, ierr);
}

static Engine* as_engine(pk_lbd_t h) {
  Engine* e = reinterpret_cast<Engine*>(h);
  if (!e || e->magic != Engine::kMagic) {
    set_error("not a live engine handle");
    return nullptr;
  }
  return e;
}

CsrOp* as_csr(pk_csr_t h) {
  CsrOp* op = reinterpret_cast<CsrOp*>(h);
  if (!op || op->magic != CsrOp::kMagic) {
    set_error("not a live CSR operator handle");
    return nullptr;
  }
  return op;
}

static void engine_free(Engine* e) {
  for (const KTimer& t : e->prof.pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t ev : e->prof.pool) (void)hipEventDestroy(ev);
  for (int i = 0; i < 2; ++i) {
    (void)hipFree(e->U[i]);
    (void)hipFree(e->V[i]);
  }
  (void)hipFree(e->part);
  (void)hipFree(e->red);
  (void)hipFree(e->dcols);
  (void)hipFree(e->dB);
  (void)hipFree(e->vec);
  (void)hipFree(e->sw_y);
  (void)hipFree(e->sw_l);
  (void)hipFree(e->kres);
  (void)hipFree(e->dctl);
  (void)hipFree(e->darr);
  (void)hipHostFree(e->hctl);
  (void)hipHostFree(e->harr);
  (void)hipHostFree(e->hred);
  (void)hipHostFree(e->hcols);
  (void)hipHostFree(e->hB);
  if (e->evB) (void)hipEventDestroy(e->evB);
  (void)hipHostFree(e->hx);
  (void)hipHostFree(e->hy);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->magic = 0;
  delete e;
}

Engine* engine_from_handle(pk_lbd_t h) { return as_engine(h); }

}  // namespace pk

using namespace pk;

#define PK_ENGINE(h)                  \
  Engine* e = as_engine(h);           \
  if (!e) return PK_ERR_HANDLE;

extern "C" {

const char* pk_last_error(void) { return pk::last_error(); }

int pk_lbd_create(int prec, int m, int n, int kmax, pk_csr_t op, pk_lbd_t* out) {
  if (!out || m < 1 || n < 1 || kmax < 1 ||
      (prec != PK_F64 && prec != PK_F32 && !prec_is_complex(prec))) {
    set_error("pk_lbd_create: bad arguments m=%d n=%d kmax=%d prec=%d", m, n, kmax, prec);
    return PK_ERR_ARG;
  }
  CsrOp* cop = nullptr;
  if (op) {
    cop = as_csr(op);
    if (!cop) return PK_ERR_HANDLE;
    if (cop->m != m || cop->n != n) {
      set_error("pk_lbd_create: operator is %dx%d, engine %dx%d", cop->m, cop->n, m, n);
      return PK_ERR_ARG;
    }
    if (cop->val_prec != prec) {
      set_error("pk_lbd_create: operator values and basis must share a precision");
      return PK_ERR_NOTSUP;
    }
  }
  Engine* e = new Engine();
  e->magic = Engine::kMagic;
  e->prec = prec;
  e->m = m;
  e->n = n;
  e->kmax = kmax;
  e->tbytes = prec_bytes(prec);
  e->ldu = round_up(m, 64);
  e->ldv = round_up(n, 64);
  e->op = cop;
  e->ccap = std::max(kMaxCols, kmax + 2);
  const long long mx = std::max(m, n);
  size_t nblk = static_cast<size_t>(std::max(stream_blocks(mx), kMaxBlocks));
  if (cop) {
    // the SpMV writes two partials per row block
    const size_t need = 2 * static_cast<size_t>(std::max(cop->nrb, cop->nrbt));
    nblk = std::max(nblk, (need + e->ccap - 1) / e->ccap);
  }
  // complex bases: three partials per SpMV slot and two per projected
  // column, the projection (2 ccap values) and the CGS norm in red
  const size_t cx = prec_is_complex(prec) ? 2 : 1;
  const size_t nb = static_cast<size_t>(kmax + 1) * (kmax + 1);
  hipError_t err = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (r != hipSuccess && err == hipSuccess) err = r;
  };
  chk(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  chk(hipMalloc(&e->U[0], e->tbytes * e->ldu * (kmax + 1)));
  chk(hipMalloc(&e->V[0], e->tbytes * e->ldv * kmax));
  chk(hipMalloc(&e->part, sizeof(double) * nblk * e->ccap * cx));
  chk(hipMalloc(&e->red, sizeof(double) * (e->ccap * cx + 16)));
  chk(hipMalloc(&e->dcols, sizeof(int) * e->ccap));
  chk(hipMalloc(&e->dB, sizeof(double) * nb));
  chk(hipMalloc(&e->vec, e->tbytes * mx * 2));
  if (cop && prec_is_complex(prec)) {
    // the panels' FP64 row sums of a multi-panel complex direction
    e->sw_y_len = zcsr_ywork(cop);
    if (e->sw_y_len > 0) chk(hipMalloc(&e->sw_y, sizeof(double) * e->sw_y_len));
  } else if (cop) {
    e->sw_y_len = csr_ywork(cop);
    chk(hipMalloc(&e->sw_y, sizeof(double) * e->sw_y_len));
    chk(hipMalloc(&e->sw_l, sizeof(double) * csr_lwork(cop)));
  }
  chk(hipHostMalloc(&e->hred, sizeof(double) * (e->ccap * cx + 16), hipHostMallocDefault));
  chk(hipHostMalloc(&e->hcols, sizeof(int) * e->ccap, hipHostMallocDefault));
  chk(hipHostMalloc(&e->hB, sizeof(double) * nb, hipHostMallocDefault));
  const size_t ab = static_cast<size_t>(CtlArrays::bytes(kmax));
  chk(hipMalloc(&e->dctl, sizeof(LbdCtl)));
  chk(hipMalloc(&e->darr, ab));
  chk(hipHostMalloc(&e->hctl, sizeof(LbdCtlRef), hipHostMallocDefault));
  chk(hipHostMalloc(&e->harr, ab, hipHostMallocDefault));
  if (e->hctl && e->harr) {
    std::memset(static_cast<void*>(e->hctl), 0, sizeof(LbdCtlRef));
    e->hctl->cap = kmax;
    e->hctl->bind(e->harr);
  }
  if (const char* md = std::getenv("PK_LBD_MODE")) e->mode = (md[0] == 'h' || md[0] == 'H') ? 1 : 0;
  if (err == hipSuccess) {
    // The bases start zeroed so that unwritten columns never hold NaN garbage.
    chk(hipMemsetAsync(e->U[0], 0, e->tbytes * e->ldu * (kmax + 1), e->stream));
    chk(hipMemsetAsync(e->V[0], 0, e->tbytes * e->ldv * kmax, e->stream));
    chk(hipStreamSynchronize(e->stream));
  }
  if (err != hipSuccess) {
    set_error("pk_lbd_create: %s", hipGetErrorString(err));
    engine_free(e);
    return err == hipErrorOutOfMemory ? PK_ERR_OOM : PK_ERR_HIP;
  }
  *out = reinterpret_cast<pk_lbd_t>(e);
  return PK_OK;
}

int pk_lbd_destroy(pk_lbd_t h) {
  PK_ENGINE(h);
  (void)hipStreamSynchronize(e->stream);
  engine_free(e);
  return PK_OK;
}

int pk_lbd_set_callback(pk_lbd_t h, void* aprod, int aprod_prec, void* dparm, int* iparm) {
  PK_ENGINE(h);
  if (!aprod || (prec_is_complex(e->prec) ? aprod_prec != e->prec
                                          : (aprod_prec != PK_F64 && aprod_prec != PK_F32))) {
    set_error("pk_lbd_set_callback: a complex engine takes a callback of its own precision, a "
              "real engine one of PK_F64 or PK_F32");
    return PK_ERR_ARG;
  }
  e->aprod = aprod;
  e->aprod_prec = aprod_prec;
  e->dparm = dparm;
  e->iparm = iparm;
  if (!e->hx) {
    const size_t b = std::max<size_t>(sizeof(double), e->tbytes) * std::max(e->m, e->n);
    if (hipHostMalloc(&e->hx, b, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&e->hy, b, hipHostMallocDefault) != hipSuccess) {
      set_error("pk_lbd_set_callback: pinned allocation failed");
      return PK_ERR_OOM;
    }
  }
  return PK_OK;
}

int pk_lbd_put(pk_lbd_t h, char which, int col, const void* x, int prec) {
  PK_ENGINE(h);
  const bool isu = (which == 'u' || which == 'U');
  const int maxc = isu ? e->kmax + 1 : e->kmax;
  if (col < 1 || col > maxc || !x) return PK_ERR_ARG;
  const long long L = isu ? e->m : e->n;
  char* dst = static_cast<char*>(isu ? e->u() : e->v()) +
              e->tbytes * static_cast<size_t>(col - 1) * (isu ? e->ldu : e->ldv);
  if (prec_is_complex(prec) != prec_is_complex(e->prec)) {
    set_error("pk_lbd_put: real and complex arrays do not mix");
    return PK_ERR_ARG;
  }
  const size_t sb = prec_bytes(prec);
  // a complex column converts as its real view of 2L values
  const long long LC = prec_is_complex(prec) ? 2 * L : L;
  // hipMemcpyDefault: x may be host or device memory (the runtime reads the
  // direction from the pointer); a device-resident x costs a device copy.
  if (sb == e->tbytes) {
    PK_HIP(hipMemcpyAsync(dst, x, sb * L, hipMemcpyDefault, e->stream));
  } else {
    PK_HIP(hipMemcpyAsync(e->vec, x, sb * L, hipMemcpyDefault, e->stream));
    int r = (e->prec == PK_F64 || e->prec == PK_C128)
                ? launch_convert<float, double>(e->stream, LC, static_cast<const float*>(e->vec),
                                                reinterpret_cast<double*>(dst))
                : launch_convert<double, float>(e->stream, LC, static_cast<const double*>(e->vec),
                                                reinterpret_cast<float*>(dst));
    if (r < 0) return PK_ERR_HIP;
  }
  PK_HIP(hipStreamSynchronize(e->stream));
  return PK_OK;
}

int pk_lbd_get(pk_lbd_t h, char which, int ncols, void* X, int ldx, int prec) {
  PK_ENGINE(h);
  const bool isu = (which == 'u' || which == 'U');
  const int maxc = isu ? e->kmax + 1 : e->kmax;
  const long long L = isu ? e->m : e->n;
  if (ncols < 0 || ncols > maxc || (ncols > 0 && !X) || ldx < L) return PK_ERR_ARG;
  if (prec_is_complex(prec) != prec_is_complex(e->prec)) {
    set_error("pk_lbd_get: real and complex arrays do not mix");
    return PK_ERR_ARG;
  }
  const size_t sb = prec_bytes(prec);
  const long long LC = prec_is_complex(prec) ? 2 * L : L;
  const char* src = static_cast<const char*>(isu ? e->u() : e->v());
  const long long ld = isu ? e->ldu : e->ldv;
  for (int j = 0; j < ncols; ++j) {
    const char* cs = src + e->tbytes * static_cast<size_t>(j) * ld;
    char* cd = static_cast<char*>(X) + sb * static_cast<size_t>(j) * ldx;
    if (sb == e->tbytes) {
      PK_HIP(hipMemcpyAsync(cd, cs, sb * L, hipMemcpyDeviceToHost, e->stream));
    } else {
      int r = (e->prec == PK_F64 || e->prec == PK_C128)
                  ? launch_convert<double, float>(e->stream, LC, reinterpret_cast<const double*>(cs),
                                                  static_cast<float*>(e->vec))
                  : launch_convert<float, double>(e->stream, LC, reinterpret_cast<const float*>(cs),
                                                  static_cast<double*>(e->vec));
      if (r < 0) return PK_ERR_HIP;
      PK_HIP(hipMemcpyAsync(cd, e->vec, sb * L, hipMemcpyDeviceToHost, e->stream));
      PK_HIP(hipStreamSynchronize(e->stream));
    }
  }
  PK_HIP(hipStreamSynchronize(e->stream));
  return PK_OK;
}

int pk_lbd_basis(pk_lbd_t h, char which, void** dptr, long long* ld) {
  PK_ENGINE(h);
  const bool isu = (which == 'u' || which == 'U');
  if (dptr) *dptr = isu ? e->u() : e->v();
  if (ld) *ld = isu ? e->ldu : e->ldv;
  return PK_OK;
}

int pk_lbd_nrm2(pk_lbd_t h, char which, int col, double* nrm) {
  PK_ENGINE(h);
  const bool isu = (which == 'u' || which == 'U');
  const int maxc = isu ? e->kmax + 1 : e->kmax;
  if (col < 1 || col > maxc || !nrm) return PK_ERR_ARG;
  if (prec_is_complex(e->prec)) {
    if (e->prec == PK_C128) {
      ZImpl<double> im(e);
      return im.nrm2(isu ? im.ucol(col) : im.vcol(col), isu ? e->m : e->n, nrm);
    }
    ZImpl<float> im(e);
    return im.nrm2(isu ? im.ucol(col) : im.vcol(col), isu ? e->m : e->n, nrm);
  }
  if (e->prec == PK_F64) {
    Impl<double> im(e);
    return im.nrm2(isu ? im.ucol(col) : im.vcol(col), isu ? e->m : e->n, nrm);
  }
  Impl<float> im(e);
  return im.nrm2(isu ? im.ucol(col) : im.vcol(col), isu ? e->m : e->n, nrm);
}

int pk_lbd_getu0(pk_lbd_t h, char transa, int j, int ntry, double* u0norm, int icgs,
                 double* anormest, int* ierr) {
  PK_ENGINE(h);
  (void)icgs;  // the device always uses block classical Gram-Schmidt
  const bool tn = (transa == 'n' || transa == 'N');
  const int maxc = tn ? e->kmax + 1 : e->kmax;
  if (j < 0 || j + 1 > maxc || ntry < 1 || !u0norm || !anormest || !ierr) return PK_ERR_ARG;
  if (e->prec == PK_C128) {
    ZImpl<double> im(e);
    return im.getu0(transa, j, ntry, tn ? im.ucol(j + 1) : im.vcol(j + 1), u0norm,
                    tn ? im.U() : im.V(), tn ? e->ldu : e->ldv, anormest, ierr);
  }
  if (e->prec == PK_C64) {
    ZImpl<float> im(e);
    return im.getu0(transa, j, ntry, tn ? im.ucol(j + 1) : im.vcol(j + 1), u0norm,
                    tn ? im.U() : im.V(), tn ? e->ldu : e->ldv, anormest, ierr);
  }
  if (e->prec == PK_F64) {
    Impl<double> im(e);
    return im.getu0(transa, j, ntry, tn ? im.ucol(j + 1) : im.vcol(j + 1), u0norm,
                    tn ? im.U() : im.V(), tn ? e->ldu : e->ldv, anormest, ierr);
  }
  Impl<float> im(e);
  return im.getu0(transa, j, ntry, tn ? im.ucol(j + 1) : im.vcol(j + 1), u0norm,
                  tn ? im.U() : im.V(), tn ? e->ldu : e->ldv, anormest, ierr);
}

int pk_lbd_lanbpro(pk_lbd_t h, int k0, int* k, double* alpha, double* beta, double* rnorm,
                   double* doption, const int* ioption, int* ierr) {
  PK_ENGINE(h);
  if (!k || !alpha || !beta || !rnorm || !doption || !ioption || !ierr) return PK_ERR_ARG;
  if (e->prec == PK_C128)
    return ZImpl<double>(e).lanbpro(k0, k, alpha, beta, rnorm, doption, ioption, ierr);
  if (e->prec == PK_C64)
    return ZImpl<float>(e).lanbpro(k0, k, alpha, beta, rnorm, doption, ioption, ierr);
  if (e->prec == PK_F64) return Impl<double>(e).lanbpro(k0, k, alpha, beta, rnorm, doption, ioption, ierr);
  return Impl<float>(e).lanbpro(k0, k, alpha, beta, rnorm, doption, ioption, ierr);
}

int pk_lbd_update(pk_lbd_t h, char which, char transb, int kdim, int ncols, const double* B,
                  int ldb) {
  PK_ENGINE(h);
  if (!B) return PK_ERR_ARG;
  if (e->prec == PK_C128) return ZImpl<double>(e).update(which, transb, kdim, ncols, B, ldb);
  if (e->prec == PK_C64) return ZImpl<float>(e).update(which, transb, kdim, ncols, B, ldb);
  if (e->prec == PK_F64) return Impl<double>(e).update(which, transb, kdim, ncols, B, ldb);
  return Impl<float>(e).update(which, transb, kdim, ncols, B, ldb);
}

int pk_lbd_stats(pk_lbd_t h, long long* out) {
  PK_ENGINE(h);
  if (!out) return PK_ERR_ARG;
  out[0] = e->cnt.nopx;
  out[1] = e->cnt.nreorth;
  out[2] = e->cnt.ndot;
  out[3] = e->cnt.ncols_touched;
  out[4] = e->cnt.nsync;
  out[5] = e->cnt.nretry;
  return PK_OK;
}

int pk_lbd_path_stats(pk_lbd_t h, long long* out) {
  PK_ENGINE(h);
  if (!out) return PK_ERR_ARG;
  for (int i = 0; i < 6; ++i) out[i] = e->cnt.rare[i];
  out[6] = e->cnt.nwarn;
  out[7] = 0;
  return PK_OK;
}

int pk_lbd_reset_stats(pk_lbd_t h) {
  PK_ENGINE(h);
  e->cnt = Counters();
  return PK_OK;
}

int pk_lbd_residuals(pk_lbd_t h, pk_lbd_t src_h, int k, const double* sigma, double* res) {
  PK_ENGINE(h);
  Engine* sr = as_engine(src_h);
  if (!sr) return PK_ERR_HANDLE;
  if (prec_is_complex(e->prec) || prec_is_complex(sr->prec)) {
    set_error("pk_lbd_residuals: real engines only");
    return PK_ERR_NOTSUP;
  }
  if (!e->op || sr->m != e->m || sr->n != e->n || k < 0 || k > sr->kmax ||
      (k > 0 && (!sigma || !res))) {
    set_error("pk_lbd_residuals: needs an engine with a device operator, a source engine of "
              "its shape and 0 <= k <= the source's kmax");
    return PK_ERR_ARG;
  }
  if (e->prec == PK_F64) return Impl<double>(e).residuals(sr, k, sigma, res);
  return Impl<float>(e).residuals(sr, k, sigma, res);
}

int pk_lbd_promote(pk_lbd_t dst_h, pk_lbd_t src_h, int ncols) {
  Engine* d = as_engine(dst_h);
  Engine* sr = as_engine(src_h);
  if (!d || !sr) return PK_ERR_HANDLE;
  if (prec_is_complex(d->prec) || prec_is_complex(sr->prec)) {
    set_error("pk_lbd_promote: real engines only");
    return PK_ERR_NOTSUP;
  }
  if (d->m != sr->m || d->n != sr->n || ncols < 0 || ncols > std::min(d->kmax, sr->kmax)) {
    set_error("pk_lbd_promote: engines of one shape and 0 <= ncols <= both kmax");
    return PK_ERR_ARG;
  }
  PK_HIP(hipStreamSynchronize(sr->stream));
  for (int isu = 0; isu < 2; ++isu) {
    const long long L = isu ? d->m : d->n;
    for (int c = 0; c < ncols; ++c) {
      const char* s = static_cast<const char*>(isu ? sr->u() : sr->v()) +
                      sr->tbytes * static_cast<size_t>(c) * (isu ? sr->ldu : sr->ldv);
      char* t = static_cast<char*>(isu ? d->u() : d->v()) +
                d->tbytes * static_cast<size_t>(c) * (isu ? d->ldu : d->ldv);
      int r;
      if (sr->prec == PK_F32 && d->prec == PK_F64)
        r = launch_convert<float, double>(d->stream, L, reinterpret_cast<const float*>(s),
                                          reinterpret_cast<double*>(t));
      else if (sr->prec == PK_F64 && d->prec == PK_F32)
        r = launch_convert<double, float>(d->stream, L, reinterpret_cast<const double*>(s),
                                          reinterpret_cast<float*>(t));
      else if (sr->prec == PK_F64)
        r = launch_convert<double, double>(d->stream, L, reinterpret_cast<const double*>(s),
                                           reinterpret_cast<double*>(t));
      else
        r = launch_convert<float, float>(d->stream, L, reinterpret_cast<const float*>(s),
                                         reinterpret_cast<float*>(t));
      if (r < 0) return PK_ERR_HIP;
    }
  }
  PK_HIP(hipStreamSynchronize(d->stream));
  return PK_OK;
}

int pk_lbd_seed(pk_lbd_t dst_h, pk_lbd_t src_h, int ncols) {
  Engine* d = as_engine(dst_h);
  Engine* sr = as_engine(src_h);
  if (!d || !sr) return PK_ERR_HANDLE;
  if (d->prec != PK_F64 || prec_is_complex(sr->prec) || d->m != sr->m || ncols < 1 ||
      ncols > sr->kmax + 1) {
    set_error("pk_lbd_seed: needs a double engine of the same shape and 1 <= ncols <= kmax+1");
    return PK_ERR_ARG;
  }
  PK_HIP(hipStreamSynchronize(sr->stream));
  int r = sr->prec == PK_F32
              ? launch_colsum<float, double>(d->stream, d->m, static_cast<const float*>(sr->u()),
                                             sr->ldu, ncols, static_cast<double*>(d->u()))
              : launch_colsum<double, double>(d->stream, d->m,
                                              static_cast<const double*>(sr->u()), sr->ldu,
                                              ncols, static_cast<double*>(d->u()));
  if (r < 0) return PK_ERR_HIP;
  PK_HIP(hipStreamSynchronize(d->stream));
  return PK_OK;
}

int pk_lbd_set_mode(pk_lbd_t h, int mode, int tries) {
  PK_ENGINE(h);
  if ((mode != 0 && mode != 1) || tries < 1 || tries > 5) return PK_ERR_ARG;
  e->mode = mode;
  e->dev_tries = tries;
  e->tries_auto = false;
  return PK_OK;
}

int pk_lbd_profile(pk_lbd_t h, int classes) {
  PK_ENGINE(h);
  e->prof.mask = static_cast<unsigned>(classes) & ((1u << KC_COUNT) - 1u);
  e->prof.on = e->prof.mask != 0;
  return PK_OK;
}

int pk_lbd_profile_get(pk_lbd_t h, double* out) {
  PK_ENGINE(h);
  if (!out) return PK_ERR_ARG;
  PK_HIP(hipStreamSynchronize(e->stream));
  profile_resolve(e);
  for (int c = 0; c < KC_COUNT; ++c) {
    out[4 * c + 0] = static_cast<double>(e->prof.count[c]);
    out[4 * c + 1] = e->prof.ms[c];
    out[4 * c + 2] = e->prof.bytes[c];
    out[4 * c + 3] = e->prof.flops[c];
  }
  return PK_OK;
}

int pk_lbd_profile_reset(pk_lbd_t h) {
  PK_ENGINE(h);
  for (int c = 0; c < KC_COUNT; ++c) {
    e->prof.count[c] = 0;
    e->prof.ms[c] = e->prof.bytes[c] = e->prof.flops[c] = 0.0;
  }
  return PK_OK;
}

int pk_lbd_shape(pk_lbd_t h, int* m, int* n, int* kmax, int* prec) {
  PK_ENGINE(h);
  if (m) *m = e->m;
  if (n) *n = e->n;
  if (kmax) *kmax = e->kmax;
  if (prec) *prec = e->prec;
  return PK_OK;
}

int pk_lbd_set_rounding(pk_lbd_t h, int model) {
  PK_ENGINE(h);
  if (model != 0 && model != 1) return PK_ERR_ARG;
  e->rounding = model;
  return PK_OK;
}

int pk_lbd_eps(pk_lbd_t h, double* eps, double* leff) {
  PK_ENGINE(h);
  if (eps) *eps = (e->prec == PK_F64 || e->prec == PK_C128) ? 0x1p-53 : 0x1p-24;
  if (leff) *leff = e->rounding ? 4.0 : static_cast<double>(std::max(e->m, e->n));
  return PK_OK;
}

int pk_lbd_sync(pk_lbd_t h) {
  PK_ENGINE(h);
  PK_HIP(stream_wait(e->stream));
  profile_resolve(e);
  return PK_OK;
}

}  // extern "C"

// Host LAPACK (the small bidiagonal SVDs of the drivers) runs single-threaded:
// select MKL's sequential layer before its first use, so that the library
// never brings a second OpenMP runtime into the caller's process.
extern "C" int MKL_Set_Threading_Layer(int);
__attribute__((constructor)) static void pk_select_sequential_lapack() {
  MKL_Set_Threading_Layer(1 /* MKL_THREADING_SEQUENTIAL */);
}

// ----------------------------------------------------------------------------
// Host scalar routines of the engine, exported for verification against the
// reference (arrays 1-based as in the reference: element 0 is ignored).
extern "C" {

int pk_larnv_pm1(int prec, long long n, void* x) {
  if (n < 0 || !x) return PK_ERR_ARG;
  if (prec == PK_F64) larnv_pm1<double>(n, static_cast<double*>(x));
  else if (prec == PK_F32) larnv_pm1<float>(n, static_cast<float*>(x));
  else return PK_ERR_ARG;
  return PK_OK;
}

int pk_larnv_pm1_device(int prec, long long n, void* x) {
  if (n < 0 || (n > 0 && !x) || (prec != PK_F64 && prec != PK_F32)) return PK_ERR_ARG;
  const int r = prec == PK_F64 ? launch_larnv_pm1<double>(nullptr, n, static_cast<double*>(x))
                               : launch_larnv_pm1<float>(nullptr, n, static_cast<float*>(x));
  if (r < 0) return PK_ERR_HIP;
  PK_HIP(hipStreamSynchronize(nullptr));
  return PK_OK;
}

void pkh_update_mu(double* mumax, double* mu, const double* nu, int j, const double* alpha,
                   const double* beta, double anorm, double eps1) {
  hd_update_mu(mumax, mu, nu, j, alpha, beta, anorm, eps1);
}

void pkh_update_nu(double* numax, const double* mu, double* nu, int j, const double* alpha,
                   const double* beta, double anorm, double eps1) {
  hd_update_nu(numax, mu, nu, j, alpha, beta, anorm, eps1);
}

void pkh_compute_int(const double* mu, int j, double delta, double eta, int* index) {
  hd_compute_int(mu, j, delta, eta, index);
}

void pkh_set_mu(int k, double* mu, const int* index, double val) { hd_set_mu(k, mu, index, val); }

}  // extern "C"

// Progress line per outer iteration of the drivers when PK_PROGRESS is set
// in the environment (Fortran-side helper, arguments by reference).
#include <chrono>
extern "C" void pk_f_progress_(const int* iter, const int* nconv, const int* neig,
                               const double* s1, const double* rnorm) {
  static const bool on = std::getenv("PK_PROGRESS") != nullptr;
  if (!on) return;
  static const auto t0 = std::chrono::steady_clock::now();
  const double t =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(stderr, "[propack_amd] t=%.3fs iter=%d converged=%d/%d sigma1=%.6e rnorm=%.3e\n",
               t, *iter, *nconv, *neig, *s1, *rnorm);
  std::fflush(stderr);
}
