// Test harness (CPU only): drives the engine's step control
// (propack_amd/csrc/pk_lbd.h, the same functions the GPU control kernels
// run) with plain C++ vector operations, so that the phase machine can be
// checked against the oracle's DLANBPRO on the CPU.  Built by
// tests/test_control_cpu.py; never part of the library.
//
// Only the k0 = 0 call is covered (the setup mirrors double/dlanbpro.F:
// 155-230); the rare host-only phases are reported as an error.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../propack_amd/csrc/pk_lbd.h"

using namespace pk;

namespace {

struct Csr {
  int m, n;
  const int *rp, *ci, *rpt, *cit;
  const double *va, *vat;
};

// y = op(A) x - c * yprev; s2 = ||y||^2, sd = yprev.y
void apply(const Csr& A, bool tr, const double* x, double* y, const double* yprev, double c,
           double* s2, double* sd) {
  const int rows = tr ? A.n : A.m;
  const int* rp = tr ? A.rpt : A.rp;
  const int* ci = tr ? A.cit : A.ci;
  const double* va = tr ? A.vat : A.va;
  double a = 0, b = 0;
  for (int i = 0; i < rows; ++i) {
    double acc = 0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) acc += va[p] * x[ci[p]];
    const double yp = yprev ? yprev[i] : 0.0;
    y[i] = acc - c * yp;
    a += y[i] * y[i];
    b += yp * y[i];
  }
  *s2 = a;
  *sd = b;
}

double nrm2(int n, const double* x) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += x[i] * x[i];
  return std::sqrt(s);
}

// one CGS pass of w against V(:, cols): returns ||w'||^2
double cgs_pass(int L, const double* V, int ld, double* w, const int* cols, int nc) {
  std::vector<double> h(nc);
  for (int c = 0; c < nc; ++c) {
    const double* v = V + static_cast<long long>(cols[c]) * ld;
    double s = 0;
    for (int i = 0; i < L; ++i) s += v[i] * w[i];
    h[c] = s;
  }
  double s2 = 0;
  for (int i = 0; i < L; ++i) {
    double acc = w[i];
    for (int c = 0; c < nc; ++c) acc -= V[static_cast<long long>(cols[c]) * ld + i] * h[c];
    w[i] = acc;
    s2 += acc * acc;
  }
  return s2;
}

void larnv_pm1(int n, double* x) {
  const uint64_t mask = (1ULL << 48) - 1, a = 33952834046453ULL;
  uint64_t s = (1ULL << 36) + (3ULL << 24) + (5ULL << 12) + 7ULL;
  const double r = 1.0 / 4096.0;
  for (int g = 0; g < n; ++g) {
    s = static_cast<uint64_t>((static_cast<unsigned __int128>(s) * a) & mask);
    const double u = r * ((double)(s >> 36) + r * ((double)((s >> 24) & 4095) +
                                                   r * ((double)((s >> 12) & 4095) +
                                                        r * (double)(s & 4095))));
    x[g] = 2.0 * u - 1.0;
  }
}

}  // namespace

extern "C" int lbd_cpu_lanbpro(int m, int n, const int* rp, const int* ci, const double* va,
                               const int* rpt, const int* cit, const double* vat, double* U,
                               double* V, int k, double* alpha, double* beta, double* rnorm,
                               double delta, double eta, int elr, long long* counters) {
  Csr A{m, n, rp, ci, rpt, cit, va, vat};
  LbdCtlRef c;
  std::memset(static_cast<void*>(&c), 0, sizeof c);
  c.cap = k;
  std::vector<double> arrays(static_cast<size_t>(CtlArrays::bytes(k) / 8 + 1), 0.0);
  c.bind(arrays.data());
  const double eps = 0x1p-53;
  const int mx = m > n ? m : n;
  auto ucol = [&](int j) { return U + static_cast<long long>(j - 1) * m; };
  auto vcol = [&](int j) { return V + static_cast<long long>(j - 1) * n; };
  // setup for k0 = 0 (dlanbpro.F:155-230): ||A|| estimate from a random vector
  std::vector<double> x(mx), y(mx);
  double s2, sd;
  const bool tn = n > m;
  larnv_pm1(tn ? n : m, x.data());
  const double nx = nrm2(tn ? n : m, x.data());
  apply(A, !tn, x.data(), y.data(), nullptr, 0.0, &s2, &sd);
  const double anormest = std::sqrt(s2) / nx;
  c.m = m;
  c.n = n;
  c.k = k;
  c.elr = elr;
  c.delta = delta;
  c.eta = eta;
  c.eps = eps;
  c.epsn = mx * eps;
  c.epsn2 = std::sqrt(static_cast<double>(mx)) * eps;
  c.sfmin = 0x1p-1022;
  c.kappa = 0.717;
  c.fudge = 1.01;
  c.full_reorth = (delta <= eta || delta == 0.0);
  c.anorm = c.fudge * anormest;
  c.beta = *rnorm;
  const double inv = 1.0 / c.beta;
  for (int i = 0; i < m; ++i) U[i] = inv * U[i];
  c.mu[1] = 1.0;
  c.nu[1] = 1.0;
  c.j = 1;
  c.phase = PH_V_SPMV;
  long long extra_nopx = 1;
  while (c.phase != PH_DONE) {
    const int j = c.j;
    switch (c.phase) {
      case PH_V_SPMV:
        if (c.uscale) {
          for (int i = 0; i < m; ++i) ucol(j)[i] = c.uscale_f * ucol(j)[i];
          c.uscale = 0;
        }
        apply(A, true, ucol(j), vcol(j), j > 1 ? vcol(j - 1) : nullptr, j > 1 ? c.beta : 0.0,
              &s2, &sd);
        ctl_v_spmv(c, s2, sd);
        break;
      case PH_V_ELR: {
        double a = 0, b = 0;
        for (int i = 0; i < n; ++i) {
          vcol(j)[i] = vcol(j)[i] - c.elr_s * vcol(j - 1)[i];
          a += vcol(j)[i] * vcol(j)[i];
          b += vcol(j - 1)[i] * vcol(j)[i];
        }
        ctl_v_elr(c, a, b);
        break;
      }
      case PH_V_REO:
        ctl_v_reo(c, cgs_pass(n, V, n, vcol(j), c.cols, c.reo_ncols));
        break;
      case PH_U_SPMV:
        if (c.vscale) {
          for (int i = 0; i < n; ++i) vcol(j)[i] = c.vscale_f * vcol(j)[i];
          c.vscale = 0;
        }
        apply(A, false, vcol(j), ucol(j + 1), ucol(j), c.alpha, &s2, &sd);
        ctl_u_spmv(c, s2, sd);
        break;
      case PH_U_ELR: {
        double a = 0, b = 0;
        for (int i = 0; i < m; ++i) {
          ucol(j + 1)[i] = ucol(j + 1)[i] - c.elr_s * ucol(j)[i];
          a += ucol(j + 1)[i] * ucol(j + 1)[i];
          b += ucol(j)[i] * ucol(j + 1)[i];
        }
        ctl_u_elr(c, a, b);
        break;
      }
      case PH_U_REO:
        ctl_u_reo(c, cgs_pass(m, U, m, ucol(j + 1), c.cols, c.reo_ncols));
        break;
      default:
        return -c.phase;  // a host-only rare path: not covered by this harness
    }
  }
  if (c.uscale) {
    for (int i = 0; i < m; ++i) ucol(c.j)[i] = c.uscale_f * ucol(c.j)[i];
    c.uscale = 0;
  }
  for (int i = 1; i <= k; ++i) {
    alpha[i - 1] = c.B1[i];
    beta[i - 1] = c.B2[i];
  }
  *rnorm = c.rnorm;
  counters[0] = c.nopx + extra_nopx;
  counters[1] = c.nreorth;
  counters[2] = c.ndot;
  counters[3] = c.ierr;
  return 0;
}
