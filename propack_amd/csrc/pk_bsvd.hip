// Batched singular values of small upper bidiagonal matrices on the device:
// the bidiagonal SVD of the restarted driver's error-bound step
// (double/dlansvd_irl.F:226-229, DBDSQR with one row of left vectors) for
// every problem of a batch in one launch.
//
// Method: one-sided Jacobi (Hestenes) on G = B^T, one workgroup per
// problem.  G's columns are made pairwise orthogonal by plane rotations,
// G Y = W; then B^T = W Y^T, so Y holds the left singular vectors of B and
// the column norms of W are its singular values.  Only the last row of Y is
// needed (it scales the error bounds), so the rotations update that row and
// nothing else of Y.  Pairs are visited in the circle (round-robin) order:
// n/2 disjoint pairs per round, n-1 rounds per sweep, each pair handled by
// one wavefront; sweeps repeat until none rotates.  The rotation of a pair
// with Gram entries a = |g_p|^2, b = |g_q|^2, c = g_p.g_q follows
// Rutishauser's formulas (zeta = (b - a) / 2c, t = sign(zeta) / (|zeta| +
// sqrt(1 + zeta^2)), cs = 1 / sqrt(1 + t^2), sn = cs t), applied when
// |c| > tol sqrt(a) sqrt(b), tol = sqrt(n) eps.  The matrix is scaled by a
// power of two to unit size first, so sums of squares neither overflow nor
// underflow for any finite input of moderate range.
#include <hip/hip_runtime.h>

#include <cmath>

#include "pk_internal.h"

namespace pk {

void set_error(const char* fmt, ...);

namespace {

constexpr int kBsvdThreads = 256;
constexpr int kBsvdWaves = kBsvdThreads / kWave;
constexpr int kBsvdMaxSweeps = 60;

// Sum over the wave of three values at once.
__device__ __forceinline__ void wave_sum3(double& a, double& b, double& c) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, kWave);
    b += __shfl_xor(b, off, kWave);
    c += __shfl_xor(c, off, kWave);
  }
}

// Problem i: order n[i] (<= nmax <= kBsvdMax), diagonal d[i*ld + j],
// superdiagonal e[i*ld + j] (j < n-1).  Out: sigma[i*ld + r] in decreasing
// order, bot[i*ld + r] the last component of the matching left singular
// vector.  Dynamic LDS: G (np x np doubles, column-major, np = n rounded up
// to even), the row r (np doubles), the column norms (np doubles).
__global__ __launch_bounds__(kBsvdThreads) void k_bsvd(const int* __restrict__ nv,
                                                       const double* __restrict__ d,
                                                       const double* __restrict__ e, int ld,
                                                       double* __restrict__ sigma,
                                                       double* __restrict__ bot) {
  extern __shared__ double lds[];
  __shared__ int rotated[kBsvdWaves];
  __shared__ double smax;
  const int i = blockIdx.x;
  const int n = nv[i];
  if (n <= 0) return;
  const int np = (n + 1) & ~1;  // a zero column pads an odd order
  double* G = lds;              // G[col * np + row]
  double* r = G + static_cast<size_t>(np) * np;
  double* nrm = r + np;
  const double* di = d + static_cast<size_t>(i) * ld;
  const double* ei = e + static_cast<size_t>(i) * ld;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;

  // scale: the largest magnitude, rounded to a power of two
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int j = 0; j < n; ++j) m = fmax(m, fabs(di[j]));
    for (int j = 0; j + 1 < n; ++j) m = fmax(m, fabs(ei[j]));
    smax = m;
  }
  __syncthreads();
  double scale = 1.0;
  if (smax > 0.0 && isfinite(smax)) {
    int ex = 0;
    (void)frexp(smax, &ex);  // smax = f 2^ex, f in [0.5, 1)
    scale = ldexp(1.0, -ex);
  }
  // G = B^T: column j holds d_j at row j and e_j at row j+1
  for (int q = threadIdx.x; q < np * np; q += kBsvdThreads) {
    const int col = q / np, row = q % np;
    double v = 0.0;
    if (col < n && row == col) v = di[col] * scale;
    if (col + 1 < n && row == col + 1) v = ei[col] * scale;
    G[q] = v;
  }
  for (int q = threadIdx.x; q < np; q += kBsvdThreads) r[q] = q == n - 1 ? 1.0 : 0.0;
  __syncthreads();

  const double tol = sqrt(static_cast<double>(n)) * 2.220446049250313e-16;
  const int m = np - 1;  // players 1..m rotate around player 0
  for (int sweep = 0; sweep < kBsvdMaxSweeps; ++sweep) {
    int any = 0;
    for (int t = 0; t < m; ++t) {
      for (int k = wid; k < np / 2; k += kBsvdWaves) {
        int p, q;
        if (k == 0) {
          p = 0;
          q = 1 + t % m;
        } else {
          p = 1 + (t + k) % m;
          q = 1 + (t - k + m) % m;
        }
        if (p > q) {
          const int s = p;
          p = q;
          q = s;
        }
        double* gp = G + static_cast<size_t>(p) * np;
        double* gq = G + static_cast<size_t>(q) * np;
        double a = 0.0, b = 0.0, c = 0.0;
        for (int row = lane; row < np; row += kWave) {
          const double x = gp[row], y = gq[row];
          a += x * x;
          b += y * y;
          c += x * y;
        }
        wave_sum3(a, b, c);
        if (fabs(c) > tol * sqrt(a) * sqrt(b)) {
          any = 1;
          const double zeta = (b - a) / (2.0 * c);
          const double az = fabs(zeta);
          const double root = az > 1e100 ? az : sqrt(1.0 + zeta * zeta);
          const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (az + root);
          const double cs = 1.0 / sqrt(1.0 + tt * tt);
          const double sn = cs * tt;
          for (int row = lane; row < np; row += kWave) {
            const double x = gp[row], y = gq[row];
            gp[row] = cs * x - sn * y;
            gq[row] = sn * x + cs * y;
          }
          if (lane == 0) {
            const double x = r[p], y = r[q];
            r[p] = cs * x - sn * y;
            r[q] = sn * x + cs * y;
          }
        }
      }
      __syncthreads();
    }
    if (lane == 0) rotated[wid] = any;
    __syncthreads();
    int done = 1;
    for (int w = 0; w < kBsvdWaves; ++w) done &= rotated[w] == 0;
    __syncthreads();
    if (done) break;
  }
  // column norms, then each value's rank in decreasing order (ties by index)
  for (int col = wid; col < np; col += kBsvdWaves) {
    const double* gc = G + static_cast<size_t>(col) * np;
    double s = 0.0, z0 = 0.0, z1 = 0.0;
    for (int row = lane; row < np; row += kWave) s += gc[row] * gc[row];
    wave_sum3(s, z0, z1);
    if (lane == 0) nrm[col] = sqrt(s);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += kBsvdThreads) {
    const double sj = nrm[j];
    int rank = 0;
    for (int l = 0; l < n; ++l) {
      const double sl = nrm[l];
      rank += (sl > sj || (sl == sj && l < j)) ? 1 : 0;
    }
    sigma[static_cast<size_t>(i) * ld + rank] = sj / scale;
    bot[static_cast<size_t>(i) * ld + rank] = r[j];
  }
}

}  // namespace

size_t bsvd_lds_bytes(int nmax) {
  const size_t np = static_cast<size_t>((nmax + 1) & ~1);
  return sizeof(double) * (np * np + 2 * np);
}

int bsvd_launch(hipStream_t s, int nb, int nmax, const int* dn, const double* dd,
                const double* de, int ld, double* dsig, double* dbot) {
  if (nb <= 0) return PK_OK;
  hipLaunchKernelGGL(k_bsvd, dim3(nb), dim3(kBsvdThreads), bsvd_lds_bytes(nmax), s, dn, dd, de,
                     ld, dsig, dbot);
  return hipGetLastError() == hipSuccess ? PK_OK : PK_ERR_HIP;
}

// Host arrays in, host arrays out, through device buffers that grow as
// needed (buf holds them between calls; the caller serializes its calls).
int bsvd_host(BsvdBuf& buf, hipStream_t s, int nb, const int* n, const double* d,
              const double* e, int ld, double* sigma, double* bot) {
  if (nb < 0 || ld < 1) {
    set_error("bidiagonal SVD: nb=%d ld=%d", nb, ld);
    return PK_ERR_ARG;
  }
  int nmax = 0;
  for (int i = 0; i < nb; ++i) {
    if (n[i] > ld || n[i] > kBsvdMax) {
      set_error("bidiagonal SVD: order %d of problem %d exceeds ld=%d or the limit %d", n[i], i,
                ld, kBsvdMax);
      return PK_ERR_ARG;
    }
    nmax = n[i] > nmax ? n[i] : nmax;
  }
  if (nb == 0 || nmax == 0) return PK_OK;
  const size_t vals = static_cast<size_t>(nb) * ld;
  if (buf.nb < nb || buf.ld != ld) {
    (void)hipFree(buf.dn);
    (void)hipFree(buf.dv);
    buf.dn = nullptr;
    buf.dv = nullptr;
    buf.nb = 0;
    if (hipMalloc(&buf.dn, sizeof(int) * nb) != hipSuccess ||
        hipMalloc(&buf.dv, sizeof(double) * vals * 4) != hipSuccess) {
      set_error("bidiagonal SVD: device buffers for %d problems of %d values", nb, ld);
      return PK_ERR_OOM;
    }
    buf.nb = nb;
    buf.ld = ld;
  }
  double* dd = buf.dv;
  double* de = dd + vals;
  double* dsig = de + vals;
  double* dbot = dsig + vals;
  if (hipMemcpyAsync(buf.dn, n, sizeof(int) * nb, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dd, d, sizeof(double) * vals, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(de, e, sizeof(double) * vals, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(dsig, 0, sizeof(double) * vals * 2, s) != hipSuccess)
    return PK_ERR_HIP;
  if (bsvd_launch(s, nb, nmax, buf.dn, dd, de, ld, dsig, dbot) != PK_OK) {
    set_error("bidiagonal SVD: launch failed");
    return PK_ERR_HIP;
  }
  if (hipMemcpyAsync(sigma, dsig, sizeof(double) * vals, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(bot, dbot, sizeof(double) * vals, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return PK_ERR_HIP;
  return PK_OK;
}

void bsvd_free(BsvdBuf& buf) {
  (void)hipFree(buf.dn);
  (void)hipFree(buf.dv);
  buf = BsvdBuf{};
}

}  // namespace pk

using namespace pk;

extern "C" int pk_bsvd_batch(int nb, const int* n, const double* d, const double* e, int ld,
                             double* sigma, double* bottom) {
  BsvdBuf buf;
  const int rc = bsvd_host(buf, nullptr, nb, n, d, e, ld, sigma, bottom);
  bsvd_free(buf);
  return rc;
}
