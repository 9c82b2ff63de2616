// Complex-valued kernels of the complex precisions (ZLANSVD / ZLANSVD_IRL,
// complex16/, and the CLANSVD twins, complex8/).
//
// PROPACK's complex drivers keep the bidiagonal B real
// (complex16/zlanbpro.F:111-112); only the vector work is complex: APROD
// applies A or A^H (TRANSA = 'c', complex16/zlanbpro.F:297), dot products
// conjugate their first argument (pzdotc, complex16/zblasext.F:167-194) and
// the Gram-Schmidt projections are V^H w (complex16/zreorth.F:169-170).
// Bases and operator values are stored as Fortran COMPLEX: (re, im) pairs.
// What is real-linear in the pairs (norms, scaling by a real factor, the
// restart / Ritz GEMMs with a real right factor) runs the real kernels on
// the 2L-long real view; the kernels here are the ones that need complex
// arithmetic.  Sums are formed in FP64 in a fixed order (deterministic), the
// results rounded once to the storage precision:
//   k_zspmv          y = op(A) x - c yprev, op = A or A^H (CSR of A, and of
//                    the explicit transpose A^T read with conjugated values),
//                    one group of G lanes per row, grid-stride over row
//                    chunks; partials ||y||^2, Re / Im of yprev^H y
//   k_zspmv_cp       the same over column panels, every panel in one launch:
//                    each panel's FP64 row sums to an array of its own, the
//                    workgroups of a panel on the XCDs that hold its slice
//                    of x; k_zcombine adds the panels' sums in panel order
//                    and applies the epilogue
//   k_zaxpy_norm_dot w -= s v for complex s; partials ||w||^2, Re / Im v^H w
//   k_zgemvt         h_j = V(:, cols[j])^H w, two partials per column
//   k_zgemv_update   w -= V(:, cols) h; partial ||w||^2
// Partial sums follow the real kernels' layout: value v of block b at
// part[v * nblk + b], reduced by launch_reduce.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pk_internal.h"

namespace pk {

void set_error(const char* fmt, ...);

namespace {

constexpr int kZRPT = 4;                   // rows per thread of the streaming kernels
constexpr int kZRows = kBlock * kZRPT;     // rows per workgroup slice
constexpr int kZCols = 4;                  // basis columns per pass of the Gram-Schmidt kernels

template <typename R>
struct V2T;
template <>
struct V2T<double> {
  using type = double2;
};
template <>
struct V2T<float> {
  using type = float2;
};

// ----------------------------------------------------------------- SpMV ----
// The rows come in chunks of kBlock / G, one row per group of G lanes;
// workgroup b takes chunks b, b + nblk, ... (a fixed assignment, so the
// partials are deterministic), and every lane of a wave runs the same trip
// count, so the group shuffles see whole groups.  Lane l of a group adds the
// row's nonzeros l, l + G, ..., two at a time so that two gathers of x are
// in flight; the group sum and the epilogue follow.
template <typename R, int G>
__global__ __launch_bounds__(kBlock) void k_zspmv(long long rows, const int* __restrict__ rp,
                                                  const int* __restrict__ ci,
                                                  const typename V2T<R>::type* __restrict__ va,
                                                  int conj,
                                                  const typename V2T<R>::type* __restrict__ x,
                                                  typename V2T<R>::type* __restrict__ y,
                                                  const typename V2T<R>::type* __restrict__ yprev,
                                                  double c, double* __restrict__ part) {
  using C2 = typename V2T<R>::type;
  __shared__ double lds[(kBlock / kWave) * 3];
  constexpr int ngrp = kBlock / G;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const double sg = conj ? -1.0 : 1.0;
  double acc[3] = {0.0, 0.0, 0.0};
  const long long stride = static_cast<long long>(gridDim.x) * ngrp;
  for (long long base = static_cast<long long>(blockIdx.x) * ngrp; base < rows; base += stride) {
    const long long row = base + grp;
    double sr = 0.0, si = 0.0;
    if (row < rows) {
      const int p1 = rp[row + 1];
      int p = rp[row] + lane;
      for (; p + G < p1; p += 2 * G) {
        const C2 a0 = va[p], a1 = va[p + G];
        const C2 x0 = x[ci[p]], x1 = x[ci[p + G]];
        const double ar0 = static_cast<double>(a0.x), ai0 = sg * static_cast<double>(a0.y);
        const double ar1 = static_cast<double>(a1.x), ai1 = sg * static_cast<double>(a1.y);
        sr += ar0 * static_cast<double>(x0.x) - ai0 * static_cast<double>(x0.y);
        si += ar0 * static_cast<double>(x0.y) + ai0 * static_cast<double>(x0.x);
        sr += ar1 * static_cast<double>(x1.x) - ai1 * static_cast<double>(x1.y);
        si += ar1 * static_cast<double>(x1.y) + ai1 * static_cast<double>(x1.x);
      }
      if (p < p1) {
        const C2 a = va[p];
        const C2 xv = x[ci[p]];
        const double ar = static_cast<double>(a.x), ai = sg * static_cast<double>(a.y);
        sr += ar * static_cast<double>(xv.x) - ai * static_cast<double>(xv.y);
        si += ar * static_cast<double>(xv.y) + ai * static_cast<double>(xv.x);
      }
    }
    sr = group_sum<G>(sr);
    si = group_sum<G>(si);
    if (lane == 0 && row < rows) {
      double yr = sr, yi = si, pr = 0.0, pi = 0.0;
      if (yprev) {
        const C2 pv = yprev[row];
        pr = static_cast<double>(pv.x);
        pi = static_cast<double>(pv.y);
        yr = sr - c * pr;
        yi = si - c * pi;
      }
      C2 out;
      out.x = static_cast<R>(yr);
      out.y = static_cast<R>(yi);
      y[row] = out;
      const double qr = static_cast<double>(out.x), qi = static_cast<double>(out.y);
      acc[0] += qr * qr + qi * qi;
      acc[1] += pr * qr + pi * qi;  // conj(yprev) * y
      acc[2] += pr * qi - pi * qr;
    }
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc[0];
    part[gridDim.x + blockIdx.x] = acc[1];
    part[2 * gridDim.x + blockIdx.x] = acc[2];
  }
}

// Every column panel in one launch.  Workgroup w runs panel w mod np (the
// dispatcher sends workgroup w to XCD w mod 8, so with np = 8 each XCD's L2
// holds one panel's slice of x, with np = 4 two XCDs share one), taking that
// panel's row chunks w / np, w / np + nblk / np, ...; the panel's FP64 row
// sums go to ypart[p * rows + row] as (re, im).
struct ZPanels {
  const int* rp[kMaxZPanels];
  const int* ci[kMaxZPanels];
  const void* va[kMaxZPanels];
};

template <typename R, int G>
__global__ __launch_bounds__(kBlock) void k_zspmv_cp(long long rows, int np, ZPanels tab,
                                                     int conj,
                                                     const typename V2T<R>::type* __restrict__ x,
                                                     double2* __restrict__ ypart) {
  using C2 = typename V2T<R>::type;
  constexpr int ngrp = kBlock / G;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G;
  const int pn = static_cast<int>(blockIdx.x % np);
  const int* __restrict__ rp = tab.rp[pn];
  const int* __restrict__ ci = tab.ci[pn];
  const C2* __restrict__ va = static_cast<const C2*>(tab.va[pn]);
  double2* __restrict__ yp = ypart + static_cast<long long>(pn) * rows;
  const double sg = conj ? -1.0 : 1.0;
  const long long stride = static_cast<long long>(gridDim.x / np) * ngrp;
  for (long long base = static_cast<long long>(blockIdx.x / np) * ngrp; base < rows;
       base += stride) {
    const long long row = base + grp;
    double sr = 0.0, si = 0.0;
    if (row < rows) {
      const int p1 = rp[row + 1];
      int p = rp[row] + lane;
      for (; p + G < p1; p += 2 * G) {
        const C2 a0 = va[p], a1 = va[p + G];
        const C2 x0 = x[ci[p]], x1 = x[ci[p + G]];
        const double ar0 = static_cast<double>(a0.x), ai0 = sg * static_cast<double>(a0.y);
        const double ar1 = static_cast<double>(a1.x), ai1 = sg * static_cast<double>(a1.y);
        sr += ar0 * static_cast<double>(x0.x) - ai0 * static_cast<double>(x0.y);
        si += ar0 * static_cast<double>(x0.y) + ai0 * static_cast<double>(x0.x);
        sr += ar1 * static_cast<double>(x1.x) - ai1 * static_cast<double>(x1.y);
        si += ar1 * static_cast<double>(x1.y) + ai1 * static_cast<double>(x1.x);
      }
      if (p < p1) {
        const C2 a = va[p];
        const C2 xv = x[ci[p]];
        const double ar = static_cast<double>(a.x), ai = sg * static_cast<double>(a.y);
        sr += ar * static_cast<double>(xv.x) - ai * static_cast<double>(xv.y);
        si += ar * static_cast<double>(xv.y) + ai * static_cast<double>(xv.x);
      }
    }
    sr = group_sum<G>(sr);
    si = group_sum<G>(si);
    if (lane == 0 && row < rows) yp[row] = make_double2(sr, si);
  }
}

// y = T(sum over panels, in panel order, of ypart - c yprev); partials
// ||y||^2, Re / Im of yprev^H y.  NP > 0: the panel count is known at
// compile time and every thread's kZRPT x NP loads (clamped to row rows - 1;
// a row past the end is loaded but neither stored nor counted) are issued
// before the first add; NP = 0 loops over np.  Both add in the same order.
template <typename R, int NP>
__global__ __launch_bounds__(kBlock) void k_zcombine(long long rows, int np,
                                                     const double2* __restrict__ ypart,
                                                     const typename V2T<R>::type* __restrict__ yprev,
                                                     double c,
                                                     typename V2T<R>::type* __restrict__ y,
                                                     double* __restrict__ part) {
  using C2 = typename V2T<R>::type;
  __shared__ double lds[(kBlock / kWave) * 3];
  const long long base = static_cast<long long>(blockIdx.x) * kZRows + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  double2 sum[kZRPT] = {};
  if (rows <= 0) {
    // no rows: zero partials below
  } else if constexpr (NP > 0) {
    double2 t[kZRPT][NP];
#pragma unroll
    for (int r = 0; r < kZRPT; ++r) {
      const long long row = min(base + static_cast<long long>(r) * kBlock, rows - 1);
#pragma unroll
      for (int q = 0; q < NP; ++q) t[r][q] = ypart[static_cast<long long>(q) * rows + row];
    }
#pragma unroll
    for (int r = 0; r < kZRPT; ++r) {
      sum[r] = t[r][0];
#pragma unroll
      for (int q = 1; q < NP; ++q) {
        sum[r].x += t[r][q].x;
        sum[r].y += t[r][q].y;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kZRPT; ++r) {
      const long long row = min(base + static_cast<long long>(r) * kBlock, rows - 1);
      sum[r] = ypart[row];
      for (int q = 1; q < np; ++q) {
        const double2 t = ypart[static_cast<long long>(q) * rows + row];
        sum[r].x += t.x;
        sum[r].y += t.y;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kZRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < rows) {
      double yr = sum[r].x, yi = sum[r].y, pr = 0.0, pi = 0.0;
      if (yprev) {
        const C2 pv = yprev[row];
        pr = static_cast<double>(pv.x);
        pi = static_cast<double>(pv.y);
        yr = sum[r].x - c * pr;
        yi = sum[r].y - c * pi;
      }
      C2 out;
      out.x = static_cast<R>(yr);
      out.y = static_cast<R>(yi);
      y[row] = out;
      const double qr = static_cast<double>(out.x), qi = static_cast<double>(out.y);
      acc[0] += qr * qr + qi * qi;
      acc[1] += pr * qr + pi * qi;
      acc[2] += pr * qi - pi * qr;
    }
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc[0];
    part[gridDim.x + blockIdx.x] = acc[1];
    part[2 * gridDim.x + blockIdx.x] = acc[2];
  }
}

// ------------------------------------------------------------- level 1 ----
template <typename R>
__global__ __launch_bounds__(kBlock) void k_zaxpy_norm_dot(long long L,
                                                           typename V2T<R>::type* __restrict__ w,
                                                           const typename V2T<R>::type* __restrict__ v,
                                                           double sr, double si,
                                                           double* __restrict__ part) {
  using C2 = typename V2T<R>::type;
  __shared__ double lds[(kBlock / kWave) * 3];
  const long long base = static_cast<long long>(blockIdx.x) * kZRows + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < kZRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      const C2 vv = v[row];
      const C2 wv = w[row];
      const double vr = static_cast<double>(vv.x), vi = static_cast<double>(vv.y);
      C2 out;
      out.x = static_cast<R>(static_cast<double>(wv.x) - (sr * vr - si * vi));
      out.y = static_cast<R>(static_cast<double>(wv.y) - (sr * vi + si * vr));
      w[row] = out;
      const double qr = static_cast<double>(out.x), qi = static_cast<double>(out.y);
      acc[0] += qr * qr + qi * qi;
      acc[1] += vr * qr + vi * qi;
      acc[2] += vr * qi - vi * qr;
    }
  }
  block_sum<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc[0];
    part[gridDim.x + blockIdx.x] = acc[1];
    part[2 * gridDim.x + blockIdx.x] = acc[2];
  }
}

// ------------------------------------------------------- Gram-Schmidt ----
// Slice b of h = V(:, cols)^H w: the slice of w stays in registers while
// the columns stream past in groups of kZCols; partials of column j at
// part[(2j) * nblk + b] (real) and part[(2j + 1) * nblk + b] (imaginary).
// Every load is unconditional: rows past L read row L-1 against a zero w,
// columns past ncols read the last listed column and are not stored, so a
// group's loads are all in flight before its sums.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_zgemvt(long long L,
                                                   const typename V2T<R>::type* __restrict__ V,
                                                   long long ldv, const int* __restrict__ cols,
                                                   int ncols,
                                                   const typename V2T<R>::type* __restrict__ w,
                                                   double* __restrict__ part) {
  using C2 = typename V2T<R>::type;
  __shared__ double lds[(kBlock / kWave) * 2 * kZCols];
  const size_t nb = gridDim.x;
  if (L <= 0) {  // empty vectors: zero projections, no loads
    for (int j = threadIdx.x; j < 2 * ncols; j += kBlock) part[j * nb + blockIdx.x] = 0.0;
    return;
  }
  const long long base = static_cast<long long>(blockIdx.x) * kZRows + threadIdx.x;
  double wr[kZRPT], wi[kZRPT];
  long long rowc[kZRPT];
#pragma unroll
  for (int r = 0; r < kZRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    rowc[r] = row < L ? row : L - 1;
    const C2 wv = w[rowc[r]];
    wr[r] = row < L ? static_cast<double>(wv.x) : 0.0;
    wi[r] = row < L ? static_cast<double>(wv.y) : 0.0;
  }
  for (int j0 = 0; j0 < ncols; j0 += kZCols) {
    C2 a[kZCols][kZRPT];
#pragma unroll
    for (int q = 0; q < kZCols; ++q) {
      const int jq = j0 + q < ncols ? j0 + q : ncols - 1;
      const C2* col = V + static_cast<long long>(cols[jq]) * ldv;
#pragma unroll
      for (int r = 0; r < kZRPT; ++r) a[q][r] = col[rowc[r]];
    }
    double acc[2 * kZCols];
#pragma unroll
    for (int q = 0; q < kZCols; ++q) {
      acc[2 * q] = acc[2 * q + 1] = 0.0;
#pragma unroll
      for (int r = 0; r < kZRPT; ++r) {
        const double ar = static_cast<double>(a[q][r].x), ai = static_cast<double>(a[q][r].y);
        acc[2 * q] += ar * wr[r] + ai * wi[r];  // conj(a) * w
        acc[2 * q + 1] += ar * wi[r] - ai * wr[r];
      }
    }
    block_sum<2 * kZCols>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < kZCols; ++q)
        if (j0 + q < ncols) {
          part[static_cast<size_t>(2 * (j0 + q)) * nb + blockIdx.x] = acc[2 * q];
          part[static_cast<size_t>(2 * (j0 + q) + 1) * nb + blockIdx.x] = acc[2 * q + 1];
        }
    }
  }
}

// w -= V(:, cols) h with h[2j], h[2j+1] the real and imaginary parts of h_j;
// the update of a row is summed in FP64 over the columns in order, then
// subtracted and rounded once.  Columns go kZCols at a time with their loads
// unconditional (past ncols: the last listed column with a zero weight;
// past L: row L-1, not stored).  Partial: ||w||^2.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_zgemv_update(long long L,
                                                         const typename V2T<R>::type* __restrict__ V,
                                                         long long ldv,
                                                         const int* __restrict__ cols, int ncols,
                                                         const double* __restrict__ h,
                                                         typename V2T<R>::type* __restrict__ w,
                                                         double* __restrict__ part) {
  using C2 = typename V2T<R>::type;
  __shared__ double lds[kBlock / kWave];
  if (L <= 0) {
    if (threadIdx.x == 0) part[blockIdx.x] = 0.0;
    return;
  }
  const long long base = static_cast<long long>(blockIdx.x) * kZRows + threadIdx.x;
  long long rowc[kZRPT];
  double ur[kZRPT], ui[kZRPT];
#pragma unroll
  for (int r = 0; r < kZRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    rowc[r] = row < L ? row : L - 1;
    ur[r] = ui[r] = 0.0;
  }
  for (int j0 = 0; j0 < ncols; j0 += kZCols) {
    C2 a[kZCols][kZRPT];
    double hr[kZCols], hi[kZCols];
#pragma unroll
    for (int q = 0; q < kZCols; ++q) {
      const bool in = j0 + q < ncols;
      const int jq = in ? j0 + q : ncols - 1;
      const C2* col = V + static_cast<long long>(cols[jq]) * ldv;
      hr[q] = in ? h[2 * jq] : 0.0;
      hi[q] = in ? h[2 * jq + 1] : 0.0;
#pragma unroll
      for (int r = 0; r < kZRPT; ++r) a[q][r] = col[rowc[r]];
    }
#pragma unroll
    for (int q = 0; q < kZCols; ++q)
#pragma unroll
      for (int r = 0; r < kZRPT; ++r) {
        const double ar = static_cast<double>(a[q][r].x), ai = static_cast<double>(a[q][r].y);
        ur[r] += ar * hr[q] - ai * hi[q];
        ui[r] += ar * hi[q] + ai * hr[q];
      }
  }
  double acc[1] = {0.0};
#pragma unroll
  for (int r = 0; r < kZRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      const C2 wv = w[row];
      C2 out;
      out.x = static_cast<R>(static_cast<double>(wv.x) - ur[r]);
      out.y = static_cast<R>(static_cast<double>(wv.y) - ui[r]);
      w[row] = out;
      const double qr = static_cast<double>(out.x), qi = static_cast<double>(out.y);
      acc[0] += qr * qr + qi * qi;
    }
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

}  // namespace

int zstream_blocks(long long L) { return L > 0 ? cdiv(L, kZRows) : 1; }

// Workgroups of the complex SpMV: one per chunk of kBlock / g rows, at most
// kMaxBlocks (eight per CU: a full wave complement), the chunks handed out
// grid-stride.
int zspmv_blocks(long long rows, int g) {
  const long long chunks = rows > 0 ? (rows + kBlock / g - 1) / (kBlock / g) : 1;
  return static_cast<int>(std::min<long long>(chunks, kMaxBlocks));
}

// Lanes per row of the complex SpMV: the power of two nearest half the mean
// row length (two nonzeros per lane), clamped to [2, 64].
int zspmv_lanes(long long nnz, long long rows) {
  const double avg = rows > 0 ? static_cast<double>(nnz) / static_cast<double>(rows) : 1.0;
  int g = 2;
  while (g < 64 && g * 3.0 < avg) g *= 2;
  return g;
}

template <typename R>
int launch_zspmv(hipStream_t s, long long rows, const int* rp, const int* ci, const void* va,
                 int conj, int g, const void* x, void* y, const void* yprev, double c,
                 double* part) {
  using C2 = typename V2T<R>::type;
  const int nblk = zspmv_blocks(rows, g);
#define PK_ZSPMV(G)                                                                            \
  hipLaunchKernelGGL((k_zspmv<R, G>), dim3(nblk), dim3(kBlock), 0, s, rows, rp, ci,             \
                     static_cast<const C2*>(va), conj, static_cast<const C2*>(x),               \
                     static_cast<C2*>(y), static_cast<const C2*>(yprev), c, part)
  switch (g) {
    case 2: PK_ZSPMV(2); break;
    case 4: PK_ZSPMV(4); break;
    case 8: PK_ZSPMV(8); break;
    case 16: PK_ZSPMV(16); break;
    case 32: PK_ZSPMV(32); break;
    default: PK_ZSPMV(64); break;
  }
#undef PK_ZSPMV
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

template <typename R>
int launch_zaxpy_norm_dot(hipStream_t s, long long L, void* w, const void* v, double sr,
                          double si, double* part) {
  using C2 = typename V2T<R>::type;
  const int nblk = zstream_blocks(L);
  hipLaunchKernelGGL(k_zaxpy_norm_dot<R>, dim3(nblk), dim3(kBlock), 0, s, L, static_cast<C2*>(w),
                     static_cast<const C2*>(v), sr, si, part);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

template <typename R>
int launch_zgemvt(hipStream_t s, long long L, const void* V, long long ldv, const int* cols,
                  int ncols, const void* w, double* part) {
  using C2 = typename V2T<R>::type;
  const int nblk = zstream_blocks(L);
  hipLaunchKernelGGL(k_zgemvt<R>, dim3(nblk), dim3(kBlock), 0, s, L, static_cast<const C2*>(V),
                     ldv, cols, ncols, static_cast<const C2*>(w), part);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

template <typename R>
int launch_zgemv_update(hipStream_t s, long long L, const void* V, long long ldv,
                        const int* cols, int ncols, const double* h, void* w, double* part) {
  using C2 = typename V2T<R>::type;
  const int nblk = zstream_blocks(L);
  hipLaunchKernelGGL(k_zgemv_update<R>, dim3(nblk), dim3(kBlock), 0, s, L,
                     static_cast<const C2*>(V), ldv, cols, ncols, h, static_cast<C2*>(w), part);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

long long zcsr_ywork(const CsrOp* op) {
  long long w = 0;
  for (const CsrOp::Dir* D : {&op->da, &op->dt})
    if (D->np > 1) w = std::max(w, 2LL * D->np * D->rows);
  return w;
}

template <typename R>
static int launch_zspmv_cp(hipStream_t s, const CsrOp::Dir& D, int conj, const void* x,
                           void* y, const void* yprev, double c, double* part, double* ywork) {
  using C2 = typename V2T<R>::type;
  ZPanels tab{};
  for (int p = 0; p < D.np; ++p) {
    tab.rp[p] = D.pan[p].rp;
    tab.ci[p] = D.pan[p].ci;
    tab.va[p] = D.pan[p].va;
  }
  const int g = D.pan[0].g, np = D.np;
  const long long rows = D.rows;
  // per panel as many workgroups as the one-panel kernel would use, the
  // grid a multiple of np
  const int per = std::max(1, std::min(zspmv_blocks(rows, g), kMaxBlocks / np));
  const int grid = per * np;
  double2* yp = reinterpret_cast<double2*>(ywork);
#define PK_ZSPMV_CP(G)                                                                       \
  hipLaunchKernelGGL((k_zspmv_cp<R, G>), dim3(grid), dim3(kBlock), 0, s, rows, np, tab, conj, \
                     static_cast<const C2*>(x), yp)
  switch (g) {
    case 2: PK_ZSPMV_CP(2); break;
    case 4: PK_ZSPMV_CP(4); break;
    case 8: PK_ZSPMV_CP(8); break;
    case 16: PK_ZSPMV_CP(16); break;
    case 32: PK_ZSPMV_CP(32); break;
    default: PK_ZSPMV_CP(64); break;
  }
#undef PK_ZSPMV_CP
  if (hipGetLastError() != hipSuccess) return -1;
  const int nblk = zstream_blocks(rows);
#define PK_ZCB(NP)                                                                         \
  hipLaunchKernelGGL((k_zcombine<R, NP>), dim3(nblk), dim3(kBlock), 0, s, rows, np, yp,      \
                     static_cast<const C2*>(yprev), c, static_cast<C2*>(y), part)
  switch (np) {
    case 2: PK_ZCB(2); break;
    case 4: PK_ZCB(4); break;
    case 8: PK_ZCB(8); break;
    default: PK_ZCB(0); break;
  }
#undef PK_ZCB
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// y = op(A) x - c * yprev for a complex operator (trans 'n': A; 't': A^T;
// 'c': A^H); returns the partial-slot count (three values per slot).
template <typename R>
int zcsr_apply(const CsrOp* op, hipStream_t s, char trans, const void* x, void* y,
               const void* yprev, double c, double* part, double* ywork) {
  const bool tr = trans != 'n' && trans != 'N';
  const bool cj = trans == 'c' || trans == 'C';
  const CsrOp::Dir& D = tr ? op->dt : op->da;
  if (D.np > 1) {
    if (!ywork) return -1;
    return launch_zspmv_cp<R>(s, D, cj ? 1 : 0, x, y, yprev, c, part, ywork);
  }
  const CsrOp::Panel& P = D.pan[0];
  return launch_zspmv<R>(s, D.rows, P.rp, P.ci, P.va, cj ? 1 : 0, P.g, x, y, yprev, c, part);
}

#define PK_INST_Z(R)                                                                            \
  template int launch_zspmv<R>(hipStream_t, long long, const int*, const int*, const void*, int, \
                               int, const void*, void*, const void*, double, double*);          \
  template int launch_zaxpy_norm_dot<R>(hipStream_t, long long, void*, const void*, double,     \
                                        double, double*);                                       \
  template int launch_zgemvt<R>(hipStream_t, long long, const void*, long long, const int*, int, \
                                const void*, double*);                                          \
  template int launch_zgemv_update<R>(hipStream_t, long long, const void*, long long,           \
                                      const int*, int, const double*, void*, double*);          \
  template int zcsr_apply<R>(const CsrOp*, hipStream_t, char, const void*, void*, const void*,  \
                             double, double*, double*);
PK_INST_Z(double)
PK_INST_Z(float)
#undef PK_INST_Z

}  // namespace pk
