// Hot-path kernels of the Lanczos bidiagonalization engine, written for
// gfx950 (CDNA4, wave64).
//
//   k_spmv_rb       y = op(A) x - c * yprev over a CSR operator (A or its
//                   explicit transpose) in row blocks staged through LDS,
//                   fused with the Lanczos recurrence update and the partial
//                   sums for ||y||^2 and yprev.y (the norm and the ELR dot
//                   product); column panels accumulate in FP64 across launches.
//   k_gemvt         h = V(:,I)^T w for the iterated classical Gram-Schmidt
//                   step; each workgroup owns a contiguous row slice, keeps
//                   its slice of w in registers and streams column tiles of V.
//   k_gemv_update   w -= V(:,I) h with h broadcast from LDS, fused with ||w||^2.
//   k_axpy_norm_dot, k_scale, k_norm   level-1 updates with fused reductions.
//   k_reduce        second reduction stage (fixed order => deterministic).
//   k_gemm_tall     C(:,1:n) = A(:,1:k) B for restarts / Ritz vectors (spare buffer).
#include "pk_kernels.h"

namespace pk {

// Rows per thread of the dense streaming kernels; each workgroup owns the
// contiguous slice [b * kRowsPerBlock, (b + 1) * kRowsPerBlock).
constexpr int kRPT = 4;
constexpr int kRowsPerBlock = kBlock * kRPT;

int stream_blocks(long long L) { return L > 0 ? cdiv(L, kRowsPerBlock) : 1; }

// ------------------------------------------------------ row-block SpMV ----
// The rows are partitioned into workgroup descriptors {x, y, z, w}:
//   * y >= 0: a row block, rows [x, y) holding at most kSpmvCap nonzeros
//     and kSpmvRows rows.  All 256 threads turn its nonzeros into products
//     val * x[col] in LDS -- index and value streams read fully coalesced,
//     kUnroll independent gathers in flight per thread -- then groups of G
//     lanes sum each row from LDS (consecutive, conflict-free reads;
//     shuffle reduction) in a fixed order and apply the fused epilogue.
//   * y < 0: chunk -(y+1) of the long row x, nonzeros [z, w): its partial
//     dot product goes to lsum[-(y+1)]; k_spmv_longfix finishes the row.
// Partials: part[b] = sum y^2, part[nbt + b] = sum yprev.y (zero for
// chunks); the long rows' contributions follow the blocks' slots.
// Column panels: an operator split into column panels runs one launch per
// panel; the row sums of the panels before the last one go to the FP64
// accumulator yacc (yacc_in = null for the first panel), and only the last
// launch (last = 1) adds its own sum and applies the epilogue and partials.
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_rb(const int4* __restrict__ desc,
                                                    const int* __restrict__ rp,
                                                    const int* __restrict__ ci,
                                                    const VT* __restrict__ va,
                                                    const T* __restrict__ x, T* __restrict__ y,
                                                    const T* __restrict__ yprev, Scal coef,
                                                    double* __restrict__ part, int nbt,
                                                    double* __restrict__ lsum,
                                                    const double* __restrict__ yacc_in,
                                                    double* __restrict__ yacc_out, int last,
                                                    Gate gate) {
  if (gate.off()) return;
  __shared__ double prod[kSpmvCap];
  __shared__ double lds[(kBlock / kWave) * 2];
  const int4 d = desc[blockIdx.x];
  if (d.y < 0) {
    double acc[1] = {0.0};
    for (int p = d.z + static_cast<int>(threadIdx.x); p < d.w; p += kBlock)
      acc[0] += static_cast<double>(va[p]) * static_cast<double>(x[ci[p]]);
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) {
      lsum[-(d.y + 1)] = acc[0];
      if (last) {
        part[blockIdx.x] = 0.0;
        part[nbt + blockIdx.x] = 0.0;
      }
    }
    return;
  }
  const double c = coef.get();
  const int r0 = d.x, r1 = d.y;
  const int p0 = rp[r0];
  const int nz = rp[r1] - p0;
  double s2 = 0.0, sd = 0.0;
  constexpr int kUnroll = 8;
  for (int base = 0; base < nz; base += kBlock * kUnroll) {
    int idx[kUnroll];
    double v[kUnroll], xv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int t = base + u * kBlock + static_cast<int>(threadIdx.x);
      idx[u] = t < nz ? ci[p0 + t] : -1;
      v[u] = t < nz ? static_cast<double>(va[p0 + t]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) xv[u] = idx[u] >= 0 ? static_cast<double>(x[idx[u]]) : 0.0;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int t = base + u * kBlock + static_cast<int>(threadIdx.x);
      if (t < nz) prod[t] = v[u] * xv[u];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & (G - 1);
  const int group = threadIdx.x / G;
  constexpr int ngroups = kBlock / G;
  const int nrows = r1 - r0;
  const int iters = (nrows + ngroups - 1) / ngroups;  // uniform across the block
  for (int it = 0; it < iters; ++it) {
    const int r = r0 + group + it * ngroups;
    const bool valid = r < r1;
    int a = 0, b = 0;
    if (valid) {
      a = rp[r] - p0;
      b = rp[r + 1] - p0;
    }
    double acc = 0.0;
    for (int q = a + lane; q < b; q += G) acc += prod[q];
    acc = group_sum<G>(acc);
    if (valid && lane == 0) {
      if (yacc_in) acc = yacc_in[r] + acc;
      if (!last) {
        yacc_out[r] = acc;
      } else {
        const double yp = yprev ? static_cast<double>(yprev[r]) : 0.0;
        const T yv = static_cast<T>(acc - c * yp);
        y[r] = yv;
        s2 += static_cast<double>(yv) * static_cast<double>(yv);
        sd += yp * static_cast<double>(yv);
      }
    }
  }
  if (!last) return;
  double v2[2] = {s2, sd};
  block_sum<2>(v2, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v2[0];
    part[nbt + blockIdx.x] = v2[1];
  }
}

// Long rows: ld[i] = {row, first chunk, number of chunks, -}; one thread per
// row sums its chunks in order and applies the epilogue; its two partials
// go to slots nbmain + i.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_longfix(int nlong, const int4* __restrict__ ld,
                                                         const double* __restrict__ lsum,
                                                         T* __restrict__ y,
                                                         const T* __restrict__ yprev, Scal coef,
                                                         double* __restrict__ part, int nbt,
                                                         int nbmain,
                                                         const double* __restrict__ yacc_in,
                                                         double* __restrict__ yacc_out, int last,
                                                         Gate gate) {
  if (gate.off()) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nlong) return;
  const int4 d = ld[i];
  double acc = 0.0;
  for (int q = 0; q < d.z; ++q) acc += lsum[d.y + q];
  if (yacc_in) acc = yacc_in[d.x] + acc;
  if (!last) {
    yacc_out[d.x] = acc;
    return;
  }
  const double yp = yprev ? static_cast<double>(yprev[d.x]) : 0.0;
  const T yv = static_cast<T>(acc - coef.get() * yp);
  y[d.x] = yv;
  part[nbmain + i] = static_cast<double>(yv) * static_cast<double>(yv);
  part[nbt + nbmain + i] = yp * static_cast<double>(yv);
}

template <typename VT, typename T>
int launch_spmv_rb(hipStream_t s, int nblk, const int4* desc, int nlong, const int4* ldesc,
                   double* lsum, const int* rp, const int* ci, const VT* va, const T* x, T* y,
                   const T* yprev, Scal coef, double* part, int g, const double* yacc_in,
                   double* yacc_out, int last, Gate gate) {
  const int nbt = nblk + nlong;
  if (nblk > 0) {
#define PK_RB_CASE(G)                                                                          \
  case G:                                                                                      \
    hipLaunchKernelGGL((k_spmv_rb<VT, T, G>), dim3(nblk), dim3(kBlock), 0, s, desc, rp, ci, va, \
                       x, y, yprev, coef, part, nbt, lsum, yacc_in, yacc_out, last, gate);     \
    break;
    switch (g) {
      PK_RB_CASE(2)
      PK_RB_CASE(4)
      PK_RB_CASE(8)
      PK_RB_CASE(16)
      PK_RB_CASE(32)
      PK_RB_CASE(64)
      default:
        return -1;
    }
#undef PK_RB_CASE
  }
  if (nlong > 0)
    hipLaunchKernelGGL(k_spmv_longfix<T>, dim3(cdiv(nlong, kBlock)), dim3(kBlock), 0, s, nlong,
                       ldesc, lsum, y, yprev, coef, part, nbt, nblk, yacc_in, yacc_out, last,
                       gate);
  return hipGetLastError() == hipSuccess ? nbt : -1;
}

// -------------------------------------------------------------- reduce ----
__global__ __launch_bounds__(kBlock) void k_reduce(int nblk, const double* __restrict__ part,
                                                   double* __restrict__ out,
                                                   const int* nval_p, Gate gate) {
  if (gate.off()) return;
  if (nval_p && static_cast<int>(blockIdx.x) >= *nval_p) return;
  __shared__ double lds[kBlock / kWave];
  const double v = reduce_partials(part + static_cast<size_t>(blockIdx.x) * nblk, nblk, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

int launch_reduce(hipStream_t s, int nval, int nblk, const double* part, double* out,
                  Gate gate, const int* nval_p) {
  if (nval <= 0) return 0;
  hipLaunchKernelGGL(k_reduce, dim3(nval), dim3(kBlock), 0, s, nblk, part, out, nval_p, gate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------- gemvt ----
// Column tile width: CT accumulators per thread, CT column loads in flight
// per row held.
constexpr int kCT = 8;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemvt(long long L, const T* __restrict__ V,
                                                  long long ldv, const int* __restrict__ cols,
                                                  int ncols_v, const int* ncols_p,
                                                  const T* __restrict__ w,
                                                  double* __restrict__ part, Gate gate) {
  if (gate.off()) return;
  const int ncols = ncols_p ? *ncols_p : ncols_v;
  __shared__ int scol[kMaxCols];
  __shared__ double lds[2][(kBlock / kWave) * kCT];
  for (int i = threadIdx.x; i < ncols; i += kBlock) scol[i] = cols[i];
  __syncthreads();
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
  double wr[kRPT];
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    wr[r] = row < L ? static_cast<double>(w[row]) : 0.0;
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  int buf = 0;
  for (int c0 = 0; c0 < ncols; c0 += kCT) {
    double acc[kCT];
#pragma unroll
    for (int c = 0; c < kCT; ++c) acc[c] = 0.0;
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const long long row = base + static_cast<long long>(r) * kBlock;
      if (row < L) {
#pragma unroll
        for (int c = 0; c < kCT; ++c) {
          if (c0 + c < ncols)
            acc[c] += static_cast<double>(V[static_cast<long long>(scol[c0 + c]) * ldv + row]) *
                      wr[r];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kCT; ++c) acc[c] = wave_sum(acc[c]);
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < kCT; ++c) lds[buf][wid * kCT + c] = acc[c];
    }
    __syncthreads();
    if (threadIdx.x < kCT && c0 + threadIdx.x < ncols) {
      double sum = 0.0;
      for (int ww = 0; ww < kBlock / kWave; ++ww) sum += lds[buf][ww * kCT + threadIdx.x];
      part[static_cast<size_t>(c0 + threadIdx.x) * gridDim.x + blockIdx.x] = sum;
    }
    buf ^= 1;  // the other buffer is free: its readers passed the barrier above
  }
}

template <typename T>
int launch_gemvt(hipStream_t s, long long L, const T* V, long long ldv, const int* cols,
                 int ncols, const int* ncols_p, const T* w, double* part, Gate gate) {
  if (ncols > kMaxCols) return -1;
  const int nblk = stream_blocks(L);
  hipLaunchKernelGGL(k_gemvt<T>, dim3(nblk), dim3(kBlock), 0, s, L, V, ldv, cols, ncols,
                     ncols_p, w, part, gate);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// --------------------------------------------------------- gemv update ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_update(long long L, const T* __restrict__ V,
                                                        long long ldv,
                                                        const int* __restrict__ cols,
                                                        int ncols_v, const int* ncols_p,
                                                        const double* __restrict__ h,
                                                        T* __restrict__ w,
                                                        double* __restrict__ part,
                                                        Gate gate) {
  if (gate.off()) return;
  const int ncols = ncols_p ? *ncols_p : ncols_v;
  __shared__ int scol[kMaxCols];
  __shared__ double sh[kMaxCols];
  __shared__ double lds[kBlock / kWave];
  for (int i = threadIdx.x; i < ncols; i += kBlock) {
    scol[i] = cols[i];
    sh[i] = h[i];
  }
  __syncthreads();
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
  double acc[kRPT];
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    acc[r] = row < L ? static_cast<double>(w[row]) : 0.0;
  }
  for (int c = 0; c < ncols; ++c) {
    const T* vc = V + static_cast<long long>(scol[c]) * ldv;
    const double hc = sh[c];
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const long long row = base + static_cast<long long>(r) * kBlock;
      if (row < L) acc[r] -= static_cast<double>(vc[row]) * hc;
    }
  }
  double v[1] = {0.0};
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      const T o = static_cast<T>(acc[r]);
      w[row] = o;
      v[0] += static_cast<double>(o) * static_cast<double>(o);
    }
  }
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

template <typename T>
int launch_gemv_update(hipStream_t s, long long L, const T* V, long long ldv, const int* cols,
                       int ncols, const int* ncols_p, const double* h, T* w, double* part,
                       Gate gate) {
  if (ncols > kMaxCols) return -1;
  const int nblk = stream_blocks(L);
  hipLaunchKernelGGL(k_gemv_update<T>, dim3(nblk), dim3(kBlock), 0, s, L, V, ldv, cols, ncols,
                     ncols_p, h, w, part, gate);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// ------------------------------------------------------------- level 1 ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(long long L, T* __restrict__ x, Scal a,
                                                  Gate gate) {
  if (gate.off()) return;
  const double av = a.get();
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) x[row] = static_cast<T>(av * static_cast<double>(x[row]));
  }
}

template <typename T>
int launch_scale(hipStream_t s, long long L, T* x, Scal a, Gate gate) {
  hipLaunchKernelGGL(k_scale<T>, dim3(stream_blocks(L)), dim3(kBlock), 0, s, L, x, a, gate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_norm_dot(long long L, T* __restrict__ w,
                                                          const T* __restrict__ v, Scal c,
                                                          double* __restrict__ part,
                                                          Gate gate) {
  if (gate.off()) return;
  __shared__ double lds[(kBlock / kWave) * 2];
  const double cv = c.get();
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      const double vv = static_cast<double>(v[row]);
      const T o = static_cast<T>(static_cast<double>(w[row]) - cv * vv);
      w[row] = o;
      acc[0] += static_cast<double>(o) * static_cast<double>(o);
      acc[1] += vv * static_cast<double>(o);
    }
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc[0];
    part[gridDim.x + blockIdx.x] = acc[1];
  }
}

template <typename T>
int launch_axpy_norm_dot(hipStream_t s, long long L, T* w, const T* v, Scal c, double* part,
                         Gate gate) {
  const int nblk = stream_blocks(L);
  hipLaunchKernelGGL(k_axpy_norm_dot<T>, dim3(nblk), dim3(kBlock), 0, s, L, w, v, c, part,
                     gate);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm(long long L, const T* __restrict__ x,
                                                 double* __restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
  double acc[1] = {0.0};
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) acc[0] += static_cast<double>(x[row]) * static_cast<double>(x[row]);
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

template <typename T>
int launch_norm(hipStream_t s, long long L, const T* x, double* part) {
  const int nblk = stream_blocks(L);
  hipLaunchKernelGGL(k_norm<T>, dim3(nblk), dim3(kBlock), 0, s, L, x, part);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_convert(long long L, const S* __restrict__ src,
                                                    D* __restrict__ dst) {
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) dst[row] = static_cast<D>(src[row]);
  }
}

template <typename S, typename D>
int launch_convert(hipStream_t s, long long L, const S* src, D* dst) {
  hipLaunchKernelGGL((k_convert<S, D>), dim3(stream_blocks(L)), dim3(kBlock), 0, s, L, src,
                     dst);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_zero_bytes(size_t n, unsigned char* p) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

int launch_fill_zero_bytes(hipStream_t s, void* p, size_t bytes) {
  if (bytes == 0) return 0;
  const int nblk = static_cast<int>(std::min<size_t>((bytes + 255) / 256, 4096));
  hipLaunchKernelGGL(k_zero_bytes, dim3(nblk), dim3(256), 0, s, bytes,
                     static_cast<unsigned char*>(p));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------- GEMM ----
// C(:,0:ncols) = A(:,0:kdim) * B (B kdim x ncols, double, column-major).
// Out of place: the engine writes into its spare basis buffer and swaps, so
// restarts and Ritz vectors need no in-place hazard handling.  One thread
// owns one row and up to NC output columns (register accumulators); B is
// read with wave-uniform addresses (scalar loads).  FP64 accumulation.
constexpr int kGemmNC = 64;

template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_gemm_tall(long long L, const T* __restrict__ A,
                                                      long long lda, int kdim,
                                                      const double* __restrict__ B, int ldb,
                                                      int ncols, T* __restrict__ C,
                                                      long long ldc) {
  const long long row = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (row >= L) return;
  double acc[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) acc[j] = 0.0;
  for (int l = 0; l < kdim; ++l) {
    const double a = static_cast<double>(A[static_cast<long long>(l) * lda + row]);
    const double* bl = B + l;
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (j < ncols) acc[j] += a * bl[static_cast<long long>(j) * ldb];
  }
#pragma unroll
  for (int j = 0; j < NC; ++j)
    if (j < ncols) C[static_cast<long long>(j) * ldc + row] = static_cast<T>(acc[j]);
}

template <typename T>
int launch_gemm_tall(hipStream_t s, long long L, const T* A, long long lda, int kdim,
                     const double* B, int ldb, int ncols, T* C, long long ldc) {
  if (L <= 0 || ncols <= 0 || kdim <= 0) return 0;
  const int nblk = cdiv(L, kBlock);
  for (int j0 = 0; j0 < ncols; j0 += kGemmNC) {
    const int nc = std::min(kGemmNC, ncols - j0);
    const double* Bj = B + static_cast<long long>(j0) * ldb;
    T* Cj = C + static_cast<long long>(j0) * ldc;
    if (nc <= 16)
      hipLaunchKernelGGL((k_gemm_tall<T, 16>), dim3(nblk), dim3(kBlock), 0, s, L, A, lda, kdim,
                         Bj, ldb, nc, Cj, ldc);
    else if (nc <= 32)
      hipLaunchKernelGGL((k_gemm_tall<T, 32>), dim3(nblk), dim3(kBlock), 0, s, L, A, lda, kdim,
                         Bj, ldb, nc, Cj, ldc);
    else
      hipLaunchKernelGGL((k_gemm_tall<T, 64>), dim3(nblk), dim3(kBlock), 0, s, L, A, lda, kdim,
                         Bj, ldb, nc, Cj, ldc);
    if (hipGetLastError() != hipSuccess) return -1;
  }
  return 0;
}

// ------------------------------------------------------------- control ----
// One workgroup: reduce the partials of the preceding vector kernel (same
// order as k_reduce), stage the control state in LDS, let thread 0 run the
// step's scalar logic, write the state back.  Runs only while the state is
// in (want_phase, want_j); otherwise the whole chain for this step is idle.
__global__ __launch_bounds__(kBlock) void k_ctl(LbdCtl* __restrict__ gc, int kind, int want_phase,
                                                int want_j, const double* __restrict__ part,
                                                int nb) {
  __shared__ LbdCtl c;
  __shared__ double lds[kBlock / kWave];
  if (gc->phase != want_phase || gc->j != want_j) return;
  const bool two = (kind == CTL_V_SPMV || kind == CTL_V_ELR || kind == CTL_U_SPMV ||
                    kind == CTL_U_ELR);
  const double a = reduce_partials(part, nb, lds);
  const double b = two ? reduce_partials(part + nb, nb, lds) : 0.0;
  constexpr int nw = sizeof(LbdCtl) / sizeof(int);
  static_assert(sizeof(LbdCtl) % sizeof(int) == 0, "state must be whole words");
  const int* src = reinterpret_cast<const int*>(gc);
  int* dst = reinterpret_cast<int*>(&c);
  for (int i = threadIdx.x; i < nw; i += kBlock) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    switch (kind) {
      case CTL_V_SPMV: ctl_v_spmv(c, a, b); break;
      case CTL_V_ELR: ctl_v_elr(c, a, b); break;
      case CTL_V_REO: ctl_v_reo(c, a); break;
      case CTL_U_SPMV: ctl_u_spmv(c, a, b); break;
      case CTL_U_ELR: ctl_u_elr(c, a, b); break;
      case CTL_U_REO: ctl_u_reo(c, a); break;
      default: break;
    }
  }
  __syncthreads();
  int* back = reinterpret_cast<int*>(gc);
  for (int i = threadIdx.x; i < nw; i += kBlock) back[i] = dst[i];
}

int launch_ctl(hipStream_t s, LbdCtl* c, int kind, int want_phase, int want_j,
               const double* part, int nb) {
  hipLaunchKernelGGL(k_ctl, dim3(1), dim3(kBlock), 0, s, c, kind, want_phase, want_j, part, nb);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ MFMA GEMM ----
// C(:,0:ncols) = A(:,0:kdim) * B on FP64 matrix cores
// (v_mfma_f64_16x16x4_f64).  Each wave owns 16 rows of C and up to 64
// columns (four 16-wide column tiles).  The product is formed transposed,
// C^T tile = B^T(16 x 4) * A^T(4 x 16), so that:
//   * the A^T operand of lane l is A[i0 + l%16, k + l/16]: four 128-byte
//     row segments of column-major A per instruction (coalesced);
//   * the B^T operand of lane l is B[k + l/16, j0 + l%16] (small, cached);
//   * accumulator r of lane l holds row l/16 + 4r of the 16x16 result, i.e.
//     C[i0 + l%16, j0 + l/16 + 4r], so each group of 16 lanes stores 128
//     contiguous bytes of one C column.
// Out-of-range rows and the k padding to a multiple of 4 read as zero.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kMfmaRowsPerWave = 16;
constexpr int kMfmaWaves = 4;

template <typename T>
__global__ __launch_bounds__(kWave * kMfmaWaves) void k_gemm_mfma(
    long long L, const T* __restrict__ A, long long lda, int kdim, const double* __restrict__ B,
    int ldb, int ncols, T* __restrict__ C, long long ldc) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const long long i0 =
      (static_cast<long long>(blockIdx.x) * kMfmaWaves + wid) * kMfmaRowsPerWave;
  if (i0 >= L) return;  // whole wave out of range (wave-uniform)
  const int li = lane % 16, lk = lane / 16;
  const long long row = i0 + li;
  const bool row_ok = row < L;
  const int ntiles = (ncols + 15) / 16;
  v4d acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < kdim; k0 += 4) {
    const int k = k0 + lk;
    const double a = (row_ok && k < kdim) ? static_cast<double>(A[static_cast<long long>(k) * lda + row]) : 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t < ntiles) {
        const int j = t * 16 + li;
        const double b = (k < kdim && j < ncols) ? B[static_cast<long long>(j) * ldb + k] : 0.0;
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, acc[t], 0, 0, 0);
      }
    }
  }
  if (!row_ok) return;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (t < ntiles) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = t * 16 + lk + 4 * r;
        if (j < ncols) C[static_cast<long long>(j) * ldc + row] = static_cast<T>(acc[t][r]);
      }
    }
  }
}

template <typename T>
int launch_gemm_mfma(hipStream_t s, long long L, const T* A, long long lda, int kdim,
                     const double* B, int ldb, int ncols, T* C, long long ldc) {
  if (L <= 0 || ncols <= 0 || kdim <= 0) return 0;
  const int rows_per_block = kMfmaRowsPerWave * kMfmaWaves;
  const int nblk = cdiv(L, rows_per_block);
  for (int j0 = 0; j0 < ncols; j0 += 64) {
    const int nc = std::min(64, ncols - j0);
    hipLaunchKernelGGL(k_gemm_mfma<T>, dim3(nblk), dim3(kWave * kMfmaWaves), 0, s, L, A, lda,
                       kdim, B + static_cast<long long>(j0) * ldb, ldb, nc,
                       C + static_cast<long long>(j0) * ldc, ldc);
    if (hipGetLastError() != hipSuccess) return -1;
  }
  return 0;
}

// ------------------------------------------------------------ seeding ----
// dst = sum_{c < ncols} src(:, c), accumulated in double (start vector of an
// FP64 polish from the Ritz vectors of an FP32-basis solve).
template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_colsum(long long L, const S* __restrict__ src,
                                                   long long ld, int ncols, D* __restrict__ dst) {
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      double acc = 0.0;
      for (int c = 0; c < ncols; ++c) acc += static_cast<double>(src[c * ld + row]);
      dst[row] = static_cast<D>(acc);
    }
  }
}

template <typename S, typename D>
int launch_colsum(hipStream_t s, long long L, const S* src, long long ld, int ncols, D* dst) {
  hipLaunchKernelGGL((k_colsum<S, D>), dim3(stream_blocks(L)), dim3(kBlock), 0, s, L, src, ld,
                     ncols, dst);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int launch_colsum<float, double>(hipStream_t, long long, const float*, long long, int,
                                          double*);
template int launch_colsum<double, double>(hipStream_t, long long, const double*, long long, int,
                                           double*);

// ------------------------------------------------------ instantiations ----
#define PK_INST_T(T)                                                                          \
  template int launch_gemvt<T>(hipStream_t, long long, const T*, long long, const int*, int,  \
                               const int*, const T*, double*, Gate);                          \
  template int launch_gemv_update<T>(hipStream_t, long long, const T*, long long, const int*, \
                                     int, const int*, const double*, T*, double*, Gate);      \
  template int launch_scale<T>(hipStream_t, long long, T*, Scal, Gate);                       \
  template int launch_axpy_norm_dot<T>(hipStream_t, long long, T*, const T*, Scal, double*,   \
                                       Gate);                                                 \
  template int launch_norm<T>(hipStream_t, long long, const T*, double*);                     \
  template int launch_gemm_tall<T>(hipStream_t, long long, const T*, long long, int,         \
                                   const double*, int, int, T*, long long);                  \
  template int launch_gemm_mfma<T>(hipStream_t, long long, const T*, long long, int,         \
                                   const double*, int, int, T*, long long);
PK_INST_T(double)
PK_INST_T(float)
#undef PK_INST_T

#define PK_INST_RB(VT, T)                                                                   \
  template int launch_spmv_rb<VT, T>(hipStream_t, int, const int4*, int, const int4*, double*, \
                                     const int*, const int*, const VT*, const T*, T*,         \
                                     const T*, Scal, double*, int, const double*, double*,   \
                                     int, Gate);
PK_INST_RB(double, double)
PK_INST_RB(float, float)
PK_INST_RB(double, float)
PK_INST_RB(float, double)
#undef PK_INST_RB

template int launch_convert<double, float>(hipStream_t, long long, const double*, float*);
template int launch_convert<float, double>(hipStream_t, long long, const float*, double*);
template int launch_convert<double, double>(hipStream_t, long long, const double*, double*);
template int launch_convert<float, float>(hipStream_t, long long, const float*, float*);

}  // namespace pk
