// Hot-path kernels of the Lanczos bidiagonalization engine, written for
// gfx950 (CDNA4, wave64).
//
//   k_spmv          y = op(A) x - c * yprev over a CSR operator (A or its
//                   explicit transpose), lanes-per-row sub-waves, fused with
//                   the Lanczos recurrence update and the partial sums for
//                   ||y||^2 and yprev.y (the norm and the ELR dot product).
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

// ---------------------------------------------------------------- SpMV ----
template <typename VT, typename T, int LPR>
__global__ __launch_bounds__(kBlock) void k_spmv(int rows, const int* __restrict__ rp,
                                                 const int* __restrict__ ci,
                                                 const VT* __restrict__ va,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 const T* __restrict__ yprev, Scal coef,
                                                 double* __restrict__ part, const int* gate) {
  if (gated_off(gate)) return;
  __shared__ double lds[(kBlock / kWave) * 2];
  const double c = coef.get();
  const int lane = threadIdx.x & (LPR - 1);
  const long long group = (static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x) / LPR;
  const long long ngroups = static_cast<long long>(gridDim.x) * kBlock / LPR;
  const long long iters = (rows + ngroups - 1) / ngroups;  // uniform trip count
  double s2 = 0.0, sd = 0.0;
  for (long long it = 0; it < iters; ++it) {
    const long long row = group + it * ngroups;
    const bool valid = row < rows;
    int p0 = 0, p1 = 0;
    if (valid) {
      p0 = rp[row];
      p1 = rp[row + 1];
    }
    double acc = 0.0;
    for (int p = p0 + lane; p < p1; p += LPR)
      acc += static_cast<double>(va[p]) * static_cast<double>(x[ci[p]]);
    acc = group_sum<LPR>(acc);
    if (valid && lane == 0) {
      const double yp = yprev ? static_cast<double>(yprev[row]) : 0.0;
      const T yv = static_cast<T>(acc - c * yp);
      y[row] = yv;
      s2 += static_cast<double>(yv) * static_cast<double>(yv);
      sd += yp * static_cast<double>(yv);
    }
  }
  double v[2] = {s2, sd};
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[gridDim.x + blockIdx.x] = v[1];
  }
}

template <typename VT, typename T>
int launch_spmv(hipStream_t s, int rows, const int* rp, const int* ci, const VT* va,
                const T* x, T* y, const T* yprev, Scal coef, double* part, int lpr,
                const int* gate) {
  const int nblk = kMaxBlocks;
#define PK_SPMV_CASE(L)                                                                    \
  case L:                                                                                  \
    hipLaunchKernelGGL((k_spmv<VT, T, L>), dim3(nblk), dim3(kBlock), 0, s, rows, rp, ci, va, \
                       x, y, yprev, coef, part, gate);                                     \
    break;
  switch (lpr) {
    PK_SPMV_CASE(1)
    PK_SPMV_CASE(2)
    PK_SPMV_CASE(4)
    PK_SPMV_CASE(8)
    PK_SPMV_CASE(16)
    PK_SPMV_CASE(32)
    PK_SPMV_CASE(64)
    default:
      return -1;
  }
#undef PK_SPMV_CASE
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// -------------------------------------------------------------- reduce ----
__global__ __launch_bounds__(kBlock) void k_reduce(int nblk, const double* __restrict__ part,
                                                   double* __restrict__ out, const int* gate) {
  if (gated_off(gate)) return;
  __shared__ double lds[kBlock / kWave];
  const double* p = part + static_cast<size_t>(blockIdx.x) * nblk;
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nblk; i += kBlock) v[0] += p[i];
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

int launch_reduce(hipStream_t s, int nval, int nblk, const double* part, double* out,
                  const int* gate) {
  if (nval <= 0) return 0;
  hipLaunchKernelGGL(k_reduce, dim3(nval), dim3(kBlock), 0, s, nblk, part, out, gate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------- gemvt ----
// Column tile width: CT accumulators per thread, CT column loads in flight
// per row held.
constexpr int kCT = 8;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemvt(long long L, const T* __restrict__ V,
                                                  long long ldv, const int* __restrict__ cols,
                                                  int ncols, const T* __restrict__ w,
                                                  double* __restrict__ part, const int* gate) {
  if (gated_off(gate)) return;
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
                 int ncols, const T* w, double* part, const int* gate) {
  if (ncols > kMaxCols) return -1;
  const int nblk = stream_blocks(L);
  hipLaunchKernelGGL(k_gemvt<T>, dim3(nblk), dim3(kBlock), 0, s, L, V, ldv, cols, ncols, w,
                     part, gate);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// --------------------------------------------------------- gemv update ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_update(long long L, const T* __restrict__ V,
                                                        long long ldv,
                                                        const int* __restrict__ cols, int ncols,
                                                        const double* __restrict__ h,
                                                        T* __restrict__ w,
                                                        double* __restrict__ part,
                                                        const int* gate) {
  if (gated_off(gate)) return;
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
                       int ncols, const double* h, T* w, double* part, const int* gate) {
  if (ncols > kMaxCols) return -1;
  const int nblk = stream_blocks(L);
  hipLaunchKernelGGL(k_gemv_update<T>, dim3(nblk), dim3(kBlock), 0, s, L, V, ldv, cols, ncols,
                     h, w, part, gate);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// ------------------------------------------------------------- level 1 ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(long long L, T* __restrict__ x, Scal a,
                                                  const int* gate) {
  if (gated_off(gate)) return;
  const double av = a.get();
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) x[row] = static_cast<T>(av * static_cast<double>(x[row]));
  }
}

template <typename T>
int launch_scale(hipStream_t s, long long L, T* x, Scal a, const int* gate) {
  hipLaunchKernelGGL(k_scale<T>, dim3(stream_blocks(L)), dim3(kBlock), 0, s, L, x, a, gate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_norm_dot(long long L, T* __restrict__ w,
                                                          const T* __restrict__ v, Scal c,
                                                          double* __restrict__ part,
                                                          const int* gate) {
  if (gated_off(gate)) return;
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
                         const int* gate) {
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

// ------------------------------------------------------ instantiations ----
#define PK_INST_T(T)                                                                          \
  template int launch_gemvt<T>(hipStream_t, long long, const T*, long long, const int*, int,  \
                               const T*, double*, const int*);                                \
  template int launch_gemv_update<T>(hipStream_t, long long, const T*, long long, const int*, \
                                     int, const double*, T*, double*, const int*);            \
  template int launch_scale<T>(hipStream_t, long long, T*, Scal, const int*);                 \
  template int launch_axpy_norm_dot<T>(hipStream_t, long long, T*, const T*, Scal, double*,   \
                                       const int*);                                           \
  template int launch_norm<T>(hipStream_t, long long, const T*, double*);                     \
  template int launch_gemm_tall<T>(hipStream_t, long long, const T*, long long, int,         \
                                   const double*, int, int, T*, long long);
PK_INST_T(double)
PK_INST_T(float)
#undef PK_INST_T

template int launch_spmv<double, double>(hipStream_t, int, const int*, const int*,
                                         const double*, const double*, double*, const double*,
                                         Scal, double*, int, const int*);
template int launch_spmv<float, float>(hipStream_t, int, const int*, const int*, const float*,
                                       const float*, float*, const float*, Scal, double*, int,
                                       const int*);
template int launch_spmv<double, float>(hipStream_t, int, const int*, const int*,
                                        const double*, const float*, float*, const float*,
                                        Scal, double*, int, const int*);
template int launch_spmv<float, double>(hipStream_t, int, const int*, const int*,
                                        const float*, const double*, double*, const double*,
                                        Scal, double*, int, const int*);
template int launch_convert<double, float>(hipStream_t, long long, const double*, float*);
template int launch_convert<float, double>(hipStream_t, long long, const float*, double*);
template int launch_convert<double, double>(hipStream_t, long long, const double*, double*);
template int launch_convert<float, float>(hipStream_t, long long, const float*, float*);

}  // namespace pk
