// Launch wrappers for the hot-path kernels (implemented in pk_kernels.hip).
// Every wrapper enqueues on the given stream and returns the launch status.
#pragma once

#include "pk_common.h"

namespace pk {

// Partial-sum buffer layout: part[v * nblk + b] for value v of block b.
// stream_blocks(L) is the fixed grid every streaming kernel over L rows uses.
int stream_blocks(long long L);

// y = op(A) x  - coef * yprev   (yprev may be null), CSR rows [0, rows).
// Writes 2 partials per block: sum y^2, sum yprev.y  (of the stored values).
// lpr = lanes per row (1..64, power of two).  Returns nblk used.
template <typename VT, typename T>
int launch_spmv(hipStream_t s, int rows, const int* rp, const int* ci, const VT* va,
                const T* x, T* y, const T* yprev, Scal coef, double* part, int lpr,
                const int* gate);

// part (nval x nblk) -> out[v] = sum_b part[v*nblk+b] in index order.
int launch_reduce(hipStream_t s, int nval, int nblk, const double* part, double* out,
                  const int* gate);

// h_part(ncols x nblk) = partial V(:,cols)^T w over row blocks.
template <typename T>
int launch_gemvt(hipStream_t s, long long L, const T* V, long long ldv, const int* cols,
                 int ncols, const T* w, double* part, const int* gate);

// w -= V(:,cols) h ; partials sum w^2 (1 value per block).
template <typename T>
int launch_gemv_update(hipStream_t s, long long L, const T* V, long long ldv,
                       const int* cols, int ncols, const double* h, T* w, double* part,
                       const int* gate);

// x *= a
template <typename T>
int launch_scale(hipStream_t s, long long L, T* x, Scal a, const int* gate);

// w -= c * v ; partials: sum w^2, sum v.w  (2 values per block)
template <typename T>
int launch_axpy_norm_dot(hipStream_t s, long long L, T* w, const T* v, Scal c, double* part,
                         const int* gate);

// partials: sum x^2 (1 value per block)
template <typename T>
int launch_norm(hipStream_t s, long long L, const T* x, double* part);

// dst = src (typed copy with conversion), L elements
template <typename S, typename D>
int launch_convert(hipStream_t s, long long L, const S* src, D* dst);

// C(:, 0:ncols) = A(:, 0:kdim) * B   with B (kdim x ncols, ldb) in double.
template <typename T>
int launch_gemm_tall(hipStream_t s, long long L, const T* A, long long lda, int kdim,
                     const double* B, int ldb, int ncols, T* C, long long ldc);

int launch_fill_zero_bytes(hipStream_t s, void* p, size_t bytes);

}  // namespace pk
