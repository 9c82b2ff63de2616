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
//   k_gemm_mfma     C(:,1:n) = A(:,1:k) B on FP64 matrix cores for restarts /
//                   Ritz vectors (out of place into the spare basis buffer).
#include <algorithm>
#include <vector>

#include "pk_bodies.h"

namespace pk {

int stream_blocks(long long L) { return L > 0 ? cdiv(L, kRowsPerBlock) : 1; }
int cgs_blocks(long long L) { return L > 0 ? cdiv(L, kCgsRows) : 1; }

// ------------------------------------------------------ row-block SpMV ----
// One workgroup per descriptor (row block or long-row chunk, see
// spmv_rb_body); partial slots b and nbt + b.  The long rows' two partials
// follow the blocks' slots (nbmain + i).
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
  spmv_rb_body<VT, T, G>(desc[blockIdx.x], rp, ci, va, x, y, yprev, coef.get(),
                         part + blockIdx.x, part + nbt + blockIdx.x, lsum, yacc_in, yacc_out,
                         last, prod, lds);
}

// Row-wise form over the same descriptors (spmv_vec_body): no LDS staging.
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_vec(const int4* __restrict__ desc,
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
  __shared__ double lds[(kBlock / kWave) * 2];
  spmv_vec_body<VT, T, G>(desc[blockIdx.x], rp, ci, va, x, y, yprev, coef.get(),
                          part + blockIdx.x, part + nbt + blockIdx.x, lsum, yacc_in, yacc_out,
                          last, lds);
}

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
  const int i = blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x) / kWave;
  if (i >= nlong) return;  // wave-uniform
  spmv_longfix_body<T>(ld[i], lsum, y, yprev, coef.get(), part + nbmain + i,
                       part + nbt + nbmain + i, yacc_in, yacc_out, last,
                       static_cast<int>(threadIdx.x) % kWave);
}

template <typename VT, typename T>
int launch_spmv_rb(hipStream_t s, int nblk, const int4* desc, int nlong, const int4* ldesc,
                   double* lsum, const int* rp, const int* ci, const VT* va, const T* x, T* y,
                   const T* yprev, Scal coef, double* part, int g, const double* yacc_in,
                   double* yacc_out, int last, Gate gate, int rowwise) {
  const int nbt = nblk + nlong;
  if (nblk > 0) {
#define PK_RB_CASE(G)                                                                           \
  case G:                                                                                       \
    if (rowwise)                                                                                \
      hipLaunchKernelGGL((k_spmv_vec<VT, T, G>), dim3(nblk), dim3(kBlock), 0, s, desc, rp, ci,  \
                         va, x, y, yprev, coef, part, nbt, lsum, yacc_in, yacc_out, last, gate); \
    else                                                                                        \
      hipLaunchKernelGGL((k_spmv_rb<VT, T, G>), dim3(nblk), dim3(kBlock), 0, s, desc, rp, ci,   \
                         va, x, y, yprev, coef, part, nbt, lsum, yacc_in, yacc_out, last, gate); \
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
    hipLaunchKernelGGL(k_spmv_longfix<T>, dim3(cdiv(nlong, kBlock / kWave)), dim3(kBlock), 0,
                       s, nlong, ldesc, lsum, y, yprev, coef, part, nbt, nblk, yacc_in, yacc_out,
                       last, gate);
  return hipGetLastError() == hipSuccess ? nbt : -1;
}

// ------------------------------------------------- column-sorted SpMV ----
// One kSortThreads workgroup per descriptor of the sorted layout
// (spmv_sorted_body); the products of a block take cap doubles of dynamic
// LDS.  Partial slots and the long-row fix-up as in k_spmv_rb.
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kSortThreads) void k_spmv_sorted(
    const int4* __restrict__ desc, const int* __restrict__ rp, const int* __restrict__ sci,
    const VT* __restrict__ sva, const unsigned short* __restrict__ spos, int col0, int pbits,
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ yprev, Scal coef,
    double* __restrict__ part, int nbt, double* __restrict__ lsum,
    const double* __restrict__ yacc_in, double* __restrict__ yacc_out, int last, Gate gate) {
  if (gate.off()) return;
  extern __shared__ double prod_dyn[];
  __shared__ double lds[(kSortThreads / kWave) * 2];
  spmv_sorted_body<VT, T, G>(desc[blockIdx.x], rp, sci, sva, spos, col0, pbits, x, y, yprev,
                             coef.get(),
                             part + blockIdx.x, part + nbt + blockIdx.x, lsum, yacc_in,
                             yacc_out, last, prod_dyn, lds);
}

template <typename VT, typename T>
int launch_spmv_sorted(hipStream_t s, int nblk, const int4* desc, int nlong, const int4* ldesc,
                       double* lsum, const int* rp, const int* sci, const VT* sva,
                       const unsigned short* spos, int col0, int pbits, int cap, const T* x,
                       T* y, const T* yprev, Scal coef, double* part, int g,
                       const double* yacc_in, double* yacc_out, int last, Gate gate) {
  // packed (spos null): 0 < pbits <= 16; unpacked: global columns, no base or shift
  if (cap <= 0 || cap > kSortCapMax || pbits < 0 || pbits > 16 || (!spos && pbits == 0) ||
      (spos && (pbits != 0 || col0 != 0)))
    return -1;
  const int nbt = nblk + nlong;
  const size_t shm = sizeof(double) * static_cast<size_t>(cap);
  if (nblk > 0) {
#define PK_SO_CASE(G)                                                                          \
  case G:                                                                                      \
    hipLaunchKernelGGL((k_spmv_sorted<VT, T, G>), dim3(nblk), dim3(kSortThreads), shm, s, desc, \
                       rp, sci, sva, spos, col0, pbits, x, y, yprev, coef, part, nbt, lsum,    \
                       yacc_in, yacc_out, last, gate);                                         \
    break;
    switch (g) {
      PK_SO_CASE(2)
      PK_SO_CASE(4)
      PK_SO_CASE(8)
      PK_SO_CASE(16)
      PK_SO_CASE(32)
      PK_SO_CASE(64)
      default:
        return -1;
    }
#undef PK_SO_CASE
  }
  if (nlong > 0)
    hipLaunchKernelGGL(k_spmv_longfix<T>, dim3(cdiv(nlong, kBlock / kWave)), dim3(kBlock), 0,
                       s, nlong, ldesc, lsum, y, yprev, coef, part, nbt, nblk, yacc_in, yacc_out,
                       last, gate);
  return hipGetLastError() == hipSuccess ? nbt : -1;
}

// -------------------------------------------- sorted SpMV, panels at once --
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kSortThreads) void k_spmv_sorted_cp(
    const SortPanels* __restrict__ table, const int2* __restrict__ map, long long rows,
    const T* __restrict__ x, double* __restrict__ ypart, Gate gate) {
  if (gate.off()) return;
  const int2 mp = map[blockIdx.x];
  if (mp.x < 0) return;
  extern __shared__ double prod_dyn[];
  __shared__ double lds[(kSortThreads / kWave) * 2];
  const int p = mp.x;
  spmv_sorted_body<VT, T, G>(table->desc[p][mp.y], table->rp[p], table->sci[p],
                             static_cast<const VT*>(table->sva[p]), table->spos[p],
                             table->col0[p], table->pbits[p], x, nullptr, nullptr, 0.0,
                             nullptr, nullptr, nullptr, nullptr,
                             ypart + static_cast<long long>(p) * rows, 0, prod_dyn, lds);
}

template <typename VT, typename T>
int launch_spmv_sorted_cp(hipStream_t s, const SortPanels* table, const int2* map, int nmap,
                          int cap, int g, long long rows, const T* x, double* ypart,
                          Gate gate) {
  if (cap <= 0 || cap > kSortCapMax || nmap <= 0) return -1;
  const size_t shm = sizeof(double) * static_cast<size_t>(cap);
  switch (g) {
#define PK_CP_CASE(G)                                                                       \
  case G:                                                                                   \
    hipLaunchKernelGGL((k_spmv_sorted_cp<VT, T, G>), dim3(nmap), dim3(kSortThreads), shm, s, \
                       table, map, rows, x, ypart, gate);                                   \
    break;
    PK_CP_CASE(2)
    PK_CP_CASE(4)
    PK_CP_CASE(8)
    PK_CP_CASE(16)
    PK_CP_CASE(32)
    PK_CP_CASE(64)
#undef PK_CP_CASE
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// y = (sum of the np panels' row sums, added in panel order) - c yprev and
// the partials of ||y||^2 and yprev.y.  NP > 0: the panel count is known at
// compile time, so every thread's kRPT x NP loads (clamped to row L - 1; a
// row past L is loaded but neither stored nor counted) are issued before the
// first add; NP = 0 loops over np.  Both forms add in the same order.
template <typename T, int NP>
__global__ __launch_bounds__(kBlock) void k_spmv_combine(long long L, int np,
                                                         const double* __restrict__ ypart,
                                                         const T* __restrict__ yprev,
                                                         Scal coef, T* __restrict__ y,
                                                         double* __restrict__ part,
                                                         Gate gate) {
  if (gate.off()) return;
  __shared__ double lds[(kBlock / kWave) * 2];
  const double c = coef.get();
  const long long base = static_cast<long long>(blockIdx.x) * kRowsPerBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  double acc[kRPT] = {};
  if (L <= 0) {
    // no rows: zero partials below
  } else if constexpr (NP > 0) {
    double t[kRPT][NP];
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const long long row = min(base + static_cast<long long>(r) * kBlock, L - 1);
#pragma unroll
      for (int p = 0; p < NP; ++p) t[r][p] = ypart[static_cast<long long>(p) * L + row];
    }
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      acc[r] = t[r][0];
#pragma unroll
      for (int p = 1; p < NP; ++p) acc[r] = acc[r] + t[r][p];
    }
  } else {
#pragma unroll
    for (int r = 0; r < kRPT; ++r) {
      const long long row = min(base + static_cast<long long>(r) * kBlock, L - 1);
      acc[r] = ypart[row];
      for (int p = 1; p < np; ++p) acc[r] = acc[r] + ypart[static_cast<long long>(p) * L + row];
    }
  }
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      const double yp = yprev ? static_cast<double>(yprev[row]) : 0.0;
      const T yv = static_cast<T>(acc[r] - c * yp);
      y[row] = yv;
      v[0] += static_cast<double>(yv) * static_cast<double>(yv);
      v[1] += yp * static_cast<double>(yv);
    }
  }
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[gridDim.x + blockIdx.x] = v[1];
  }
}

template <typename T>
int launch_spmv_combine(hipStream_t s, long long rows, int np, const double* ypart,
                        const T* yprev, Scal coef, T* y, double* part, Gate gate) {
  if (np < 1) return -1;
  const int nblk = stream_blocks(rows);
#define PK_CB_CASE(NP)                                                                     \
  hipLaunchKernelGGL((k_spmv_combine<T, NP>), dim3(nblk), dim3(kBlock), 0, s, rows, np, ypart, \
                     yprev, coef, y, part, gate)
  switch (np) {
    case 2: PK_CB_CASE(2); break;
    case 4: PK_CB_CASE(4); break;
    case 8: PK_CB_CASE(8); break;
    default: PK_CB_CASE(0); break;
  }
#undef PK_CB_CASE
  return hipGetLastError() == hipSuccess ? nblk : -1;
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
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemvt(long long L, const T* __restrict__ V,
                                                  long long ldv, const int* __restrict__ cols,
                                                  int ncols_v, const int* ncols_p,
                                                  const T* __restrict__ w,
                                                  double* __restrict__ part, Gate gate) {
  if (gate.off()) return;
  const int ncols = ncols_p ? *ncols_p : ncols_v;
  __shared__ double lds[2][(kBlock / kWave) * kCT];
  gemvt_body<T>(blockIdx.x, L, V, ldv, cols, ncols, w, part, gridDim.x, blockIdx.x, lds);
}

template <typename T>
int launch_gemvt(hipStream_t s, long long L, const T* V, long long ldv, const int* cols,
                 int ncols, const int* ncols_p, const T* w, double* part, Gate gate) {
  const int nblk = cgs_blocks(L);
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
  __shared__ double lds[kBlock / kWave];
  gemv_update_body<T>(blockIdx.x, L, V, ldv, cols, h, ncols, w, part + blockIdx.x, lds);
}

template <typename T>
int launch_gemv_update(hipStream_t s, long long L, const T* V, long long ldv, const int* cols,
                       int ncols, const int* ncols_p, const double* h, T* w, double* part,
                       Gate gate) {
  const int nblk = cgs_blocks(L);
  hipLaunchKernelGGL(k_gemv_update<T>, dim3(nblk), dim3(kBlock), 0, s, L, V, ldv, cols, ncols,
                     ncols_p, h, w, part, gate);
  return hipGetLastError() == hipSuccess ? nblk : -1;
}

// ------------------------------------------------------------- level 1 ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(long long L, T* __restrict__ x, Scal a,
                                                  Gate gate) {
  if (gate.off()) return;
  scale_body<T>(blockIdx.x, L, x, a.get());
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
  axpy_norm_dot_body<T>(blockIdx.x, L, w, v, c.get(), part + blockIdx.x,
                        part + gridDim.x + blockIdx.x, lds);
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

// ------------------------------------------------------------- control ----
// One workgroup: reduce the partials of the preceding vector kernel (same
// order as k_reduce), then run the step's control stage (ctl_stage_body).
// Runs only while the state is in (want_phase, want_j); otherwise the whole
// chain for this step is idle.
__global__ __launch_bounds__(kBlock) void k_ctl(LbdCtl* __restrict__ gc, char* garr, int cap,
                                                int kind, int want_phase, int want_j,
                                                const double* __restrict__ part, int nb) {
  __shared__ LbdCtlLds c;
  __shared__ LbdCtlRef cg;
  __shared__ double lds[(kBlock / kWave) * 2];
  if (gc->phase != want_phase || gc->j != want_j) return;
  const bool two = (kind == CTL_V_SPMV || kind == CTL_V_ELR || kind == CTL_U_SPMV ||
                    kind == CTL_U_ELR);
  ctl_stage_body(gc, garr, cap, want_j, kind, part, nb, two ? part + nb : nullptr, nb, c, cg,
                 lds);
}

int launch_ctl(hipStream_t s, LbdCtl* c, void* arrays, int cap, int kind, int want_phase,
               int want_j, const double* part, int nb) {
  hipLaunchKernelGGL(k_ctl, dim3(1), dim3(kBlock), 0, s, c, static_cast<char*>(arrays), cap,
                     kind, want_phase, want_j, part, nb);
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

// The same product with B staged in LDS: each workgroup copies B once
// (zero-padded to NT 16-column tiles and a k extent kpad = round_up(kdim, 4)
// + 1, odd, so that the 16 columns one MFMA operand reads fall in distinct
// banks) and then strides over blocks of kGemmRows rows (measured faster
// than a contiguous range of blocks per workgroup).  Each wave owns
// kGemmRowTiles 16-row tiles and all NT column tiles, so one B operand read
// from LDS feeds kGemmRowTiles MFMAs; the A operands of the next two k-steps
// are in flight while the current one multiplies, in three register stages
// of a loop unrolled by three, and the next row block's first two k-steps
// load while a block's C is stored (measured on 2,000,000 x 121 times 121 x
// 51: 0.62 ms; without the next-block loads 0.63 ms; with the stages rotated
// by register copies, which made every k-step wait for all its loads, 0.67
// ms; four stages 0.62 ms, within the spread between boxes; two row tiles per wave with six waves 0.98 ms, 16-byte
// loads of row pairs 0.87 ms).  The launch bound of two waves per SIMD keeps
// the accumulators and stages within 256 registers.  Rows past L load row
// L - 1 and are not stored; k past kdim loads column kdim - 1, which meets
// zero rows of the staged B.
constexpr int kGemmRowTiles = 4;
constexpr int kGemmWaves = 4;
constexpr int kGemmRows = 16 * kGemmRowTiles * kGemmWaves;
constexpr int kGemmMaxKpad = 317;  // 64 columns x 317 x 8 bytes fit the 160 KiB of LDS

template <typename T, int NT>
__global__ __launch_bounds__(kWave * kGemmWaves, 2) void k_gemm_lds(
    long long L, const T* __restrict__ A, long long lda, int kdim, int kpad,
    const double* __restrict__ B, int ldb, int ncols, T* __restrict__ C, long long ldc,
    long long nblocks) {
  extern __shared__ double Bs[];  // Bs[j * kpad + k] = B(k, j), zero outside kdim x ncols
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int li = lane % 16, lk = lane / 16;
  for (int e = threadIdx.x; e < NT * 16 * kpad; e += blockDim.x) {
    const int j = e / kpad, k = e - j * kpad;
    Bs[e] = (j < ncols && k < kdim) ? B[static_cast<long long>(j) * ldb + k] : 0.0;
  }
  __syncthreads();
  const int nks = (kdim + 3) / 4;
  const long long wrow = static_cast<long long>(wid) * (16 * kGemmRowTiles);
  // this wave's blocks: static_cast<long long>(blockIdx.x), + static_cast<long long>(gridDim.x), ... while blk < nblocks and the wave's
  // first row is inside L (the first rows grow with blk: a prefix)
  long long blk = static_cast<long long>(blockIdx.x);
  if (blk >= nblocks || blk * kGemmRows + wrow >= L) return;
  auto rows = [&](long long bk, long long (&rc)[kGemmRowTiles]) {
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) {
      const long long row = bk * kGemmRows + wrow + rt * 16 + li;
      rc[rt] = row < L ? row : L - 1;
    }
  };
  long long rowc[kGemmRowTiles];
  rows(blk, rowc);
  double a0[kGemmRowTiles], a1[kGemmRowTiles], a2[kGemmRowTiles];
  auto load = [&](int ks, const long long (&rc)[kGemmRowTiles], double (&a)[kGemmRowTiles]) {
    // clamped: the loads stay in bounds, and a step past the end is never used
    const long long k = min(ks * 4 + lk, kdim - 1);
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) a[rt] = static_cast<double>(A[k * lda + rc[rt]]);
  };
  load(0, rowc, a0);
  load(1, rowc, a1);
  for (;;) {
    v4d acc[kGemmRowTiles][NT];
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = v4d{0.0, 0.0, 0.0, 0.0};
    auto step = [&](int ks, const double (&a)[kGemmRowTiles]) {
      const int k = ks * 4 + lk;
      double b[NT];
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) b[ct] = Bs[(ct * 16 + li) * kpad + k];
#pragma unroll
      for (int rt = 0; rt < kGemmRowTiles; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
          acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ct], a[rt], acc[rt][ct], 0, 0, 0);
    };
    // Three register stages, the loop unrolled by three so that each stage
    // keeps its registers: rotating the operands by copies made the compiler
    // wait for every load in flight at the end of each k-step.  Loop entry:
    // a0 holds step ks, a1 step ks + 1.
    int ks = 0;
    for (; ks + 3 <= nks; ks += 3) {
      load(ks + 2, rowc, a2);
      step(ks, a0);
      load(ks + 3, rowc, a0);
      step(ks + 1, a1);
      load(ks + 4, rowc, a1);
      step(ks + 2, a2);
    }
    if (ks < nks) step(ks, a0);
    if (ks + 1 < nks) step(ks + 1, a1);
    // the next block's first two steps load while this block's C is stored
    const long long nb = blk + static_cast<long long>(gridDim.x);
    const bool more = nb < nblocks && nb * kGemmRows + wrow < L;
    long long rown[kGemmRowTiles];
    if (more) {
      rows(nb, rown);
      load(0, rown, a0);
      load(1, rown, a1);
    }
    const long long i0 = blk * kGemmRows + wrow;
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) {
      const long long row = i0 + rt * 16 + li;
      if (row >= L) continue;
#pragma unroll
      for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = ct * 16 + lk + 4 * r;
          if (j < ncols) C[static_cast<long long>(j) * ldc + row] = static_cast<T>(acc[rt][ct][r]);
        }
    }
    if (!more) break;
    blk = nb;
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) rowc[rt] = rown[rt];
  }
}

// Many products in one launch (the restart GEMMs of a batch): workgroup w
// runs item w, a range of row blocks of one product's 64-column chunk, with
// that chunk's B staged.  The body is k_gemm_lds's, statement for statement,
// so every row block comes out bit for bit as that kernel computes it (kept
// as its own text: sharing one device function between the two kernels
// changed k_gemm_lds's register allocation and slowed it by a fifth,
// measured on the 2M x 121 * 121 x 51 shape).
template <typename T, int NT>
__global__ __launch_bounds__(kWave * kGemmWaves, 2) void k_gemm_lds_group(
    const GemmItem* __restrict__ items) {
  const GemmItem it = items[blockIdx.x];
  const long long L = it.L, lda = it.lda, ldc = it.ldc;
  const int kdim = it.kdim, kpad = it.kpad, ldb = it.ldb, ncols = it.ncols;
  const T* __restrict__ A = static_cast<const T*>(it.A);
  const double* __restrict__ B = it.B;
  T* __restrict__ C = static_cast<T*>(it.C);
  extern __shared__ double Bs[];  // Bs[j * kpad + k] = B(k, j), zero outside kdim x ncols
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int li = lane % 16, lk = lane / 16;
  for (int e = threadIdx.x; e < NT * 16 * kpad; e += blockDim.x) {
    const int j = e / kpad, k = e - j * kpad;
    Bs[e] = (j < ncols && k < kdim) ? B[static_cast<long long>(j) * ldb + k] : 0.0;
  }
  __syncthreads();
  const int nks = (kdim + 3) / 4;
  const long long wrow = static_cast<long long>(wid) * (16 * kGemmRowTiles);
  // this wave's blocks: it.blk0, + 1, ... while blk < it.blk1 and the wave's
  // first row is inside L (the first rows grow with blk: a prefix)
  long long blk = it.blk0;
  if (blk >= it.blk1 || blk * kGemmRows + wrow >= L) return;
  auto rows = [&](long long bk, long long (&rc)[kGemmRowTiles]) {
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) {
      const long long row = bk * kGemmRows + wrow + rt * 16 + li;
      rc[rt] = row < L ? row : L - 1;
    }
  };
  long long rowc[kGemmRowTiles];
  rows(blk, rowc);
  double a0[kGemmRowTiles], a1[kGemmRowTiles], a2[kGemmRowTiles];
  auto load = [&](int ks, const long long (&rc)[kGemmRowTiles], double (&a)[kGemmRowTiles]) {
    // clamped: the loads stay in bounds, and a step past the end is never used
    const long long k = min(ks * 4 + lk, kdim - 1);
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) a[rt] = static_cast<double>(A[k * lda + rc[rt]]);
  };
  load(0, rowc, a0);
  load(1, rowc, a1);
  for (;;) {
    v4d acc[kGemmRowTiles][NT];
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt)
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = v4d{0.0, 0.0, 0.0, 0.0};
    auto step = [&](int ks, const double (&a)[kGemmRowTiles]) {
      const int k = ks * 4 + lk;
      double b[NT];
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) b[ct] = Bs[(ct * 16 + li) * kpad + k];
#pragma unroll
      for (int rt = 0; rt < kGemmRowTiles; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
          acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ct], a[rt], acc[rt][ct], 0, 0, 0);
    };
    // Three register stages, the loop unrolled by three so that each stage
    // keeps its registers: rotating the operands by copies made the compiler
    // wait for every load in flight at the end of each k-step.  Loop entry:
    // a0 holds step ks, a1 step ks + 1.
    int ks = 0;
    for (; ks + 3 <= nks; ks += 3) {
      load(ks + 2, rowc, a2);
      step(ks, a0);
      load(ks + 3, rowc, a0);
      step(ks + 1, a1);
      load(ks + 4, rowc, a1);
      step(ks + 2, a2);
    }
    if (ks < nks) step(ks, a0);
    if (ks + 1 < nks) step(ks + 1, a1);
    // the next block's first two steps load while this block's C is stored
    const long long nb = blk + 1;
    const bool more = nb < it.blk1 && nb * kGemmRows + wrow < L;
    long long rown[kGemmRowTiles];
    if (more) {
      rows(nb, rown);
      load(0, rown, a0);
      load(1, rown, a1);
    }
    const long long i0 = blk * kGemmRows + wrow;
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) {
      const long long row = i0 + rt * 16 + li;
      if (row >= L) continue;
#pragma unroll
      for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = ct * 16 + lk + 4 * r;
          if (j < ncols) C[static_cast<long long>(j) * ldc + row] = static_cast<T>(acc[rt][ct][r]);
        }
    }
    if (!more) break;
    blk = nb;
#pragma unroll
    for (int rt = 0; rt < kGemmRowTiles; ++rt) rowc[rt] = rown[rt];
  }
}

template <typename T>
int launch_gemm_group(hipStream_t s, const GemmJob* jobs, int njobs, GemmItem** ditems,
                      size_t* dcap) {
  // items per column-tile count (NT = 1..4), row blocks cut into ranges of
  // at most kGroupBlocks
  constexpr long long kGroupBlocks = 8;
  std::vector<GemmItem> items[4];
  int kmax_pad[4] = {0, 0, 0, 0};
  for (int q = 0; q < njobs; ++q) {
    const GemmJob& jb = jobs[q];
    if (jb.L <= 0 || jb.ncols <= 0 || jb.kdim <= 0) continue;
    const int kpad = (jb.kdim + 3) / 4 * 4 + 1;
    if (kpad > kGemmMaxKpad) return -2;  // the grouped path stages B in LDS only
    const long long nblocks = (jb.L + kGemmRows - 1) / kGemmRows;
    for (int j0 = 0; j0 < jb.ncols; j0 += 64) {
      const int nc = std::min(64, jb.ncols - j0);
      const int nt = (nc + 15) / 16;
      kmax_pad[nt - 1] = std::max(kmax_pad[nt - 1], kpad);
      for (long long b0 = 0; b0 < nblocks; b0 += kGroupBlocks) {
        GemmItem it;
        it.A = jb.A;
        it.C = static_cast<char*>(jb.C) + sizeof(T) * static_cast<size_t>(j0) * jb.ldc;
        it.B = jb.B + static_cast<long long>(j0) * jb.ldb;
        it.lda = jb.lda;
        it.ldc = jb.ldc;
        it.L = jb.L;
        it.kdim = jb.kdim;
        it.kpad = kpad;
        it.ldb = jb.ldb;
        it.ncols = nc;
        it.blk0 = b0;
        it.blk1 = std::min(nblocks, b0 + kGroupBlocks);
        items[nt - 1].push_back(it);
      }
    }
  }
  size_t total = 0;
  for (auto& v : items) total += v.size();
  if (total == 0) return 0;
  if (*dcap < total) {
    (void)hipFree(*ditems);
    *ditems = nullptr;
    *dcap = 0;
    if (hipMalloc(ditems, sizeof(GemmItem) * total) != hipSuccess) return -1;
    *dcap = total;
  }
  std::vector<GemmItem> all;
  all.reserve(total);
  for (auto& v : items) all.insert(all.end(), v.begin(), v.end());
  if (hipMemcpyAsync(*ditems, all.data(), sizeof(GemmItem) * total, hipMemcpyHostToDevice, s) !=
      hipSuccess)
    return -1;
  size_t off = 0;
  for (int nt = 1; nt <= 4; ++nt) {
    const size_t cnt = items[nt - 1].size();
    if (cnt == 0) continue;
    const size_t shm = sizeof(double) * 16 * nt * kmax_pad[nt - 1];
    const GemmItem* d = *ditems + off;
#define PK_GEMM_GRP(NTV)                                                                     \
  hipLaunchKernelGGL((k_gemm_lds_group<T, NTV>), dim3(static_cast<unsigned>(cnt)),            \
                     dim3(kWave * kGemmWaves), shm, s, d)
    switch (nt) {
      case 1: PK_GEMM_GRP(1); break;
      case 2: PK_GEMM_GRP(2); break;
      case 3: PK_GEMM_GRP(3); break;
      default: PK_GEMM_GRP(4); break;
    }
#undef PK_GEMM_GRP
    if (hipGetLastError() != hipSuccess) return -1;
    off += cnt;
  }
  // the host vector of items is read by the copy above: wait for it
  return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

template <typename T>
int launch_gemm_mfma(hipStream_t s, long long L, const T* A, long long lda, int kdim,
                     const double* B, int ldb, int ncols, T* C, long long ldc) {
  if (L <= 0 || ncols <= 0 || kdim <= 0) return 0;
  const int kpad = (kdim + 3) / 4 * 4 + 1;
  if (kpad <= kGemmMaxKpad) {
    const long long nblocks = (L + kGemmRows - 1) / kGemmRows;
    for (int j0 = 0; j0 < ncols; j0 += 64) {
      const int nc = std::min(64, ncols - j0);
      const int nt = (nc + 15) / 16;
      const size_t shm = sizeof(double) * 16 * nt * kpad;
      // two workgroups per CU when B fits 80 KiB, else one
      const int per_cu = shm <= 80 * 1024 ? 2 : 1;
      const int grid = static_cast<int>(std::min<long long>(nblocks, 256LL * per_cu));
      const double* Bj = B + static_cast<long long>(j0) * ldb;
      T* Cj = C + static_cast<long long>(j0) * ldc;
#define PK_GEMM_LDS(NTV)                                                                        \
  hipLaunchKernelGGL((k_gemm_lds<T, NTV>), dim3(grid), dim3(kWave * kGemmWaves), shm, s, L, A, \
                     lda, kdim, kpad, Bj, ldb, nc, Cj, ldc, nblocks)
      switch (nt) {
        case 1: PK_GEMM_LDS(1); break;
        case 2: PK_GEMM_LDS(2); break;
        case 3: PK_GEMM_LDS(3); break;
        default: PK_GEMM_LDS(4); break;
      }
#undef PK_GEMM_LDS
      if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
  }
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

// ------------------------------------------------------- random start ----
// Thread t writes elements t, t + S, t + 2S, ... (S threads in the grid, so
// every store instruction is coalesced); element g uses the state
// s0 a^(g+1), and stepping by S multiplies the state by a^S.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_larnv_pm1(long long n, unsigned long long first,
                                                      unsigned long long step,
                                                      T* __restrict__ x) {
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  const long long S = static_cast<long long>(gridDim.x) * kBlock;
  if (t >= n) return;
  unsigned long long s =
      lcg_mulmod(kLcgSeed, lcg_pow(first + static_cast<unsigned long long>(t) + 1));
  for (long long g = t; g < n; g += S) {
    x[g] = lcg_pm1<T>(s);
    s = lcg_mulmod(s, step);
  }
}

template <typename T>
int launch_larnv_pm1(hipStream_t s, long long n, T* x, long long first) {
  if (n <= 0) return 0;
  const int nblk = static_cast<int>(std::min<long long>(cdiv(n, kBlock), 2048));
  const unsigned long long step = lcg_pow(static_cast<unsigned long long>(nblk) * kBlock);
  hipLaunchKernelGGL(k_larnv_pm1<T>, dim3(nblk), dim3(kBlock), 0, s, n,
                     static_cast<unsigned long long>(first), step, x);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int launch_larnv_pm1<double>(hipStream_t, long long, double*, long long);
template int launch_larnv_pm1<float>(hipStream_t, long long, float*, long long);

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
  template int launch_gemm_mfma<T>(hipStream_t, long long, const T*, long long, int,         \
                                   const double*, int, int, T*, long long);                   \
  template int launch_gemm_group<T>(hipStream_t, const GemmJob*, int, GemmItem**, size_t*);
PK_INST_T(double)
PK_INST_T(float)
#undef PK_INST_T

#define PK_INST_RB(VT, T)                                                                   \
  template int launch_spmv_rb<VT, T>(hipStream_t, int, const int4*, int, const int4*, double*, \
                                     const int*, const int*, const VT*, const T*, T*,         \
                                     const T*, Scal, double*, int, const double*, double*,   \
                                     int, Gate, int);
PK_INST_RB(double, double)
PK_INST_RB(float, float)
PK_INST_RB(double, float)
PK_INST_RB(float, double)
#undef PK_INST_RB

#define PK_INST_SO(VT, T)                                                                    \
  template int launch_spmv_sorted<VT, T>(hipStream_t, int, const int4*, int, const int4*,    \
                                         double*, const int*, const int*, const VT*,         \
                                         const unsigned short*, int, int, int, const T*, T*, \
                                         const T*,                                           \
                                         Scal, double*, int, const double*, double*, int,    \
                                         Gate);
PK_INST_SO(double, double)
PK_INST_SO(float, float)
PK_INST_SO(double, float)
PK_INST_SO(float, double)
#undef PK_INST_SO
#define PK_INST_CP(VT, T)                                                                    \
  template int launch_spmv_sorted_cp<VT, T>(hipStream_t, const SortPanels*, const int2*, int, \
                                            int, int, long long, const T*, double*, Gate);
PK_INST_CP(double, double)
PK_INST_CP(float, float)
PK_INST_CP(double, float)
PK_INST_CP(float, double)
#undef PK_INST_CP
template int launch_spmv_combine<double>(hipStream_t, long long, int, const double*,
                                         const double*, Scal, double*, double*, Gate);
template int launch_spmv_combine<float>(hipStream_t, long long, int, const double*,
                                        const float*, Scal, float*, double*, Gate);

template int launch_convert<double, float>(hipStream_t, long long, const double*, float*);
template int launch_convert<float, double>(hipStream_t, long long, const float*, double*);
template int launch_convert<double, double>(hipStream_t, long long, const double*, double*);
template int launch_convert<float, float>(hipStream_t, long long, const float*, float*);

}  // namespace pk
