// Launch wrappers for the hot-path kernels (implemented in pk_kernels.hip).
// Every wrapper enqueues on the given stream and returns the launch status.
#pragma once

#include "pk_common.h"
#include "pk_lbd.h"

namespace pk {

// Partial-sum buffer layout: part[v * nblk + b] for value v of block b.
// stream_blocks(L) is the fixed grid every streaming kernel over L rows uses;
// cgs_blocks(L) the grid of the Gram-Schmidt kernels (larger row slices).
int stream_blocks(long long L);
int cgs_blocks(long long L);

// Row-block SpMV over nblk workgroup descriptors (row blocks of at most
// kSpmvRows rows / kSpmvCap nonzeros, or chunks of kLongChunk nonzeros of a
// long row) plus nlong long rows finished by a fix-up kernel; g lanes per row
// in the LDS reduction.  y = op(A) x - coef * yprev with fused partials
// (2 per slot, nblk + nlong slots).  Column panels: yacc_in (FP64 row sums
// of the earlier panels, null for the first), yacc_out (where a panel that
// is not the last stores its running sums), last (1: apply the epilogue and
// write the partials).  Returns the slot count.
constexpr int kSpmvCap = 1024;
constexpr int kSpmvRows = 1024;
constexpr int kLongChunk = 8192;
// Column-sorted layout: workgroup size, most nonzeros and rows per block.
constexpr int kSortThreads = 1024;
constexpr int kSortCapMax = 16384;
constexpr int kSortRows = 16384;
// rowwise = 1: the same descriptors summed row by row without LDS staging
// (k_spmv_vec; bit-identical row sums).
template <typename VT, typename T>
int launch_spmv_rb(hipStream_t s, int nblk, const int4* desc, int nlong, const int4* ldesc,
                   double* lsum, const int* rp, const int* ci, const VT* va, const T* x, T* y,
                   const T* yprev, Scal coef, double* part, int g, const double* yacc_in,
                   double* yacc_out, int last, Gate gate, int rowwise = 0);

// The same over the column-sorted layout (pk_bodies.h): descriptors of
// blocks of at most cap (<= kSortCapMax) nonzeros, sorted column indices
// sci, values sva and row-major offsets spos (or, spos null, sci packed:
// column col0 + (sci >> pbits), offset sci & (2^pbits - 1)); same outputs
// and partial slots.
template <typename VT, typename T>
int launch_spmv_sorted(hipStream_t s, int nblk, const int4* desc, int nlong, const int4* ldesc,
                       double* lsum, const int* rp, const int* sci, const VT* sva,
                       const unsigned short* spos, int col0, int pbits, int cap, const T* x,
                       T* y, const T* yprev, Scal coef, double* part, int g,
                       const double* yacc_in, double* yacc_out, int last, Gate gate);

// Column-sorted layout with every column panel in one launch: the panels'
// arrays (device table), workgroup w runs descriptor map[w].y of panel
// map[w].x (x < 0: nothing) and stores that panel's FP64 row sums in
// ypart[panel * rows + row] (the panels' blocks must be free of long-row
// chunks and share one lanes-per-row count g).  The combine then forms, per
// row, ((ypart[0] + ypart[1]) + ...) in panel order -- the sums the
// one-launch-per-panel path forms in its running sums -- and the epilogue
// y = T(sum - coef * yprev) with partials of ||y||^2 and yprev.y over the
// stream_blocks(rows) grid (returned).
constexpr int kMaxSortPanels = 16;
struct SortPanels {
  const int4* desc[kMaxSortPanels];
  const int* rp[kMaxSortPanels];
  const int* sci[kMaxSortPanels];
  const void* sva[kMaxSortPanels];
  const unsigned short* spos[kMaxSortPanels];
  int col0[kMaxSortPanels];
  int pbits[kMaxSortPanels];
};
template <typename VT, typename T>
int launch_spmv_sorted_cp(hipStream_t s, const SortPanels* table, const int2* map, int nmap,
                          int cap, int g, long long rows, const T* x, double* ypart,
                          Gate gate);
template <typename T>
int launch_spmv_combine(hipStream_t s, long long rows, int np, const double* ypart,
                        const T* yprev, Scal coef, T* y, double* part, Gate gate);

// part (nval x nblk) -> out[v] = sum_b part[v*nblk+b] in index order.
int launch_reduce(hipStream_t s, int nval, int nblk, const double* part, double* out,
                  Gate gate, const int* nval_p = nullptr);

// h_part(ncols x nblk) = partial V(:,cols)^T w over row blocks
// (nblk = cgs_blocks(L), returned).
template <typename T>
int launch_gemvt(hipStream_t s, long long L, const T* V, long long ldv, const int* cols,
                 int ncols, const int* ncols_p, const T* w, double* part, Gate gate);

// w -= V(:,cols) h ; partials sum w^2 (1 value per block).
template <typename T>
int launch_gemv_update(hipStream_t s, long long L, const T* V, long long ldv,
                       const int* cols, int ncols, const int* ncols_p, const double* h, T* w,
                       double* part, Gate gate);

// x *= a
template <typename T>
int launch_scale(hipStream_t s, long long L, T* x, Scal a, Gate gate);

// w -= c * v ; partials: sum w^2, sum v.w  (2 values per block)
template <typename T>
int launch_axpy_norm_dot(hipStream_t s, long long L, T* w, const T* v, Scal c, double* part,
                         Gate gate);

// partials: sum x^2 (1 value per block)
template <typename T>
int launch_norm(hipStream_t s, long long L, const T* x, double* part);

// x[0:n) = elements first .. first+n-1 of the xLARNV(IDIST=2) sequence from
// the fixed seed (1,3,5,7) (a DGETU0 call's try t draws elements
// (t-1)*n .. t*n-1: xLARNV advances its seed by n), generated on the
// device; bit-identical to the host restatement for T = double and float.
template <typename T>
int launch_larnv_pm1(hipStream_t s, long long n, T* x, long long first = 0);

// dst = src (typed copy with conversion), L elements
template <typename S, typename D>
int launch_convert(hipStream_t s, long long L, const S* src, D* dst);

// C(:, 0:ncols) = A(:, 0:kdim) * B with B (kdim x ncols, ldb) in double, on
// FP64 matrix cores (v_mfma_f64_16x16x4_f64).
template <typename T>
int launch_gemm_mfma(hipStream_t s, long long L, const T* A, long long lda, int kdim,
                     const double* B, int ldb, int ncols, T* C, long long ldc);

// Many products C_q(:, 0:ncols_q) = A_q(:, 0:kdim_q) * B_q in one launch per
// column-tile count (the restart GEMMs of a batch): every row block computed
// as launch_gemm_mfma computes it (bit for bit), kdim_q <= 316 (B staged in
// LDS).  ditems / dcap: a device buffer of items the call grows as needed.
// Returns 0, -1 on a HIP error, -2 for a kdim past the LDS limit.  Waits for
// its stream before returning.
struct GemmJob {
  const void* A;
  void* C;
  const double* B;  // device, kdim x ncols, leading dimension ldb
  long long lda, ldc, L;
  int kdim, ldb, ncols;
};
struct GemmItem {
  const void* A;
  void* C;
  const double* B;
  long long lda, ldc, L;
  int kdim, kpad, ldb, ncols;
  long long blk0, blk1;
};
template <typename T>
int launch_gemm_group(hipStream_t s, const GemmJob* jobs, int njobs, GemmItem** ditems,
                      size_t* dcap);

// Control kernel of the device-resident step (see pk_lbd.h).
enum CtlKind { CTL_V_SPMV = 0, CTL_V_ELR, CTL_V_REO, CTL_U_SPMV, CTL_U_ELR, CTL_U_REO };
int launch_ctl(hipStream_t s, LbdCtl* c, void* arrays, int cap, int kind, int want_phase,
               int want_j, const double* part, int nb);

// dst = sum of the first ncols columns of src (double accumulation).
template <typename S, typename D>
int launch_colsum(hipStream_t s, long long L, const S* src, long long ld, int ncols, D* dst);

int launch_fill_zero_bytes(hipStream_t s, void* p, size_t bytes);

}  // namespace pk
