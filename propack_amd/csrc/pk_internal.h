// Internal structures of libpropack_amd: the device CSR operator and the
// Lanczos bidiagonalization engine context.
#pragma once

#include <functional>
#include <vector>

#include "../../include/propack_amd.h"
#include "pk_kernels.h"

namespace pk {

struct CsrOp {
  static constexpr unsigned kMagic = 0x504b4353u;  // "PKCS"
  static constexpr int kMaxPanels = 16;
  unsigned magic = 0;
  int m = 0, n = 0;
  long long nnz = 0;
  int val_prec = PK_F64;
  size_t val_bytes = 8;
  // SpMV workgroup descriptors (row blocks and long-row chunks), long-row
  // table and its chunk partial sums, for one CSR array set
  struct Part {
    int4* desc = nullptr;
    int ndesc = 0;
    int4* ldesc = nullptr;
    int nlong = 0;
    double* lsum = nullptr;
    int slots() const { return ndesc + nlong; }
  };
  // One direction (A or the explicit A^T), stored as np column panels:
  // panel p holds the entries with column in [p * width, (p + 1) * width)
  // as its own CSR over all rows (global column indices, entries of a row in
  // their original order).  np = 1 is the plain CSR.
  struct Panel {
    int* rp = nullptr;
    int* ci = nullptr;
    void* va = nullptr;
    long long nnz = 0;
    int g = 8;  // lanes per row in the LDS row reduction
    Part part;
    // column-sorted layout (built on request): per block, the entries in
    // increasing column order and their row-major offsets (spos, or packed
    // into sci when spbits > 0: column scol0 + (sci >> spbits)); own
    // descriptors
    int* sci = nullptr;
    void* sva = nullptr;
    unsigned short* spos = nullptr;
    int scol0 = 0, spbits = 0;
    Part spart;
  };
  struct Dir {
    int rows = 0, cols = 0;
    int np = 1, width = 0;
    Panel pan[kMaxPanels];
    double* yacc = nullptr;  // FP64 running row sums between panels (np > 1)
    int slots = 0;           // partial slots of the widest panel
    int lane_shift = 0;      // lanes per row = built-in choice * 2^lane_shift
    int sort_cap = 0;        // > 0: column-sorted layout, blocks of <= sort_cap nonzeros
    int rowwise = 0;         // row-block arrays summed row-wise, no LDS staging (k_spmv_vec)
    // column-sorted layout with every panel in one launch (np > 1): per-panel
    // FP64 row sums (np * rows doubles) combined in panel order afterwards
    int concurrent = 0;
    SortPanels* ctable = nullptr;  // device copy of the panels' sorted arrays
    int2* cmap = nullptr;          // (panel, descriptor) of each workgroup
    int cmap_n = 0;
    double* ypart = nullptr;       // the operator's own per-panel row sums
  } da, dt;
  int nrb = 0, nrbt = 0;  // partial slots per direction
  // bumped by every layout change (lanes per row, sorted layout): a batch
  // built on the earlier layout refuses to run
  unsigned layout_epoch = 0;
};

// Runs f(i) for i in [0, n) on up to host_threads() threads (the calling
// thread among them), indices handed out one at a time; returns when all
// are done.  host_threads(): PK_HOST_THREADS when set, else the hardware
// threads, at most 16.
int host_threads();
void host_parallel_for(int n, const std::function<void(int)>& f);

// Batched bidiagonal SVD on the device (pk_bsvd.hip): orders up to
// kBsvdMax; device buffers kept in a BsvdBuf between calls.
constexpr int kBsvdMax = 128;
struct BsvdBuf {
  int* dn = nullptr;
  double* dv = nullptr;
  int nb = 0, ld = 0;
};
int bsvd_host(BsvdBuf& buf, hipStream_t s, int nb, const int* n, const double* d,
              const double* e, int ld, double* sigma, double* bot);
void bsvd_free(BsvdBuf& buf);

// panels_a / panels_t: column panels of A and A^T (powers of two up to
// CsrOp::kMaxPanels), or <= 0 for the built-in choice.
int csr_create(int m, int n, long long nnz, const int* rowptr, const int* colind,
               const void* val, int val_prec, int where, CsrOp** out, int panels_a = 0,
               int panels_t = 0);
void csr_destroy(CsrOp* op);
// Lanes per row of every panel of D: the built-in choice times 2^shift,
// clamped to [2, 64]; shift in [-2, 2].
void csr_set_lanes(CsrOp::Dir& D, int shift);
// Switches direction D of op to the column-sorted layout with blocks of at
// most cap nonzeros (kSpmvCap <= cap <= kSortCapMax; built on the device),
// or back to the row-block layout (cap = 0, the sorted copy freed), or to
// the row-block arrays read row-wise (cap = -1).
// cap <= -kSpmvCap: the column-sorted layout with blocks of at most -cap
// nonzeros and every column panel in one launch (concurrent panels; kept
// sequential when the direction has one panel or long-row chunks).
int csr_set_sort(CsrOp* op, CsrOp::Dir& D, int cap);
// The layout code pk_csr_sort reports for a direction: -1 row-wise, -cap
// concurrent panels, cap sorted, 0 row-block.
int csr_sort_code(const CsrOp::Dir& D);
// Whether a direction runs its panels in one launch (concurrent layout, one
// lanes-per-row count across its panels), and the running-sum scratch an
// application needs (doubles): np * rows then, rows otherwise.
bool csr_concurrent_ok(const CsrOp::Dir& D);
long long csr_ywork(const CsrOp* op);
// ywork (rows of the direction, doubles) and lwork (csr_lwork(op) doubles),
// when given, replace the operator's own running row sums and long-row
// partial sums, so that callers with their own scratch (every engine) may
// apply one operator at the same time on different streams.
// ywork_len (doubles; 0: rows) bounds ywork: a concurrent direction whose
// np * rows does not fit runs its panels one launch each instead.
template <typename T>
int csr_apply(const CsrOp* op, hipStream_t s, char trans, const T* x, T* y, const T* yprev,
              Scal coef, double* part, Gate gate, double* ywork = nullptr,
              double* lwork = nullptr, long long ywork_len = 0);
// Long-row partial sums one panel can need, for any layout of op: a row of
// len > kSpmvCap nonzeros has ceil(len / kLongChunk) <= len / kSpmvCap chunks.
inline long long csr_lwork(const CsrOp* op) { return op->nnz / 1024 + 2; }

// ---- complex precisions (pk_cplx.hip) ----
// A complex operator (val_prec PK_C128 / PK_C64) keeps A and the explicit
// A^T as column panels of plain CSR (values as (re, im) pairs, one
// lanes-per-row count per direction in pan[p].g), at most kMaxZPanels per
// direction; the sorted layouts and lane shifts of the real operators do not
// apply to it.
constexpr int kMaxZPanels = 8;
inline bool prec_is_complex(int prec) { return prec == PK_C128 || prec == PK_C64; }
inline size_t prec_bytes(int prec) {
  return prec == PK_F64 ? 8 : prec == PK_F32 ? 4 : prec == PK_C128 ? 16 : 8;
}
int zstream_blocks(long long L);
int zspmv_lanes(long long nnz, long long rows);
int zspmv_blocks(long long rows, int g);  // <= kMaxBlocks
// y = op(A) x - c yprev (trans 'n', 't' or 'c' = A^H) with partials
// ||y||^2, Re and Im of yprev^H y (three values per slot); returns the slot
// count.  R is the real type of the storage precision.  A direction with
// several panels runs them in one launch and needs ywork of
// zcsr_ywork(op) doubles (the panels' FP64 row sums).
template <typename R>
int zcsr_apply(const CsrOp* op, hipStream_t s, char trans, const void* x, void* y,
               const void* yprev, double c, double* part, double* ywork);
long long zcsr_ywork(const CsrOp* op);
// w -= s v (complex s = sr + i si); partials ||w||^2, Re and Im of v^H w.
template <typename R>
int launch_zaxpy_norm_dot(hipStream_t s, long long L, void* w, const void* v, double sr,
                          double si, double* part);
// Partials of h = V(:, cols)^H w: real and imaginary parts of h_j at
// part[(2j) * nblk + b] and part[(2j+1) * nblk + b].
template <typename R>
int launch_zgemvt(hipStream_t s, long long L, const void* V, long long ldv, const int* cols,
                  int ncols, const void* w, double* part);
// w -= V(:, cols) h (h as 2 ncols doubles, re / im pairs); partial ||w||^2.
template <typename R>
int launch_zgemv_update(hipStream_t s, long long L, const void* V, long long ldv,
                        const int* cols, int ncols, const double* h, void* w, double* part);

// Fortran APROD callback: (TRANSA, M, N, X, Y, DPARM, IPARM, len(TRANSA)).
typedef void (*aprod_d_fn)(const char*, const int*, const int*, const double*, double*,
                           double*, int*, size_t);
typedef void (*aprod_s_fn)(const char*, const int*, const int*, const float*, float*, float*,
                           int*, size_t);

// Live per-kernel-class timing with HIP events on the engine's stream.
enum KClass { KC_SPMV_N = 0, KC_SPMV_T, KC_GEMVT, KC_GEMV, KC_GEMM, KC_VEC, KC_COUNT };
struct KTimer {
  int cls;
  hipEvent_t a, b;
  double bytes, flops;
};
struct Profile {
  bool on = false;
  unsigned mask = 0;  // kernel classes timed (bit c: class c)
  std::vector<hipEvent_t> pool;
  std::vector<KTimer> pending;
  double ms[KC_COUNT] = {}, bytes[KC_COUNT] = {}, flops[KC_COUNT] = {};
  long long count[KC_COUNT] = {};
};

struct Counters {
  long long nopx = 0, nreorth = 0, ndot = 0, ncols_touched = 0, nsync = 0, nretry = 0;
  // host-executed rare phases (PH_V_INVARIANT, PH_V_TINY, PH_V_ZERO,
  // PH_U_INVARIANT, PH_U_TINY, PH_U_ZERO) and near-invariance warnings
  long long rare[6] = {0, 0, 0, 0, 0, 0};
  long long nwarn = 0;
};

struct Engine {
  static constexpr unsigned kMagic = 0x504b4c42u;  // "PKLB"
  unsigned magic = 0;
  int prec = PK_F64;                 // storage type of bases (and of the operator values):
                                     // PK_F64, PK_F32, or complex PK_C128 / PK_C64
  int m = 0, n = 0, kmax = 0;
  long long ldu = 0, ldv = 0;
  void* U[2] = {nullptr, nullptr};   // current / spare left basis  (ldu x (kmax+1))
  void* V[2] = {nullptr, nullptr};   // current / spare right basis (ldv x kmax)
  int ucur = 0, vcur = 0;
  size_t tbytes = 8;
  // operator: device CSR, or a host APROD callback
  CsrOp* op = nullptr;
  void* aprod = nullptr;
  int aprod_prec = PK_F64;           // precision of the callback's X/Y arrays
  void* dparm = nullptr;
  int* iparm = nullptr;
  void* hx = nullptr;                // pinned staging for the callback path
  void* hy = nullptr;
  // device work
  hipStream_t stream = nullptr;
  int ccap = 0;                      // column capacity of one Gram-Schmidt pass: max(kMaxCols, kmax + 2)
  double* part = nullptr;            // partial sums (max_blocks * ccap)
  double* red = nullptr;             // reduced sums (ccap + 16)
  double* hred = nullptr;            // pinned host mirror of red
  int* dcols = nullptr;              // reorthogonalization column list
  int* hcols = nullptr;              // pinned host staging of the list
  double* dB = nullptr;              // small dense matrix for the tall GEMM
  double* hB = nullptr;              // pinned host staging
  hipEvent_t evB = nullptr;          // the last copy out of hB (reuse guard)
  void* vec = nullptr;               // scratch vector, max(m, n)
  double* sw_y = nullptr;            // SpMV running row sums (csr_ywork(op) at creation)
  long long sw_y_len = 0;
  double* sw_l = nullptr;            // SpMV long-row partial sums (csr_lwork(op))
  void* kres = nullptr;              // residual check: three vectors of max(m, n), on first use
  LbdCtl* dctl = nullptr;            // step control state: scalars (device)
  void* darr = nullptr;              // its arrays, capacity kmax (CtlArrays layout)
  LbdCtlRef* hctl = nullptr;         // pinned host copy: scalars, pointers into harr
  void* harr = nullptr;              // pinned host copy of the arrays
  int mode = 0;                      // 0: device-resident steps, 1: host executor
  int dev_tries = 1;                 // CGS passes per side held by a device chain
  // dev_tries adapts: a chain that ran out of passes (a second CGS pass was
  // needed, finished on the host) raises it to 2 for the engine's later
  // calls.  pk_lbd_set_mode fixes it.
  bool tries_auto = true;
  // Rounding model of the thresholds epsn = leff * eps, epsn2 = sqrt(leff) * eps:
  // 0 = the reference's (leff = max(m, n): accumulation in the storage
  // precision); 1 = storage rounding only (leff = 4: FP32 storage with FP64
  // accumulation, where max(m, n) * 2^-24 would exceed the Lanczos residuals
  // of large problems and flag false invariant subspaces).
  int rounding = 0;
  // complex bases (prec PK_C128 / PK_C64): the imaginary part of the
  // pending ELR coefficient (its real part is LbdCtl::elr_s)
  double elr_im = 0.0;
  Counters cnt;
  Profile prof;
  int last_error = 0;

  void* u() const { return U[ucur]; }
  void* v() const { return V[vcur]; }
};

// A DLANBPRO call in three pieces (setup into e->hctl, the steps, the
// outputs), so that several engines can run their steps together.
int engine_lanbpro_begin(Engine* e, int k0, int k, double* alpha, double* beta, double* rnorm,
                         double* doption, const int* ioption, int* ierr, bool* finished);
int engine_lanbpro_loop(Engine* e);
int engine_lanbpro_end(Engine* e, int k0, int* kp, double* alpha, double* beta, double* rnorm,
                       double* doption, int* ierr);
Engine* engine_from_handle(pk_lbd_t h);

}  // namespace pk
