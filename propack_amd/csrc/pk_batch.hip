// Batched Lanczos steps: many independent engines (one truncated SVD each,
// of uneven sizes) advance together, one launch per kernel per step for
// the whole batch.  Every workgroup looks up its problem in a table and runs
// the same per-workgroup body as the single-problem kernels (pk_bodies.h),
// on the same rows, slots and summation order, so a problem gives
// bit-identical results in a batch and alone.  Each problem's control state
// gates its own workgroups (a problem that needs no reorthogonalization, or
// has finished, idles through that part of the step).
//
// A problem whose state leaves the device-resident phases (invariant
// subspace, tiny scale, failed Gram-Schmidt) drops out of the batched chain
// and finishes its call on its own engine afterwards.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <system_error>
#include <thread>
#include <vector>

#include "pk_bodies.h"
#include "pk_internal.h"

namespace pk {
void set_error(const char* fmt, ...);

namespace {

constexpr int kPanelsMax = CsrOp::kMaxPanels;

// Per-problem view, rebuilt for every batched call (restarts swap bases).
struct BProb {
  LbdCtl* ctl;
  char* arr;     // its arrays (CtlArrays layout, capacity cap)
  int cap;
  void* U;
  void* V;
  long long ldu, ldv;
  int m, n;
  int sm0, nsm;  // workgroup slices over m rows: first global slice, count
  int sn0, nsn;  // over n rows
  int cm0, ncm;  // Gram-Schmidt slices (cgs_blocks) over m rows
  int cn0, ncn;  // over n rows
  int qa0, nqa;  // partial slots of A v (last panel): first global slot, count
  int qt0, nqt;  // of A^T u
  int npa, npt;  // column panels of A, A^T
  double* yacc_a;
  double* yacc_t;
};

// A panel of one problem's direction in the layout its operator uses: the
// row-block arrays (spos null) or the column-sorted ones.
struct BPanel {
  const int* rp;
  const int* ci;
  const void* va;
  double* lsum;
  const unsigned short* spos;
  int col0, pbits;  // packed sorted layout (CsrOp::Panel::scol0, spbits)
};

// Work item of a batched SpMV launch: descriptor of problem pid, local
// partial slot loc (meaningful in the problem's last panel).
struct BWork {
  int pid, loc;
};

template <typename T>
__device__ __forceinline__ T* bcol(const BProb& P, bool vside, int j) {
  return vside ? static_cast<T*>(P.V) + static_cast<long long>(j - 1) * P.ldv
               : static_cast<T*>(P.U) + static_cast<long long>(j - 1) * P.ldu;
}

// ---------------------------------------------------------------- kernels --
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bscale(const BProb* __restrict__ probs,
                                                   const int* __restrict__ slice_pid, int vside,
                                                   int J, int phase, int flag) {
  const BProb P = probs[slice_pid[blockIdx.x]];
  if (Gate{P.ctl, phase, J, flag}.off()) return;
  const long long b = static_cast<long long>(blockIdx.x) - (vside ? P.sn0 : P.sm0);
  const double f = flag == 2 ? P.ctl->uscale_f : P.ctl->vscale_f;
  scale_body<T>(b, vside ? P.n : P.m, bcol<T>(P, vside, J), f);
}

// tr = 1: v_J = A^T u_J - beta v_{J-1}; tr = 0: u_{J+1} = A v_J - alpha u_J.
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kBlock) void k_bspmv(const BProb* __restrict__ probs,
                                                  const BPanel* __restrict__ panels,
                                                  const int4* __restrict__ desc,
                                                  const BWork* __restrict__ work, int tr, int p,
                                                  int J, int phase, double* __restrict__ part,
                                                  int Q) {
  const BWork w = work[blockIdx.x];
  const BProb P = probs[w.pid];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  __shared__ double prod[kSpmvCap];
  __shared__ double lds[(kBlock / kWave) * 2];
  const BPanel pan = panels[(w.pid * 2 + tr) * kPanelsMax + p];
  const int np = tr ? P.npt : P.npa;
  const int last = p == np - 1;
  double* yacc = tr ? P.yacc_t : P.yacc_a;
  const T* x = tr ? bcol<T>(P, false, J) : bcol<T>(P, true, J);
  T* y = tr ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  const T* yprev = tr ? (J > 1 ? bcol<T>(P, true, J - 1) : nullptr) : bcol<T>(P, false, J);
  const double c = tr ? P.ctl->beta : P.ctl->alpha;
  const int q = (tr ? P.qt0 : P.qa0) + w.loc;
  spmv_rb_body<VT, T, G>(desc[blockIdx.x], pan.rp, pan.ci, static_cast<const VT*>(pan.va), x, y,
                         yprev, c, part + q, part + Q + q, pan.lsum, p == 0 ? nullptr : yacc,
                         last ? nullptr : yacc, last, prod, lds);
}

// The row-wise form over the row-block descriptors (spmv_vec_body).
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kBlock) void k_bspmv_vec(const BProb* __restrict__ probs,
                                                      const BPanel* __restrict__ panels,
                                                      const int4* __restrict__ desc,
                                                      const BWork* __restrict__ work, int tr,
                                                      int p, int J, int phase,
                                                      double* __restrict__ part, int Q) {
  const BWork w = work[blockIdx.x];
  const BProb P = probs[w.pid];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  __shared__ double lds[(kBlock / kWave) * 2];
  const BPanel pan = panels[(w.pid * 2 + tr) * kPanelsMax + p];
  const int np = tr ? P.npt : P.npa;
  const int last = p == np - 1;
  double* yacc = tr ? P.yacc_t : P.yacc_a;
  const T* x = tr ? bcol<T>(P, false, J) : bcol<T>(P, true, J);
  T* y = tr ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  const T* yprev = tr ? (J > 1 ? bcol<T>(P, true, J - 1) : nullptr) : bcol<T>(P, false, J);
  const double c = tr ? P.ctl->beta : P.ctl->alpha;
  const int q = (tr ? P.qt0 : P.qa0) + w.loc;
  spmv_vec_body<VT, T, G>(desc[blockIdx.x], pan.rp, pan.ci, static_cast<const VT*>(pan.va), x,
                          y, yprev, c, part + q, part + Q + q, pan.lsum,
                          p == 0 ? nullptr : yacc, last ? nullptr : yacc, last, lds);
}

// The same over the column-sorted layout (spmv_sorted_body): one
// kSortThreads workgroup per descriptor, the launch's largest block cap of
// dynamic LDS.
template <typename VT, typename T, int G>
__global__ __launch_bounds__(kSortThreads) void k_bspmv_sorted(
    const BProb* __restrict__ probs, const BPanel* __restrict__ panels,
    const int4* __restrict__ desc, const BWork* __restrict__ work, int tr, int p, int J,
    int phase, double* __restrict__ part, int Q) {
  const BWork w = work[blockIdx.x];
  const BProb P = probs[w.pid];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  extern __shared__ double prod_dyn[];
  __shared__ double lds[(kSortThreads / kWave) * 2];
  const BPanel pan = panels[(w.pid * 2 + tr) * kPanelsMax + p];
  const int np = tr ? P.npt : P.npa;
  const int last = p == np - 1;
  double* yacc = tr ? P.yacc_t : P.yacc_a;
  const T* x = tr ? bcol<T>(P, false, J) : bcol<T>(P, true, J);
  T* y = tr ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  const T* yprev = tr ? (J > 1 ? bcol<T>(P, true, J - 1) : nullptr) : bcol<T>(P, false, J);
  const double c = tr ? P.ctl->beta : P.ctl->alpha;
  const int q = (tr ? P.qt0 : P.qa0) + w.loc;
  spmv_sorted_body<VT, T, G>(desc[blockIdx.x], pan.rp, pan.ci, static_cast<const VT*>(pan.va),
                             pan.spos, pan.col0, pan.pbits, x, y, yprev, c, part + q,
                             part + Q + q, pan.lsum,
                             p == 0 ? nullptr : yacc, last ? nullptr : yacc, last, prod_dyn,
                             lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_blongfix(const BProb* __restrict__ probs,
                                                     const BPanel* __restrict__ panels,
                                                     const int4* __restrict__ ldesc,
                                                     const BWork* __restrict__ work, int nlong,
                                                     int tr, int p, int J, int phase,
                                                     double* __restrict__ part, int Q) {
  const int i = blockIdx.x * (kBlock / kWave) + static_cast<int>(threadIdx.x) / kWave;
  if (i >= nlong) return;  // one wave per long row
  const BWork w = work[i];
  const BProb P = probs[w.pid];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  const BPanel pan = panels[(w.pid * 2 + tr) * kPanelsMax + p];
  const int np = tr ? P.npt : P.npa;
  const int last = p == np - 1;
  double* yacc = tr ? P.yacc_t : P.yacc_a;
  T* y = tr ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  const T* yprev = tr ? (J > 1 ? bcol<T>(P, true, J - 1) : nullptr) : bcol<T>(P, false, J);
  const double c = tr ? P.ctl->beta : P.ctl->alpha;
  const int q = (tr ? P.qt0 : P.qa0) + w.loc;
  spmv_longfix_body<T>(ldesc[i], pan.lsum, y, yprev, c, part + q, part + Q + q,
                       p == 0 ? nullptr : yacc, last ? nullptr : yacc, last,
                       static_cast<int>(threadIdx.x) % kWave);
}

// One workgroup per problem: the control stage of `kind` over the problem's
// own partial range.
__global__ __launch_bounds__(kBlock) void k_bctl(const BProb* __restrict__ probs, int kind,
                                                 int phase, int J, const double* __restrict__ part,
                                                 int Qa, int Qt, int Sm, int Sn) {
  const BProb P = probs[blockIdx.x];
  __shared__ LbdCtlLds c;
  __shared__ LbdCtlRef cg;
  __shared__ double lds[(kBlock / kWave) * 2];
  if (P.ctl->phase != phase || P.ctl->j != J) return;
  const double* pa;
  const double* pb = nullptr;
  int na;
  switch (kind) {
    case CTL_V_SPMV: pa = part + P.qt0; pb = part + Qt + P.qt0; na = P.nqt; break;
    case CTL_U_SPMV: pa = part + P.qa0; pb = part + Qa + P.qa0; na = P.nqa; break;
    case CTL_V_ELR: pa = part + P.sn0; pb = part + Sn + P.sn0; na = P.nsn; break;
    case CTL_U_ELR: pa = part + P.sm0; pb = part + Sm + P.sm0; na = P.nsm; break;
    case CTL_V_REO: pa = part + P.cn0; na = P.ncn; break;
    default: pa = part + P.cm0; na = P.ncm; break;  // CTL_U_REO
  }
  ctl_stage_body(P.ctl, P.arr, P.cap, J, kind, pa, na, pb, na, c, cg, lds);
}

// ELR update: v side w = v_J, v = v_{J-1}; u side w = u_{J+1}, v = u_J.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_baxpy(const BProb* __restrict__ probs,
                                                  const int* __restrict__ slice_pid, int vside,
                                                  int J, int phase, double* __restrict__ part,
                                                  int S) {
  const BProb P = probs[slice_pid[blockIdx.x]];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  __shared__ double lds[(kBlock / kWave) * 2];
  const int s0 = vside ? P.sn0 : P.sm0;
  const long long b = static_cast<long long>(blockIdx.x) - s0;
  T* w = vside ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  const T* v = vside ? bcol<T>(P, true, J - 1) : bcol<T>(P, false, J);
  axpy_norm_dot_body<T>(b, vside ? P.n : P.m, w, v, P.ctl->elr_s, part + blockIdx.x,
                        part + S + blockIdx.x, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bgemvt(const BProb* __restrict__ probs,
                                                   const int* __restrict__ slice_pid, int vside,
                                                   int J, int phase, double* __restrict__ part,
                                                   int S) {
  const BProb P = probs[slice_pid[blockIdx.x]];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  __shared__ double lds[2][(kBlock / kWave) * kCT];
  const int ncols = P.ctl->reo_ncols;
  const int* scol = CtlArrays::cols(P.arr, P.cap);
  const long long b = static_cast<long long>(blockIdx.x) - (vside ? P.cn0 : P.cm0);
  const T* basis = static_cast<const T*>(vside ? P.V : P.U);
  const T* w = vside ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  gemvt_body<T>(b, vside ? P.n : P.m, basis, vside ? P.ldv : P.ldu, scol, ncols, w, part, S,
                blockIdx.x, lds);
}

// red[pid * rs + c] = sum over the problem's slices of the projection
// partials of column c (grid: problems x columns).
__global__ __launch_bounds__(kBlock) void k_breduce(const BProb* __restrict__ probs, int vside,
                                                    int J, int phase,
                                                    const double* __restrict__ part, int S,
                                                    double* __restrict__ red, int rs) {
  const BProb P = probs[blockIdx.x];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  const int c = blockIdx.y;
  if (c >= P.ctl->reo_ncols) return;
  __shared__ double lds[kBlock / kWave];
  const int s0 = vside ? P.cn0 : P.cm0, ns = vside ? P.ncn : P.ncm;
  const double v = reduce_partials(part + static_cast<long long>(c) * S + s0, ns, lds);
  if (threadIdx.x == 0) red[static_cast<long long>(blockIdx.x) * rs + c] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bgemv(const BProb* __restrict__ probs,
                                                  const int* __restrict__ slice_pid, int vside,
                                                  int J, int phase,
                                                  const double* __restrict__ red, int rs,
                                                  double* __restrict__ part) {
  const int pid = slice_pid[blockIdx.x];
  const BProb P = probs[pid];
  if (Gate{P.ctl, phase, J, 0}.off()) return;
  __shared__ double lds[kBlock / kWave];
  const int ncols = P.ctl->reo_ncols;
  const int* scol = CtlArrays::cols(P.arr, P.cap);
  const double* sh = red + static_cast<long long>(pid) * rs;
  const long long b = static_cast<long long>(blockIdx.x) - (vside ? P.cn0 : P.cm0);
  const T* basis = static_cast<const T*>(vside ? P.V : P.U);
  T* w = vside ? bcol<T>(P, true, J) : bcol<T>(P, false, J + 1);
  gemv_update_body<T>(b, vside ? P.n : P.m, basis, vside ? P.ldv : P.ldu, scol, sh, ncols, w,
                      part + blockIdx.x, lds);
}

}  // namespace

// ------------------------------------------------------------------ batch --
struct Batch {
  static constexpr unsigned kMagic = 0x504b4254u;  // "PKBT"
  unsigned magic = 0;
  int nprob = 0, prec = PK_F64, val_prec = PK_F64;
  std::vector<Engine*> eng;
  std::vector<unsigned> epoch;  // each operator's layout_epoch at creation
  hipStream_t stream = nullptr;
  // static tables (operators do not change)
  int Sm = 0, Sn = 0, Qa = 0, Qt = 0, maxk = 0;
  int rstride = 0;             // per-problem stride of red: the largest column capacity
  int Cm = 0, Cn = 0;          // Gram-Schmidt slices over m, n rows
  int* slice_pid_m = nullptr;  // global slice over m -> problem
  int* slice_pid_n = nullptr;
  int* cslice_pid_m = nullptr;  // Gram-Schmidt slice over m -> problem
  int* cslice_pid_n = nullptr;
  BPanel* panels = nullptr;    // (pid * 2 + tr) * kPanelsMax + p
  struct Launch {              // one batched SpMV launch
    int tr, p, g, cap, count;  // cap > 0: column-sorted blocks of at most cap nonzeros;
                               // cap < 0: row-wise over the row-block arrays
    int4* desc;
    BWork* work;
  };
  struct Long {
    int tr, p, count;
    int4* ldesc;
    BWork* work;
  };
  std::vector<Launch> spmv;
  std::vector<Long> lng;
  std::vector<BProb> hprob;    // host template (pointers filled per call)
  BProb* dprob = nullptr;
  double* part = nullptr;
  double* red = nullptr;
  size_t part_len = 0;
  // The kernel chain of a batched call depends only on its step range, the
  // ELR and CGS pass counts and the tables above (per-problem data reach
  // the kernels through dprob, rewritten before every launch), so it is
  // captured once per key into a hipGraph and replayed (a restarted solve
  // replays the same chain every restart).
  struct Graph {
    int j0, j1, elr, tries;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  bool use_graph = true;
  // the restarted driver's bidiagonal SVDs on the device (pk_bsvd.hip)
  bool use_bsvd = false;
  BsvdBuf bsvd;
  // the restart GEMMs of all problems in one launch (launch_gemm_group)
  GemmItem* gitems = nullptr;
  size_t gcap = 0;
};
constexpr size_t kMaxBatchGraphs = 8;

namespace {

template <typename T>
void enqueue_batch_step(Batch& B, int J, int elr, int tries) {
  hipStream_t s = B.stream;
  const int np = B.nprob;
  // ---- v_J = A^T u_J - beta v_{J-1} ----
  hipLaunchKernelGGL(k_bscale<T>, dim3(B.Sm), dim3(kBlock), 0, s, B.dprob, B.slice_pid_m, 0, J,
                     static_cast<int>(PH_V_SPMV), 2);
  auto spmv = [&](int tr, int phase) {
    for (int p = 0; p < kPanelsMax; ++p) {
      bool any = false;
      for (const auto& L : B.spmv) {
        if (L.tr != tr || L.p != p || L.count == 0) continue;
        any = true;
        const int Q = tr ? B.Qt : B.Qa;
#define PK_BS_CASE(G)                                                                           \
  case G:                                                                                       \
    if (B.val_prec == PK_F64)                                                                   \
      hipLaunchKernelGGL((k_bspmv<double, T, G>), dim3(L.count), dim3(kBlock), 0, s, B.dprob,   \
                         B.panels, L.desc, L.work, tr, p, J, phase, B.part, Q);                 \
    else                                                                                        \
      hipLaunchKernelGGL((k_bspmv<float, T, G>), dim3(L.count), dim3(kBlock), 0, s, B.dprob,    \
                         B.panels, L.desc, L.work, tr, p, J, phase, B.part, Q);                 \
    break;
#define PK_BV_CASE(G)                                                                           \
  case G:                                                                                       \
    if (B.val_prec == PK_F64)                                                                   \
      hipLaunchKernelGGL((k_bspmv_vec<double, T, G>), dim3(L.count), dim3(kBlock), 0, s,        \
                         B.dprob, B.panels, L.desc, L.work, tr, p, J, phase, B.part, Q);        \
    else                                                                                        \
      hipLaunchKernelGGL((k_bspmv_vec<float, T, G>), dim3(L.count), dim3(kBlock), 0, s,         \
                         B.dprob, B.panels, L.desc, L.work, tr, p, J, phase, B.part, Q);        \
    break;
#define PK_BSS_CASE(G)                                                                          \
  case G:                                                                                       \
    if (B.val_prec == PK_F64)                                                                   \
      hipLaunchKernelGGL((k_bspmv_sorted<double, T, G>), dim3(L.count), dim3(kSortThreads), shm, \
                         s, B.dprob, B.panels, L.desc, L.work, tr, p, J, phase, B.part, Q);     \
    else                                                                                        \
      hipLaunchKernelGGL((k_bspmv_sorted<float, T, G>), dim3(L.count), dim3(kSortThreads), shm,  \
                         s, B.dprob, B.panels, L.desc, L.work, tr, p, J, phase, B.part, Q);     \
    break;
        if (L.cap > 0) {
          const size_t shm = sizeof(double) * static_cast<size_t>(L.cap);
          switch (L.g) {
            PK_BSS_CASE(2)
            PK_BSS_CASE(4)
            PK_BSS_CASE(8)
            PK_BSS_CASE(16)
            PK_BSS_CASE(32)
            PK_BSS_CASE(64)
            default: break;
          }
        } else if (L.cap < 0) {
          switch (L.g) {
            PK_BV_CASE(2)
            PK_BV_CASE(4)
            PK_BV_CASE(8)
            PK_BV_CASE(16)
            PK_BV_CASE(32)
            PK_BV_CASE(64)
            default: break;
          }
        } else {
          switch (L.g) {
            PK_BS_CASE(2)
            PK_BS_CASE(4)
            PK_BS_CASE(8)
            PK_BS_CASE(16)
            PK_BS_CASE(32)
            PK_BS_CASE(64)
            default: break;
          }
        }
#undef PK_BS_CASE
#undef PK_BV_CASE
#undef PK_BSS_CASE
      }
      for (const auto& L : B.lng) {
        if (L.tr != tr || L.p != p || L.count == 0) continue;
        any = true;
        hipLaunchKernelGGL(k_blongfix<T>, dim3(cdiv(L.count, kBlock / kWave)), dim3(kBlock), 0,
                           s, B.dprob, B.panels, L.ldesc, L.work, L.count, tr, p, J, phase,
                           B.part, tr ? B.Qt : B.Qa);
      }
      if (!any) break;
    }
  };
  spmv(1, PH_V_SPMV);
  hipLaunchKernelGGL(k_bctl, dim3(np), dim3(kBlock), 0, s, B.dprob, static_cast<int>(CTL_V_SPMV),
                     static_cast<int>(PH_V_SPMV), J, B.part, B.Qa, B.Qt, B.Sm, B.Sn);
  if (J > 1) {
    for (int i = 0; i < elr; ++i) {
      hipLaunchKernelGGL(k_baxpy<T>, dim3(B.Sn), dim3(kBlock), 0, s, B.dprob, B.slice_pid_n, 1,
                         J, static_cast<int>(PH_V_ELR), B.part, B.Sn);
      hipLaunchKernelGGL(k_bctl, dim3(np), dim3(kBlock), 0, s, B.dprob,
                         static_cast<int>(CTL_V_ELR), static_cast<int>(PH_V_ELR), J, B.part, B.Qa,
                         B.Qt, B.Sm, B.Sn);
    }
    for (int t = 0; t < tries; ++t) {
      hipLaunchKernelGGL(k_bgemvt<T>, dim3(B.Cn), dim3(kBlock), 0, s, B.dprob, B.cslice_pid_n, 1,
                         J, static_cast<int>(PH_V_REO), B.part, B.Cn);
      hipLaunchKernelGGL(k_breduce, dim3(np, J - 1), dim3(kBlock), 0, s, B.dprob, 1, J,
                         static_cast<int>(PH_V_REO), B.part, B.Cn, B.red, B.rstride);
      hipLaunchKernelGGL(k_bgemv<T>, dim3(B.Cn), dim3(kBlock), 0, s, B.dprob, B.cslice_pid_n, 1,
                         J, static_cast<int>(PH_V_REO), B.red, B.rstride, B.part);
      hipLaunchKernelGGL(k_bctl, dim3(np), dim3(kBlock), 0, s, B.dprob,
                         static_cast<int>(CTL_V_REO), static_cast<int>(PH_V_REO), J, B.part, B.Qa,
                         B.Qt, B.Sm, B.Sn);
    }
  }
  // ---- u_{J+1} = A v_J - alpha u_J ----
  hipLaunchKernelGGL(k_bscale<T>, dim3(B.Sn), dim3(kBlock), 0, s, B.dprob, B.slice_pid_n, 1, J,
                     static_cast<int>(PH_U_SPMV), 1);
  spmv(0, PH_U_SPMV);
  hipLaunchKernelGGL(k_bctl, dim3(np), dim3(kBlock), 0, s, B.dprob, static_cast<int>(CTL_U_SPMV),
                     static_cast<int>(PH_U_SPMV), J, B.part, B.Qa, B.Qt, B.Sm, B.Sn);
  for (int i = 0; i < elr; ++i) {
    hipLaunchKernelGGL(k_baxpy<T>, dim3(B.Sm), dim3(kBlock), 0, s, B.dprob, B.slice_pid_m, 0, J,
                       static_cast<int>(PH_U_ELR), B.part, B.Sm);
    hipLaunchKernelGGL(k_bctl, dim3(np), dim3(kBlock), 0, s, B.dprob, static_cast<int>(CTL_U_ELR),
                       static_cast<int>(PH_U_ELR), J, B.part, B.Qa, B.Qt, B.Sm, B.Sn);
  }
  for (int t = 0; t < tries; ++t) {
    hipLaunchKernelGGL(k_bgemvt<T>, dim3(B.Cm), dim3(kBlock), 0, s, B.dprob, B.cslice_pid_m, 0, J,
                       static_cast<int>(PH_U_REO), B.part, B.Cm);
    hipLaunchKernelGGL(k_breduce, dim3(np, J), dim3(kBlock), 0, s, B.dprob, 0, J,
                       static_cast<int>(PH_U_REO), B.part, B.Cm, B.red, B.rstride);
    hipLaunchKernelGGL(k_bgemv<T>, dim3(B.Cm), dim3(kBlock), 0, s, B.dprob, B.cslice_pid_m, 0, J,
                       static_cast<int>(PH_U_REO), B.red, B.rstride, B.part);
    hipLaunchKernelGGL(k_bctl, dim3(np), dim3(kBlock), 0, s, B.dprob, static_cast<int>(CTL_U_REO),
                       static_cast<int>(PH_U_REO), J, B.part, B.Qa, B.Qt, B.Sm, B.Sn);
  }
}

void batch_free(Batch* b) {
  if (!b) return;
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
  bsvd_free(b->bsvd);
  for (auto& L : b->spmv) {
    (void)hipFree(L.desc);
    (void)hipFree(L.work);
  }
  for (auto& L : b->lng) {
    (void)hipFree(L.ldesc);
    (void)hipFree(L.work);
  }
  (void)hipFree(b->slice_pid_m);
  (void)hipFree(b->slice_pid_n);
  (void)hipFree(b->cslice_pid_m);
  (void)hipFree(b->cslice_pid_n);
  (void)hipFree(b->panels);
  (void)hipFree(b->dprob);
  (void)hipFree(b->part);
  (void)hipFree(b->red);
  (void)hipFree(b->gitems);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->magic = 0;
  delete b;
}

template <typename X>
bool upload_vec(const std::vector<X>& h, X** d) {
  *d = nullptr;
  if (h.empty()) return true;
  return hipMalloc(d, sizeof(X) * h.size()) == hipSuccess &&
         hipMemcpy(*d, h.data(), sizeof(X) * h.size(), hipMemcpyHostToDevice) == hipSuccess;
}

int batch_create(int nprob, const pk_lbd_t* handles, Batch** out) {
  if (nprob < 1 || !handles || !out) return PK_ERR_ARG;
  Batch* B = new Batch();
  B->magic = Batch::kMagic;
  B->nprob = nprob;
  int rc = PK_OK;
  std::vector<int> spm, spn, cpm, cpn;
  std::vector<BPanel> hpan(static_cast<size_t>(nprob) * 2 * kPanelsMax, BPanel{});
  // SpMV work lists keyed by (tr, panel, lanes per row, sorted-layout cap)
  struct Key {
    int tr, p, g, cap;
    std::vector<int4> desc;
    std::vector<BWork> work;
  };
  std::vector<Key> keys;
  struct LKey {
    int tr, p;
    std::vector<int4> ld;
    std::vector<BWork> work;
  };
  std::vector<LKey> lkeys;
  B->hprob.resize(nprob);
  for (int i = 0; i < nprob && rc == PK_OK; ++i) {
    Engine* e = engine_from_handle(handles[i]);
    if (!e || !e->op) {
      set_error("pk_batch_create: problem %d has no device operator", i);
      rc = PK_ERR_ARG;
      break;
    }
    if (prec_is_complex(e->prec)) {
      set_error("pk_batch_create: problem %d is complex; batches take real engines", i);
      rc = PK_ERR_NOTSUP;
      break;
    }
    if (i == 0) {
      B->prec = e->prec;
      B->val_prec = e->op->val_prec;
    } else if (e->prec != B->prec || e->op->val_prec != B->val_prec) {
      set_error("pk_batch_create: problem %d has another precision", i);
      rc = PK_ERR_ARG;
      break;
    }
    // An engine holds one problem's state and scratch (its SpMV running row
    // sums and long-row sums among them): it may appear once in a batch.
    // Engines on one operator may: the operator's arrays are only read.
    bool shared = false;
    for (const Engine* o : B->eng) shared = shared || o == e;
    if (shared) {
      set_error("pk_batch_create: problem %d shares its engine with another", i);
      rc = PK_ERR_ARG;
      break;
    }
    B->eng.push_back(e);
    B->epoch.push_back(e->op->layout_epoch);
    BProb& P = B->hprob[i];
    std::memset(&P, 0, sizeof(BProb));
    P.m = e->m;
    P.n = e->n;
    P.sm0 = B->Sm;
    P.nsm = stream_blocks(e->m);
    P.sn0 = B->Sn;
    P.nsn = stream_blocks(e->n);
    B->Sm += P.nsm;
    B->Sn += P.nsn;
    for (int s = 0; s < P.nsm; ++s) spm.push_back(i);
    for (int s = 0; s < P.nsn; ++s) spn.push_back(i);
    P.cm0 = B->Cm;
    P.ncm = cgs_blocks(e->m);
    P.cn0 = B->Cn;
    P.ncn = cgs_blocks(e->n);
    B->Cm += P.ncm;
    B->Cn += P.ncn;
    for (int s = 0; s < P.ncm; ++s) cpm.push_back(i);
    for (int s = 0; s < P.ncn; ++s) cpn.push_back(i);
    B->maxk = std::max(B->maxk, e->kmax);
    B->rstride = std::max(B->rstride, e->ccap);
    for (int tr = 0; tr < 2 && rc == PK_OK; ++tr) {
      const CsrOp::Dir& D = tr ? e->op->dt : e->op->da;
      (tr ? P.npt : P.npa) = D.np;
      // the engine's own scratch: the SpMVs of both directions run one after
      // the other on the batch's stream
      (tr ? P.yacc_t : P.yacc_a) = e->sw_y;
      // the layout the operator uses for this direction (row-block,
      // row-wise over the row-block arrays (cap -1), or column-sorted),
      // exactly as its own SpMV runs it
      const int cap = D.rowwise ? -1 : D.sort_cap;
      for (int p = 0; p < D.np; ++p) {
        const CsrOp::Panel& pan = D.pan[p];
        const CsrOp::Part& part = cap > 0 ? pan.spart : pan.part;
        hpan[(static_cast<size_t>(i) * 2 + tr) * kPanelsMax + p] =
            cap > 0 ? BPanel{pan.rp, pan.sci, pan.sva, e->sw_l, pan.spos, pan.scol0, pan.spbits}
                    : BPanel{pan.rp, pan.ci, pan.va, e->sw_l, nullptr, 0, 0};
        std::vector<int4> hd(part.ndesc), hl(part.nlong);
        if ((!hd.empty() && hipMemcpy(hd.data(), part.desc, sizeof(int4) * hd.size(),
                                      hipMemcpyDeviceToHost) != hipSuccess) ||
            (!hl.empty() && hipMemcpy(hl.data(), part.ldesc, sizeof(int4) * hl.size(),
                                      hipMemcpyDeviceToHost) != hipSuccess)) {
          rc = PK_ERR_HIP;
          break;
        }
        auto kit = std::find_if(keys.begin(), keys.end(), [&](const Key& k) {
          return k.tr == tr && k.p == p && k.g == pan.g && k.cap == cap;
        });
        if (kit == keys.end()) {
          keys.push_back(Key{tr, p, pan.g, cap, {}, {}});
          kit = keys.end() - 1;
        }
        for (int b = 0; b < static_cast<int>(hd.size()); ++b) {
          kit->desc.push_back(hd[b]);
          kit->work.push_back(BWork{i, b});
        }
        auto lit = std::find_if(lkeys.begin(), lkeys.end(),
                                [&](const LKey& k) { return k.tr == tr && k.p == p; });
        if (lit == lkeys.end()) {
          lkeys.push_back(LKey{tr, p, {}, {}});
          lit = lkeys.end() - 1;
        }
        for (int l = 0; l < static_cast<int>(hl.size()); ++l) {
          lit->ld.push_back(hl[l]);
          lit->work.push_back(BWork{i, part.ndesc + l});
        }
        if (p == D.np - 1) {
          const int slots = part.slots();
          if (tr) {
            P.qt0 = B->Qt;
            P.nqt = slots;
            B->Qt += slots;
          } else {
            P.qa0 = B->Qa;
            P.nqa = slots;
            B->Qa += slots;
          }
        }
      }
    }
  }
  if (rc == PK_OK && hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
    rc = PK_ERR_HIP;
  if (rc == PK_OK) {
    bool ok = upload_vec(spm, &B->slice_pid_m) && upload_vec(spn, &B->slice_pid_n) &&
              upload_vec(cpm, &B->cslice_pid_m) && upload_vec(cpn, &B->cslice_pid_n) &&
              upload_vec(hpan, &B->panels);
    for (auto& k : keys) {
      Batch::Launch L{k.tr, k.p, k.g, k.cap, static_cast<int>(k.desc.size()), nullptr, nullptr};
      ok = ok && upload_vec(k.desc, &L.desc) && upload_vec(k.work, &L.work);
      B->spmv.push_back(L);
    }
    for (auto& k : lkeys) {
      Batch::Long L{k.tr, k.p, static_cast<int>(k.ld.size()), nullptr, nullptr};
      ok = ok && upload_vec(k.ld, &L.ldesc) && upload_vec(k.work, &L.work);
      B->lng.push_back(L);
    }
    const size_t S = static_cast<size_t>(std::max(B->Sm, B->Sn));
    const size_t Q = static_cast<size_t>(std::max(B->Qa, B->Qt));
    B->part_len = std::max(2 * Q, static_cast<size_t>(std::max(B->maxk + 1, 2)) * S);
    ok = ok && hipMalloc(&B->part, sizeof(double) * B->part_len) == hipSuccess &&
         hipMalloc(&B->red, sizeof(double) * static_cast<size_t>(nprob) * B->rstride) ==
             hipSuccess &&
         hipMalloc(&B->dprob, sizeof(BProb) * nprob) == hipSuccess;
    if (!ok) rc = PK_ERR_HIP;
  }
  if (rc != PK_OK) {
    batch_free(B);
    return rc;
  }
  *out = B;
  return PK_OK;
}

// One batched DLANBPRO call: problem i (active[i] != 0) goes from k0[i] to
// k[i] steps.  alpha[i] / beta[i]: its B(:,1), B(:,2); doption 4 per problem.
int batch_lanbpro(Batch& B, const int* active, const int* k0, int* k, double* const* alpha,
                  double* const* beta, double* rnorm, double* doption, const int* ioption,
                  int* ierr) {
  const int np = B.nprob;
  // the batch's descriptor tables and panel pointers follow each operator's
  // layout at creation; a layout changed since then would be read stale
  for (int i = 0; i < np; ++i)
    if (B.eng[i]->op->layout_epoch != B.epoch[i]) {
      set_error("pk_batch_lanbpro: the SpMV layout of problem %d's operator changed after "
                "pk_batch_create; create the batch again",
                i);
      return PK_ERR_ARG;
    }
  std::vector<char> run(np, 0), fin(np, 0);
  std::vector<int> rcs(np, PK_OK);
  int Jmin = 1 << 30, Jmax = 0, elr = 0, tries = 1;
  // each problem's setup (its own engine and stream) on the host threads
  host_parallel_for(np, [&](int i) {
    if (!active[i]) return;
    Engine* e = B.eng[i];
    bool finished = false;
    int rc = engine_lanbpro_begin(e, k0[i], k[i], alpha[i], beta[i], &rnorm[i],
                                  doption + 4 * i, ioption, &ierr[i], &finished);
    if (rc == PK_OK && finished)
      fin[i] = 1;
    else if (rc == PK_OK && hipStreamSynchronize(e->stream) != hipSuccess)
      rc = PK_ERR_HIP;
    rcs[i] = rc;
  });
  for (int i = 0; i < np; ++i)
    if (rcs[i] != PK_OK) return rcs[i];
  for (int i = 0; i < np; ++i) {
    if (!active[i] || fin[i]) continue;
    Engine* e = B.eng[i];
    const LbdCtl& c = *e->hctl;
    if (e->mode == 0 && c.phase == PH_V_SPMV) {
      run[i] = 1;
      Jmin = std::min(Jmin, c.j);
      Jmax = std::max(Jmax, c.k);
      elr = std::max(elr, c.elr);
      tries = std::max(tries, e->dev_tries);
    }
  }
  bool any = false;
  for (int i = 0; i < np; ++i) {
    BProb& P = B.hprob[i];
    Engine* e = B.eng[i];
    P.ctl = e->dctl;
    P.arr = static_cast<char*>(e->darr);
    P.cap = e->kmax;
    P.U = e->u();
    P.V = e->v();
    P.ldu = e->ldu;
    P.ldv = e->ldv;
    if (run[i]) {
      any = true;
      if (hipMemcpyAsync(e->dctl, static_cast<LbdCtl*>(e->hctl), sizeof(LbdCtl),
                         hipMemcpyHostToDevice, B.stream) != hipSuccess ||
          hipMemcpyAsync(e->darr, e->harr, static_cast<size_t>(CtlArrays::bytes(e->kmax)),
                         hipMemcpyHostToDevice, B.stream) != hipSuccess)
        return PK_ERR_HIP;
    } else if (hipMemsetAsync(&e->dctl->phase, 0, sizeof(int), B.stream) != hipSuccess) {
      return PK_ERR_HIP;  // phase 0 gates the problem off
    }
  }
  if (any) {
    if (hipMemcpyAsync(B.dprob, B.hprob.data(), sizeof(BProb) * np, hipMemcpyHostToDevice,
                       B.stream) != hipSuccess)
      return PK_ERR_HIP;
    auto enqueue = [&] {
      for (int J = Jmin; J <= Jmax; ++J) {
        if (B.prec == PK_F64)
          enqueue_batch_step<double>(B, J, elr, tries);
        else
          enqueue_batch_step<float>(B, J, elr, tries);
      }
    };
    hipGraphExec_t exec = nullptr;
    if (B.use_graph) {
      for (const auto& g : B.graphs)
        if (g.j0 == Jmin && g.j1 == Jmax && g.elr == elr && g.tries == tries) exec = g.exec;
      if (!exec && B.graphs.size() < kMaxBatchGraphs) {
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
          return PK_ERR_HIP;
        enqueue();
        const hipError_t ce = hipStreamEndCapture(B.stream, &graph);
        if (ce != hipSuccess || hipGetLastError() != hipSuccess) {
          if (graph) (void)hipGraphDestroy(graph);
          set_error("pk_batch_lanbpro: capture of the batched steps failed");
          return PK_ERR_HIP;
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
          set_error("pk_batch_lanbpro: hipGraphInstantiate: %s", hipGetErrorString(ie));
          return PK_ERR_HIP;
        }
        B.graphs.push_back(Batch::Graph{Jmin, Jmax, elr, tries, exec});
      }
    }
    if (exec) {
      if (hipGraphLaunch(exec, B.stream) != hipSuccess) return PK_ERR_HIP;
    } else {
      enqueue();
    }
    if (hipGetLastError() != hipSuccess) return PK_ERR_HIP;
    for (int i = 0; i < np; ++i)
      if (run[i] && (hipMemcpyAsync(static_cast<LbdCtl*>(B.eng[i]->hctl), B.eng[i]->dctl,
                                    sizeof(LbdCtl), hipMemcpyDeviceToHost, B.stream) !=
                         hipSuccess ||
                     hipMemcpyAsync(B.eng[i]->harr, B.eng[i]->darr,
                                    static_cast<size_t>(CtlArrays::bytes(B.eng[i]->kmax)),
                                    hipMemcpyDeviceToHost, B.stream) != hipSuccess))
        return PK_ERR_HIP;
  }
  if (hipStreamSynchronize(B.stream) != hipSuccess) return PK_ERR_HIP;
  // whatever the chain left (rare phases, further CGS passes) runs on each
  // problem's own engine, then its outputs
  host_parallel_for(np, [&](int i) {
    if (!active[i] || fin[i]) return;
    Engine* e = B.eng[i];
    if (run[i]) e->cnt.nsync++;
    int rc = engine_lanbpro_loop(e);
    if (rc == PK_OK) rc = engine_lanbpro_end(e, k0[i], &k[i], alpha[i], beta[i], &rnorm[i],
                                             doption + 4 * i, &ierr[i]);
    rcs[i] = rc;
  });
  for (int i = 0; i < np; ++i)
    if (rcs[i] != PK_OK) return rcs[i];
  return PK_OK;
}

Batch* as_batch(pk_batch_t h) {
  Batch* b = reinterpret_cast<Batch*>(h);
  if (!b || b->magic != Batch::kMagic) {
    set_error("not a live batch handle");
    return nullptr;
  }
  return b;
}

}  // namespace

int host_threads() {
  if (const char* v = std::getenv("PK_HOST_THREADS")) {
    const int t = std::atoi(v);
    if (t >= 1) return t;
  }
  const int hw = static_cast<int>(std::thread::hardware_concurrency());
  return std::max(1, std::min(hw, 16));
}

void host_parallel_for(int n, const std::function<void(int)>& f) {
  const int nt = std::min(n, host_threads());
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next{0};
  auto work = [&] {
    for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
  };
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  try {
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  } catch (const std::system_error&) {
    // fewer threads than asked for: the ones started and this thread share
    // the remaining indices
  }
  work();
  for (auto& th : pool) th.join();
}

}  // namespace pk

using namespace pk;

extern "C" {

// fn(ctx, i) for i in [0, n) on the host threads (the batched Fortran driver's
// per-problem restart work).
void pkx_parallel_for(int n, void (*fn)(void*, int), void* ctx) {
  host_parallel_for(n, [&](int i) { fn(ctx, i); });
}

int pk_batch_create(int nprob, const pk_lbd_t* engines, pk_batch_t* out) {
  Batch* b = nullptr;
  const int rc = batch_create(nprob, engines, &b);
  if (rc == PK_OK) *out = reinterpret_cast<pk_batch_t>(b);
  return rc;
}

int pk_batch_destroy(pk_batch_t h) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  return PK_OK;
}

int pk_batch_set_graph(pk_batch_t h, int on) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  b->use_graph = on != 0;
  return PK_OK;
}

int pk_batch_set_bsvd(pk_batch_t h, int on) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  b->use_bsvd = on != 0;
  return PK_OK;
}

int pk_batch_get_bsvd(pk_batch_t h, int* on) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  if (!on) return PK_ERR_ARG;
  *on = b->use_bsvd ? 1 : 0;
  return PK_OK;
}

int pk_batch_bsvd(pk_batch_t h, int nb, const int* n, const double* d, const double* e, int ld,
                  double* sigma, double* bottom) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  if (nb > 0 && (!n || !d || !e || !sigma || !bottom)) return PK_ERR_ARG;
  return bsvd_host(b->bsvd, b->stream, nb, n, d, e, ld, sigma, bottom);
}

// DGEMM_OVWR_LEFT (double/dgemm_ovwr.F:56-89) of problems sel[0..nsel) in
// one grouped launch: X_i(:, 1:ncols[q]) = X_i(:, 1:kdim[q]) op(B_q) into
// each engine's spare buffer, then the buffers swap, as pk_lbd_update does
// for one engine (and bit for bit the same products).
int pk_batch_update(pk_batch_t h, int nsel, const int* sel, char which, char transb,
                    const int* kdim, const int* ncols, const double* const* bm, const int* ldb) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  if (nsel < 0 || (nsel > 0 && (!sel || !kdim || !ncols || !bm || !ldb))) return PK_ERR_ARG;
  const bool isu = which == 'u' || which == 'U';
  const bool tb = transb == 't' || transb == 'T';
  // each problem at most once: one problem's staging buffers (hB, dB) and
  // spare basis hold one product at a time
  std::vector<char> seen(b->nprob, 0);
  for (int q = 0; q < nsel; ++q) {
    const int i = sel[q];
    if (i < 0 || i >= b->nprob) {
      set_error("pk_batch_update: problem %d outside the batch", i);
      return PK_ERR_ARG;
    }
    if (seen[i]) {
      set_error("pk_batch_update: problem %d selected twice", i);
      return PK_ERR_ARG;
    }
    seen[i] = 1;
  }
  // the grouped kernel stages B in LDS (kdim <= 316); past that, one update
  // per engine (the long-kdim GEMM)
  for (int q = 0; q < nsel; ++q)
    if (kdim[q] > 316) {
      for (int r = 0; r < nsel; ++r) {
        const int rc = pk_lbd_update(reinterpret_cast<pk_lbd_t>(b->eng[sel[r]]), which, transb,
                                     kdim[r], ncols[r], bm[r], ldb[r]);
        if (rc != PK_OK) return rc;
      }
      return PK_OK;
    }
  std::vector<GemmJob> jobs;
  jobs.reserve(nsel);
  for (int q = 0; q < nsel; ++q) {
    const int i = sel[q];
    Engine* e = b->eng[i];
    const int maxk = isu ? e->kmax + 1 : e->kmax;
    if (kdim[q] < 1 || ncols[q] < 1 || kdim[q] > maxk || ncols[q] > maxk ||
        ldb[q] < (tb ? ncols[q] : kdim[q]) || !bm[q]) {
      set_error("pk_batch_update: problem %d: kdim=%d ncols=%d ldb=%d outside [1, %d]", i,
                kdim[q], ncols[q], ldb[q], maxk);
      return PK_ERR_ARG;
    }
    void** buf = isu ? e->U : e->V;
    const int cur = isu ? e->ucur : e->vcur;
    const long long ld = isu ? e->ldu : e->ldv;
    if (!buf[1 - cur] && hipMalloc(&buf[1 - cur], e->tbytes * ld * maxk) != hipSuccess) {
      set_error("pk_batch_update: spare basis of problem %d", i);
      return PK_ERR_OOM;
    }
    // the engine's earlier work (a copy out of hB among it) is done first
    if (hipStreamSynchronize(e->stream) != hipSuccess) return PK_ERR_HIP;
    const double* B = bm[q];
    for (int j = 0; j < ncols[q]; ++j)
      for (int l = 0; l < kdim[q]; ++l)
        e->hB[static_cast<size_t>(j) * kdim[q] + l] =
            tb ? B[static_cast<size_t>(l) * ldb[q] + j] : B[static_cast<size_t>(j) * ldb[q] + l];
    if (hipMemcpyAsync(e->dB, e->hB, sizeof(double) * kdim[q] * ncols[q], hipMemcpyHostToDevice,
                       b->stream) != hipSuccess)
      return PK_ERR_HIP;
    jobs.push_back(GemmJob{buf[cur], buf[1 - cur], e->dB, ld, ld, isu ? e->m : e->n, kdim[q],
                           kdim[q], ncols[q]});
  }
  if (jobs.empty()) return PK_OK;
  const int rc = b->prec == PK_F64
                     ? launch_gemm_group<double>(b->stream, jobs.data(),
                                                 static_cast<int>(jobs.size()), &b->gitems,
                                                 &b->gcap)
                     : launch_gemm_group<float>(b->stream, jobs.data(),
                                                static_cast<int>(jobs.size()), &b->gitems,
                                                &b->gcap);
  if (rc != 0) {
    set_error("pk_batch_update: grouped GEMM launch failed");
    return PK_ERR_HIP;
  }
  for (int q = 0; q < nsel; ++q) {
    Engine* e = b->eng[sel[q]];
    int& cur = isu ? e->ucur : e->vcur;
    cur = 1 - cur;
  }
  return PK_OK;
}

int pk_batch_lanbpro(pk_batch_t h, const int* active, const int* k0, int* k,
                     double* const* alpha, double* const* beta, double* rnorm, double* doption,
                     const int* ioption, int* ierr) {
  Batch* b = as_batch(h);
  if (!b) return PK_ERR_HANDLE;
  if (!active || !k0 || !k || !alpha || !beta || !rnorm || !doption || !ioption || !ierr)
    return PK_ERR_ARG;
  return batch_lanbpro(*b, active, k0, k, alpha, beta, rnorm, doption, ioption, ierr);
}

}  // extern "C"
