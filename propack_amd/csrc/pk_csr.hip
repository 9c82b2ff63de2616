// Device-resident CSR operator: A and its explicitly built transpose, both
// in CSR with int32 row pointers / column indices, values in double or float.
//
// The transpose is built on the device by a stable radix sort of the
// nonzero indices keyed by column (rocPRIM), so every row of A^T lists its
// entries in increasing original-row order: the build is deterministic and
// the A^T SpMV sums in a fixed order.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "pk_internal.h"

namespace pk {

namespace {

// rowidx[p] = the row holding nonzero p: one thread per nonzero, a binary
// search for the last row i with rp[i] <= p (empty rows are passed over), so
// a row of 10^5 nonzeros costs no more than any other.
__global__ void k_row_index(int m, long long nnz, const int* __restrict__ rp,
                            int* __restrict__ rowidx) {
  for (long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz;
       p += static_cast<long long>(gridDim.x) * blockDim.x) {
    int lo = 0, hi = m;  // rp[lo] <= p < rp[hi]
    while (hi - lo > 1) {
      const int mid = lo + (hi - lo) / 2;
      if (rp[mid] <= p)
        lo = mid;
      else
        hi = mid;
    }
    rowidx[p] = lo;
  }
}

__global__ void k_iota(long long n, int* __restrict__ v) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    v[i] = static_cast<int>(i);
}

template <typename VT>
__global__ void k_gather_t(long long nnz, const int* __restrict__ perm,
                           const int* __restrict__ rowidx, const VT* __restrict__ va,
                           int* __restrict__ cit, VT* __restrict__ vat) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int p = perm[q];
    cit[q] = rowidx[p];
    vat[q] = va[p];
  }
}

// rpt[j] = first position q with key[q] >= j, for j in [0, n]; keys sorted.
__global__ void k_rowptr_from_keys(long long nnz, int n, const int* __restrict__ keys,
                                   int* __restrict__ rpt) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q <= nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int lo = q == 0 ? -1 : keys[q - 1];
    const int hi = q == nnz ? n : keys[q];
    for (int j = lo + 1; j <= hi; ++j) rpt[j] = static_cast<int>(q);
  }
}

int bits_for(int n) {
  int b = 1;
  while ((1LL << b) < n) ++b;
  return b;
}

// Lanes per row for the sub-wave SpMV: the power of two nearest half the
// mean row length (two nonzeros per lane; measured faster on MI355X than one
// per lane for the uniform and the batched workloads), clamped to [2, 64],
// times 2^shift (clamped again).
int choose_lpr(long long nnz, int rows, int shift = 0) {
  const double avg = rows > 0 ? static_cast<double>(nnz) / rows : 1.0;
  int l = 2;
  while (l < 64 && l * 3.0 < avg) l *= 2;
  for (; shift > 0 && l < 64; --shift) l *= 2;
  for (; shift < 0 && l > 2; ++shift) l /= 2;
  return l;
}

// Workgroup descriptors of the row-block SpMV: greedy blocks of at most
// max_rows rows and cap nonzeros (kSpmvRows and kSpmvCap for the row-block
// layout, larger for the column-sorted one); a row longer than kSpmvCap
// becomes ceil(len / kLongChunk) chunk descriptors plus an entry of the
// long-row table {row, first chunk, chunks, 0}.
struct Partition {
  std::vector<int4> desc, ldesc;
  int nchunks = 0;
};

Partition partition(const std::vector<int>& rp, int rows, int cap = kSpmvCap,
                    int max_rows = kSpmvRows) {
  Partition P;
  P.desc.reserve(static_cast<size_t>(rows) / 64 + 2);
  int r = 0;
  while (r < rows) {
    const long long len0 = static_cast<long long>(rp[r + 1]) - rp[r];
    if (len0 > kSpmvCap) {
      const int nch = static_cast<int>((len0 + kLongChunk - 1) / kLongChunk);
      for (int c = 0; c < nch; ++c) {
        const int p0 = rp[r] + c * kLongChunk;
        const int p1 = static_cast<int>(std::min<long long>(p0 + static_cast<long long>(kLongChunk), rp[r + 1]));
        P.desc.push_back(make_int4(r, -(P.nchunks + c + 1), p0, p1));
      }
      P.ldesc.push_back(make_int4(r, P.nchunks, nch, 0));
      P.nchunks += nch;
      ++r;
      continue;
    }
    const int start = r;
    long long nz = 0;
    while (r < rows && r - start < max_rows) {
      const long long len = static_cast<long long>(rp[r + 1]) - rp[r];
      if (len > kSpmvCap || nz + len > cap) break;
      nz += len;
      ++r;
    }
    P.desc.push_back(make_int4(start, r, -1, 0));
  }
  return P;
}

bool upload(const Partition& P, CsrOp::Part& d) {
  d.ndesc = static_cast<int>(P.desc.size());
  d.nlong = static_cast<int>(P.ldesc.size());
  const size_t nd = std::max<size_t>(P.desc.size(), 1), nl = std::max<size_t>(P.ldesc.size(), 1);
  return hipMalloc(&d.desc, sizeof(int4) * nd) == hipSuccess &&
         hipMalloc(&d.ldesc, sizeof(int4) * nl) == hipSuccess &&
         hipMalloc(&d.lsum, sizeof(double) * std::max(P.nchunks, 1)) == hipSuccess &&
         (P.desc.empty() || hipMemcpy(d.desc, P.desc.data(), sizeof(int4) * P.desc.size(),
                                      hipMemcpyHostToDevice) == hipSuccess) &&
         (P.ldesc.empty() || hipMemcpy(d.ldesc, P.ldesc.data(), sizeof(int4) * P.ldesc.size(),
                                       hipMemcpyHostToDevice) == hipSuccess);
}

}  // namespace

// A^T of the CSR (m x n, nnz) as CSR (n x m): stable radix sort of the
// nonzero positions keyed by column.
template <typename VT>
static int build_transpose(int m, int n, long long nnz, const int* rp, const int* ci,
                           const void* va, int** rpt_out, int** cit_out, void** vat_out,
                           hipStream_t s) {
  int *rowidx = nullptr, *keys_in = nullptr, *keys_out = nullptr, *perm_in = nullptr,
      *perm_out = nullptr, *rpt = nullptr, *cit = nullptr;
  void *vat = nullptr, *tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = PK_OK;
  const size_t ib = sizeof(int) * static_cast<size_t>(std::max<long long>(nnz, 1));
#define PKT(x)                 \
  if ((x) != hipSuccess) {     \
    rc = PK_ERR_HIP;           \
    goto out;                  \
  }
  PKT(hipMalloc(&rowidx, ib));
  PKT(hipMalloc(&keys_in, ib));
  PKT(hipMalloc(&keys_out, ib));
  PKT(hipMalloc(&perm_in, ib));
  PKT(hipMalloc(&perm_out, ib));
  PKT(hipMalloc(&rpt, sizeof(int) * (static_cast<size_t>(n) + 1)));
  PKT(hipMalloc(&cit, ib));
  PKT(hipMalloc(&vat, sizeof(VT) * static_cast<size_t>(std::max<long long>(nnz, 1))));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_row_index, dim3(2048), dim3(256), 0, s, m, nnz, rp, rowidx);
    PKT(hipMemcpyAsync(keys_in, ci, sizeof(int) * nnz, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_iota, dim3(2048), dim3(256), 0, s, nnz, perm_in);
    PKT(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, keys_out, perm_in, perm_out,
                                  static_cast<size_t>(nnz), 0, bits_for(n), s));
    PKT(hipMalloc(&tmp, tmp_bytes));
    PKT(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, perm_in, perm_out,
                                  static_cast<size_t>(nnz), 0, bits_for(n), s));
    hipLaunchKernelGGL((k_gather_t<VT>), dim3(2048), dim3(256), 0, s, nnz, perm_out, rowidx,
                       static_cast<const VT*>(va), cit, static_cast<VT*>(vat));
  }
  hipLaunchKernelGGL(k_rowptr_from_keys, dim3(2048), dim3(256), 0, s, nnz, n, keys_out, rpt);
  PKT(hipGetLastError());
  PKT(hipStreamSynchronize(s));
out:
#undef PKT
  (void)hipFree(rowidx);
  (void)hipFree(keys_in);
  (void)hipFree(keys_out);
  (void)hipFree(perm_in);
  (void)hipFree(perm_out);
  (void)hipFree(tmp);
  if (rc != PK_OK) {
    (void)hipFree(rpt);
    (void)hipFree(cit);
    (void)hipFree(vat);
    rpt = cit = nullptr;
    vat = nullptr;
  }
  *rpt_out = rpt;
  *cit_out = cit;
  *vat_out = vat;
  return rc;
}

namespace {

__global__ void k_panel_key(long long nnz, const int* __restrict__ ci, int width,
                            int* __restrict__ key) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz;
       q += static_cast<long long>(gridDim.x) * blockDim.x)
    key[q] = ci[q] / width;
}

template <typename VT>
__global__ void k_gather_panel(long long cnt, const int* __restrict__ perm,
                               const int* __restrict__ rowidx, const int* __restrict__ ci,
                               const VT* __restrict__ va, int* __restrict__ ridx,
                               int* __restrict__ cip, VT* __restrict__ vap) {
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < cnt;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int p = perm[q];
    ridx[q] = rowidx[p];
    cip[q] = ci[p];
    vap[q] = va[p];
  }
}

// Column panels of an operator direction (rows x cols, nnz): the smallest
// power of two that cuts the gathered vector into slices of at most
// kPanelSliceBytes (2 MiB, half of one XCD's L2: with the index and value
// streams read non-temporally the slice stays in L2, so the random gathers
// hit L2 instead of going to the Infinity Cache; a 4 MiB slice measured 2x
// the algorithmic fabric traffic), up to 8; one panel when rows are short
// (mean below 8 nonzeros: the FP64 running row sums would cost more traffic
// than the gathers save).  PK_SPMV_PANELS ("p" for both directions or
// "pA,pAt") overrides.
constexpr double kPanelSliceBytes = 2.0 * 1024 * 1024;

int panels_for(bool transposed, int rows, int cols, long long nnz, size_t vbytes) {
  int want = 1;
  if (const char* e = std::getenv("PK_SPMV_PANELS"); e && *e) {
    int a = 1, b = 1;
    const int got = std::sscanf(e, "%d,%d", &a, &b);
    if (got >= 1) want = got == 2 && transposed ? b : a;
  } else if (rows > 0 && static_cast<double>(nnz) / rows >= 8.0) {
    const double xbytes = static_cast<double>(cols) * static_cast<double>(vbytes);
    while (want < 8 && xbytes / want > kPanelSliceBytes) want *= 2;
  }
  int q = 1;
  while (q * 2 <= want && q * 2 <= CsrOp::kMaxPanels) q *= 2;
  return q;
}

}  // namespace

void csr_set_lanes(CsrOp::Dir& D, int shift) {
  D.lane_shift = shift;
  for (int p = 0; p < D.np; ++p) D.pan[p].g = choose_lpr(D.pan[p].nnz, D.rows, shift);
}

namespace {

int pow2_floor(int p) {
  int q = 1;
  while (q * 2 <= p && q * 2 <= CsrOp::kMaxPanels) q *= 2;
  return q;
}

bool finish_panel(CsrOp::Panel& P, int rows) {
  std::vector<int> hrp(static_cast<size_t>(rows) + 1);
  if (hipMemcpy(hrp.data(), P.rp, sizeof(int) * (static_cast<size_t>(rows) + 1),
                hipMemcpyDeviceToHost) != hipSuccess)
    return false;
  P.g = choose_lpr(P.nnz, rows);
  return upload(partition(hrp, rows), P.part);
}

// Column-sorted copy of a panel, one workgroup per segment [sb, se) (a
// descriptor's nonzeros): perm holds the panel positions in sorted order,
// sci the sorted columns.  pbits > 0: sci[q] becomes the packed word
// (column - col0) << pbits | offset (offset bits past pbits dropped: only
// long-row chunks have such offsets, and their offsets are not used);
// otherwise the offset goes to spos.
template <typename VT>
__global__ void k_sorted_finish(const int* __restrict__ sb, const int* __restrict__ se,
                                const int* __restrict__ perm, const VT* __restrict__ va,
                                VT* __restrict__ sva, int* __restrict__ sci,
                                unsigned short* __restrict__ spos, int col0, int pbits) {
  const int b = sb[blockIdx.x], e = se[blockIdx.x];
  for (int q = b + static_cast<int>(threadIdx.x); q < e; q += static_cast<int>(blockDim.x)) {
    const int p = perm[q];
    sva[q] = va[p];
    if (pbits > 0) {
      const unsigned off = static_cast<unsigned>(sci[q] - col0);
      const unsigned pos = static_cast<unsigned>(p - b) & ((1u << pbits) - 1u);
      sci[q] = static_cast<int>((off << pbits) | pos);
    } else {
      spos[q] = static_cast<unsigned short>(p - b);
    }
  }
}

void free_sorted(CsrOp::Panel& P) {
  (void)hipFree(P.sci);
  (void)hipFree(P.sva);
  (void)hipFree(P.spos);
  (void)hipFree(P.spart.desc);
  (void)hipFree(P.spart.ldesc);
  (void)hipFree(P.spart.lsum);
  P.sci = nullptr;
  P.sva = nullptr;
  P.spos = nullptr;
  P.scol0 = P.spbits = 0;
  P.spart = CsrOp::Part{};
}

// Sorted layout of one panel (rows x cols): the descriptors of partition()
// with blocks of at most cap nonzeros and kSortRows rows; inside each
// descriptor's range a stable segmented radix sort of the positions keyed
// by column (rocPRIM), so ties keep their row-major order.
// The offsets of a block (< cap) and the columns of a panel relative to
// its first column (< width) share one 32-bit word when their bits fit.
template <typename VT>
int build_sorted_panel(CsrOp::Panel& P, int rows, int cols, int col0, int width, int cap,
                       int max_slots, hipStream_t s) {
  std::vector<int> hrp(static_cast<size_t>(rows) + 1);
  if (hipMemcpy(hrp.data(), P.rp, sizeof(int) * hrp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return PK_ERR_HIP;
  const Partition part = partition(hrp, rows, cap, kSortRows);
  if (static_cast<int>(part.desc.size() + part.ldesc.size()) > max_slots) return PK_ERR_NOTSUP;
  const int nseg = static_cast<int>(part.desc.size());
  std::vector<int> sb(std::max(nseg, 1)), se(std::max(nseg, 1));
  for (int i = 0; i < nseg; ++i) {
    const int4 d = part.desc[i];
    sb[i] = d.y < 0 ? d.z : hrp[d.x];
    se[i] = d.y < 0 ? d.w : hrp[d.y];
  }
  const long long nnz = P.nnz;
  const size_t nz = static_cast<size_t>(std::max<long long>(nnz, 1));
  int *dsb = nullptr, *dse = nullptr, *perm_in = nullptr, *perm_out = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = PK_OK;
#define PKT(x)             \
  if ((x) != hipSuccess) { \
    rc = PK_ERR_HIP;       \
    goto out;              \
  }
  {
    // unpacked: sci holds global columns (base 0, no shift)
    const int pbits = bits_for(cap), wbits = bits_for(std::max(width, 2));
    P.spbits = pbits + wbits <= 32 ? pbits : 0;
    P.scol0 = P.spbits > 0 ? col0 : 0;
  }
  PKT(hipMalloc(&P.sci, sizeof(int) * nz));
  PKT(hipMalloc(&P.sva, sizeof(VT) * nz));
  if (P.spbits == 0) PKT(hipMalloc(&P.spos, sizeof(unsigned short) * nz));
  PKT(hipMalloc(&dsb, sizeof(int) * sb.size()));
  PKT(hipMalloc(&dse, sizeof(int) * se.size()));
  PKT(hipMalloc(&perm_in, sizeof(int) * nz));
  PKT(hipMalloc(&perm_out, sizeof(int) * nz));
  PKT(hipMemcpyAsync(dsb, sb.data(), sizeof(int) * sb.size(), hipMemcpyHostToDevice, s));
  PKT(hipMemcpyAsync(dse, se.data(), sizeof(int) * se.size(), hipMemcpyHostToDevice, s));
  if (nnz > 0 && nseg > 0) {
    hipLaunchKernelGGL(k_iota, dim3(2048), dim3(256), 0, s, nnz, perm_in);
    PKT(rocprim::segmented_radix_sort_pairs(nullptr, tmp_bytes, P.ci, P.sci, perm_in, perm_out,
                                            static_cast<unsigned>(nnz),
                                            static_cast<unsigned>(nseg), dsb, dse, 0,
                                            bits_for(cols), s));
    PKT(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    PKT(rocprim::segmented_radix_sort_pairs(tmp, tmp_bytes, P.ci, P.sci, perm_in, perm_out,
                                            static_cast<unsigned>(nnz),
                                            static_cast<unsigned>(nseg), dsb, dse, 0,
                                            bits_for(cols), s));
    hipLaunchKernelGGL((k_sorted_finish<VT>), dim3(nseg), dim3(256), 0, s, dsb, dse, perm_out,
                       static_cast<const VT*>(P.va), static_cast<VT*>(P.sva), P.sci, P.spos,
                       P.scol0, P.spbits);
  }
  PKT(hipGetLastError());
  PKT(hipStreamSynchronize(s));
  if (!upload(part, P.spart)) rc = PK_ERR_HIP;
out:
#undef PKT
  (void)hipFree(dsb);
  (void)hipFree(dse);
  (void)hipFree(perm_in);
  (void)hipFree(perm_out);
  (void)hipFree(tmp);
  if (rc != PK_OK) free_sorted(P);
  return rc;
}

}  // namespace

namespace {

void free_concurrent(CsrOp::Dir& D) {
  (void)hipFree(D.ctable);
  (void)hipFree(D.cmap);
  (void)hipFree(D.ypart);
  D.ctable = nullptr;
  D.cmap = nullptr;
  D.ypart = nullptr;
  D.cmap_n = 0;
  D.concurrent = 0;
}

// The concurrent launch's workgroup order.  Workgroups are dispatched to the
// GPU's eight XCDs in turn (workgroup w to XCD w mod 8), each XCD with its
// own L2: XCD x is given whole panels -- panel x / (8 / np) with every
// (8 / np)-th of its blocks when np <= 8, panels x, x + 8, ... one after the
// other when np = 16 -- so that each L2 holds the slice of x of one panel at
// a time.  The order decides only where blocks run, never what they sum.
int build_concurrent(CsrOp::Dir& D) {
  constexpr int kX = 8;
  std::vector<std::vector<int2>> lists(kX);
  for (int x = 0; x < kX; ++x) {
    if (D.np <= kX) {
      const int g = kX / D.np, p = x / g, j = x % g;
      for (int i = j; i < D.pan[p].spart.ndesc; i += g) lists[x].push_back(int2{p, i});
    } else {
      for (int p = x; p < D.np; p += kX)
        for (int i = 0; i < D.pan[p].spart.ndesc; ++i) lists[x].push_back(int2{p, i});
    }
  }
  size_t w = 0;
  for (const auto& l : lists) w = std::max(w, l.size());
  std::vector<int2> map(w * kX, int2{-1, -1});
  for (size_t k = 0; k < w; ++k)
    for (int x = 0; x < kX; ++x)
      if (k < lists[x].size()) map[k * kX + x] = lists[x][k];
  SortPanels t{};
  for (int p = 0; p < D.np; ++p) {
    const CsrOp::Panel& P = D.pan[p];
    t.desc[p] = P.spart.desc;
    t.rp[p] = P.rp;
    t.sci[p] = P.sci;
    t.sva[p] = P.sva;
    t.spos[p] = P.spos;
    t.col0[p] = P.scol0;
    t.pbits[p] = P.spbits;
  }
  const size_t ny = static_cast<size_t>(D.np) * std::max(D.rows, 1);
  if (hipMalloc(&D.ctable, sizeof(SortPanels)) != hipSuccess ||
      hipMalloc(&D.cmap, sizeof(int2) * std::max<size_t>(map.size(), 1)) != hipSuccess ||
      hipMalloc(&D.ypart, sizeof(double) * ny) != hipSuccess ||
      hipMemcpy(D.ctable, &t, sizeof t, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(D.cmap, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice) !=
          hipSuccess) {
    free_concurrent(D);
    return PK_ERR_HIP;
  }
  D.cmap_n = static_cast<int>(map.size());
  D.concurrent = 1;
  return PK_OK;
}

}  // namespace

int csr_sort_code(const CsrOp::Dir& D) {
  if (D.rowwise) return -1;
  return D.concurrent ? -D.sort_cap : D.sort_cap;
}

bool csr_concurrent_ok(const CsrOp::Dir& D) {
  if (!D.concurrent || D.np <= 1 || D.sort_cap <= 0 || D.cmap_n <= 0) return false;
  for (int p = 1; p < D.np; ++p)
    if (D.pan[p].g != D.pan[0].g) return false;
  return true;
}

long long csr_ywork(const CsrOp* op) {
  long long w = std::max<long long>(std::max(op->m, op->n), 1);
  for (const CsrOp::Dir* D : {&op->da, &op->dt})
    if (D->concurrent) w = std::max(w, static_cast<long long>(D->np) * D->rows);
  return w;
}

int csr_set_sort(CsrOp* op, CsrOp::Dir& D, int cap) {
  ++op->layout_epoch;
  const bool conc = cap <= -kSpmvCap;
  if (conc) cap = -cap;
  if (cap != 0 && cap != -1 && (cap < kSpmvCap || cap > kSortCapMax)) return PK_ERR_ARG;
  free_concurrent(D);
  for (int p = 0; p < D.np; ++p) free_sorted(D.pan[p]);
  D.sort_cap = 0;
  D.rowwise = cap == -1;
  if (cap <= 0) return PK_OK;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PK_ERR_HIP;
  int rc = PK_OK;
  for (int p = 0; p < D.np && rc == PK_OK; ++p) {
    // panel p holds columns [p * width, min((p + 1) * width, cols))
    const int col0 = D.np > 1 ? p * D.width : 0;
    const int width = D.np > 1 ? std::min(D.width, D.cols - col0) : D.cols;
    rc = op->val_prec == PK_F64
             ? build_sorted_panel<double>(D.pan[p], D.rows, D.cols, col0, width, cap, D.slots, s)
             : build_sorted_panel<float>(D.pan[p], D.rows, D.cols, col0, width, cap, D.slots, s);
  }
  (void)hipStreamDestroy(s);
  if (rc != PK_OK) {
    for (int p = 0; p < D.np; ++p) free_sorted(D.pan[p]);
    return rc;
  }
  D.sort_cap = cap;
  if (conc && D.np > 1) {
    bool chunks = false;
    for (int p = 0; p < D.np; ++p) chunks = chunks || D.pan[p].spart.nlong > 0;
    if (!chunks) {
      rc = build_concurrent(D);
      if (rc != PK_OK) {
        for (int p = 0; p < D.np; ++p) free_sorted(D.pan[p]);
        D.sort_cap = 0;
        return rc;
      }
    }
  }
  return PK_OK;
}

// Lays one direction out as np column panels.  Takes ownership of the
// plain CSR arrays (rp, ci, va): kept as the single panel when np = 1,
// freed after the panels are built otherwise.
template <typename VT>
static int build_dir(CsrOp::Dir& D, int rows, int cols, long long nnz, int np, int* rp,
                     int* ci, void* va, hipStream_t s) {
  D.rows = rows;
  D.cols = cols;
  if (np <= 1 || cols < np) {
    D.np = 1;
    D.width = cols;
    D.pan[0].rp = rp;
    D.pan[0].ci = ci;
    D.pan[0].va = va;
    D.pan[0].nnz = nnz;
    if (!finish_panel(D.pan[0], rows)) return PK_ERR_HIP;
    D.slots = D.pan[0].part.slots();
    return PK_OK;
  }
  D.np = np;
  D.width = (cols + np - 1) / np;
  int *rowidx = nullptr, *keys_in = nullptr, *keys_out = nullptr, *perm_in = nullptr,
      *perm_out = nullptr, *poff_d = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = PK_OK;
  std::vector<int> poff(static_cast<size_t>(np) + 1, 0);
  const size_t ib = sizeof(int) * static_cast<size_t>(std::max<long long>(nnz, 1));
#define PKT(x)                 \
  if ((x) != hipSuccess) {     \
    rc = PK_ERR_HIP;           \
    goto out;                  \
  }
  PKT(hipMalloc(&rowidx, ib));
  PKT(hipMalloc(&keys_in, ib));
  PKT(hipMalloc(&keys_out, ib));
  PKT(hipMalloc(&perm_in, ib));
  PKT(hipMalloc(&perm_out, ib));
  PKT(hipMalloc(&poff_d, sizeof(int) * (static_cast<size_t>(np) + 1)));
  PKT(hipMalloc(&D.yacc, sizeof(double) * std::max(rows, 1)));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_row_index, dim3(2048), dim3(256), 0, s, rows, nnz, rp, rowidx);
    hipLaunchKernelGGL(k_panel_key, dim3(2048), dim3(256), 0, s, nnz, ci, D.width, keys_in);
    hipLaunchKernelGGL(k_iota, dim3(2048), dim3(256), 0, s, nnz, perm_in);
    PKT(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, keys_out, perm_in, perm_out,
                                  static_cast<size_t>(nnz), 0, bits_for(np), s));
    PKT(hipMalloc(&tmp, tmp_bytes));
    PKT(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, perm_in, perm_out,
                                  static_cast<size_t>(nnz), 0, bits_for(np), s));
  }
  hipLaunchKernelGGL(k_rowptr_from_keys, dim3(64), dim3(256), 0, s, nnz, np, keys_out, poff_d);
  PKT(hipGetLastError());
  PKT(hipMemcpyAsync(poff.data(), poff_d, sizeof(int) * (np + 1), hipMemcpyDeviceToHost, s));
  PKT(hipStreamSynchronize(s));
  for (int p = 0; p < np; ++p) {
    CsrOp::Panel& P = D.pan[p];
    const long long cnt = static_cast<long long>(poff[p + 1]) - poff[p];
    const size_t cb = static_cast<size_t>(std::max<long long>(cnt, 1));
    P.nnz = cnt;
    PKT(hipMalloc(&P.rp, sizeof(int) * (static_cast<size_t>(rows) + 1)));
    PKT(hipMalloc(&P.ci, sizeof(int) * cb));
    PKT(hipMalloc(&P.va, sizeof(VT) * cb));
    if (cnt > 0)
      hipLaunchKernelGGL((k_gather_panel<VT>), dim3(2048), dim3(256), 0, s, cnt,
                         perm_out + poff[p], rowidx, ci, static_cast<const VT*>(va), keys_in,
                         P.ci, static_cast<VT*>(P.va));
    // entries of a panel stay in row-major order (stable sort): row
    // pointers from the gathered row indices
    hipLaunchKernelGGL(k_rowptr_from_keys, dim3(2048), dim3(256), 0, s, cnt, rows, keys_in,
                       P.rp);
    PKT(hipGetLastError());
    PKT(hipStreamSynchronize(s));
    if (!finish_panel(P, rows)) {
      rc = PK_ERR_HIP;
      goto out;
    }
    D.slots = std::max(D.slots, P.part.slots());
  }
out:
#undef PKT
  (void)hipFree(rowidx);
  (void)hipFree(keys_in);
  (void)hipFree(keys_out);
  (void)hipFree(perm_in);
  (void)hipFree(perm_out);
  (void)hipFree(poff_d);
  (void)hipFree(tmp);
  (void)hipFree(rp);
  (void)hipFree(ci);
  (void)hipFree(va);
  return rc;
}

static void free_dir(CsrOp::Dir& D) {
  for (int p = 0; p < CsrOp::kMaxPanels; ++p) {
    CsrOp::Panel& P = D.pan[p];
    free_sorted(P);
    (void)hipFree(P.rp);
    (void)hipFree(P.ci);
    (void)hipFree(P.va);
    (void)hipFree(P.part.desc);
    (void)hipFree(P.part.ldesc);
    (void)hipFree(P.part.lsum);
    P = CsrOp::Panel{};
  }
  (void)hipFree(D.yacc);
  D.yacc = nullptr;
  free_concurrent(D);
}

int csr_create(int m, int n, long long nnz, const int* rowptr, const int* colind,
               const void* val, int val_prec, int where, CsrOp** out, int panels_a,
               int panels_t) {
  if (!out || m < 0 || n < 0 || nnz < 0 || nnz >= (1LL << 31)) return PK_ERR_ARG;
  if (val_prec != PK_F64 && val_prec != PK_F32 && !prec_is_complex(val_prec)) return PK_ERR_ARG;
  if (where != PK_HOST && where != PK_DEVICE) return PK_ERR_ARG;
  if (!rowptr || (nnz > 0 && (!colind || !val))) return PK_ERR_ARG;
  if (where == PK_HOST) {
    if (rowptr[0] != 0 || rowptr[m] != nnz) return PK_ERR_ARG;
    for (int i = 0; i < m; ++i)
      if (rowptr[i + 1] < rowptr[i]) return PK_ERR_ARG;
    for (long long p = 0; p < nnz; ++p)
      if (colind[p] < 0 || colind[p] >= n) return PK_ERR_ARG;
  }
  CsrOp* op = new CsrOp();
  op->m = m;
  op->n = n;
  op->nnz = nnz;
  op->val_prec = val_prec;
  op->val_bytes = prec_bytes(val_prec);
  op->magic = CsrOp::kMagic;
  const hipMemcpyKind kind = where == PK_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  hipStream_t s = nullptr;
  int rc = PK_OK;
  const size_t nz = static_cast<size_t>(std::max<long long>(nnz, 1));
  int *rp = nullptr, *ci = nullptr, *rpt = nullptr, *cit = nullptr;
  void *va = nullptr, *vat = nullptr;
  if (hipStreamCreate(&s) != hipSuccess) {
    delete op;
    return PK_ERR_HIP;
  }
  if (hipMalloc(&rp, sizeof(int) * (static_cast<size_t>(m) + 1)) != hipSuccess ||
      hipMalloc(&ci, sizeof(int) * nz) != hipSuccess ||
      hipMalloc(&va, op->val_bytes * nz) != hipSuccess ||
      hipMemcpyAsync(rp, rowptr, sizeof(int) * (static_cast<size_t>(m) + 1), kind, s) !=
          hipSuccess ||
      (nnz > 0 && hipMemcpyAsync(ci, colind, sizeof(int) * nnz, kind, s) != hipSuccess) ||
      (nnz > 0 && hipMemcpyAsync(va, val, op->val_bytes * nnz, kind, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess) {
    rc = PK_ERR_HIP;
  }
  if (rc == PK_OK && prec_is_complex(val_prec)) {
    // complex values: A and the explicit A^T as one plain CSR panel each
    // (the transpose keeps the values; A^H conjugates them as it reads)
    rc = val_prec == PK_C128
             ? build_transpose<double2>(m, n, nnz, rp, ci, va, &rpt, &cit, &vat, s)
             : build_transpose<float2>(m, n, nnz, rp, ci, va, &rpt, &cit, &vat, s);
    if (rc == PK_OK) {
      // column panels as for real values (each panel's slice of x sized for
      // one XCD's L2), at most kMaxZPanels; build_dir owns the arrays
      auto cap_np = [](int q) { return q > kMaxZPanels ? kMaxZPanels : q; };
      const int pa = cap_np(panels_a > 0 ? pow2_floor(panels_a)
                                         : panels_for(false, m, n, nnz, op->val_bytes));
      const int pt = cap_np(panels_t > 0 ? pow2_floor(panels_t)
                                         : panels_for(true, n, m, nnz, op->val_bytes));
      rc = val_prec == PK_C128 ? build_dir<double2>(op->da, m, n, nnz, pa, rp, ci, va, s)
                               : build_dir<float2>(op->da, m, n, nnz, pa, rp, ci, va, s);
      rp = ci = nullptr;
      va = nullptr;
      const int rc2 = val_prec == PK_C128
                          ? build_dir<double2>(op->dt, n, m, nnz, pt, rpt, cit, vat, s)
                          : build_dir<float2>(op->dt, n, m, nnz, pt, rpt, cit, vat, s);
      rpt = cit = nullptr;
      vat = nullptr;
      if (rc == PK_OK) rc = rc2;
      if (rc == PK_OK) {
        // one lanes-per-row count per direction (the kernel's template
        // parameter), from the mean row length of a panel
        for (CsrOp::Dir* D : {&op->da, &op->dt})
          for (int p = 0; p < D->np; ++p)
            D->pan[p].g = zspmv_lanes(nnz / D->np, D->rows);
      }
    }
  } else if (rc == PK_OK) {
    rc = val_prec == PK_F64
             ? build_transpose<double>(m, n, nnz, rp, ci, va, &rpt, &cit, &vat, s)
             : build_transpose<float>(m, n, nnz, rp, ci, va, &rpt, &cit, &vat, s);
  }
  if (rc == PK_OK && !prec_is_complex(val_prec)) {
    // build_dir owns the arrays from here on
    const int pa = panels_a > 0 ? pow2_floor(panels_a) : panels_for(false, m, n, nnz, op->val_bytes);
    const int pt = panels_t > 0 ? pow2_floor(panels_t) : panels_for(true, n, m, nnz, op->val_bytes);
    rc = val_prec == PK_F64 ? build_dir<double>(op->da, m, n, nnz, pa, rp, ci, va, s)
                            : build_dir<float>(op->da, m, n, nnz, pa, rp, ci, va, s);
    rp = ci = nullptr;
    va = nullptr;
    const int rc2 = val_prec == PK_F64
                        ? build_dir<double>(op->dt, n, m, nnz, pt, rpt, cit, vat, s)
                        : build_dir<float>(op->dt, n, m, nnz, pt, rpt, cit, vat, s);
    rpt = cit = nullptr;
    vat = nullptr;
    if (rc == PK_OK) rc = rc2;
  }
  (void)hipFree(rp);
  (void)hipFree(ci);
  (void)hipFree(va);
  (void)hipFree(rpt);
  (void)hipFree(cit);
  (void)hipFree(vat);
  (void)hipStreamDestroy(s);
  if (rc != PK_OK) {
    csr_destroy(op);
    return rc;
  }
  op->nrb = op->da.slots;
  op->nrbt = op->dt.slots;
  *out = op;
  return PK_OK;
}

void csr_destroy(CsrOp* op) {
  if (!op) return;
  free_dir(op->da);
  free_dir(op->dt);
  op->magic = 0;
  delete op;
}

// y = op(A) x - coef * yprev with fused partial sums; x, y, yprev device
// vectors of basis type T.  One launch per column panel (the running row
// sums stay in FP64 between them).  Returns the partial count or < 0.
template <typename T>
int csr_apply(const CsrOp* op, hipStream_t s, char trans, const T* x, T* y, const T* yprev,
              Scal coef, double* part, Gate gate, double* ywork, double* lwork,
              long long ywork_len) {
  const bool tr = (trans == 't' || trans == 'T');
  const CsrOp::Dir& D = tr ? op->dt : op->da;
  if (csr_concurrent_ok(D)) {
    const long long need = static_cast<long long>(D.np) * D.rows;
    double* yp = ywork ? (std::max<long long>(ywork_len, D.rows) >= need ? ywork : nullptr)
                       : D.ypart;
    if (yp) {
      const int rc =
          op->val_prec == PK_F64
              ? launch_spmv_sorted_cp<double, T>(s, D.ctable, D.cmap, D.cmap_n, D.sort_cap,
                                                 D.pan[0].g, D.rows, x, yp, gate)
              : launch_spmv_sorted_cp<float, T>(s, D.ctable, D.cmap, D.cmap_n, D.sort_cap,
                                                D.pan[0].g, D.rows, x, yp, gate);
      if (rc < 0) return rc;
      return launch_spmv_combine<T>(s, D.rows, D.np, yp, yprev, coef, y, part, gate);
    }
  }
  double* yacc = ywork ? ywork : D.yacc;
  int nb = -1;
  for (int p = 0; p < D.np; ++p) {
    const CsrOp::Panel& P = D.pan[p];
    const double* yin = p == 0 ? nullptr : yacc;
    const int last = p == D.np - 1;
    double* yout = last ? nullptr : yacc;
    if (D.sort_cap > 0) {
      const CsrOp::Part& Q = P.spart;
      double* lsum = lwork ? lwork : Q.lsum;
      if (op->val_prec == PK_F64)
        nb = launch_spmv_sorted<double, T>(s, Q.ndesc, Q.desc, Q.nlong, Q.ldesc, lsum, P.rp,
                                           P.sci, static_cast<const double*>(P.sva), P.spos,
                                           P.scol0, P.spbits, D.sort_cap, x, y, yprev, coef,
                                           part, P.g, yin, yout, last, gate);
      else
        nb = launch_spmv_sorted<float, T>(s, Q.ndesc, Q.desc, Q.nlong, Q.ldesc, lsum, P.rp,
                                          P.sci, static_cast<const float*>(P.sva), P.spos,
                                          P.scol0, P.spbits, D.sort_cap, x, y, yprev, coef,
                                          part, P.g, yin, yout, last, gate);
      if (nb < 0) return nb;
      continue;
    }
    const CsrOp::Part& Q = P.part;
    double* lsum = lwork ? lwork : Q.lsum;
    if (op->val_prec == PK_F64)
      nb = launch_spmv_rb<double, T>(s, Q.ndesc, Q.desc, Q.nlong, Q.ldesc, lsum, P.rp, P.ci,
                                     static_cast<const double*>(P.va), x, y, yprev, coef, part,
                                     P.g, yin, yout, last, gate, D.rowwise);
    else
      nb = launch_spmv_rb<float, T>(s, Q.ndesc, Q.desc, Q.nlong, Q.ldesc, lsum, P.rp, P.ci,
                                    static_cast<const float*>(P.va), x, y, yprev, coef, part,
                                    P.g, yin, yout, last, gate, D.rowwise);
    if (nb < 0) return nb;
  }
  return nb;
}

template int csr_apply<double>(const CsrOp*, hipStream_t, char, const double*, double*,
                               const double*, Scal, double*, Gate, double*, double*, long long);
template int csr_apply<float>(const CsrOp*, hipStream_t, char, const float*, float*,
                              const float*, Scal, double*, Gate, double*, double*, long long);

}  // namespace pk

// ------------------------------------------------------------------ C ABI --
namespace pk {
void set_error(const char* fmt, ...);
CsrOp* as_csr(pk_csr_t h);
}  // namespace pk

using namespace pk;

namespace {

// y = op(A) x for caller vectors on the host or the device, precision prec.
template <typename T>
int spmv_any(CsrOp* op, char trans, const T* x, T* y, int where) {
  const bool tr = (trans == 't' || trans == 'T');
  const long long rin = tr ? op->m : op->n, rout = tr ? op->n : op->m;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PK_ERR_HIP;
  T *dx = nullptr, *dy = nullptr;
  double* part = nullptr;
  int rc = PK_OK;
  const size_t npart = 2 * static_cast<size_t>(std::max(std::max(op->nrb, op->nrbt), kMaxBlocks));
  if (hipMalloc(&part, sizeof(double) * npart) != hipSuccess) rc = PK_ERR_OOM;
  if (rc == PK_OK && where == PK_HOST) {
    if (hipMalloc(&dx, sizeof(T) * std::max<long long>(rin, 1)) != hipSuccess ||
        hipMalloc(&dy, sizeof(T) * std::max<long long>(rout, 1)) != hipSuccess ||
        hipMemcpyAsync(dx, x, sizeof(T) * rin, hipMemcpyHostToDevice, s) != hipSuccess)
      rc = PK_ERR_HIP;
  } else if (rc == PK_OK) {
    dx = const_cast<T*>(x);
    dy = y;
  }
  if (rc == PK_OK) {
    if (op->val_prec == (sizeof(T) == 8 ? PK_F64 : PK_F32)) {
      if (csr_apply<T>(op, s, trans, dx, dy, nullptr, sval(0.0), part, Gate{}) < 0)
        rc = PK_ERR_HIP;
    } else {
      set_error("pk_csr_spmv: vector precision differs from the operator's");
      rc = PK_ERR_NOTSUP;
    }
  }
  if (rc == PK_OK && where == PK_HOST &&
      hipMemcpyAsync(y, dy, sizeof(T) * rout, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = PK_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && rc == PK_OK) rc = PK_ERR_HIP;
  if (where == PK_HOST) {
    (void)hipFree(dx);
    (void)hipFree(dy);
  }
  (void)hipFree(part);
  (void)hipStreamDestroy(s);
  return rc;
}

// y = op(A) x for a complex operator (trans 'n', 't' or 'c'), vectors in
// its precision, on the host or the device.
int zspmv_any(CsrOp* op, char trans, const void* x, void* y, int where) {
  const bool tr = trans != 'n' && trans != 'N';
  const long long rin = tr ? op->m : op->n, rout = tr ? op->n : op->m;
  const size_t eb = op->val_bytes;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PK_ERR_HIP;
  void *dx = nullptr, *dy = nullptr;
  double *part = nullptr, *yw = nullptr;
  int rc = PK_OK;
  // zspmv_blocks and the combine's zstream_blocks stay within kMaxBlocks slots
  const size_t npart = 3 * static_cast<size_t>(std::max(kMaxBlocks, zstream_blocks(std::max(op->m, op->n))));
  const long long nyw = zcsr_ywork(op);
  if (hipMalloc(&part, sizeof(double) * npart) != hipSuccess ||
      (nyw > 0 && hipMalloc(&yw, sizeof(double) * nyw) != hipSuccess))
    rc = PK_ERR_OOM;
  if (rc == PK_OK && where == PK_HOST) {
    if (hipMalloc(&dx, eb * std::max<long long>(rin, 1)) != hipSuccess ||
        hipMalloc(&dy, eb * std::max<long long>(rout, 1)) != hipSuccess ||
        hipMemcpyAsync(dx, x, eb * rin, hipMemcpyHostToDevice, s) != hipSuccess)
      rc = PK_ERR_HIP;
  } else if (rc == PK_OK) {
    dx = const_cast<void*>(x);
    dy = y;
  }
  if (rc == PK_OK) {
    const int nb = op->val_prec == PK_C128
                       ? zcsr_apply<double>(op, s, trans, dx, dy, nullptr, 0.0, part, yw)
                       : zcsr_apply<float>(op, s, trans, dx, dy, nullptr, 0.0, part, yw);
    if (nb < 0) rc = PK_ERR_HIP;
  }
  if (rc == PK_OK && where == PK_HOST &&
      hipMemcpyAsync(y, dy, eb * rout, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = PK_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && rc == PK_OK) rc = PK_ERR_HIP;
  if (where == PK_HOST) {
    (void)hipFree(dx);
    (void)hipFree(dy);
  }
  (void)hipFree(part);
  (void)hipFree(yw);
  (void)hipStreamDestroy(s);
  return rc;
}

CsrOp* csr_from_iparm(const int* iparm) {
  if (!iparm) return nullptr;
  const uint64_t lo = static_cast<uint32_t>(iparm[0]), hi = static_cast<uint32_t>(iparm[1]);
  return as_csr(reinterpret_cast<pk_csr_t>(static_cast<uintptr_t>(lo | (hi << 32))));
}

}  // namespace

namespace {

// Times `reps` applications of op(A) (after one warm-up) on a device vector
// of ones; microseconds per application.
template <typename T>
int time_apply(const CsrOp* op, char trans, int reps, double* us) {
  const bool tr = trans == 't';
  const long long rin = tr ? op->m : op->n, rout = tr ? op->n : op->m;
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PK_ERR_HIP;
  T *x = nullptr, *y = nullptr;
  double* part = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PK_OK;
  const size_t npart = 2 * static_cast<size_t>(std::max(std::max(op->nrb, op->nrbt), kMaxBlocks));
  std::vector<T> ones(static_cast<size_t>(std::max<long long>(rin, 1)), T(1));
  if (hipMalloc(&x, sizeof(T) * ones.size()) != hipSuccess ||
      hipMalloc(&y, sizeof(T) * std::max<long long>(rout, 1)) != hipSuccess ||
      hipMalloc(&part, sizeof(double) * npart) != hipSuccess ||
      hipMemcpy(x, ones.data(), sizeof(T) * ones.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = PK_ERR_HIP;
  } else {
    float ms = 0;
    for (int i = 0; i <= reps && rc == PK_OK; ++i) {
      if (i == 1 && hipEventRecord(e0, s) != hipSuccess) rc = PK_ERR_HIP;
      if (csr_apply<T>(op, s, trans, x, y, nullptr, sval(0.0), part, Gate{}) < 0) rc = PK_ERR_HIP;
    }
    if (rc == PK_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = PK_ERR_HIP;
    *us = 1000.0 * ms / std::max(reps, 1);
  }
  (void)hipFree(x);
  (void)hipFree(y);
  (void)hipFree(part);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return rc;
}

}  // namespace

// ------------------------------------------------------ no-allocation SpMV --
// Scratch of pk_csr_spmv_noalloc: the SpMV's partial sums, its running row
// sums between column panels and its long-row chunk sums, laid out as
// [part | yacc | lsum] (doubles).  A file-scope device array: every GPU the
// library runs on holds its own copy (48 MiB) until the process ends.
constexpr size_t kNoAllocDoubles = size_t{6} << 20;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_lock;
double* g_noalloc_ptr[kNoAllocDevices] = {};

// Device address of `scratch` on the current device, looked up on the
// first call on that device (with it current) and kept per device.  In host
// code the array's name is only a host-side stand-in: kernels get this
// address, never the name.
int noalloc_scratch(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) {
    set_error("pk_csr_spmv_noalloc: no usable current device (%d)", dev);
    return PK_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_lock);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t r = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !ptr) {
      set_error("pk_csr_spmv_noalloc: hipGetSymbolAddress of the scratch array failed on "
                "device %d: %s", dev, hipGetErrorString(r));
      return PK_ERR_HIP;
    }
    g_noalloc_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_noalloc_ptr[dev];
  return PK_OK;
}

bool dir_has_long_rows(const CsrOp::Dir& D) {
  for (int p = 0; p < D.np; ++p)
    if ((D.sort_cap > 0 ? D.pan[p].spart.nlong : D.pan[p].part.nlong) > 0) return true;
  return false;
}

}  // namespace

extern "C" {

int pk_csr_create(int m, int n, long long nnz, const int* rowptr, const int* colind,
                  const void* val, int val_prec, int where, pk_csr_t* op) {
  CsrOp* o = nullptr;
  const int rc = csr_create(m, n, nnz, rowptr, colind, val, val_prec, where, &o);
  if (rc != PK_OK) {
    set_error("pk_csr_create failed (code %d)", rc);
    return rc;
  }
  *op = reinterpret_cast<pk_csr_t>(o);
  return PK_OK;
}

int pk_csr_create_panels(int m, int n, long long nnz, const int* rowptr, const int* colind,
                         const void* val, int val_prec, int where, int panels_a, int panels_t,
                         pk_csr_t* op) {
  CsrOp* o = nullptr;
  const int rc = csr_create(m, n, nnz, rowptr, colind, val, val_prec, where, &o,
                            panels_a > 0 ? panels_a : 1, panels_t > 0 ? panels_t : 1);
  if (rc != PK_OK) {
    set_error("pk_csr_create_panels failed (code %d)", rc);
    return rc;
  }
  *op = reinterpret_cast<pk_csr_t>(o);
  return PK_OK;
}

int pk_csr_panels(pk_csr_t h, int* panels_a, int* panels_t) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (panels_a) *panels_a = op->da.np;
  if (panels_t) *panels_t = op->dt.np;
  return PK_OK;
}

// SpMV autotuner: builds the operator with 1, 2, 4 and 8 column panels per
// direction (up to 16 when sorted layouts are tried) and, when lanes_a /
// lanes_t are given, tries lanes per row of half, equal to and twice the
// built-in choice on each (and a quarter when sorted layouts are tried: the
// sorted kernel favours fewer lanes per row); when sort_a / sort_t are
// given, each of these in
// the row-block layout and in the column-sorted layout with blocks of 8192
// and 16384 nonzeros.  Times each direction's application and returns the
// fastest setting of each direction and its time.  The choice changes how
// the products are grouped (not what is summed), so an operator built with
// it gives results equal to the built-in layout's up to rounding.
int pk_csr_autotune(int m, int n, long long nnz, const int* rowptr, const int* colind,
                    const void* val, int val_prec, int where, int* panels_a, int* panels_t,
                    double* us_a, double* us_t, int* lanes_a, int* lanes_t, int* sort_a,
                    int* sort_t) {
  if (!panels_a || !panels_t) return PK_ERR_ARG;
  if (prec_is_complex(val_prec)) {
    set_error("pk_csr_autotune: complex operators take column panels only (pk_csr_create_panels)");
    return PK_ERR_NOTSUP;
  }
  const bool lanes = lanes_a && lanes_t;
  const bool sorts = sort_a && sort_t;
  static const int kCaps[6] = {0, -1, 8192, 16384, -8192, -16384};
  const int ncaps = sorts ? 6 : 1;
  // PK_AUTOTUNE_LOG=1: one line per timed setting on stderr
  const char* lg = std::getenv("PK_AUTOTUNE_LOG");
  const bool verbose = lg && *lg && *lg != '0';
  const int pmax = sorts ? CsrOp::kMaxPanels : 8;
  double best_a = 1e300, best_t = 1e300;
  int pa = 1, pt = 1, sa = 0, st = 0, ca = 0, ct = 0;
  for (int P = 1; P <= pmax; P *= 2) {
    CsrOp* o = nullptr;
    int rc = csr_create(m, n, nnz, rowptr, colind, val, val_prec, where, &o, P, P);
    if (rc != PK_OK) return rc;
    const int npa = o->da.np, npt = o->dt.np;
    for (int ic = 0; ic < ncaps && rc == PK_OK; ++ic) {
      const int cap = kCaps[ic];
      if (cap < -1 && npa == 1 && npt == 1) continue;  // one panel: the sorted layout itself
      const int rca = csr_set_sort(o, o->da, cap), rct = csr_set_sort(o, o->dt, cap);
      if (rca == PK_ERR_NOTSUP || rct == PK_ERR_NOTSUP) continue;  // more slots than built
      rc = rca != PK_OK ? rca : rct;
      for (int sh = lanes ? (sorts ? -2 : -1) : 0; sh <= (lanes ? 1 : 0) && rc == PK_OK; ++sh) {
        csr_set_lanes(o->da, sh);
        csr_set_lanes(o->dt, sh);
        double ta = 0, tt = 0;
        if (val_prec == PK_F64) {
          rc = time_apply<double>(o, 'n', 5, &ta);
          if (rc == PK_OK) rc = time_apply<double>(o, 't', 5, &tt);
        } else {
          rc = time_apply<float>(o, 'n', 5, &ta);
          if (rc == PK_OK) rc = time_apply<float>(o, 't', 5, &tt);
        }
        if (rc != PK_OK) break;
        if (verbose)
          std::fprintf(stderr, "pk_csr_autotune: panels %d/%d cap %d shift %d: A %.1f us, A^T %.1f us\n",
                       npa, npt, cap, sh, ta, tt);
        if (npa == P && ta < best_a) {
          best_a = ta;
          pa = P;
          sa = sh;
          ca = csr_sort_code(o->da);
        }
        if (npt == P && tt < best_t) {
          best_t = tt;
          pt = P;
          st = sh;
          ct = csr_sort_code(o->dt);
        }
      }
    }
    csr_destroy(o);
    if (rc != PK_OK) return rc;
  }
  *panels_a = pa;
  *panels_t = pt;
  if (lanes) {
    *lanes_a = sa;
    *lanes_t = st;
  }
  if (sorts) {
    *sort_a = ca;
    *sort_t = ct;
  }
  if (us_a) *us_a = best_a;
  if (us_t) *us_t = best_t;
  return PK_OK;
}

int pk_csr_set_sort(pk_csr_t h, int cap_a, int cap_t) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (prec_is_complex(op->val_prec)) {
    set_error("pk_csr_set_sort: complex operators take column panels only (pk_csr_create_panels)");
    return PK_ERR_NOTSUP;
  }
  for (const int c : {cap_a, cap_t}) {
    const int a = c < -1 ? -c : c;
    if (c != 0 && c != -1 && (a < kSpmvCap || a > kSortCapMax)) {
      set_error("pk_csr_set_sort: a block cap must be 0, -1 or lie in [%d, %d] (negated: "
                "every column panel in one launch)", kSpmvCap, kSortCapMax);
      return PK_ERR_ARG;
    }
  }
  int rc = csr_set_sort(op, op->da, cap_a);
  if (rc == PK_OK) rc = csr_set_sort(op, op->dt, cap_t);
  if (rc != PK_OK) {
    (void)csr_set_sort(op, op->da, 0);
    (void)csr_set_sort(op, op->dt, 0);
    set_error("pk_csr_set_sort failed (code %d); the operator is back in the row-block layout",
              rc);
  }
  return rc;
}

int pk_csr_sort(pk_csr_t h, int* cap_a, int* cap_t) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (cap_a) *cap_a = csr_sort_code(op->da);
  if (cap_t) *cap_t = csr_sort_code(op->dt);
  return PK_OK;
}

int pk_csr_set_lanes(pk_csr_t h, int shift_a, int shift_t) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (prec_is_complex(op->val_prec)) {
    set_error("pk_csr_set_lanes: complex operators take column panels only (pk_csr_create_panels)");
    return PK_ERR_NOTSUP;
  }
  if (shift_a < -2 || shift_a > 2 || shift_t < -2 || shift_t > 2) {
    set_error("pk_csr_set_lanes: shifts must lie in [-2, 2]");
    return PK_ERR_ARG;
  }
  csr_set_lanes(op->da, shift_a);
  csr_set_lanes(op->dt, shift_t);
  ++op->layout_epoch;
  return PK_OK;
}

int pk_csr_lanes(pk_csr_t h, int* shift_a, int* shift_t) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (shift_a) *shift_a = op->da.lane_shift;
  if (shift_t) *shift_t = op->dt.lane_shift;
  return PK_OK;
}

int pk_csr_destroy(pk_csr_t h) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  csr_destroy(op);
  return PK_OK;
}

int pk_csr_dims(pk_csr_t h, int* m, int* n, long long* nnz) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (m) *m = op->m;
  if (n) *n = op->n;
  if (nnz) *nnz = op->nnz;
  return PK_OK;
}

int pk_csr_spmv(pk_csr_t h, char trans, const void* x, void* y, int prec, int where) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (!x || !y || (where != PK_HOST && where != PK_DEVICE)) return PK_ERR_ARG;
  if (prec_is_complex(prec) || prec_is_complex(op->val_prec)) {
    if (prec != op->val_prec) {
      set_error("pk_csr_spmv: vector precision differs from the operator's");
      return PK_ERR_NOTSUP;
    }
    return zspmv_any(op, trans, x, y, where);
  }
  if (prec == PK_F64)
    return spmv_any<double>(op, trans, static_cast<const double*>(x), static_cast<double*>(y),
                            where);
  if (prec == PK_F32)
    return spmv_any<float>(op, trans, static_cast<const float*>(x), static_cast<float*>(y), where);
  return PK_ERR_ARG;
}

int pk_csr_spmv_noalloc(pk_csr_t h, char trans, const void* x, void* y, int prec,
                        void* stream) {
  CsrOp* op = as_csr(h);
  if (!op) return PK_ERR_HANDLE;
  if (!x || !y || (prec != PK_F64 && prec != PK_F32)) return PK_ERR_ARG;
  if (prec_is_complex(op->val_prec)) {
    set_error("pk_csr_spmv_noalloc: not available for complex operators");
    return PK_ERR_NOTSUP;
  }
  if (op->val_prec != prec) {
    set_error("pk_csr_spmv_noalloc: vector precision differs from the operator's");
    return PK_ERR_NOTSUP;
  }
  const bool tr = (trans == 't' || trans == 'T');
  const CsrOp::Dir& D = tr ? op->dt : op->da;
  const size_t npart =
      2 * static_cast<size_t>(std::max(std::max(op->nrb, op->nrbt), kMaxBlocks));
  const size_t nyacc = D.np > 1 ? static_cast<size_t>(D.rows) : 0;
  const size_t nlong = dir_has_long_rows(D) ? static_cast<size_t>(csr_lwork(op)) : 0;
  if (npart + nyacc + nlong > kNoAllocDoubles) {
    set_error("pk_csr_spmv_noalloc: this operator needs %zu doubles of scratch (%zu partial "
              "sums, %zu running row sums, %zu long-row sums); the fixed scratch holds %zu",
              npart + nyacc + nlong, npart, nyacc, nlong, kNoAllocDoubles);
    return PK_ERR_NOTSUP;
  }
  double* base = nullptr;
  const int rc = noalloc_scratch(&base);
  if (rc != PK_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb =
      prec == PK_F64
          ? csr_apply<double>(op, s, trans, static_cast<const double*>(x), static_cast<double*>(y),
                              nullptr, sval(0.0), base, Gate{}, base + npart,
                              base + npart + nyacc, static_cast<long long>(nyacc))
          : csr_apply<float>(op, s, trans, static_cast<const float*>(x), static_cast<float*>(y),
                             nullptr, sval(0.0), base, Gate{}, base + npart, base + npart + nyacc,
                             static_cast<long long>(nyacc));
  if (nb < 0) {
    set_error("pk_csr_spmv_noalloc: kernel launch failed");
    return PK_ERR_HIP;
  }
  return PK_OK;
}

int pk_csr_to_iparm(pk_csr_t h, int* iparm) {
  if (!as_csr(h) || !iparm) return PK_ERR_ARG;
  const uint64_t v = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(h));
  iparm[0] = static_cast<int>(static_cast<uint32_t>(v & 0xffffffffu));
  iparm[1] = static_cast<int>(static_cast<uint32_t>(v >> 32));
  return PK_OK;
}

void pk_dcsr_aprod_(const char* transa, const int* m, const int* n, const double* x, double* y,
                    double* dparm, int* iparm, size_t len_transa) {
  (void)m; (void)n; (void)dparm; (void)len_transa;
  CsrOp* op = csr_from_iparm(iparm);
  if (op) spmv_any<double>(op, *transa, x, y, PK_HOST);
}

void pk_scsr_aprod_(const char* transa, const int* m, const int* n, const float* x, float* y,
                    float* dparm, int* iparm, size_t len_transa) {
  (void)m; (void)n; (void)dparm; (void)len_transa;
  CsrOp* op = csr_from_iparm(iparm);
  if (op) spmv_any<float>(op, *transa, x, y, PK_HOST);
}

void pk_zcsr_aprod_(const char* transa, const int* m, const int* n, const void* x, void* y,
                    void* zparm, int* iparm, size_t len_transa) {
  (void)m; (void)n; (void)zparm; (void)len_transa;
  CsrOp* op = csr_from_iparm(iparm);
  if (op && op->val_prec == PK_C128) zspmv_any(op, *transa, x, y, PK_HOST);
}

void pk_ccsr_aprod_(const char* transa, const int* m, const int* n, const void* x, void* y,
                    void* cparm, int* iparm, size_t len_transa) {
  (void)m; (void)n; (void)cparm; (void)len_transa;
  CsrOp* op = csr_from_iparm(iparm);
  if (op && op->val_prec == PK_C64) zspmv_any(op, *transa, x, y, PK_HOST);
}

// Fortran-side helper (implicit interface, all arguments by reference):
// kind = 1 when APROD is the library's pk_dcsr_aprod_, 2 for pk_scsr_aprod_,
// 3 for pk_zcsr_aprod_, 4 for pk_ccsr_aprod_, with a live operator handle
// in IPARM(1:2) returned in *handle; kind = 0 for any other (user) APROD.
void pk_f_classify_(void* aprod, const int* iparm, int* kind, long long* handle) {
  *kind = 0;
  *handle = 0;
  int k = 0;
  if (aprod == reinterpret_cast<void*>(&pk_dcsr_aprod_)) k = 1;
  if (aprod == reinterpret_cast<void*>(&pk_scsr_aprod_)) k = 2;
  if (aprod == reinterpret_cast<void*>(&pk_zcsr_aprod_)) k = 3;
  if (aprod == reinterpret_cast<void*>(&pk_ccsr_aprod_)) k = 4;
  if (k == 0) return;
  CsrOp* op = csr_from_iparm(iparm);
  if (!op) return;
  *kind = k;
  *handle = static_cast<long long>(reinterpret_cast<uintptr_t>(op));
}

// Fortran-side helpers: register a user APROD (double / real arrays) on an
// engine given by its handle value.
void pk_f_set_callback_d_(const long long* eng, void* aprod, double* dparm, int* iparm,
                          int* info) {
  *info = pk_lbd_set_callback(reinterpret_cast<pk_lbd_t>(static_cast<uintptr_t>(*eng)), aprod,
                              PK_F64, dparm, iparm);
}
void pk_f_set_callback_s_(const long long* eng, void* aprod, float* sparm, int* iparm,
                          int* info) {
  *info = pk_lbd_set_callback(reinterpret_cast<pk_lbd_t>(static_cast<uintptr_t>(*eng)), aprod,
                              PK_F32, sparm, iparm);
}
void pk_f_set_callback_z_(const long long* eng, void* aprod, void* zparm, int* iparm,
                          int* info) {
  *info = pk_lbd_set_callback(reinterpret_cast<pk_lbd_t>(static_cast<uintptr_t>(*eng)), aprod,
                              PK_C128, zparm, iparm);
}
void pk_f_set_callback_c_(const long long* eng, void* aprod, void* cparm, int* iparm,
                          int* info) {
  *info = pk_lbd_set_callback(reinterpret_cast<pk_lbd_t>(static_cast<uintptr_t>(*eng)), aprod,
                              PK_C64, cparm, iparm);
}

}  // extern "C"
