// Per-workgroup bodies of the streaming kernels.  The single-problem
// kernels (pk_kernels.hip) and the batched kernels (pk_batch.hip) both call
// these, so one problem gives bit-identical results whether it runs alone
// or inside a batch: the same rows per workgroup, the same summation order,
// the same partial slots.
#pragma once

#include "pk_kernels.h"

namespace pk {

// Rows per thread of the dense streaming kernels; workgroup b of a vector of
// length L owns the contiguous slice [b * kRowsPerBlock, (b + 1) * kRowsPerBlock).
constexpr int kRPT = 4;
constexpr int kRowsPerBlock = kBlock * kRPT;
// Gram-Schmidt kernels (projection and update) own slices of kCgsRows rows,
// kCgsRPT rows per thread, and stream kCT-column tiles of the basis.
constexpr int kCgsRPT = 8;
constexpr int kCgsRows = kBlock * kCgsRPT;
constexpr int kCT = 4;

// ------------------------------------------------------ row-block SpMV ----
// Chunk -(d.y+1) of a long row, nonzeros [d.z, d.w), summed by an NT-thread
// workgroup (thread-strided, then the block reduction); the sum goes to
// lsum, the fix-up kernel finishes the row.  Column of entry p:
// col0 + (ci[p] >> pbits) (packed sorted layout; col0 = pbits = 0: ci[p]).
template <typename VT, typename T, int NT>
__device__ __forceinline__ void spmv_chunk(const int4 d, const int* __restrict__ ci,
                                           const VT* __restrict__ va, const T* __restrict__ x,
                                           double* q_s2, double* q_sd,
                                           double* __restrict__ lsum, int last, double* lds,
                                           int col0 = 0, int pbits = 0) {
  double acc[1] = {0.0};
  for (int p = d.z + static_cast<int>(threadIdx.x); p < d.w; p += NT) {
    const int col = col0 + static_cast<int>(static_cast<unsigned>(ci[p]) >> pbits);
    acc[0] += static_cast<double>(va[p]) * static_cast<double>(x[col]);
  }
  block_sum<1, NT>(acc, lds);
  if (threadIdx.x == 0) {
    lsum[-(d.y + 1)] = acc[0];
    if (last) {
      *q_s2 = 0.0;
      *q_sd = 0.0;
    }
  }
}

// Row sums of rows [r0, r1) from their products in LDS (prod[q] for the
// block's q-th nonzero in row-major order): groups of G lanes, lane l of a
// row adds prod[a + l], prod[a + l + G], ... then a shuffle tree; then the
// epilogue.  The order depends only on G and the row, not on the block.
template <typename T, int G, int NT>
__device__ __forceinline__ void spmv_rows(int r0, int r1, int p0, const int* __restrict__ rp,
                                          const double* prod, T* __restrict__ y,
                                          const T* __restrict__ yprev, double c, double* q_s2,
                                          double* q_sd, const double* __restrict__ yacc_in,
                                          double* __restrict__ yacc_out, int last,
                                          double* lds) {
  double s2 = 0.0, sd = 0.0;
  const int lane = threadIdx.x & (G - 1);
  const int group = threadIdx.x / G;
  constexpr int ngroups = NT / G;
  const int nrows = r1 - r0;
  const int iters = (nrows + ngroups - 1) / ngroups;  // uniform across the block
  // The row bounds, yprev and the running sums of kPre of a group's rows
  // are loaded together ahead of their sums (one memory round trip per kPre
  // rows instead of two per row); loads of rows past r1 read row r0 and
  // are not used.  The sums themselves are unchanged.
  constexpr int kPre = 4;
  for (int it0 = 0; it0 < iters; it0 += kPre) {
    int a[kPre], b[kPre];
    double yp[kPre], yin[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
      const int r = r0 + group + (it0 + u) * ngroups;
      const int rr = (it0 + u < iters && r < r1) ? r : r0;
      a[u] = rp[rr] - p0;
      b[u] = rp[rr + 1] - p0;
      yp[u] = (last && yprev) ? static_cast<double>(yprev[rr]) : 0.0;
      yin[u] = yacc_in ? yacc_in[rr] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
      if (it0 + u >= iters) break;  // uniform across the block
      const int r = r0 + group + (it0 + u) * ngroups;
      const bool valid = r < r1;
      double acc = 0.0;
      if (valid)
        for (int q = a[u] + lane; q < b[u]; q += G) acc += prod[q];
      acc = group_sum<G>(acc);
      if (valid && lane == 0) {
        if (yacc_in) acc = yin[u] + acc;
        if (!last) {
          yacc_out[r] = acc;
        } else {
          const T yv = static_cast<T>(acc - c * yp[u]);
          y[r] = yv;
          s2 += static_cast<double>(yv) * static_cast<double>(yv);
          sd += yp[u] * static_cast<double>(yv);
        }
      }
    }
  }
  if (!last) return;
  double v2[2] = {s2, sd};
  block_sum<2, NT>(v2, lds);
  if (threadIdx.x == 0) {
    *q_s2 = v2[0];
    *q_sd = v2[1];
  }
}

// Descriptor d = {x, y, z, w}:
//   * y >= 0: a row block, rows [x, y) holding at most kSpmvCap nonzeros
//     and kSpmvRows rows.  All 256 threads turn its nonzeros into products
//     val * x[col] in LDS (prod) -- index and value streams read fully
//     coalesced, kUnroll independent gathers in flight per thread -- then
//     groups of G lanes sum each row from LDS (consecutive, conflict-free
//     reads; shuffle reduction) in a fixed order and apply the epilogue.
//   * y < 0: chunk -(y+1) of the long row x, nonzeros [z, w): its partial
//     dot product goes to lsum[-(y+1)]; the long-row fix-up finishes the row.
// Epilogue (last = 1): y = T(sum - c * yprev), *q_s2 = sum y^2,
// *q_sd = sum yprev.y over the block (zero for chunks).  Column panels
// (last = 0) store the FP64 running row sums in yacc_out instead; a panel
// after the first adds yacc_in.
template <typename VT, typename T, int G>
__device__ __forceinline__ void spmv_rb_body(const int4 d, const int* __restrict__ rp,
                                             const int* __restrict__ ci,
                                             const VT* __restrict__ va, const T* __restrict__ x,
                                             T* __restrict__ y, const T* __restrict__ yprev,
                                             double c, double* q_s2, double* q_sd,
                                             double* __restrict__ lsum,
                                             const double* __restrict__ yacc_in,
                                             double* __restrict__ yacc_out, int last,
                                             double* prod, double* lds) {
  if (d.y < 0) {
    spmv_chunk<VT, T, kBlock>(d, ci, va, x, q_s2, q_sd, lsum, last, lds);
    return;
  }
  const int r0 = d.x, r1 = d.y;
  const int p0 = rp[r0];
  const int nz = rp[r1] - p0;
  // Eight independent loads per thread; with kSpmvCap = 1024 the upper
  // half is masked, which measured faster than a 4-deep unroll.
  constexpr int kUnroll = 8;
  for (int base = 0; base < nz; base += kBlock * kUnroll) {
    int idx[kUnroll];
    double v[kUnroll], xv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int t = base + u * kBlock + static_cast<int>(threadIdx.x);
      // index and value streams are read once: non-temporal, so that they
      // do not evict the gathered vector slice from L2
      idx[u] = t < nz ? __builtin_nontemporal_load(ci + p0 + t) : -1;
      v[u] = t < nz ? static_cast<double>(__builtin_nontemporal_load(va + p0 + t)) : 0.0;
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
  spmv_rows<T, G, kBlock>(r0, r1, p0, rp, prod, y, yprev, c, q_s2, q_sd, yacc_in, yacc_out,
                          last, lds);
}

// --------------------------------------------------- row-wise SpMV ----
// Row-wise form of a row-block descriptor: the block's rows are taken by
// groups of G lanes exactly as spmv_rows takes them (same rows per group,
// same iterations), but each lane multiplies its entries straight from the
// index and value streams -- entries a + l, a + l + G, ... of its row, four
// of them in flight, added in that order -- so there is no LDS staging and
// no barrier before the sums.  Products are rounded to double before they
// are added (no contraction), as spmv_rb_body stores them: every row sum,
// and the block's partial sums, are bit for bit the row-block layout's for
// the same G.  Long-row chunks as in spmv_rb_body.
template <typename VT, typename T, int G>
__device__ __forceinline__ void spmv_vec_body(const int4 d, const int* __restrict__ rp,
                                              const int* __restrict__ ci,
                                              const VT* __restrict__ va, const T* __restrict__ x,
                                              T* __restrict__ y, const T* __restrict__ yprev,
                                              double c, double* q_s2, double* q_sd,
                                              double* __restrict__ lsum,
                                              const double* __restrict__ yacc_in,
                                              double* __restrict__ yacc_out, int last,
                                              double* lds) {
  if (d.y < 0) {
    spmv_chunk<VT, T, kBlock>(d, ci, va, x, q_s2, q_sd, lsum, last, lds);
    return;
  }
  const int r0 = d.x, r1 = d.y;
  double s2 = 0.0, sd = 0.0;
  const int lane = threadIdx.x & (G - 1);
  const int group = threadIdx.x / G;
  constexpr int ngroups = kBlock / G;
  const int iters = (r1 - r0 + ngroups - 1) / ngroups;  // uniform across the block
  for (int it = 0; it < iters; ++it) {
    const int r = r0 + group + it * ngroups;
    const bool valid = r < r1;
    int a = 0, b = 0;
    if (valid) {
      a = rp[r];
      b = rp[r + 1];
    }
    double acc = 0.0;
    int q = a + lane;
    for (; q + 3 * G < b; q += 4 * G) {
      int cc[4];
      double vv[4], xx[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cc[u] = __builtin_nontemporal_load(ci + q + u * G);
        vv[u] = static_cast<double>(__builtin_nontemporal_load(va + q + u * G));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) xx[u] = static_cast<double>(x[cc[u]]);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        // rounded product, then the sum (no FMA): as the staged products
#pragma clang fp contract(off)
        const double pr = vv[u] * xx[u];
        acc += pr;
      }
    }
    for (; q < b; q += G) {
#pragma clang fp contract(off)
      const double pr = static_cast<double>(__builtin_nontemporal_load(va + q)) *
                        static_cast<double>(x[__builtin_nontemporal_load(ci + q)]);
      acc += pr;
    }
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
  block_sum<2, kBlock>(v2, lds);
  if (threadIdx.x == 0) {
    *q_s2 = v2[0];
    *q_sd = v2[1];
  }
}

// ------------------------------------------------- column-sorted SpMV ----
// Sorted layout of a panel: row blocks of at most cap nonzeros (a
// per-operator choice, <= kSortCapMax) and kSortRows rows, rows longer than
// kSpmvCap split into long-row chunks as in the row-block layout.  Within a
// block the nonzeros are stored in increasing column order (sci, sva; ties
// in row-major order) with spos[q] = the entry's offset in row-major order
// within the block.  Consecutive lanes read consecutive sorted entries, so
// the gathers of nearby columns fall into the same cache lines and merge;
// each product is written to LDS at its row-major offset and the rows are
// summed from there by spmv_rows -- for the same G every row sum is bit for
// bit the row-block layout's.  Long-row chunks keep the sorted order too
// (the chunk sum is a fixed-order block reduction either way).  Packed form
// (spos null): sci[q] = (column - col0) << pbits | offset, one 32-bit word
// per entry, used when the panel's width and the cap fit in 32 bits.
template <typename VT, typename T, int G>
__device__ __forceinline__ void spmv_sorted_body(
    const int4 d, const int* __restrict__ rp, const int* __restrict__ sci,
    const VT* __restrict__ sva, const unsigned short* __restrict__ spos, int col0, int pbits,
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ yprev, double c,
    double* q_s2, double* q_sd, double* __restrict__ lsum, const double* __restrict__ yacc_in,
    double* __restrict__ yacc_out, int last, double* prod, double* lds) {
  if (d.y < 0) {
    spmv_chunk<VT, T, kSortThreads>(d, sci, sva, x, q_s2, q_sd, lsum, last, lds, col0, pbits);
    return;
  }
  const unsigned pmask = (1u << pbits) - 1u;
  const int r0 = d.x, r1 = d.y;
  const int p0 = rp[r0];
  const int nz = rp[r1] - p0;
  constexpr int kUnroll = 8;
  for (int base = 0; base < nz; base += kSortThreads * kUnroll) {
    int idx[kUnroll], pos[kUnroll];
    double v[kUnroll], xv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int t = base + u * kSortThreads + static_cast<int>(threadIdx.x);
      const unsigned wv =
          t < nz ? static_cast<unsigned>(__builtin_nontemporal_load(sci + p0 + t)) : 0u;
      idx[u] = t < nz ? col0 + static_cast<int>(wv >> pbits) : -1;
      v[u] = t < nz ? static_cast<double>(__builtin_nontemporal_load(sva + p0 + t)) : 0.0;
      if (spos)
        pos[u] = t < nz ? static_cast<int>(__builtin_nontemporal_load(spos + p0 + t)) : 0;
      else
        pos[u] = static_cast<int>(wv & pmask);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) xv[u] = idx[u] >= 0 ? static_cast<double>(x[idx[u]]) : 0.0;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (idx[u] >= 0) prod[pos[u]] = v[u] * xv[u];
  }
  __syncthreads();
  spmv_rows<T, G, kSortThreads>(r0, r1, p0, rp, prod, y, yprev, c, q_s2, q_sd, yacc_in,
                                yacc_out, last, lds);
}

// Long row ld = {row, first chunk, number of chunks, -}: one wave sums the
// chunks (lane l takes chunks l, l + 64, ... in order, then a fixed shuffle
// tree) and lane 0 applies the epilogue (or stores the running sum).
template <typename T>
__device__ __forceinline__ void spmv_longfix_body(const int4 d, const double* __restrict__ lsum,
                                                  T* __restrict__ y,
                                                  const T* __restrict__ yprev, double c,
                                                  double* q_s2, double* q_sd,
                                                  const double* __restrict__ yacc_in,
                                                  double* __restrict__ yacc_out, int last,
                                                  int lane) {
  double acc = 0.0;
  for (int q = lane; q < d.z; q += kWave) acc += lsum[d.y + q];
  acc = wave_sum(acc);
  if (lane != 0) return;
  if (yacc_in) acc = yacc_in[d.x] + acc;
  if (!last) {
    yacc_out[d.x] = acc;
    return;
  }
  const double yp = yprev ? static_cast<double>(yprev[d.x]) : 0.0;
  const T yv = static_cast<T>(acc - c * yp);
  y[d.x] = yv;
  *q_s2 = static_cast<double>(yv) * static_cast<double>(yv);
  *q_sd = yp * static_cast<double>(yv);
}

// ----------------------------------------------- Gram-Schmidt projection ----
// Sums of kCT = 4 per-lane partials over the wave by a butterfly: the
// exchanges at lane distances 32 and 16 halve the values a lane holds (lane
// keeps column 2*bit5 + bit4 of its index), distances 8..1 finish the sum.
// Lanes with (lane & 15) == 0 return the sum of column lane >> 4.
__device__ __forceinline__ double wave_sum4(double (&v)[4], int lane) {
  static_assert(kCT == 4, "the butterfly is written for four columns");
  const bool h5 = lane & 32, h4 = lane & 16;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double send = h5 ? v[k] : v[k + 2];
    const double keep = h5 ? v[k + 2] : v[k];
    v[k] = keep + __shfl_xor(send, 32, kWave);
  }
  {
    const double send = h4 ? v[0] : v[1];
    const double keep = h4 ? v[1] : v[0];
    v[0] = keep + __shfl_xor(send, 16, kWave);
  }
  double t = v[0];
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) t += __shfl_xor(t, off, kWave);
  return t;
}

// Slice b (kCgsRows rows) of h = V(:, scol[0:ncols])^T w: the slice of w
// stays in registers, kCT-column tiles of V stream through (a butterfly over
// the wave, then LDS across the waves, per tile), the next tile's loads
// issued before the current tile is reduced.  Partial of column c goes to
// part[c * stride + slot].  scol: global column list, read with uniform
// indices (scalar loads, no LDS staging).  Rows past L load row L - 1 (every
// load unconditional, so a tile's loads issue back to back) and are weighted
// by 0; columns past ncols load the last column and are not written.
template <typename T>
__device__ __forceinline__ void gemvt_body(long long b, long long L, const T* __restrict__ V,
                                           long long ldv, const int* __restrict__ scol,
                                           int ncols, const T* __restrict__ w,
                                           double* __restrict__ part,
                                           long long stride, long long slot,
                                           double (*lds)[(kBlock / kWave) * kCT]) {
  if (ncols <= 0 || L <= 0) return;
  const long long base = b * kCgsRows + threadIdx.x;
  double wr[kCgsRPT];
  long long rowc[kCgsRPT];
#pragma unroll
  for (int r = 0; r < kCgsRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    rowc[r] = row < L ? row : L - 1;
    const double wv = static_cast<double>(w[rowc[r]]);
    wr[r] = row < L ? wv : 0.0;
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  int buf = 0;
  // Software pipeline over column tiles in two register buffers, t0 and t1,
  // with no copy between them (a rotation by copies made the compiler wait
  // for the next tile's loads before reducing the current one): the loads of
  // the next tile are in flight while the current tile is reduced.  Columns
  // past ncols load the last column and are not written.
  T t0[kCT][kCgsRPT], t1[kCT][kCgsRPT];
  auto load = [&](int c0, T (&t)[kCT][kCgsRPT]) {
#pragma unroll
    for (int c = 0; c < kCT; ++c) {
      const T* vc = V + static_cast<long long>(scol[c0 + c < ncols ? c0 + c : ncols - 1]) * ldv;
#pragma unroll
      for (int r = 0; r < kCgsRPT; ++r) t[c][r] = vc[rowc[r]];
    }
  };
  auto reduce = [&](int c0, const T (&t)[kCT][kCgsRPT]) {
    double acc[kCT];
#pragma unroll
    for (int c = 0; c < kCT; ++c) {
      acc[c] = 0.0;
#pragma unroll
      for (int r = 0; r < kCgsRPT; ++r) acc[c] += static_cast<double>(t[c][r]) * wr[r];
    }
    const double colsum = wave_sum4(acc, lane);
    if ((lane & 15) == 0) lds[buf][wid * kCT + (lane >> 4)] = colsum;
    __syncthreads();
    if (threadIdx.x < kCT && c0 + threadIdx.x < ncols) {
      double sum = 0.0;
      for (int ww = 0; ww < kBlock / kWave; ++ww) sum += lds[buf][ww * kCT + threadIdx.x];
      part[static_cast<long long>(c0 + threadIdx.x) * stride + slot] = sum;
    }
    buf ^= 1;
  };
  // loop entry: t0 holds tile c0 (c0 < ncols)
  load(0, t0);
  int c0 = 0;
  while (c0 + kCT < ncols) {
    load(c0 + kCT, t1);
    reduce(c0, t0);
    if (c0 + 2 * kCT >= ncols) {
      reduce(c0 + kCT, t1);
      c0 = ncols;
      break;
    }
    load(c0 + 2 * kCT, t0);
    reduce(c0 + kCT, t1);
    c0 += 2 * kCT;
  }
  if (c0 < ncols) reduce(c0, t0);
}

// Slice b (kCgsRows rows) of w -= V(:, scol[0:ncols]) sh (scol, sh global,
// read with uniform indices); *q = ||w slice||^2.  kCT columns' loads are
// issued together (unconditional: rows past L load
// row L - 1 and are never stored, columns past ncols load the last column
// and are not applied); the updates keep column order.
template <typename T>
__device__ __forceinline__ void gemv_update_body(long long b, long long L,
                                                 const T* __restrict__ V, long long ldv,
                                                 const int* __restrict__ scol,
                                                 const double* __restrict__ sh, int ncols,
                                                 T* __restrict__ w, double* q, double* lds) {
  if (L <= 0) return;
  const long long base = b * kCgsRows + threadIdx.x;
  double acc[kCgsRPT];
  long long rowc[kCgsRPT];
#pragma unroll
  for (int r = 0; r < kCgsRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    rowc[r] = row < L ? row : L - 1;
    acc[r] = static_cast<double>(w[rowc[r]]);
  }
  for (int c = 0; c < ncols; c += kCT) {  // kCT columns' loads in flight together
    T x[kCT][kCgsRPT];
#pragma unroll
    for (int u = 0; u < kCT; ++u) {
      const T* vc = V + static_cast<long long>(scol[c + u < ncols ? c + u : ncols - 1]) * ldv;
#pragma unroll
      for (int r = 0; r < kCgsRPT; ++r) x[u][r] = vc[rowc[r]];
    }
#pragma unroll
    for (int u = 0; u < kCT; ++u) {
      if (c + u < ncols) {
        const double hc = sh[c + u];
#pragma unroll
        for (int r = 0; r < kCgsRPT; ++r) acc[r] -= static_cast<double>(x[u][r]) * hc;
      }
    }
  }
  double v[1] = {0.0};
#pragma unroll
  for (int r = 0; r < kCgsRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) {
      const T o = static_cast<T>(acc[r]);
      w[row] = o;
      v[0] += static_cast<double>(o) * static_cast<double>(o);
    }
  }
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) *q = v[0];
}

// ------------------------------------------------------------- level 1 ----
template <typename T>
__device__ __forceinline__ void scale_body(long long b, long long L, T* __restrict__ x,
                                           double av) {
  const long long base = b * kRowsPerBlock + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kRPT; ++r) {
    const long long row = base + static_cast<long long>(r) * kBlock;
    if (row < L) x[row] = static_cast<T>(av * static_cast<double>(x[row]));
  }
}

// Slice b of w = w - cv * v; *q0 = ||w||^2, *q1 = v.w over the slice.
template <typename T>
__device__ __forceinline__ void axpy_norm_dot_body(long long b, long long L, T* __restrict__ w,
                                                   const T* __restrict__ v, double cv,
                                                   double* q0, double* q1, double* lds) {
  const long long base = b * kRowsPerBlock + threadIdx.x;
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
    *q0 = acc[0];
    *q1 = acc[1];
  }
}

// ----------------------------------------------------------- control ----
// Maximum over the workgroup with the control code's max (exact, so the
// result does not depend on the order); valid in thread 0.
__device__ __forceinline__ double block_max(double v, double* red) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = hd_max(v, __shfl_xor(v, o));
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < kBlock / kWave; ++w) v = hd_max(v, red[w]);
  __syncthreads();
  return v;
}

// hd_update_nu (dlanbpro.F:684-723) with its elements spread over the
// workgroup: every element is computed by the serial form's expression and
// the maximum is exact, so the results are bit-identical to it.
template <class C>
__device__ inline void par_update_nu(C& c, double* red) {
#pragma clang fp contract(off)
  const int j = c.j;
  if (j <= 1) return;
  const double* alpha = c.B1;
  const double* beta = c.B2;
  const double* mu = c.mu;
  double* nu = c.nu;
  const double eps1 = c.epsn2, anorm = c.anorm;
  double mx = 0.0;
  for (int k = 1 + static_cast<int>(threadIdx.x); k <= j - 1; k += kBlock) {
    double v = beta[k] * mu[k + 1] + alpha[k] * mu[k] - beta[j - 1] * nu[k];
    const double d =
        eps1 * (hd_lapy2(alpha[k], beta[k]) + hd_lapy2(alpha[j], beta[j - 1])) + eps1 * anorm;
    v = (v + hd_sign(d, v)) / alpha[j];
    nu[k] = v;
    mx = hd_max(mx, fabs(v));
  }
  mx = block_max(mx, red);
  if (threadIdx.x == 0) {
    c.numax = mx;
    nu[j] = 1.0;
  }
}

// hd_update_mu (dlanbpro.F:628-679), spread over the workgroup likewise.
template <class C>
__device__ inline void par_update_mu(C& c, double* red) {
#pragma clang fp contract(off)
  const int j = c.j;
  if (j == 1) {
    if (threadIdx.x == 0) hd_update_mu(&c.mumax, c.mu, c.nu, j, c.B1, c.B2, c.anorm, c.epsn2);
    return;
  }
  const double* alpha = c.B1;
  const double* beta = c.B2;
  const double* nu = c.nu;
  double* mu = c.mu;
  const double eps1 = c.epsn2, anorm = c.anorm;
  double mx = 0.0;
  for (int k = 1 + static_cast<int>(threadIdx.x); k <= j; k += kBlock) {
    double v, d;
    if (k == 1) {
      v = alpha[1] * nu[1] - alpha[j] * mu[1];
      d = eps1 * (hd_lapy2(alpha[j], beta[j]) + alpha[1]) + eps1 * anorm;
    } else if (k < j) {
      v = alpha[k] * nu[k] + beta[k - 1] * nu[k - 1] - alpha[j] * mu[k];
      d = eps1 * (hd_lapy2(alpha[j], beta[j]) + hd_lapy2(alpha[k], beta[k - 1])) + eps1 * anorm;
    } else {
      v = beta[j - 1] * nu[j - 1];
      d = eps1 * (hd_lapy2(alpha[j], beta[j]) + hd_lapy2(alpha[j], beta[j - 1])) + eps1 * anorm;
    }
    v = (v + hd_sign(d, v)) / beta[j];
    mu[k] = v;
    mx = hd_max(mx, fabs(v));
  }
  mx = block_max(mx, red);
  if (threadIdx.x == 0) {
    c.mumax = mx;
    mu[j + 1] = 1.0;
  }
}

// LDS copy of a control state at a step j <= kCtlLdsJ: the scalars and the
// windows of the arrays the step reads and writes (mu, nu, B1, B2 over
// 0..j+2, INDEX over 0..2j+4); the CGS column list stays in device memory.
constexpr int kCtlWinD = (kCtlLdsJ + 4) & ~1;
constexpr int kCtlWinI = (2 * kCtlLdsJ + 8) & ~3;
struct LbdCtlLds : LbdCtl {
  double mu[kCtlWinD], nu[kCtlWinD], B1[kCtlWinD], B2[kCtlWinD];
  int index[kCtlWinI];
  int* cols;
};
constexpr int kCtlHdrWords = static_cast<int>(sizeof(LbdCtl) / sizeof(int4));
static_assert(sizeof(LbdCtl) % sizeof(int4) == 0, "state must be whole 16-byte words");
// staged 16-byte words per thread at the largest LDS step
constexpr int kCtlPer = 6;
static_assert(kCtlHdrWords + 2 * kCtlWinD + kCtlWinI / 4 <= kCtlPer * kBlock,
              "the staging holds six 16-byte words per thread");

// The control stage itself on state c (LDS or device arrays): thread 0 runs
// the stage, the workgroup a pending omega update, thread 0 the remainder.
template <class C>
__device__ __forceinline__ void ctl_run(C& c, int kind, double a, double b, double* lds) {
  if (threadIdx.x == 0) {
    switch (kind) {
      case CTL_V_SPMV: ctl_v_spmv_s(c, a, b); break;
      case CTL_V_ELR: ctl_v_elr_s(c, a, b); break;
      case CTL_V_REO: ctl_v_reo_s(c, a); break;
      case CTL_U_SPMV: ctl_u_spmv_s(c, a, b); break;
      case CTL_U_ELR: ctl_u_elr_s(c, a, b); break;
      case CTL_U_REO: ctl_u_reo_s(c, a); break;
      default: break;
    }
  }
  __syncthreads();
  const int pend = c.pending;
  if (pend != PEND_NONE) {  // uniform: every thread read the same LDS word
    if (pend == PEND_NU)
      par_update_nu(c, lds);
    else
      par_update_mu(c, lds);
    __syncthreads();
    if (threadIdx.x == 0) {
      c.pending = PEND_NONE;
      if (pend == PEND_NU)
        ctl_v_nu_post(c);
      else
        ctl_u_mu_post(c);
    }
  }
  __syncthreads();
}

// Word w of the staging of step J: its device address and its LDS address.
// Words [0, kCtlHdrWords) are the scalars, then the four double windows of
// wd doubles each, then the INDEX window.
__device__ __forceinline__ void ctl_word(int w, LbdCtl* gc, char* garr, int cap, int wd,
                                        LbdCtlLds& c, int4** src, int4** dst) {
  if (w < kCtlHdrWords) {
    *src = reinterpret_cast<int4*>(gc) + w;
    *dst = reinterpret_cast<int4*>(static_cast<LbdCtl*>(&c)) + w;
    return;
  }
  w -= kCtlHdrWords;
  const int hw = wd / 2;
  if (w < 4 * hw) {
    const int a = w / hw, o = w - a * hw;
    *src = reinterpret_cast<int4*>(CtlArrays::mu(garr, cap) + a * CtlArrays::dlen(cap)) + o;
    *dst = reinterpret_cast<int4*>(c.mu + a * kCtlWinD) + o;
    return;
  }
  w -= 4 * hw;
  *src = reinterpret_cast<int4*>(CtlArrays::index(garr, cap)) + w;
  *dst = reinterpret_cast<int4*>(c.index) + w;
}

// Control stage of step J for the state gc (scalars) / garr (arrays of
// capacity cap), both in device memory: stage the state in LDS with 16-byte
// loads issued before the partial reductions, reduce a = sum pa[0:na) (and
// b = sum pb[0:nb) when pb is set) in index order, run the stage
// (ctl_run), write the state back.  Steps past kCtlLdsJ stage the scalars
// only and work on the arrays in device memory.  lds holds
// (kBlock / kWave) * 2 doubles.
__device__ __forceinline__ void ctl_stage_body(LbdCtl* __restrict__ gc, char* garr, int cap,
                                               int J, int kind, const double* pa, int na,
                                               const double* pb, int nb, LbdCtlLds& c,
                                               LbdCtlRef& cg, double* lds) {
  const bool in_lds = J <= kCtlLdsJ;
  const int wd = (J + 4) & ~1, wi = (2 * J + 8) & ~3;
  const int nw = in_lds ? kCtlHdrWords + 2 * wd + wi / 4 : kCtlHdrWords;
  const int t = static_cast<int>(threadIdx.x);
  // Word indices past the end are clamped to the last word, so every load
  // is unconditional and all of them are in flight at once; the clamped
  // lanes store that word's own value again.  The words wait in registers
  // while the partials are reduced, so the state and the partials arrive in
  // one memory round trip.  (six named words: an array here was placed in
  // scratch memory)
  int4 *s0, *s1, *s2, *s3, *s4, *s5, *d0, *d1, *d2, *d3, *d4, *d5;
  ctl_word(min(t, nw - 1), gc, garr, cap, wd, c, &s0, &d0);
  ctl_word(min(t + kBlock, nw - 1), gc, garr, cap, wd, c, &s1, &d1);
  ctl_word(min(t + 2 * kBlock, nw - 1), gc, garr, cap, wd, c, &s2, &d2);
  ctl_word(min(t + 3 * kBlock, nw - 1), gc, garr, cap, wd, c, &s3, &d3);
  ctl_word(min(t + 4 * kBlock, nw - 1), gc, garr, cap, wd, c, &s4, &d4);
  ctl_word(min(t + 5 * kBlock, nw - 1), gc, garr, cap, wd, c, &s5, &d5);
  const int4 w0 = *s0, w1 = *s1, w2 = *s2, w3 = *s3, w4 = *s4, w5 = *s5;
  double a, b = 0.0;
  if (pb && nb == na) {
    const double* const q[2] = {pa, pb};
    double v[2];
    reduce_partials_n<2>(q, na, v, lds);
    a = v[0];
    b = v[1];
  } else {
    a = reduce_partials(pa, na, lds);
    if (pb) b = reduce_partials(pb, nb, lds);
  }
  *d0 = w0;
  *d1 = w1;
  *d2 = w2;
  *d3 = w3;
  *d4 = w4;
  *d5 = w5;
  __syncthreads();
  if (in_lds) {
    if (threadIdx.x == 0) c.cols = CtlArrays::cols(garr, cap);
    ctl_run(c, kind, a, b, lds);
  } else {
    if (threadIdx.x == 0) {
      *static_cast<LbdCtl*>(&cg) = *static_cast<LbdCtl*>(&c);
      cg.bind(garr);
    }
    __syncthreads();
    ctl_run(cg, kind, a, b, lds);
    if (threadIdx.x == 0) *static_cast<LbdCtl*>(&c) = *static_cast<LbdCtl*>(&cg);
    __syncthreads();
  }
  *s0 = *d0;
  *s1 = *d1;
  *s2 = *d2;
  *s3 = *d3;
  *s4 = *d4;
  *s5 = *d5;
}

}  // namespace pk
