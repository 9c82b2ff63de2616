// Shared definitions for the MI355X (gfx950) truncated-SVD engine.
//
// Conventions
//   * Dense vectors and Lanczos bases are column-major with a padded leading
//     dimension (a multiple of 64 elements), 0-based in C++.
//   * "T" is the storage type of the basis and of the operator values
//     (double, or float in the FP32-basis form).  Every reduction and every
//     dot product accumulates in double.
//   * Reductions are two-stage and deterministic: a kernel writes one partial
//     per workgroup (fixed grid, fixed in-block order), and a second kernel
//     sums the partials in index order.  The same inputs give bit-identical
//     results run to run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pk {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // threads per workgroup for streaming kernels
constexpr int kMaxBlocks = 2048;   // fixed upper bound on the streaming grid (8 per CU)
constexpr int kMaxCols = 512;      // most basis columns one reorthogonalization can touch

// Scalar operand of a kernel: read from device memory when p != nullptr
// (device-resident control), otherwise the by-value v (host control).
struct Scal {
  const double* p;
  double v;
  __device__ __forceinline__ double get() const { return p ? *p : v; }
};
inline Scal sval(double v) { return Scal{nullptr, v}; }
inline Scal sptr(const double* p) { return Scal{p, 0.0}; }


__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over a group of g lanes (g a power of two <= 64, groups aligned).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic block reduction of NV values held by every thread; thread 0
// receives the sums.  lds must hold (kBlock / kWave) * NV doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  constexpr int nw = kBlock / kWave;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[wid * NV + i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[w * NV + i];
      v[i] = s;
    }
  }
  __syncthreads();
}

// Sums of NV arrays of n partials each by one 256-thread workgroup in a
// fixed order (thread t adds p[t], p[t+256], ... in that order; then the
// block reduction); every caller that reduces the same partials gets the same
// bits.  Eight loads per array are issued ahead of their additions, so a
// thread waits for one memory latency per eight partials, not per partial.
// lds must hold (kBlock / kWave) * NV doubles; the sums land in thread 0.
template <int NV>
__device__ __forceinline__ void reduce_partials_n(const double* const (&p)[NV], int n,
                                                  double (&v)[NV], double* lds) {
  constexpr int U = 8;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int i = threadIdx.x; i < n; i += U * kBlock) {
    // indices past the end load the last partial (unconditional loads, all
    // in flight together) and are not added
    double t[NV][U];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u) t[k][u] = p[k][min(i + u * kBlock, n - 1)];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i + u * kBlock < n) v[k] += t[k][u];
  }
  block_sum<NV>(v, lds);
}

__device__ __forceinline__ double reduce_partials(const double* p, int n, double* lds) {
  const double* const q[1] = {p};
  double v[1];
  reduce_partials_n<1>(q, n, v, lds);
  return v[0];  // valid in thread 0
}

inline int round_up(int x, int a) { return (x + a - 1) / a * a; }
inline int cdiv(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace pk
