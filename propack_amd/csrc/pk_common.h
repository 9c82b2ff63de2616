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

// Deterministic block reduction of NV values held by every thread of an
// NT-thread workgroup; thread 0 receives the sums.  lds must hold
// (NT / kWave) * NV doubles.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  constexpr int nw = NT / kWave;
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

// xLARUV's 48-bit multiplicative congruential generator (LAPACK
// double/Lapack_Util/dlaruv.f:42-46 tabulates powers of this multiplier):
// state s' = s * a mod 2^48, seed (1,3,5,7) as 12-bit limbs.  The low 64 bits
// of a 64-bit product, masked to 48 bits, are the product mod 2^48.
constexpr unsigned long long kLcgMask = (1ULL << 48) - 1;
constexpr unsigned long long kLcgA = 33952834046453ULL;
constexpr unsigned long long kLcgSeed = (1ULL << 36) + (3ULL << 24) + (5ULL << 12) + 7ULL;
__host__ __device__ inline unsigned long long lcg_mulmod(unsigned long long x,
                                                         unsigned long long y) {
  return (x * y) & kLcgMask;
}
// a^e mod 2^48
__host__ __device__ inline unsigned long long lcg_pow(unsigned long long e) {
  unsigned long long r = 1, b = kLcgA;
  for (; e; e >>= 1) {
    if (e & 1) r = lcg_mulmod(r, b);
    b = lcg_mulmod(b, b);
  }
  return r;
}
// 2 u - 1 for the state s, u converted from the four 12-bit limbs of s in
// Horner form in T (xLARUV, then xLARNV IDIST=2); no contraction, so float
// rounds as the reference's REAL arithmetic.
template <typename T>
__host__ __device__ inline T lcg_pm1(unsigned long long s) {
#pragma clang fp contract(off)
  const T r = T(1) / T(4096);
  const T i1 = static_cast<T>(s >> 36), i2 = static_cast<T>((s >> 24) & 4095);
  const T i3 = static_cast<T>((s >> 12) & 4095), i4 = static_cast<T>(s & 4095);
  const T u = r * (i1 + r * (i2 + r * (i3 + r * i4)));
  return T(2) * u - T(1);
}

inline int round_up(int x, int a) { return (x + a - 1) / a * a; }
inline int cdiv(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace pk
