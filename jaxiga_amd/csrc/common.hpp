// Shared device helpers for the gfx950 Galerkin kernels: compile-time shape
// arithmetic, Bernstein evaluation, wave reductions, the NURBS quotient with
// the FP64 push-forward, and the Krylov building blocks (CSR row products,
// deterministic per-block partial reductions, device-resident solver state).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "../../include/jaxiga_hip.h"

namespace iga {

constexpr int kWave = 64;

// Orders the LDS accesses of one wavefront: lanes exchange data through LDS
// with no workgroup barrier (LDS executes a wave's operations in order; the
// fence keeps the compiler from moving accesses across this point).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
__host__ __device__ constexpr int pow2ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

#define IGA_CHECK_LAUNCH()                              \
  do {                                                  \
    if (hipGetLastError() != hipSuccess) return (int)IGA_ERR_HIP; \
  } while (0)

// Degree-P Bernstein polynomials on [-1,1] and their derivatives at u, built
// by the two-term recurrence B^j_k = s B^{j-1}_k + t B^{j-1}_{k-1} with
// s = (1-u)/2, t = (1+u)/2; dB_k = P/2 (B^{P-1}_{k-1} - B^{P-1}_k).
template <int P>
__device__ __forceinline__ void bernstein(double u, double* B, double* dB) {
  const double s = 0.5 * (1.0 - u), t = 0.5 * (1.0 + u);
  double low[P + 1];
  low[0] = 1.0;
#pragma unroll
  for (int j = 1; j < P; ++j) {
    low[j] = t * low[j - 1];
#pragma unroll
    for (int k = j - 1; k >= 1; --k) low[k] = s * low[k] + t * low[k - 1];
    low[0] = s * low[0];
  }
#pragma unroll
  for (int k = 0; k <= P; ++k) {
    const double a = (k >= 1) ? low[k - 1] : 0.0;
    const double b = (k < P) ? low[k] : 0.0;
    B[k] = s * b + t * a;
    dB[k] = 0.5 * P * (a - b);
  }
}

template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = W / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, W);
  return v;
}

// Block-wide FP64 sum (blockDim.x a multiple of 64, <= 1024); result valid in
// thread 0.
__device__ __forceinline__ double block_sum(double v, double* scratch32) {
  v = group_sum<kWave>(v);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) scratch32[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / kWave;
    for (int i = 0; i < nw; ++i) s += scratch32[i];
  }
  __syncthreads();
  return s;
}

__device__ __forceinline__ int lower_bound(const int32_t* a, int lo, int hi, int key) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}


// ---------------------------------------------------------------------------
// Krylov building blocks shared by the CG / FCG drivers
// ---------------------------------------------------------------------------

// LPR lanes per row; returns the row's value in every lane of the group.
template <int LPR, typename TV = double>
__device__ __forceinline__ double row_dot(int r, const int32_t* __restrict__ indptr,
                                          const int32_t* __restrict__ cols,
                                          const TV* __restrict__ vals,
                                          const double* __restrict__ x, int l) {
  const int lo = indptr[r], hi = indptr[r + 1];
  double s = 0.0;
  for (int k = lo + l; k < hi; k += LPR) s += (double)vals[k] * x[cols[k]];
  return group_sum<LPR>(s);
}

// FP32 row product (FP32 values and vector, FP32 accumulation)
template <int LPR>
__device__ __forceinline__ float row_dot_f(int r, const int32_t* __restrict__ indptr,
                                           const int32_t* __restrict__ cols,
                                           const float* __restrict__ vals,
                                           const float* __restrict__ x, int l) {
  const int lo = indptr[r], hi = indptr[r + 1];
  float s = 0.f;
  for (int k = lo + l; k < hi; k += LPR) s += vals[k] * x[cols[k]];
#pragma unroll
  for (int off = LPR / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, LPR);
  return s;
}

// Lanes per CSR row: several nonzeros per lane, so each lane has
// independent loads in flight (measured on the 2-D elasticity and Poisson
// operators: 8 or 16 lanes beat 32 and 64).
inline int pick_lpr(int n, int nnz) {
  const double avg = n > 0 ? double(nnz) / n : 0.0;
  return avg > 64 ? 16 : 8;
}


template <typename F>
inline int by_lpr(int lpr, F f) {
  switch (lpr) {
    case 8: return f(std::integral_constant<int, 8>{});
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    default: return f(std::integral_constant<int, 64>{});
  }
}

struct CGState {
  double rr, bb, tol2, pad;
  int iter, maxiter, done, pad2;
};

constexpr int kVecThreads = 256;
constexpr int kRedBlocks = 1024;  // upper bound on blocks of any Krylov kernel

struct CGBufs {
  CGState* st;
  double* P_pq;  // [kRedBlocks]
  double* P_rz;  // [2][kRedBlocks]
  double* P_rr;  // [kRedBlocks]
  double* P_bb;  // [kRedBlocks]
};

// Sum of n partials, identical in every block; valid in all threads.
__device__ __forceinline__ double reduce_partials(const double* parts, int n) {
  __shared__ double scratch[16];
  __shared__ double result;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += parts[i];
  s = block_sum(s, scratch);
  if (threadIdx.x == 0) result = s;
  __syncthreads();
  const double r = result;
  __syncthreads();
  return r;
}

// reduce_partials in two halves, so that a kernel can issue the loads of the
// partials together with its other loads and reduce them later: this thread's
// share (at most K partials per thread, n <= K * blockDim.x), then the sum of
// the shares. The summation order is reduce_partials' own, so both give the
// same bits.
template <int K>
__device__ __forceinline__ double partials_share(const double* parts, int n) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * blockDim.x;
    v[k] = i < n ? parts[i] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) s += v[k];
  return s;
}

__device__ __forceinline__ double partials_total(double s) {
  __shared__ double scratch[16];
  __shared__ double result;
  s = block_sum(s, scratch);
  if (threadIdx.x == 0) result = s;
  __syncthreads();
  const double r = result;
  __syncthreads();
  return r;
}

__device__ __forceinline__ void store_partial(double* parts, double v) {
  __shared__ double scratch[16];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) parts[blockIdx.x] = v;
}

inline int vec_blocks(int64_t n) {
  const int64_t b = (n + kVecThreads - 1) / kVecThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kRedBlocks));
}

inline int row_blocks(int64_t n, int lpr) {
  const int64_t b = (n * lpr + kVecThreads - 1) / kVecThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kRedBlocks));
}

}  // namespace iga
