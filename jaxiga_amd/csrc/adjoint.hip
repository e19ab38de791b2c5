// Implicit differentiation of the Galerkin solve, and fields at points.
//
// The reverse pass of K(theta) u = f(theta) is one adjoint solve K lam = u_bar
// on the same symmetric CSR (a Galerkin operator is symmetric, so no
// transposed matrix is formed), followed by
//   * the cotangent of the CSR values, data_bar[k] = -lam[row k] u[col k]
//     (iga_csr_outer), and
//   * the point sensitivities of the element kernels, formed from lam and
//     the full solution u (prescribed values included) at the quadrature
//     points: values and gradients of both fields from the basis data
//     (iga_field_points). The same kernel evaluates a solution and its
//     gradient at any point set for post-processing.
#include "common.hpp"

namespace iga {

// out[k] = alpha * a[rows[k]] * b[cols[k]]
__global__ void __launch_bounds__(256)
    k_csr_outer(int64_t nnz, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                const double* __restrict__ a, const double* __restrict__ b, double alpha,
                double* __restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x)
    out[k] = alpha * a[rows[k]] * b[cols[k]];
}

// One thread per point p = (element e, point q): for each field f,
//   val[f][p][c]     = sum_a R[p][a] u_f[dofs[e][a] * VEC + c]
//   grad[f][p][c][d] = sum_a dR[p][d][a] u_f[dofs[e][a] * VEC + c]
// R is (n_p, n_l), dR (n_p, DIM, n_l): each thread reads its own contiguous
// rows, so consecutive threads read consecutive memory.
template <int VEC, int DIM, int NF>
__global__ void __launch_bounds__(256)
    k_field_points(int64_t n_p, int n_q, int n_l, const double* __restrict__ R,
                   const double* __restrict__ dR, const int32_t* __restrict__ dofs,
                   const double* __restrict__ u0, const double* __restrict__ u1,
                   double* __restrict__ val, double* __restrict__ grad) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_p) return;
  const int64_t e = p / n_q;
  const double* Rp = R + p * n_l;
  const double* dRp = dR ? dR + p * DIM * n_l : nullptr;
  const int32_t* dp = dofs + e * n_l;
  double v[NF][VEC], g[NF][VEC][DIM];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int c = 0; c < VEC; ++c) {
      v[f][c] = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) g[f][c][d] = 0.0;
    }
#pragma unroll 4
  for (int a = 0; a < n_l; ++a) {
    const int64_t base = (int64_t)dp[a] * VEC;
    double w[NF][VEC];
#pragma unroll
    for (int c = 0; c < VEC; ++c) {
      w[0][c] = u0[base + c];
      if (NF > 1) w[NF - 1][c] = u1[base + c];
    }
    if (val) {
      const double r = Rp[a];
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int c = 0; c < VEC; ++c) v[f][c] += r * w[f][c];
    }
    if (dRp) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const double r = dRp[d * n_l + a];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int c = 0; c < VEC; ++c) g[f][c][d] += r * w[f][c];
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if (val)
#pragma unroll
      for (int c = 0; c < VEC; ++c) val[((int64_t)f * n_p + p) * VEC + c] = v[f][c];
    if (grad && dRp)
#pragma unroll
      for (int c = 0; c < VEC; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) grad[(((int64_t)f * n_p + p) * VEC + c) * DIM + d] = g[f][c][d];
  }
}

// The same fields through LDS, for elements whose basis data fits: a block
// takes `eb` consecutive elements, copies their R / dR rows (contiguous in
// HBM) into padded LDS rows with batched loads, gathers the coefficients of
// their functions once, and each thread forms rows (e, q) of values and
// (e, q, d) of gradients from LDS. Rows are padded to n_l + 1 doubles so
// that threads reading different rows hit different banks.
constexpr int kFpThreads = 256;
constexpr int kFpBatch = 8;

__device__ __forceinline__ void fp_stage_rows(const double* __restrict__ src, int tot, int n_l,
                                              double* __restrict__ dst) {
  const int ld = n_l + 1;
  for (int base = 0; base < tot; base += kFpThreads * kFpBatch) {
    double v[kFpBatch];
#pragma unroll
    for (int k = 0; k < kFpBatch; ++k) {
      const int i = base + k * kFpThreads + (int)threadIdx.x;
      v[k] = src[min(i, tot - 1)];
    }
#pragma unroll
    for (int k = 0; k < kFpBatch; ++k) {
      const int i = base + k * kFpThreads + (int)threadIdx.x;
      if (i < tot) dst[(i / n_l) * ld + i % n_l] = v[k];
    }
  }
}

template <int VEC, int DIM, int NF>
__global__ void __launch_bounds__(kFpThreads)
    k_field_points_lds(int n_e, int n_q, int n_l, int eb, const double* __restrict__ R,
                       const double* __restrict__ dR, const int32_t* __restrict__ dofs,
                       const double* __restrict__ u0, const double* __restrict__ u1,
                       double* __restrict__ val, double* __restrict__ grad) {
  extern __shared__ double fp_smem[];
  const int e0 = blockIdx.x * eb;
  const int ne = min(eb, n_e - e0);
  const int ld = n_l + 1;
  const int64_t n_p = (int64_t)n_e * n_q;
  double* cs = fp_smem;                          // [NF][eb * n_l][VEC]
  double* rs = cs + NF * eb * n_l * VEC;         // [eb * n_q][ld]        (values)
  double* gs = rs + (val ? eb * n_q * ld : 0);   // [eb * n_q * DIM][ld]  (gradients)
  for (int i = threadIdx.x; i < ne * n_l; i += kFpThreads) {
    const int64_t cp = dofs[(int64_t)e0 * n_l + i];
#pragma unroll
    for (int c = 0; c < VEC; ++c) {
      cs[i * VEC + c] = u0[cp * VEC + c];
      if (NF > 1) cs[(eb * n_l + i) * VEC + c] = u1[cp * VEC + c];
    }
  }
  if (val) fp_stage_rows(R + (int64_t)e0 * n_q * n_l, ne * n_q * n_l, n_l, rs);
  if (grad) fp_stage_rows(dR + (int64_t)e0 * n_q * DIM * n_l, ne * n_q * DIM * n_l, n_l, gs);
  __syncthreads();
  if (val)
    for (int r = threadIdx.x; r < ne * n_q; r += kFpThreads) {
      const int e = r / n_q;
      double acc[NF][VEC];
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc[f][c] = 0.0;
      for (int a = 0; a < n_l; ++a) {
        const double w = rs[r * ld + a];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int c = 0; c < VEC; ++c) acc[f][c] += w * cs[((f * eb + e) * n_l + a) * VEC + c];
      }
      const int64_t p = (int64_t)e0 * n_q + r;
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int c = 0; c < VEC; ++c) val[((int64_t)f * n_p + p) * VEC + c] = acc[f][c];
    }
  if (grad)
    for (int r = threadIdx.x; r < ne * n_q * DIM; r += kFpThreads) {
      const int e = r / (n_q * DIM), d = r % DIM;
      double acc[NF][VEC];
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int c = 0; c < VEC; ++c) acc[f][c] = 0.0;
      for (int a = 0; a < n_l; ++a) {
        const double w = gs[r * ld + a];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int c = 0; c < VEC; ++c) acc[f][c] += w * cs[((f * eb + e) * n_l + a) * VEC + c];
      }
      const int64_t p = (int64_t)e0 * n_q + r / DIM;
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int c = 0; c < VEC; ++c) grad[(((int64_t)f * n_p + p) * VEC + c) * DIM + d] = acc[f][c];
    }
}

// elements per block of the LDS kernel (0: the basis data of one element
// does not fit in 64 KB of LDS)
static int field_points_eb(int n_q, int n_l, int dim, int vec, int nf, bool val, bool grad) {
  const int64_t per = (int64_t)nf * n_l * vec + (val ? (int64_t)n_q * (n_l + 1) : 0) +
                      (grad ? (int64_t)n_q * dim * (n_l + 1) : 0);
  const int64_t fit = (64 * 1024) / (8 * per);
  const int rows = n_q * (grad ? dim : 1);
  const int want = std::max(1, kFpThreads / std::max(rows, 1));
  return (int)std::min<int64_t>(fit, want);
}

template <int VEC, int DIM>
static int field_points_nf(int nf, int64_t n_p, int n_q, int n_l, const double* R, const double* dR,
                           const int32_t* dofs, const double* u0, const double* u1, double* val,
                           double* grad, hipStream_t st) {
  const int n_e = (int)(n_p / n_q);
  const int eb = field_points_eb(n_q, n_l, DIM, VEC, nf, val != nullptr, grad != nullptr);
  if (eb >= 1) {
    const int64_t per = (int64_t)nf * n_l * VEC + (val ? (int64_t)n_q * (n_l + 1) : 0) +
                        (grad ? (int64_t)n_q * DIM * (n_l + 1) : 0);
    const size_t lds = (size_t)(8 * per * eb);
    const dim3 g((unsigned)((n_e + eb - 1) / eb)), b(kFpThreads);
    if (nf == 1)
      hipLaunchKernelGGL((k_field_points_lds<VEC, DIM, 1>), g, b, lds, st, n_e, n_q, n_l, eb, R, dR,
                         dofs, u0, u1, val, grad);
    else
      hipLaunchKernelGGL((k_field_points_lds<VEC, DIM, 2>), g, b, lds, st, n_e, n_q, n_l, eb, R, dR,
                         dofs, u0, u1, val, grad);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }
  const dim3 grid((unsigned)((n_p + 255) / 256)), block(256);
  if (nf == 1)
    hipLaunchKernelGGL((k_field_points<VEC, DIM, 1>), grid, block, 0, st, n_p, n_q, n_l, R, dR, dofs,
                       u0, u1, val, grad);
  else
    hipLaunchKernelGGL((k_field_points<VEC, DIM, 2>), grid, block, 0, st, n_p, n_q, n_l, R, dR, dofs,
                       u0, u1, val, grad);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

}  // namespace iga

using namespace iga;

extern "C" int iga_csr_outer(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* a,
                             const double* b, double alpha, double* out, void* stream) {
  if (nnz == 0) return IGA_OK;
  if (nnz < 0 || !rows || !cols || !a || !b || !out) return IGA_ERR_ARG;
  const int64_t blocks = std::min<int64_t>((nnz + 255) / 256, 16384);
  hipLaunchKernelGGL(k_csr_outer, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, nnz,
                     rows, cols, a, b, alpha, out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_field_points(int32_t n_e, int32_t n_q, int32_t n_l, int32_t dim, int32_t vec,
                                const double* R, const double* dR, const int32_t* dofs,
                                const double* u0, const double* u1, double* val, double* grad,
                                void* stream) {
  if (n_e == 0 || n_q == 0) return IGA_OK;
  if (n_e < 0 || n_q < 0 || n_l <= 0 || !dofs || !u0) return IGA_ERR_ARG;
  if ((val && !R) || (grad && !dR) || (!val && !grad)) return IGA_ERR_ARG;
  const int nf = u1 ? 2 : 1;
  const int64_t n_p = (int64_t)n_e * n_q;
  hipStream_t st = (hipStream_t)stream;
  const double* Ru = val ? R : nullptr;
  const double* dRu = grad ? dR : nullptr;
  if (dim == 2 && vec == 1) return field_points_nf<1, 2>(nf, n_p, n_q, n_l, Ru, dRu, dofs, u0, u1, val, grad, st);
  if (dim == 2 && vec == 2) return field_points_nf<2, 2>(nf, n_p, n_q, n_l, Ru, dRu, dofs, u0, u1, val, grad, st);
  if (dim == 3 && vec == 1) return field_points_nf<1, 3>(nf, n_p, n_q, n_l, Ru, dRu, dofs, u0, u1, val, grad, st);
  if (dim == 3 && vec == 3) return field_points_nf<3, 3>(nf, n_p, n_q, n_l, Ru, dRu, dofs, u0, u1, val, grad, st);
  return IGA_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
// Shape cotangent: point "forces" pulled back to the control points.
//
// A first variation of the geometry moves the control point k by dX_k; at
// each Gauss point the map moves by R_k dX_k, so a point functional whose
// derivative with respect to the point's position is xbar and whose
// derivative with respect to the map's gradient is carried by the tensor P
// (P[c][j] multiplies d(grad R_k)_j for a motion along c) pulls back to
//   G[dofs[e][k]][c] += sum_q ( sum_j P[e][q][c][j] dR[e][q][j][k]
//                               + R[e][q][k] xbar[e][q][c] ).
// One thread per (element, local function); FP64 atomics into G (n_cp, DIM).
// ---------------------------------------------------------------------------

namespace iga {

template <int DIM>
__global__ void __launch_bounds__(256)
    k_point_forces(int64_t n_e, int n_q, int n_l, const double* __restrict__ R,
                   const double* __restrict__ dR, const double* __restrict__ P,
                   const double* __restrict__ xbar, const int32_t* __restrict__ dofs,
                   double* __restrict__ G) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_e * n_l) return;
  const int64_t e = t / n_l;
  const int k = (int)(t % n_l);
  double g[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) g[c] = 0.0;
  for (int q = 0; q < n_q; ++q) {
    const int64_t p = e * n_q + q;
    double d[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) d[j] = dR[(p * DIM + j) * n_l + k];
    const double r = xbar ? R[p * n_l + k] : 0.0;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      double s = xbar ? r * xbar[p * DIM + c] : 0.0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) s += P[(p * DIM + c) * DIM + j] * d[j];
      g[c] += s;
    }
  }
  const int64_t cp = dofs[e * n_l + k];
#pragma unroll
  for (int c = 0; c < DIM; ++c) atomicAdd(G + cp * DIM + c, g[c]);
}

}  // namespace iga

extern "C" int iga_point_forces(int32_t n_e, int32_t n_q, int32_t n_l, int32_t dim, const double* R,
                                const double* dR, const double* P, const double* xbar,
                                const int32_t* dofs, double* G, void* stream) {
  if (!dR || !P || !dofs || !G || (xbar && !R)) return IGA_ERR_ARG;
  if (dim != 2 && dim != 3) return IGA_ERR_UNSUPPORTED;
  const int64_t n = (int64_t)n_e * n_l;
  if (n == 0) return IGA_OK;
  hipStream_t st = (hipStream_t)stream;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (dim == 2)
    hipLaunchKernelGGL(k_point_forces<2>, dim3(blocks), dim3(256), 0, st, (int64_t)n_e, n_q, n_l, R,
                       dR, P, xbar, dofs, G);
  else
    hipLaunchKernelGGL(k_point_forces<3>, dim3(blocks), dim3(256), 0, st, (int64_t)n_e, n_q, n_l, R,
                       dR, P, xbar, dofs, G);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}
