// Compressible neo-Hookean solid: element residuals and tangents for the
// Newton path of the Galerkin method.
//
// psi(F) = mu/2 (tr F^T F - d) - mu ln J + lam/2 (ln J)^2, F = I + grad u,
// first Piola stress P = mu F + c F^-T with c = lam ln J - mu, and tangent
//   dP_iJ/dF_kL = mu d_ik d_JL + lam Fi_Ji Fi_Lk - c Fi_Jk Fi_Li  (Fi = F^-1).
// With g_a = F^-T grad R_a at a point, the element blocks are
//   r[(a,i)]       = sum_q w (P_iJ d_J R_a - f_i R_a)
//   K[(a,i),(b,k)] = sum_q w (mu d_ik grad R_a . grad R_b + lam g_a[i] g_b[k]
//                             - c g_a[k] g_b[i])
// The basis data (R, grad R, w at the Gauss points) is evaluated once per
// geometry and reused every Newton step. One workgroup per element: the
// point quantities and g_a are staged in LDS, the residual rows are added
// into the free residual with FP64 atomics, and the element matrix is
// written in the row layout the CSR gather of iga_assemble reads (row
// a*DIM + i), so the tangent reuses the pattern and the gather every step.
#include "common.hpp"

namespace iga {

template <int DIM>
__global__ void __launch_bounds__(256)
    k_neohooke(int n_q, int n_l, const double* __restrict__ R, const double* __restrict__ dR,
               const double* __restrict__ w, const int32_t* __restrict__ edofs,
               const double* __restrict__ u, const int32_t* __restrict__ f2f, double mu, double lam,
               const double* __restrict__ src, double* __restrict__ rfree, double* __restrict__ ke) {
  extern __shared__ double sm[];
  const int e = blockIdx.x;
  const int NLV = n_l * DIM;
  double* sP = sm;                          // [n_q][DIM][DIM]
  double* sc = sP + n_q * DIM * DIM;        // [n_q] c
  double* sf = sc + n_q;                    // [n_q][DIM] source
  double* sg = sf + n_q * DIM;              // [n_q][n_l][DIM] g_a
  double* sd = sg + n_q * n_l * DIM;        // [n_q][DIM][n_l] grad R
  const double* dRe = dR + (size_t)e * n_q * DIM * n_l;
  for (int t = threadIdx.x; t < n_q * DIM * n_l; t += blockDim.x) sd[t] = dRe[t];
  __syncthreads();
  for (int q = threadIdx.x; q < n_q; q += blockDim.x) {
    double F[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) F[i][j] = i == j ? 1.0 : 0.0;
    for (int a = 0; a < n_l; ++a) {
      const int64_t base = (int64_t)edofs[(size_t)e * n_l + a] * DIM;
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        const double ui = u[base + i];
#pragma unroll
        for (int j = 0; j < DIM; ++j) F[i][j] += ui * sd[(q * DIM + j) * n_l + a];
      }
    }
    double Fi[DIM][DIM], J;
    if constexpr (DIM == 2) {
      J = F[0][0] * F[1][1] - F[0][1] * F[1][0];
      Fi[0][0] = F[1][1] / J;
      Fi[0][1] = -F[0][1] / J;
      Fi[1][0] = -F[1][0] / J;
      Fi[1][1] = F[0][0] / J;
    } else {
      const double c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
      const double c01 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
      const double c02 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
      J = F[0][0] * c00 + F[0][1] * c01 + F[0][2] * c02;
      Fi[0][0] = c00 / J;
      Fi[1][0] = c01 / J;
      Fi[2][0] = c02 / J;
      Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) / J;
      Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) / J;
      Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) / J;
      Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) / J;
      Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) / J;
      Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) / J;
    }
    // log of a non-positive J is NaN (or -inf): the step is then rejected
    const double c = lam * log(J) - mu;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) sP[(q * DIM + i) * DIM + j] = mu * F[i][j] + c * Fi[j][i];
    sc[q] = c;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
      sf[q * DIM + i] = src ? src[((size_t)e * n_q + q) * DIM + i] : 0.0;
    for (int a = 0; a < n_l; ++a)
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) s += Fi[j][i] * sd[(q * DIM + j) * n_l + a];
        sg[(q * n_l + a) * DIM + i] = s;
      }
  }
  __syncthreads();
  const double* Re = R + (size_t)e * n_q * n_l;
  const double* we = w + (size_t)e * n_q;
  if (rfree) {
    for (int t = threadIdx.x; t < NLV; t += blockDim.x) {
      const int a = t / DIM, i = t % DIM;
      double r = 0.0;
      for (int q = 0; q < n_q; ++q) {
        double s = -sf[q * DIM + i] * Re[q * n_l + a];
#pragma unroll
        for (int j = 0; j < DIM; ++j) s += sP[(q * DIM + i) * DIM + j] * sd[(q * DIM + j) * n_l + a];
        r += we[q] * s;
      }
      const int f = f2f[(int64_t)edofs[(size_t)e * n_l + a] * DIM + i];
      if (f >= 0) atomicAdd(rfree + f, r);
    }
  }
  if (ke) {
    double* kout = ke + (size_t)e * NLV * NLV;
    for (int t = threadIdx.x; t < NLV * NLV; t += blockDim.x) {
      const int row = t / NLV, col = t % NLV;
      const int a = row / DIM, i = row % DIM, b = col / DIM, k = col % DIM;
      double s = 0.0;
      for (int q = 0; q < n_q; ++q) {
        const double* ga = sg + (q * n_l + a) * DIM;
        const double* gb = sg + (q * n_l + b) * DIM;
        double v = lam * ga[i] * gb[k] - sc[q] * ga[k] * gb[i];
        if (i == k) {
          double dd = 0.0;
#pragma unroll
          for (int j = 0; j < DIM; ++j) dd += sd[(q * DIM + j) * n_l + a] * sd[(q * DIM + j) * n_l + b];
          v += mu * dd;
        }
        s += we[q] * v;
      }
      kout[t] = s;
    }
  }
}

}  // namespace iga

using namespace iga;

extern "C" int iga_neohooke(int32_t n_e, int32_t n_q, int32_t n_l, int32_t dim, const double* R,
                            const double* dR, const double* w, const int32_t* edofs,
                            const double* u, const int32_t* full_to_free, double mu, double lam,
                            const double* src, double* rfree, double* ke, void* stream) {
  if (!R || !dR || !w || !edofs || !u || !full_to_free || (!rfree && !ke)) return IGA_ERR_ARG;
  if (dim != 2 && dim != 3) return IGA_ERR_UNSUPPORTED;
  if (n_e == 0) return IGA_OK;
  const size_t lds = sizeof(double) * ((size_t)n_q * dim * dim + n_q + (size_t)n_q * dim +
                                       2 * (size_t)n_q * n_l * dim);
  if (lds > 64 * 1024) return IGA_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  if (dim == 2)
    hipLaunchKernelGGL(k_neohooke<2>, dim3(n_e), dim3(256), lds, st, n_q, n_l, R, dR, w, edofs, u,
                       full_to_free, mu, lam, src, rfree, ke);
  else
    hipLaunchKernelGGL(k_neohooke<3>, dim3(n_e), dim3(256), lds, st, n_q, n_l, R, dR, w, edofs, u,
                       full_to_free, mu, lam, src, rfree, ke);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}
