// The element-list path: the padded-element contract and chunked assembly.
//
// A space is given by its element connectivity elem_dofs (n_e, n_lmax), in
// which short elements are padded with the sentinel n_scalar (one past the
// last scalar basis function), and by its per-element Bezier extraction
// (n_e, n_lmax, NB), NB = (p+1)^dim, with zero rows under the padding, so a
// padded function contributes nothing. This is the layout hierarchical
// spaces use; a tensor space is the case without padding.
//
//  * iga_elem_pattern: the CSR pattern of the free-free block and the
//    element-to-slot scatter map, as sorted unique keys row * n_free + col
//    over every (element, local row, local column): keys, radix sort,
//    unique, then row pointers and slots by binary search. The sentinel and
//    prescribed dofs map to the slot nnz.
//  * iga_elem_basis: basis data of a range of elements at the tensor Gauss
//    points from the extraction rows (rational quotient, FP64 push-forward).
//  * iga_elem_assemble: element tangents and loads of a range of elements
//    from that basis data, scatter-added into the CSR values and the load
//    with FP64 atomics through the scatter map (the lifting -K_fp g
//    included).
// A caller bounds memory by running basis and assembly over element ranges
// (chunks): the basis data and nothing else scale with the chunk.
#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace iga {

constexpr int64_t kNoKey = INT64_MAX;

__device__ __forceinline__ int64_t pair_key(int64_t t, int n_lmax, int vec,
                                            const int32_t* __restrict__ elem_dofs, int n_scalar,
                                            const int32_t* __restrict__ f2f, int64_t n_free) {
  const int nlv = n_lmax * vec;
  const int64_t e = t / ((int64_t)nlv * nlv);
  const int r = (int)((t / nlv) % nlv), c = (int)(t % nlv);
  const int dr = elem_dofs[e * n_lmax + r / vec], dc = elem_dofs[e * n_lmax + c / vec];
  if (dr >= n_scalar || dc >= n_scalar) return kNoKey;
  const int fr = f2f[(int64_t)dr * vec + r % vec], fc = f2f[(int64_t)dc * vec + c % vec];
  if (fr < 0 || fc < 0) return kNoKey;
  return (int64_t)fr * n_free + fc;
}

__global__ void k_pair_keys(int64_t K, int n_lmax, int vec, const int32_t* __restrict__ elem_dofs,
                            int n_scalar, const int32_t* __restrict__ f2f, int64_t n_free,
                            int64_t* __restrict__ keys) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < K;
       t += (int64_t)gridDim.x * blockDim.x)
    keys[t] = pair_key(t, n_lmax, vec, elem_dofs, n_scalar, f2f, n_free);
}

__device__ __forceinline__ int64_t lower_bound64(const int64_t* a, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_pattern_fill(int64_t nnz, const int64_t* __restrict__ ukeys, int64_t n_free,
                               int32_t* __restrict__ rows, int32_t* __restrict__ cols) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz;
       t += (int64_t)gridDim.x * blockDim.x) {
    rows[t] = (int32_t)(ukeys[t] / n_free);
    cols[t] = (int32_t)(ukeys[t] % n_free);
  }
}

__global__ void k_pattern_indptr(int64_t n_free, const int64_t* __restrict__ ukeys, int64_t nnz,
                                 int32_t* __restrict__ indptr) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= n_free;
       r += (int64_t)gridDim.x * blockDim.x)
    indptr[r] = (int32_t)lower_bound64(ukeys, nnz, r * n_free);
}

__global__ void k_pattern_scatter(int64_t K, int n_lmax, int vec, const int32_t* __restrict__ elem_dofs,
                                  int n_scalar, const int32_t* __restrict__ f2f, int64_t n_free,
                                  const int64_t* __restrict__ ukeys, int64_t nnz,
                                  int32_t* __restrict__ scatter) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < K;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t key = pair_key(t, n_lmax, vec, elem_dofs, n_scalar, f2f, n_free);
    scatter[t] = key == kNoKey ? (int32_t)nnz : (int32_t)lower_bound64(ukeys, nnz, key);
  }
}

// ---------------------------------------------------------------------------
// basis data of elements lo .. lo + n_c - 1 at the tensor Gauss points
// ---------------------------------------------------------------------------

template <int DIM, int P>
__global__ void __launch_bounds__(256)
    k_elem_basis(int64_t lo, int64_t n_c, int n_lmax, const double* __restrict__ qp,
                 const double* __restrict__ qw, const double* __restrict__ ext,
                 const int32_t* __restrict__ elem_dofs, int n_scalar, const double* __restrict__ cpts,
                 const double* __restrict__ wgts, const double* __restrict__ boxes,
                 double* __restrict__ R, double* __restrict__ dR, double* __restrict__ x,
                 double* __restrict__ w) {
  constexpr int N1 = P + 1, NB = ipow(N1, DIM), NQ = ipow(N1, DIM);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_c * NQ) return;
  const int64_t c = t / NQ, e = lo + c;
  const int q = (int)(t % NQ);
  const int qi[3] = {q % N1, (q / N1) % N1, q / (N1 * N1)};
  double B1[DIM][N1], dB1[DIM][N1], h[DIM], wq = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    bernstein<P>(qp[qi[d]], B1[d], dB1[d]);
    h[d] = boxes[e * 2 * DIM + DIM + d] - boxes[e * 2 * DIM + d];
    wq *= qw[qi[d]] * 0.5 * h[d];
  }
  double Bt[NB], dBt[DIM][NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int bi[3] = {b % N1, (b / N1) % N1, b / (N1 * N1)};
    double v = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) v *= B1[d][bi[d]];
    Bt[b] = v;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      double g = 2.0 / h[k];
#pragma unroll
      for (int d = 0; d < DIM; ++d) g *= d == k ? dB1[d][bi[d]] : B1[d][bi[d]];
      dBt[k][b] = g;
    }
  }
  const double* Ce = ext + e * (int64_t)n_lmax * NB;
  const int32_t* de = elem_dofs + e * n_lmax;
  // pass 1: W, grad W, sum M X, sum dM X over the element's functions
  double W = 0.0, dW[DIM] = {}, X[DIM] = {}, Y[DIM][DIM] = {};
  for (int a = 0; a < n_lmax; ++a) {
    const int dof = de[a];
    if (dof >= n_scalar) continue;
    double N = 0.0, dN[DIM] = {};
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double cb = Ce[a * NB + b];
      N += cb * Bt[b];
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] += cb * dBt[k][b];
    }
    const double wa = wgts[dof];
    W += N * wa;
#pragma unroll
    for (int k = 0; k < DIM; ++k) dW[k] += dN[k] * wa;
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) {
      const double xc = cpts[(int64_t)dof * DIM + cc];
      X[cc] += N * wa * xc;
#pragma unroll
      for (int k = 0; k < DIM; ++k) Y[k][cc] += dN[k] * wa * xc;
    }
  }
  double xp[DIM], J[DIM][DIM];
#pragma unroll
  for (int cc = 0; cc < DIM; ++cc) xp[cc] = X[cc] / W;
#pragma unroll
  for (int k = 0; k < DIM; ++k)
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) J[k][cc] = (Y[k][cc] - dW[k] * xp[cc]) / W;
  // J[k][c] = d x_c / d xi_k; physical gradient solves J g = g_param
  double inv[DIM][DIM], det;
  if constexpr (DIM == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    inv[0][0] = J[1][1] / det;
    inv[0][1] = -J[0][1] / det;
    inv[1][0] = -J[1][0] / det;
    inv[1][1] = J[0][0] / det;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    inv[0][0] = c00 / det;
    inv[1][0] = c01 / det;
    inv[2][0] = c02 / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  }
  const int64_t p = c * NQ + q;
  // pass 2: R and physical gradients (zero under the padding)
  for (int a = 0; a < n_lmax; ++a) {
    const int dof = de[a];
    double Ra = 0.0, g[DIM] = {};
    if (dof < n_scalar) {
      double N = 0.0, dN[DIM] = {};
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const double cb = Ce[a * NB + b];
        N += cb * Bt[b];
#pragma unroll
        for (int k = 0; k < DIM; ++k) dN[k] += cb * dBt[k][b];
      }
      const double wa = wgts[dof];
      Ra = N * wa / W;
      double gp[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) gp[k] = (dN[k] * wa - Ra * dW[k]) / W;
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += inv[cc][k] * gp[k];
        g[cc] = s;
      }
    }
    if (R) R[p * n_lmax + a] = Ra;
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) dR[(p * DIM + cc) * n_lmax + a] = g[cc];
  }
  if (x)
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) x[p * DIM + cc] = xp[cc];
  w[p] = wq * fabs(det);
}

// ---------------------------------------------------------------------------
// element tangents and loads, scatter-added through the slot map
// ---------------------------------------------------------------------------

struct ElemAsm {
  int problem, vec, coef_kind, src_kind;
  double coef_const;
  const double* coef;  // (n_c, n_q) per point
  double cmat[36];
  double src_const[3];
  const double* src;  // (n_c, n_q, vec)
  const int32_t* f2f;
  const double* gfull;
  const int32_t* scatter;
  int64_t nnz;
  double* vals;
  double* rhs;
};

template <int DIM>
__global__ void __launch_bounds__(256)
    k_elem_assemble(int64_t lo, int n_q, int n_lmax, const double* __restrict__ R,
                    const double* __restrict__ dR, const double* __restrict__ w,
                    const int32_t* __restrict__ elem_dofs, int n_scalar, ElemAsm A) {
  const int64_t c = blockIdx.x, e = lo + c;
  const int vec = A.vec, nlv = n_lmax * vec;
  const double* dRc = dR + c * (int64_t)n_q * DIM * n_lmax;
  const double* wc = w + c * (int64_t)n_q;
  const int32_t* de = elem_dofs + e * n_lmax;
  constexpr int NS = DIM == 2 ? 3 : 6;
  for (int t = threadIdx.x; t < nlv * nlv; t += blockDim.x) {
    const int row = t / nlv, col = t % nlv;
    const int a = row / vec, i = row % vec, b = col / vec, k = col % vec;
    if (de[a] >= n_scalar || de[b] >= n_scalar) continue;
    double s = 0.0;
    if (A.problem == IGA_POISSON) {
      if (i == k)
        for (int q = 0; q < n_q; ++q) {
          double dd = 0.0;
#pragma unroll
          for (int j = 0; j < DIM; ++j) dd += dRc[(q * DIM + j) * n_lmax + a] * dRc[(q * DIM + j) * n_lmax + b];
          const double coef = A.coef_kind == IGA_COEF_POINTS ? A.coef[c * n_q + q] : A.coef_const;
          s += wc[q] * coef * dd;
        }
    } else {
      for (int q = 0; q < n_q; ++q) {
        // engineering Voigt strain-displacement columns of (a, i) and (b, k)
        double ga[DIM], gb[DIM], Ba[NS] = {}, Bb[NS] = {};
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
          ga[j] = dRc[(q * DIM + j) * n_lmax + a];
          gb[j] = dRc[(q * DIM + j) * n_lmax + b];
        }
        if constexpr (DIM == 2) {
          Ba[i] = ga[i];
          Ba[2] = ga[1 - i];
          Bb[k] = gb[k];
          Bb[2] = gb[1 - k];
        } else {
          // rows xx, yy, zz, yz, xz, xy
          Ba[i] = ga[i];
          Bb[k] = gb[k];
          const int sh[3][2] = {{4, 5}, {3, 5}, {3, 4}};  // shears containing each component
          const int other[3][2] = {{2, 1}, {2, 0}, {1, 0}};
#pragma unroll
          for (int m = 0; m < 2; ++m) {
            Ba[sh[i][m]] = ga[other[i][m]];
            Bb[sh[k][m]] = gb[other[k][m]];
          }
        }
        double v = 0.0;
#pragma unroll
        for (int p1 = 0; p1 < NS; ++p1)
#pragma unroll
          for (int p2 = 0; p2 < NS; ++p2) v += Ba[p1] * A.cmat[p1 * NS + p2] * Bb[p2];
        s += wc[q] * v;
      }
    }
    const int slot = A.scatter[(e * nlv + row) * (int64_t)nlv + col];
    if (A.vals && slot < A.nnz) atomicAdd(A.vals + slot, s);
    if (A.rhs && A.gfull) {
      // lifting: a free row against a prescribed column
      const int fr = A.f2f[(int64_t)de[a] * vec + i];
      const int64_t gc = (int64_t)de[b] * vec + k;
      if (fr >= 0 && A.f2f[gc] < 0 && A.gfull[gc] != 0.0) atomicAdd(A.rhs + fr, -s * A.gfull[gc]);
    }
  }
  if (A.rhs && A.src_kind != IGA_SRC_NONE) {
    const double* Rc = R + c * (int64_t)n_q * n_lmax;
    for (int row = threadIdx.x; row < nlv; row += blockDim.x) {
      const int a = row / vec, i = row % vec;
      if (de[a] >= n_scalar) continue;
      const int fr = A.f2f[(int64_t)de[a] * vec + i];
      if (fr < 0) continue;
      double s = 0.0;
      for (int q = 0; q < n_q; ++q) {
        const double f = A.src_kind == IGA_SRC_POINTS ? A.src[(c * n_q + q) * vec + i] : A.src_const[i];
        s += wc[q] * f * Rc[q * n_lmax + a];
      }
      atomicAdd(A.rhs + fr, s);
    }
  }
}

template <typename F>
int by_dim_degree(int dim, int degree, F f) {
  switch (dim * 10 + degree) {
    case 21: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{});
    case 22: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 2>{});
    case 23: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 3>{});
    case 24: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 4>{});
    case 31: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 1>{});
    case 32: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 2>{});
    case 33: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 3>{});
    default: return IGA_ERR_UNSUPPORTED;
  }
}

struct PatternWork {
  int64_t K, keys, sorted, uniq, count, temp, temp_bytes, total;
};

static int pattern_layout(int64_t n_e, int n_lmax, int vec, PatternWork* L) {
  const int64_t nlv = (int64_t)n_lmax * vec;
  L->K = n_e * nlv * nlv;
  size_t sort_bytes = 0, uniq_bytes = 0;
  if (hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, (int64_t*)nullptr, (int64_t*)nullptr,
                                        (int)std::max<int64_t>(L->K, 1)) != hipSuccess)
    return IGA_ERR_HIP;
  if (hipcub::DeviceSelect::Unique(nullptr, uniq_bytes, (int64_t*)nullptr, (int64_t*)nullptr,
                                   (int64_t*)nullptr, (int)std::max<int64_t>(L->K, 1)) != hipSuccess)
    return IGA_ERR_HIP;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off += (bytes + 255) & ~(int64_t)255;
    return o;
  };
  L->keys = take(L->K * 8);
  L->sorted = take(L->K * 8);
  L->uniq = take(L->K * 8);
  L->count = take(8);
  L->temp_bytes = (int64_t)std::max(sort_bytes, uniq_bytes);
  L->temp = take(L->temp_bytes);
  L->total = off;
  return IGA_OK;
}

}  // namespace iga

using namespace iga;

extern "C" int64_t iga_elem_pattern_workspace(int32_t n_e, int32_t n_lmax, int32_t vec) {
  PatternWork L;
  if (n_e < 0 || n_lmax <= 0 || vec <= 0) return -1;
  if ((int64_t)n_e * n_lmax * vec * n_lmax * vec > INT32_MAX) return -1;  // hipCUB item counts
  return pattern_layout(n_e, n_lmax, vec, &L) ? -1 : L.total;
}

extern "C" int iga_elem_pattern(int32_t n_e, int32_t n_lmax, int32_t vec, const int32_t* elem_dofs,
                                int32_t n_scalar, const int32_t* full_to_free, int32_t n_free,
                                void* work, int64_t work_bytes, int32_t* indptr, int32_t* cols,
                                int32_t* rows, int32_t* scatter, int64_t* out_nnz, void* stream) {
  if (!elem_dofs || !full_to_free || !work || !out_nnz) return IGA_ERR_ARG;
  PatternWork L;
  if (int rc = pattern_layout(n_e, n_lmax, vec, &L)) return rc;
  if (work_bytes < L.total) return IGA_ERR_SHAPE;
  hipStream_t st = (hipStream_t)stream;
  char* base = reinterpret_cast<char*>(work);
  int64_t* keys = reinterpret_cast<int64_t*>(base + L.keys);
  int64_t* sorted = reinterpret_cast<int64_t*>(base + L.sorted);
  int64_t* uniq = reinterpret_cast<int64_t*>(base + L.uniq);
  int64_t* count = reinterpret_cast<int64_t*>(base + L.count);
  const unsigned grid = (unsigned)std::min<int64_t>((L.K + 255) / 256 + 1, 65536);
  if (!cols) {
    // keys -> sorted -> unique; the count comes back to the host
    if (L.K == 0) {
      *out_nnz = 0;
      return IGA_OK;
    }
    hipLaunchKernelGGL(k_pair_keys, dim3(grid), dim3(256), 0, st, L.K, n_lmax, vec, elem_dofs,
                       n_scalar, full_to_free, (int64_t)n_free, keys);
    IGA_CHECK_LAUNCH();
    size_t tb = (size_t)L.temp_bytes;
    if (hipcub::DeviceRadixSort::SortKeys(base + L.temp, tb, keys, sorted, (int)L.K, 0, 64, st) !=
        hipSuccess)
      return IGA_ERR_HIP;
    tb = (size_t)L.temp_bytes;
    if (hipcub::DeviceSelect::Unique(base + L.temp, tb, sorted, uniq, count, (int)L.K, st) != hipSuccess)
      return IGA_ERR_HIP;
    int64_t nu = 0, last = 0;
    if (hipMemcpyAsync(&nu, count, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess)
      return IGA_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
    if (nu > 0) {
      if (hipMemcpyAsync(&last, uniq + nu - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess)
        return IGA_ERR_HIP;
      if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
    }
    *out_nnz = nu - (nu > 0 && last == kNoKey ? 1 : 0);
    return IGA_OK;
  }
  // fill: the unique keys of the counting call are in the workspace
  if (!indptr || !rows || !scatter) return IGA_ERR_ARG;
  const int64_t nnz = *out_nnz;
  if (nnz > INT32_MAX) return IGA_ERR_SHAPE;
  const unsigned g2 = (unsigned)std::min<int64_t>((std::max<int64_t>(nnz, n_free + 1) + 255) / 256, 65536);
  if (nnz > 0) {
    hipLaunchKernelGGL(k_pattern_fill, dim3(g2), dim3(256), 0, st, nnz, uniq, (int64_t)n_free, rows,
                       cols);
    IGA_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_pattern_indptr, dim3(g2), dim3(256), 0, st, (int64_t)n_free, uniq, nnz, indptr);
  IGA_CHECK_LAUNCH();
  if (L.K > 0) {
    hipLaunchKernelGGL(k_pattern_scatter, dim3(grid), dim3(256), 0, st, L.K, n_lmax, vec, elem_dofs,
                       n_scalar, full_to_free, (int64_t)n_free, uniq, nnz, scatter);
    IGA_CHECK_LAUNCH();
  }
  return IGA_OK;
}

extern "C" int iga_elem_basis(int32_t dim, int32_t degree, int32_t n_lmax, int64_t lo, int64_t n_c,
                              const iga_rule_t* rule, const double* ext, const int32_t* elem_dofs,
                              int32_t n_scalar, const double* cpts, const double* wgts,
                              const double* boxes, double* R, double* dR, double* x, double* w,
                              void* stream) {
  if (!rule || !ext || !elem_dofs || !cpts || !wgts || !boxes || !dR || !w) return IGA_ERR_ARG;
  if (rule->n_q1 != degree + 1) return IGA_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  return by_dim_degree(dim, degree, [&](auto D, auto Pc) -> int {
    constexpr int DIM = decltype(D)::value, P = decltype(Pc)::value;
    const int64_t n = n_c * ipow(P + 1, DIM);
    if (n == 0) return (int)IGA_OK;
    hipLaunchKernelGGL((k_elem_basis<DIM, P>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lo,
                       n_c, n_lmax, rule->pts, rule->wts, ext, elem_dofs, n_scalar, cpts, wgts, boxes, R,
                       dR, x, w);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

extern "C" int iga_elem_assemble(int32_t dim, int32_t n_q, int32_t n_lmax, int64_t lo, int64_t n_c,
                                 const double* R, const double* dR, const double* w,
                                 const int32_t* elem_dofs, int32_t n_scalar, const iga_asm_args_t* args,
                                 const int32_t* scatter, void* stream) {
  if (!dR || !w || !elem_dofs || !args || !args->full_to_free) return IGA_ERR_ARG;
  if (args->vals && !scatter) return IGA_ERR_ARG;
  if (args->rhs && args->src_kind != IGA_SRC_NONE && !R) return IGA_ERR_ARG;
  if (dim != 2 && dim != 3) return IGA_ERR_UNSUPPORTED;
  ElemAsm A{};
  A.problem = args->problem;
  A.vec = args->vec;
  A.coef_kind = args->coef_kind;
  A.coef_const = args->coef_const;
  A.coef = args->coef_points;
  for (int i = 0; i < 36; ++i) A.cmat[i] = args->cmat[i];
  A.src_kind = args->src_kind;
  for (int i = 0; i < 3; ++i) A.src_const[i] = args->src_const[i];
  A.src = args->src_points;
  A.f2f = args->full_to_free;
  A.gfull = args->gfull;
  A.scatter = scatter;
  A.nnz = args->nnz;
  A.vals = reinterpret_cast<double*>(args->vals);
  A.rhs = args->rhs;
  if (A.problem == IGA_ELASTICITY && A.vec != dim) return IGA_ERR_SHAPE;
  if (n_c == 0) return IGA_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dim == 2)
    hipLaunchKernelGGL(k_elem_assemble<2>, dim3((unsigned)n_c), dim3(256), 0, st, lo, n_q, n_lmax, R, dR,
                       w, elem_dofs, n_scalar, A);
  else
    hipLaunchKernelGGL(k_elem_assemble<3>, dim3((unsigned)n_c), dim3(256), 0, st, lo, n_q, n_lmax, R, dR,
                       w, elem_dofs, n_scalar, A);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}
