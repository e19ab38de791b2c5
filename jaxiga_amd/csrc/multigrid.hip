// Geometric multigrid on tensor-patch IGA operators with learned stencil
// smoothers, and the flexible CG that takes one V-cycle as its preconditioner.
//
// The levels come from dyadic coarsening of the knot vectors. The coarse
// spline spaces are nested in the fine one, so the transfer P (coarse ->
// fine coefficients) is exact and every coarse operator is the Galerkin
// product P^T K P, formed on the device by two masked sparse products.
//
// On every level but the coarsest the smoother is a per-control-point stencil
// of width ks with weights W[o][i], acting in diagonally scaled variables:
//     S v = D^-1/2 T_W (D^-1/2 v),   (T_W u)_i = sum_o W[o][i] u[i + off(o)].
// It runs once before the coarse correction and, transposed, once after it,
// so the V-cycle is symmetric. The free dofs of a level form an h x w grid
// (row-major, direction 0 fastest); the stencil runs on that grid. The
// coarsest level is solved with a dense inverse. The V-cycle runs in FP32; the
// outer flexible CG keeps FP64 vectors, residuals and dot products.
#include <algorithm>

#include "common.hpp"

namespace iga {

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------

// C = A B over C's pattern (COO rows/cols). A in CSR, B as the CSR of B^T;
// both with sorted column indices. One thread per entry, sorted merge.
__global__ void __launch_bounds__(256)
    k_masked_spgemm(int64_t nnz, const int32_t* __restrict__ c_rows,
                    const int32_t* __restrict__ c_cols, const int32_t* __restrict__ a_ptr,
                    const int32_t* __restrict__ a_cols, const double* __restrict__ a_vals,
                    const int32_t* __restrict__ b_ptr, const int32_t* __restrict__ b_cols,
                    const double* __restrict__ b_vals, double* __restrict__ c_vals) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int i = c_rows[k], j = c_cols[k];
    int pa = a_ptr[i], pb = b_ptr[j];
    const int ea = a_ptr[i + 1], eb = b_ptr[j + 1];
    double acc = 0.0;
    while (pa < ea && pb < eb) {
      const int ca = a_cols[pa], cb = b_cols[pb];
      if (ca == cb) {
        acc += a_vals[pa] * b_vals[pb];
        ++pa;
        ++pb;
      } else if (ca < cb) {
        ++pa;
      } else {
        ++pb;
      }
    }
    c_vals[k] = acc;
  }
}

// Galerkin product K_c = P^T K P, one wavefront per coarse row I. The row of
// T = P^T K is accumulated over the fine rows a of P^T row I (each a's K row
// is spread over the lanes; one a at a time, so lanes never collide) into a
// dense box of fine grid points held in LDS; then each coarse entry (I, J) of
// the row's CSR pattern is sum_b P^T[J, b] T[I, b] over P^T row J. Fixed
// order throughout: bitwise reproducible. The box (kGalBox^2 fine points)
// must cover the stencil of every fine row of P^T row I (checked on the host).
constexpr int kGalBox = 16;

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
  return v;
}

__global__ void __launch_bounds__(256)
    k_galerkin_rows(int nc, int wf, int R, const int32_t* __restrict__ k_ptr,
                    const int32_t* __restrict__ k_cols, const double* __restrict__ k_vals,
                    const int32_t* __restrict__ pt_ptr, const int32_t* __restrict__ pt_cols,
                    const double* __restrict__ pt_vals, const int32_t* __restrict__ c_ptr,
                    const int32_t* __restrict__ c_cols, double* __restrict__ c_vals,
                    float* __restrict__ c_vals32, double* __restrict__ c_dinv,
                    float* __restrict__ c_dsq) {
  __shared__ double boxes[4][kGalBox * kGalBox];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int I = blockIdx.x * 4 + wave;
  if (I >= nc) return;
  double* box = boxes[wave];
  for (int t = lane; t < kGalBox * kGalBox; t += kWave) box[t] = 0.0;
  const int a0 = pt_ptr[I], a1 = pt_ptr[I + 1];
  int rmin = 1 << 30, cmin = 1 << 30;
  for (int k = a0 + lane; k < a1; k += kWave) {
    const int a = pt_cols[k];
    rmin = min(rmin, a / wf);
    cmin = min(cmin, a % wf);
  }
  const int r0 = wave_min(rmin) - R, c0 = wave_min(cmin) - R;
  // the fine rows of P^T row I and their K-row bounds, fetched once, one per lane
  int my_q0 = 0, my_q1 = 0;
  double my_pa = 0.0;
  if (lane < a1 - a0 && lane < kWave) {
    const int a = pt_cols[a0 + lane];
    my_pa = pt_vals[a0 + lane];
    my_q0 = k_ptr[a];
    my_q1 = k_ptr[a + 1];
  }
  __builtin_amdgcn_wave_barrier();
  for (int k = 0; k < a1 - a0; ++k) {
    int q0, q1;
    double pa;
    if (k < kWave) {
      q0 = __shfl(my_q0, k, kWave);
      q1 = __shfl(my_q1, k, kWave);
      pa = __shfl(my_pa, k, kWave);
    } else {
      const int a = pt_cols[a0 + k];
      pa = pt_vals[a0 + k];
      q0 = k_ptr[a];
      q1 = k_ptr[a + 1];
    }
    for (int q = q0 + lane; q < q1; q += kWave) {
      const int b = k_cols[q];
      const int br = b / wf - r0, bc = b % wf - c0;
      box[br * kGalBox + bc] += pa * k_vals[q];
    }
    __builtin_amdgcn_wave_barrier();
  }
  // the coarse row in FP64 and FP32, and its diagonal scalings (1 where the
  // diagonal is zero or absent)
  bool diag = false;
  for (int k = c_ptr[I] + lane; k < c_ptr[I + 1]; k += kWave) {
    const int J = c_cols[k];
    double acc = 0.0;
    for (int q = pt_ptr[J]; q < pt_ptr[J + 1]; ++q) {
      const int b = pt_cols[q];
      const int br = b / wf - r0, bc = b % wf - c0;
      if (br >= 0 && br < kGalBox && bc >= 0 && bc < kGalBox) acc += pt_vals[q] * box[br * kGalBox + bc];
    }
    c_vals[k] = acc;
    c_vals32[k] = (float)acc;
    if (J == I) {
      const double di = acc != 0.0 ? 1.0 / acc : 1.0;
      c_dinv[I] = di;
      c_dsq[I] = (float)sqrt(fabs(di));
      diag = true;
    }
  }
  if (__ballot(diag) == 0 && lane == 0) {
    c_dinv[I] = 1.0;
    c_dsq[I] = 1.f;
  }
}

// vals32 = (float) vals over all entries
__global__ void k_to_f32(int64_t nnz, const double* __restrict__ vals, float* __restrict__ vals32) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x)
    vals32[k] = (float)vals[k];
}

// per row (sorted columns): dinv = 1 / K_ii (1 where K_ii = 0), dsq = sqrt(dinv)
__global__ void k_level_scalings(int n, const int32_t* __restrict__ indptr,
                                 const int32_t* __restrict__ cols, const double* __restrict__ vals,
                                 double* __restrict__ dinv, float* __restrict__ dsq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int lo = indptr[i], hi = indptr[i + 1];
  const int k = lower_bound(cols, lo, hi, i);
  const double d = (k < hi && cols[k] == i) ? vals[k] : 0.0;
  const double di = d != 0.0 ? 1.0 / d : 1.0;
  dinv[i] = di;
  dsq[i] = (float)sqrt(fabs(di));
}

__global__ void k_copy_d(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// feats[o][i] = K[i, i + off(o)] / K[i, i] on the h x w grid, o over the
// (2R+1)^2 offsets (feats cleared beforehand). 16 lanes per row (sorted
// columns: the diagonal by binary search), one entry per lane.
__global__ void __launch_bounds__(256)
    k_stencil_features(int n, int w, int R, const int32_t* __restrict__ indptr,
                       const int32_t* __restrict__ cols, const double* __restrict__ vals,
                       float* __restrict__ feats) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / 16, l = threadIdx.x % 16;
  if (i >= n) return;
  const int lo = indptr[i], hi = indptr[i + 1];
  const int kd = lower_bound(cols, lo, hi, i);
  double d = (kd < hi && cols[kd] == i) ? vals[kd] : 0.0;
  if (d == 0.0) d = 1.0;
  const int ri = i / w, ci = i - ri * w, B = 2 * R + 1;
  for (int k = lo + l; k < hi; k += 16) {
    const int j = cols[k];
    const int dr = j / w - ri, dc = j % w - ci;
    if (dr < -R || dr > R || dc < -R || dc > R) continue;
    feats[(size_t)((dr + R) * B + dc + R) * n + i] = (float)(vals[k] / d);
  }
}

__global__ void k_zero_range(int64_t n, float* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.f;
}

// dense augmented [A | I] (n x 2n) of a CSR matrix
__global__ void k_densify(int n, const int32_t* __restrict__ indptr,
                          const int32_t* __restrict__ cols, const double* __restrict__ vals,
                          double* __restrict__ A) {
  const int i = blockIdx.x;
  double* row = A + (size_t)i * 2 * n;
  for (int j = threadIdx.x; j < 2 * n; j += blockDim.x) row[j] = j == n + i ? 1.0 : 0.0;
  __syncthreads();
  for (int k = indptr[i] + threadIdx.x; k < indptr[i + 1]; k += blockDim.x) row[cols[k]] = vals[k];
}

// Gauss-Jordan elimination of [A | I] without pivoting (A is SPD), one block.
__global__ void __launch_bounds__(1024) k_gauss_jordan(int n, double* __restrict__ A) {
  extern __shared__ double fac[];  // [n]
  const int n2 = 2 * n;
  for (int k = 0; k < n; ++k) {
    const double piv = A[(size_t)k * n2 + k];
    __syncthreads();
    for (int j = threadIdx.x; j < n2; j += blockDim.x) A[(size_t)k * n2 + j] /= piv;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) fac[i] = A[(size_t)i * n2 + k];
    __syncthreads();
    for (int64_t t = threadIdx.x; t < (int64_t)n * n2; t += blockDim.x) {
      const int i = t / n2, j = t - (int64_t)i * n2;
      if (i != k) A[t] -= fac[i] * A[(size_t)k * n2 + j];
    }
    __syncthreads();
  }
}

__global__ void k_inverse_f32(int n, const double* __restrict__ A, float* __restrict__ inv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = t / n, j = t - (int64_t)i * n;
  inv[t] = (float)A[(size_t)i * 2 * n + n + j];
}

// ---------------------------------------------------------------------------
// V-cycle kernels (every kernel returns at once when *skip is set)
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool skipped(const int* skip) { return skip && *skip; }

// out (=|+=) D^-1/2 T (D^-1/2 v), T = T_W or its transpose; batch over
// blockIdx.z. A block computes a 16 x 16 tile of the grid: u = D^-1/2 v over
// the tile and its halo is staged in LDS (zero outside the grid), so the tap
// loop has no bounds tests on u; the weights stream from HBM, coalesced
// along the tile rows.
constexpr int kStTile = 16;

template <int KS, bool TR>
__global__ void __launch_bounds__(256)
    k_stencil(int h, int w, const float* __restrict__ W, const float* __restrict__ dsq,
              const float* __restrict__ v, float* __restrict__ out, int accumulate,
              const int* __restrict__ skip) {
  if (skipped(skip)) return;
  constexpr int R = KS / 2;
  constexpr int T = kStTile + 2 * R;
  __shared__ float u[T][T];
  const int n = h * w;
  const float* vb = v + (size_t)blockIdx.z * n;
  const int r0 = blockIdx.y * kStTile, c0 = blockIdx.x * kStTile;
  for (int t = threadIdx.x; t < T * T; t += blockDim.x) {
    const int rr = r0 + t / T - R, cc = c0 + t % T - R;
    u[t / T][t % T] = (rr >= 0 && rr < h && cc >= 0 && cc < w) ? dsq[rr * w + cc] * vb[rr * w + cc] : 0.f;
  }
  __syncthreads();
  const int ty = threadIdx.x / kStTile, tx = threadIdx.x % kStTile;
  const int r = r0 + ty, c = c0 + tx;
  if (r >= h || c >= w) return;
  const int i = r * w + c;
  float acc = 0.f;
#pragma unroll
  for (int dy = -R; dy <= R; ++dy) {
#pragma unroll
    for (int dx = -R; dx <= R; ++dx) {
      const int o = (dy + R) * KS + (dx + R);
      if (!TR) {
        acc += W[(size_t)o * n + i] * u[ty + dy + R][tx + dx + R];
      } else {
        const int rr = r - dy, cc = c - dx;
        const bool in = rr >= 0 && rr < h && cc >= 0 && cc < w;
        const float wt = in ? W[(size_t)o * n + rr * w + cc] : 0.f;
        acc += wt * u[ty - dy + R][tx - dx + R];
      }
    }
  }
  float* ob = out + (size_t)blockIdx.z * n;
  const float val = dsq[i] * acc;
  ob[i] = accumulate ? ob[i] + val : val;
}

// damped Jacobi (used where a level has no stencil): out (=|+=) omega D^-1 v
__global__ void k_jacobi_f(int n, float omega, const float* __restrict__ dsq,
                           const float* __restrict__ v, float* __restrict__ out, int accumulate,
                           const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float val = omega * dsq[i] * dsq[i] * v[i];
  out[i] = accumulate ? out[i] + val : val;
}

// y = A x (mode 0), y = b - A x (mode 1), y += A x (mode 2); batch over blockIdx.y
template <int LPR>
__global__ void __launch_bounds__(256)
    k_csr_f(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
            const float* __restrict__ vals, const float* __restrict__ x,
            const float* __restrict__ b, float* __restrict__ y, int mode, int64_t xs, int64_t ys,
            const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  const float* xb = x + (size_t)blockIdx.y * xs;
  float* yb = y + (size_t)blockIdx.y * ys;
  const float* bb = b ? b + (size_t)blockIdx.y * ys : nullptr;
  for (int row = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; row < n; row += groups) {
    const float s = row_dot_f<LPR>(row, indptr, cols, vals, xb, l);
    if (l == 0) {
      if (mode == 0)
        yb[row] = s;
      else if (mode == 1)
        yb[row] = bb[row] - s;
      else
        yb[row] += s;
    }
  }
}

// Jacobi pre-smoothing from x = 0 fused with the residual:
// x = w D^-1 b, r = b - A x (A x evaluated with the x it writes: row products
// recompute w D^-1 b at the neighbours instead of reading x)
template <int LPR>
__global__ void __launch_bounds__(256)
    k_jacobi_pre(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
                 const float* __restrict__ vals, float omega, const float* __restrict__ dsq,
                 const float* __restrict__ b, float* __restrict__ x, float* __restrict__ r,
                 const int* __restrict__ skip, int64_t vs) {
  if (skipped(skip)) return;
  b += blockIdx.y * vs;  // vector blockIdx.y of a batch
  x += blockIdx.y * vs;
  r += blockIdx.y * vs;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  for (int row = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; row < n; row += groups) {
    const int lo = indptr[row], hi = indptr[row + 1];
    float s = 0.f;
    for (int k = lo + l; k < hi; k += LPR) {
      const int j = cols[k];
      s += vals[k] * (omega * dsq[j] * dsq[j] * b[j]);
    }
#pragma unroll
    for (int off = LPR / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, LPR);
    if (l == 0) {
      x[row] = omega * dsq[row] * dsq[row] * b[row];
      r[row] = b[row] - s;
    }
  }
}

// Jacobi post-smoothing fused with its residual: x_out = x + w D^-1 (b - A x)
template <int LPR>
__global__ void __launch_bounds__(256)
    k_jacobi_post(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
                  const float* __restrict__ vals, float omega, const float* __restrict__ dsq,
                  const float* __restrict__ b, const float* __restrict__ x,
                  float* __restrict__ x_out, const int* __restrict__ skip, int64_t vs) {
  if (skipped(skip)) return;
  b += blockIdx.y * vs;
  x += blockIdx.y * vs;
  x_out += blockIdx.y * vs;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  for (int row = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; row < n; row += groups) {
    const float s = row_dot_f<LPR>(row, indptr, cols, vals, x, l);
    if (l == 0) x_out[row] = x[row] + omega * dsq[row] * dsq[row] * (b[row] - s);
  }
}

// x = M b for a dense n x n matrix (coarsest level), one wavefront per row
__global__ void __launch_bounds__(256)
    k_dense_gemv(int n, const float* __restrict__ M, const float* __restrict__ b,
                 float* __restrict__ x, const int* __restrict__ skip, int64_t vs) {
  if (skipped(skip)) return;
  b += blockIdx.y * vs;
  x += blockIdx.y * vs;
  const int i = blockIdx.x * 4 + threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (i >= n) return;
  const float* row = M + (size_t)i * n;
  float s = 0.f;
  for (int j = l; j < n; j += kWave) s += row[j] * b[j];
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (l == 0) x[i] = s;
}

__global__ void k_copy_fs(int n, const float* __restrict__ a, float* __restrict__ b,
                          const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_d2f(int n, const double* __restrict__ a, float* __restrict__ b,
                      const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_f2d(int n, const float* __restrict__ a, double* __restrict__ b,
                      const int* __restrict__ skip) {
  if (skipped(skip)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

// ---------------------------------------------------------------------------
// training kernels
// ---------------------------------------------------------------------------

// Gradient of y = D^-1/2 T_W^T (D^-1/2 v) with respect to W, summed over the
// batch: dW[o][j] = sum_b (D^-1/2 g_b)[j + off(o)] (D^-1/2 v_b)[j].
template <int KS>
__global__ void __launch_bounds__(256)
    k_stencil_wgrad_t(int h, int w, int B, const float* __restrict__ dsq,
                      const float* __restrict__ v, const float* __restrict__ g,
                      float* __restrict__ dW) {
  constexpr int R = KS / 2;
  const int n = h * w;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int r = j / w, c = j - r * w;
  const float dj = dsq[j];
#pragma unroll
  for (int dy = -R; dy <= R; ++dy) {
#pragma unroll
    for (int dx = -R; dx <= R; ++dx) {
      const int o = (dy + R) * KS + (dx + R);
      const int rr = r + dy, cc = c + dx;
      float acc = 0.f;
      if (rr >= 0 && rr < h && cc >= 0 && cc < w) {
        const int i = rr * w + cc;
        const float di = dsq[i];
        for (int b = 0; b < B; ++b) acc += g[(size_t)b * n + i] * v[(size_t)b * n + j];
        acc *= di * dj;
      }
      dW[(size_t)o * n + j] = acc;
    }
  }
}

// per-sample z_b . kz_b (FP64), their mean, and grad = -2 kz / B
__global__ void k_energy_sums(int64_t n, const float* __restrict__ z, const float* __restrict__ kz,
                              double* __restrict__ per) {
  __shared__ double scratch[16];
  const int b = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
    s += (double)z[(size_t)b * n + i] * kz[(size_t)b * n + i];
  s = block_sum(s, scratch);
  if (threadIdx.x == 0) per[b] = s;
}

__global__ void k_energy_grad(int B, int64_t n, const float* __restrict__ kz,
                              const double* __restrict__ per, float* __restrict__ grad,
                              double* loss) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    double l = 0.0;
    for (int b = 0; b < B; ++b) l += per[b];
    *loss = l / B;
  }
  if (t < (int64_t)B * n) grad[t] = -2.f * kz[t] / (float)B;
}

// ---------------------------------------------------------------------------
// host-side V-cycle
// ---------------------------------------------------------------------------

template <typename F>
static int by_ks(int ks, F f) {
  switch (ks) {
    case 3: return f(std::integral_constant<int, 3>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 7: return f(std::integral_constant<int, 7>{});
    case 9: return f(std::integral_constant<int, 9>{});
    default: return IGA_ERR_UNSUPPORTED;
  }
}

static inline int blocks_for(int64_t n, int t = 256) { return (int)std::max<int64_t>(1, (n + t - 1) / t); }

static int csr_f(int n, int nnz, const int32_t* ptr, const int32_t* cols, const float* vals,
                 const float* x, const float* b, float* y, int mode, int batch, int64_t xs,
                 int64_t ys, const int* skip, hipStream_t st) {
  const int lpr = pick_lpr(n, nnz);
  const int rb = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)n * lpr + 255) / 256, 4096));
  return by_lpr(lpr, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    hipLaunchKernelGGL(k_csr_f<LPR>, dim3(rb, batch), dim3(256), 0, st, n, ptr, cols, vals, x, b,
                       y, mode, xs, ys, skip);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

static int smooth(const iga_mg_t& mg, int l, const float* v, float* out, bool transpose,
                  int accumulate, const int* skip, hipStream_t st) {
  const iga_mg_level_t& L = mg.level[l];
  if (!L.stencil) {
    hipLaunchKernelGGL(k_jacobi_f, dim3(blocks_for(L.n)), dim3(256), 0, st, L.n, L.omega, L.dsq, v,
                       out, accumulate, skip);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }
  const dim3 grid((L.w + kStTile - 1) / kStTile, (L.h + kStTile - 1) / kStTile, 1);
  return by_ks(mg.ks, [&](auto K) {
    constexpr int KS = decltype(K)::value;
    if (transpose)
      hipLaunchKernelGGL((k_stencil<KS, true>), grid, dim3(256), 0, st, L.h, L.w, L.stencil, L.dsq,
                         v, out, accumulate, skip);
    else
      hipLaunchKernelGGL((k_stencil<KS, false>), grid, dim3(256), 0, st, L.h, L.w, L.stencil,
                         L.dsq, v, out, accumulate, skip);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

static int64_t level_offset(const iga_mg_t& mg, int l) {
  int64_t o = 0;
  for (int k = 0; k < l; ++k) o += 3 * (int64_t)mg.level[k].n;
  return o;
}

// x = V(b) starting at level l0 (b, x of level l0's size)
template <typename F>
static int launch_rows(int n, int nnz, F f) {
  const int lpr = pick_lpr(n, nnz);
  const int rb = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)n * lpr + 255) / 256, 4096));
  return by_lpr(lpr, [&](auto L) {
    f(L, rb);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

// x = V(b) starting at level l0 (b, x of level l0's size). Levels with a
// stencil smoother run pre-smooth / residual / restrict and prolong /
// residual / post-smooth as separate kernels; damped-Jacobi levels fuse the
// smoother into the residual kernels.
static int vcycle(const iga_mg_t& mg, int l0, const float* b0, float* x0, const int* skip,
                  hipStream_t st) {
  const int Lc = mg.n_levels - 1;
  const float* B[16];
  float* X[16];
  float* Rr[16];
  for (int l = l0; l <= Lc; ++l) {
    float* base = mg.work + level_offset(mg, l);
    X[l] = base;
    B[l] = base + mg.level[l].n;
    Rr[l] = base + 2 * mg.level[l].n;
  }
  X[l0] = x0;
  B[l0] = b0;
  int rc;
  // below the tail level the cycle is one dense product
  const int t = (mg.tail && mg.tail_level >= l0 && mg.tail_level < Lc) ? mg.tail_level : Lc;
  for (int l = l0; l < t; ++l) {
    const iga_mg_level_t& L = mg.level[l];
    if (L.stencil) {
      if ((rc = smooth(mg, l, B[l], X[l], false, 0, skip, st))) return rc;
      if ((rc = csr_f(L.n, L.nnz, L.indptr, L.cols, L.vals, X[l], B[l], Rr[l], 1, 1, 0, 0, skip, st)))
        return rc;
    } else {
      rc = launch_rows(L.n, L.nnz, [&](auto K, int rb) {
        constexpr int LPR = decltype(K)::value;
        hipLaunchKernelGGL(k_jacobi_pre<LPR>, dim3(rb), dim3(256), 0, st, L.n, L.indptr, L.cols,
                           L.vals, L.omega, L.dsq, B[l], X[l], Rr[l], skip, (int64_t)0);
      });
      if (rc) return rc;
    }
    const iga_mg_level_t& C = mg.level[l + 1];
    if ((rc = csr_f(C.n, L.pt_nnz, L.pt_ptr, L.pt_cols, L.pt_vals, Rr[l], nullptr,
                    const_cast<float*>(B[l + 1]), 0, 1, 0, 0, skip, st)))
      return rc;
  }
  {
    const iga_mg_level_t& C = mg.level[t];
    hipLaunchKernelGGL(k_dense_gemv, dim3((C.n + 3) / 4), dim3(256), 0, st, C.n,
                       t < Lc ? mg.tail : mg.coarse_inv, B[t], X[t], skip, (int64_t)0);
    IGA_CHECK_LAUNCH();
  }
  for (int l = t - 1; l >= l0; --l) {
    const iga_mg_level_t& L = mg.level[l];
    if ((rc = csr_f(L.n, L.p_nnz, L.p_ptr, L.p_cols, L.p_vals, X[l + 1], nullptr, X[l], 2, 1, 0, 0,
                    skip, st)))
      return rc;
    if (L.stencil) {
      if ((rc = csr_f(L.n, L.nnz, L.indptr, L.cols, L.vals, X[l], B[l], Rr[l], 1, 1, 0, 0, skip, st)))
        return rc;
      if ((rc = smooth(mg, l, Rr[l], X[l], true, 1, skip, st))) return rc;
    } else {
      // post-smooth into the residual buffer, then copy back only at the
      // caller's level (the parent prolongs from whichever buffer holds x)
      rc = launch_rows(L.n, L.nnz, [&](auto K, int rb) {
        constexpr int LPR = decltype(K)::value;
        hipLaunchKernelGGL(k_jacobi_post<LPR>, dim3(rb), dim3(256), 0, st, L.n, L.indptr, L.cols,
                           L.vals, L.omega, L.dsq, B[l], X[l], Rr[l], skip, (int64_t)0);
      });
      if (rc) return rc;
      std::swap(X[l], Rr[l]);
      if (l == l0) {
        hipLaunchKernelGGL(k_copy_fs, dim3(blocks_for(L.n)), dim3(256), 0, st, L.n, X[l], x0, skip);
        IGA_CHECK_LAUNCH();
      }
    }
  }
  return IGA_OK;
}

static int check_mg(const iga_mg_t* mg) {
  if (!mg || mg->n_levels < 1 || mg->n_levels > 16 || !mg->work || !mg->coarse_inv)
    return IGA_ERR_ARG;
  if (mg->tail) {
    if (mg->tail_level < 0 || mg->tail_level >= mg->n_levels - 1) return IGA_ERR_ARG;
    for (int l = mg->tail_level; l < mg->n_levels - 1; ++l)
      if (mg->level[l].stencil) return IGA_ERR_UNSUPPORTED;
  }
  for (int l = 0; l < mg->n_levels; ++l) {
    const iga_mg_level_t& L = mg->level[l];
    if (L.n <= 0 || (int64_t)L.h * L.w != L.n) return IGA_ERR_SHAPE;
    if (l + 1 < mg->n_levels) {
      if (!L.indptr || !L.cols || !L.vals || !L.dsq || !L.p_ptr || !L.p_cols || !L.p_vals ||
          !L.pt_ptr || !L.pt_cols || !L.pt_vals)
        return IGA_ERR_ARG;
      if (mg->level[l + 1].n <= 0) return IGA_ERR_SHAPE;
      if (L.stencil && (mg->ks < 3 || mg->ks > 9 || mg->ks % 2 == 0)) return IGA_ERR_UNSUPPORTED;
    }
  }
  return IGA_OK;
}

// ---------------------------------------------------------------------------
// flexible CG (Notay's FCG with one stored direction): device-resident
// scalars; G iterations are captured into a hipGraph and replayed, and the
// host reads the convergence flag once per replay
// ---------------------------------------------------------------------------

struct FState {
  double rr, bb, tol2, pq_prev;
  int iter, maxiter, done, pad;
};

struct FBufs {
  FState* st;
  double* P_a;  // [kRedBlocks] z.q_prev
  double* P_b;  // [kRedBlocks] p.r
  double* P_c;  // [kRedBlocks] p.q
  double* P_d;  // [kRedBlocks] r.r
};

__global__ void __launch_bounds__(kVecThreads)
    k_fcg_init(int n, const double* __restrict__ b, double* x, double* r, float* r32, double* p1,
               double* q1, FBufs F) {
  double bb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] = 0.0;
    r[i] = b[i];
    r32[i] = (float)b[i];
    p1[i] = 0.0;  // the "previous" direction of the first iteration
    q1[i] = 0.0;
    bb += b[i] * b[i];
  }
  store_partial(F.P_d, bb);
}

__global__ void k_fcg_check(int nblk, FBufs F, int init, double tol2, int maxiter) {
  FState* st = F.st;
  if (!init && st->done) return;
  const double rr = reduce_partials(F.P_d, nblk);
  if (threadIdx.x == 0) {
    if (init) {
      st->bb = rr;
      st->tol2 = tol2;
      st->maxiter = maxiter;
      st->iter = 0;
      st->pq_prev = 1.0;
    } else {
      st->iter += 1;
    }
    st->rr = rr;
    st->done = (!(rr > st->tol2 * st->bb) || st->iter >= st->maxiter) ? 1 : 0;
  }
}

// z = (double) z32 (the V-cycle output) and the partials of z . q_prev
__global__ void __launch_bounds__(kVecThreads)
    k_fcg_zq(int n, const float* __restrict__ z32, double* __restrict__ z,
             const double* __restrict__ qp, FBufs F) {
  if (F.st->done) return;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double zi = (double)z32[i];
    z[i] = zi;
    s += zi * qp[i];
  }
  store_partial(F.P_a, s);
}

__global__ void __launch_bounds__(kVecThreads)
    k_fcg_dir(int n, int nblk, const double* __restrict__ z, const double* __restrict__ pp,
              double* p, const double* __restrict__ r, FBufs F) {
  const FState* st = F.st;
  if (st->done) return;
  double beta = 0.0;
  if (st->iter > 0) beta = -reduce_partials(F.P_a, nblk) / st->pq_prev;
  double pr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double pi = z[i] + beta * pp[i];
    p[i] = pi;
    pr += pi * r[i];
  }
  store_partial(F.P_b, pr);
}

template <int LPR>
__global__ void __launch_bounds__(kVecThreads)
    k_fcg_spmv(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
               const double* __restrict__ vals, const double* __restrict__ p, double* q, FBufs F) {
  if (F.st->done) return;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double pq = 0.0;
  for (int r = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; r < n; r += groups) {
    const double s = row_dot<LPR>(r, indptr, cols, vals, p, l);
    if (l == 0) {
      q[r] = s;
      pq += s * p[r];
    }
  }
  store_partial(F.P_c, pq);
}

__global__ void __launch_bounds__(kVecThreads)
    k_fcg_update(int n, int nblk_c, int nblk, const double* __restrict__ p,
                 const double* __restrict__ q, double* x, double* r, float* r32, FBufs F) {
  FState* st = F.st;
  if (st->done) return;
  const double pq = reduce_partials(F.P_c, nblk_c);
  const double pr = reduce_partials(F.P_b, nblk);
  const double alpha = pr / pq;
  double rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    r32[i] = (float)ri;
    rr += ri * ri;
  }
  store_partial(F.P_d, rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->pq_prev = pq;
}

}  // namespace iga

using namespace iga;

extern "C" int iga_masked_spgemm(int64_t nnz_c, const int32_t* c_rows, const int32_t* c_cols,
                                 const int32_t* a_ptr, const int32_t* a_cols, const double* a_vals,
                                 const int32_t* bt_ptr, const int32_t* bt_cols,
                                 const double* bt_vals, double* c_vals, void* stream) {
  if (nnz_c == 0) return IGA_OK;
  if (!c_rows || !c_cols || !a_ptr || !a_cols || !a_vals || !bt_ptr || !bt_cols || !bt_vals ||
      !c_vals)
    return IGA_ERR_ARG;
  const int blocks = (int)std::min<int64_t>((nnz_c + 255) / 256, 256 * 32);
  hipLaunchKernelGGL(k_masked_spgemm, dim3(blocks), dim3(256), 0, (hipStream_t)stream, nnz_c,
                     c_rows, c_cols, a_ptr, a_cols, a_vals, bt_ptr, bt_cols, bt_vals, c_vals);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_mg_load(const iga_mg_setup_t* s, const double* vals0, void* stream) {
  if (!s || s->n_levels < 1 || s->n_levels > 16 || !vals0 || !s->coarse_inv || !s->inv_work)
    return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int L = s->n_levels;
  for (int l = 0; l < L; ++l) {
    const iga_mg_setup_level_t& V = s->level[l];
    if (V.n <= 0 || !V.indptr || !V.cols || !V.vals64 || !V.vals32 || !V.dinv || !V.dsq)
      return IGA_ERR_ARG;
    if (l + 1 < L && (!V.pt_ptr || !V.pt_cols || !V.pt_vals || V.w <= 0)) return IGA_ERR_ARG;
  }
  if (s->level[L - 1].n > 4096) return IGA_ERR_SHAPE;
  {
    const iga_mg_setup_level_t& V = s->level[0];
    if (V.vals64 != vals0)
      hipLaunchKernelGGL(k_copy_d, dim3((int)std::min<int64_t>(blocks_for(V.nnz), 4096)), dim3(256), 0,
                         st, (int64_t)V.nnz, vals0, V.vals64);
  }
  {
    // the finest level's FP32 copy and scalings; the Galerkin kernel writes
    // those of every coarser level with its values
    const iga_mg_setup_level_t& V = s->level[0];
    hipLaunchKernelGGL(k_to_f32, dim3((int)std::min<int64_t>(blocks_for(V.nnz), 4096)), dim3(256), 0,
                       st, (int64_t)V.nnz, V.vals64, V.vals32);
    hipLaunchKernelGGL(k_level_scalings, dim3(blocks_for(V.n)), dim3(256), 0, st, V.n, V.indptr,
                       V.cols, V.vals64, V.dinv, V.dsq);
  }
  for (int l = 0; l + 1 < L; ++l) {
    const iga_mg_setup_level_t& V = s->level[l];
    const iga_mg_setup_level_t& C = s->level[l + 1];
    hipLaunchKernelGGL(k_galerkin_rows, dim3((C.n + 3) / 4), dim3(256), 0, st, C.n, V.w,
                       s->radius, V.indptr, V.cols, V.vals64, V.pt_ptr, V.pt_cols, V.pt_vals,
                       C.indptr, C.cols, C.vals64, C.vals32, C.dinv, C.dsq);
  }
  IGA_CHECK_LAUNCH();
  const iga_mg_setup_level_t& C = s->level[L - 1];
  return iga_dense_inverse(C.n, C.indptr, C.cols, C.vals64, s->inv_work, s->coarse_inv, stream);
}

extern "C" int iga_stencil_features(int32_t n, int32_t w, int32_t radius, const int32_t* indptr,
                                    const int32_t* cols, const double* vals, float* feats,
                                    void* stream) {
  if (n == 0) return IGA_OK;
  if (!indptr || !cols || !vals || !feats || w <= 0 || n % w || radius < 0) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int64_t total = (int64_t)(2 * radius + 1) * (2 * radius + 1) * n;
  hipLaunchKernelGGL(k_zero_range, dim3((int)std::min<int64_t>(blocks_for(total), 4096)), dim3(256),
                     0, st, total, feats);
  hipLaunchKernelGGL(k_stencil_features, dim3(blocks_for((int64_t)n * 16)), dim3(256), 0, st, n, w,
                     radius, indptr, cols, vals, feats);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_dense_inverse(int32_t n, const int32_t* indptr, const int32_t* cols,
                                 const double* vals, double* work, float* inv, void* stream) {
  if (n <= 0 || n > 4096 || !indptr || !cols || !vals || !work || !inv) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_densify, dim3(n), dim3(256), 0, st, n, indptr, cols, vals, work);
  hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(1024), n * sizeof(double), st, n, work);
  hipLaunchKernelGGL(k_inverse_f32, dim3(blocks_for((int64_t)n * n)), dim3(256), 0, st, n, work,
                     inv);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int64_t iga_mg_work_floats(const iga_mg_t* mg) {
  if (!mg || mg->n_levels < 1 || mg->n_levels > 16) return -1;
  return level_offset(*mg, mg->n_levels);
}

__global__ void k_identity_f(int n, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)n * n) out[t] = (t / n == t % n) ? 1.f : 0.f;
}

__global__ void k_transpose_f(int n, const float* __restrict__ a, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)n * n) out[t] = a[(t % n) * n + t / n];
}

// Floats of the batched cycle from level t: b, x, r for every level, each
// n_t vectors long, plus the identity input.
static int64_t tail_floats(const iga_mg_t& mg, int t) {
  const int64_t B = mg.level[t].n;
  int64_t f = B * B;
  for (int l = t; l < mg.n_levels; ++l) f += 3 * B * mg.level[l].n;
  return f;
}

// The V-cycle from level t applied to the n_t unit vectors at once, with the
// kernels of vcycle() over a batch (vector j in blockIdx.y): row j of the
// result is V e_j; `out` receives the matrix of V (its transpose).
static int tail_matrix(const iga_mg_t& mg, int t, float* work, float* out, hipStream_t st) {
  const int Lc = mg.n_levels - 1;
  const int Bn = mg.level[t].n;
  float* I = work;
  float* X[16];
  float* Bv[16];
  float* Rr[16];
  float* p = work + (int64_t)Bn * Bn;
  for (int l = t; l <= Lc; ++l) {
    const int64_t sz = (int64_t)Bn * mg.level[l].n;
    X[l] = p;
    Bv[l] = p + sz;
    Rr[l] = p + 2 * sz;
    p += 3 * sz;
  }
  const int64_t nn = (int64_t)Bn * Bn;
  hipLaunchKernelGGL(k_identity_f, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Bn, I);
  IGA_CHECK_LAUNCH();
  Bv[t] = I;
  int rc;
  for (int l = t; l < Lc; ++l) {
    const iga_mg_level_t& L = mg.level[l];
    rc = launch_rows(L.n, L.nnz, [&](auto K, int rb) {
      constexpr int LPR = decltype(K)::value;
      hipLaunchKernelGGL(k_jacobi_pre<LPR>, dim3(rb, Bn), dim3(256), 0, st, L.n, L.indptr, L.cols,
                         L.vals, L.omega, L.dsq, Bv[l], X[l], Rr[l], nullptr, (int64_t)L.n);
    });
    if (rc) return rc;
    const iga_mg_level_t& C = mg.level[l + 1];
    if ((rc = csr_f(C.n, L.pt_nnz, L.pt_ptr, L.pt_cols, L.pt_vals, Rr[l], nullptr, Bv[l + 1], 0, Bn,
                    L.n, C.n, nullptr, st)))
      return rc;
  }
  {
    const iga_mg_level_t& C = mg.level[Lc];
    hipLaunchKernelGGL(k_dense_gemv, dim3((C.n + 3) / 4, Bn), dim3(256), 0, st, C.n,
                       mg.coarse_inv, Bv[Lc], X[Lc], nullptr, (int64_t)C.n);
    IGA_CHECK_LAUNCH();
  }
  for (int l = Lc - 1; l >= t; --l) {
    const iga_mg_level_t& L = mg.level[l];
    const iga_mg_level_t& C = mg.level[l + 1];
    if ((rc = csr_f(L.n, L.p_nnz, L.p_ptr, L.p_cols, L.p_vals, X[l + 1], nullptr, X[l], 2, Bn, C.n,
                    L.n, nullptr, st)))
      return rc;
    rc = launch_rows(L.n, L.nnz, [&](auto K, int rb) {
      constexpr int LPR = decltype(K)::value;
      hipLaunchKernelGGL(k_jacobi_post<LPR>, dim3(rb, Bn), dim3(256), 0, st, L.n, L.indptr, L.cols,
                         L.vals, L.omega, L.dsq, Bv[l], X[l], Rr[l], nullptr, (int64_t)L.n);
    });
    if (rc) return rc;
    std::swap(X[l], Rr[l]);
  }
  hipLaunchKernelGGL(k_transpose_f, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, Bn, X[t],
                     out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int64_t iga_mg_tail_floats(const iga_mg_t* mg, int32_t level) {
  if (!mg || mg->n_levels < 2 || mg->n_levels > 16 || level < 0 || level >= mg->n_levels - 1)
    return -1;
  return tail_floats(*mg, level);
}

extern "C" int iga_mg_tail(const iga_mg_t* mg, int32_t level, float* work, float* out,
                           void* stream) {
  if (!mg || !work || !out || !mg->coarse_inv) return IGA_ERR_ARG;
  if (mg->n_levels < 2 || mg->n_levels > 16 || level < 0 || level >= mg->n_levels - 1)
    return IGA_ERR_ARG;
  for (int l = level; l < mg->n_levels - 1; ++l) {
    const iga_mg_level_t& L = mg->level[l];
    if (L.stencil) return IGA_ERR_UNSUPPORTED;
    if (!L.indptr || !L.cols || !L.vals || !L.dsq || !L.p_ptr || !L.pt_ptr) return IGA_ERR_ARG;
  }
  return tail_matrix(*mg, level, work, out, (hipStream_t)stream);
}

extern "C" int iga_mg_vcycle(const iga_mg_t* mg, int32_t level, const float* b, float* x,
                             void* stream) {
  int rc = check_mg(mg);
  if (rc) return rc;
  if (level < 0 || level >= mg->n_levels || !b || !x) return IGA_ERR_ARG;
  return vcycle(*mg, level, b, x, nullptr, (hipStream_t)stream);
}

extern "C" int iga_csr_f32(int32_t n, int32_t nnz, const int32_t* indptr, const int32_t* cols,
                           const float* vals, const float* x, const float* b, float* y,
                           int32_t mode, int32_t batch, int64_t x_stride, int64_t y_stride,
                           void* stream) {
  if (n == 0 || batch == 0) return IGA_OK;
  if (!indptr || !cols || !vals || !x || !y || mode < 0 || mode > 2 || (mode == 1 && !b))
    return IGA_ERR_ARG;
  return csr_f(n, nnz, indptr, cols, vals, x, b, y, mode, batch, x_stride, y_stride, nullptr,
               (hipStream_t)stream);
}

extern "C" int iga_stencil_apply(int32_t h, int32_t w, int32_t ks, const float* stencil,
                                 const float* dsq, const float* v, float* out, int32_t transpose,
                                 int32_t accumulate, int32_t batch, void* stream) {
  if ((int64_t)h * w == 0 || batch == 0) return IGA_OK;
  if (!stencil || !dsq || !v || !out || h < 0 || w < 0) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((w + kStTile - 1) / kStTile, (h + kStTile - 1) / kStTile, batch);
  return by_ks(ks, [&](auto K) {
    constexpr int KS = decltype(K)::value;
    if (transpose)
      hipLaunchKernelGGL((k_stencil<KS, true>), grid, dim3(256), 0, st, h, w, stencil, dsq, v, out,
                         accumulate, nullptr);
    else
      hipLaunchKernelGGL((k_stencil<KS, false>), grid, dim3(256), 0, st, h, w, stencil, dsq, v,
                         out, accumulate, nullptr);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

extern "C" int iga_stencil_wgrad_t(int32_t h, int32_t w, int32_t ks, int32_t batch,
                                   const float* dsq, const float* v, const float* g, float* dW,
                                   void* stream) {
  if ((int64_t)h * w == 0) return IGA_OK;
  if (!dsq || !v || !g || !dW || batch < 1) return IGA_ERR_ARG;
  const int n = h * w;
  return by_ks(ks, [&](auto K) {
    constexpr int KS = decltype(K)::value;
    hipLaunchKernelGGL(k_stencil_wgrad_t<KS>, dim3(blocks_for(n)), dim3(256), 0,
                       (hipStream_t)stream, h, w, batch, dsq, v, g, dW);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

extern "C" int iga_energy_loss(int32_t batch, int64_t n, const float* z, const float* kz,
                               float* grad, double* per_sample, double* loss, void* stream) {
  if (batch < 1 || !z || !kz || !grad || !per_sample || !loss) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_energy_sums, dim3(batch), dim3(1024), 0, st, n, z, kz, per_sample);
  hipLaunchKernelGGL(k_energy_grad, dim3(blocks_for((int64_t)batch * n)), dim3(256), 0, st, batch,
                     n, kz, per_sample, grad, loss);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int64_t iga_fcg_workspace(int32_t n) { return 8 + 6 * (int64_t)n + 4 * kRedBlocks; }

// A captured flexible-CG solve over fixed buffers: every run re-initialises
// the state on the device and replays graphs until the device flag says
// done, so one plan serves every right-hand side written into `b` and every
// operator whose values are rewritten in place. Graphs of 8, 4, 2 and 1
// iterations are kept; a run first replays as many iterations as the last
// run needed, less one, without looking, then one iteration per look, so
// few iterations run past convergence (their kernels return at once but
// still cost a launch each).
struct FcgGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct iga_fcg_plan {
  iga_mg_t mg;
  int n, nnz, lpr, vb, rb, eb, G;
  const int32_t* indptr;
  const int32_t* cols;
  const double* vals;
  const double* b;
  double* x;
  FBufs F;
  double *r, *z, *P[2], *Q[2];
  float *r32, *z32;
  FcgGraph g8, g4, g2, g1[2];  // g8/g4/g2 start at parity 0, g1[p] at parity p
  int last_iters = 0;

  int iteration(int par, hipStream_t st) {
    double* p = P[par];
    double* pp = P[1 - par];
    double* q = Q[par];
    double* qp = Q[1 - par];
    const int* skip = &F.st->done;
    int rc = vcycle(mg, 0, r32, z32, skip, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fcg_zq, dim3(vb), dim3(kVecThreads), 0, st, n, z32, z, qp, F);
    hipLaunchKernelGGL(k_fcg_dir, dim3(vb), dim3(kVecThreads), 0, st, n, vb, z, pp, p, r, F);
    by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL(k_fcg_spmv<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols,
                         vals, p, q, F);
      return 0;
    });
    hipLaunchKernelGGL(k_fcg_update, dim3(vb), dim3(kVecThreads), 0, st, n, rb, vb, p, q, x, r,
                       r32, F);
    hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kVecThreads), 0, st, vb, F, 0, 0.0, 0);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  }

  int capture_one(FcgGraph& out, int iters, int par0, hipStream_t st) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return IGA_ERR_HIP;
    int rc = 0;
    for (int k = 0; k < iters && !rc; ++k) rc = iteration((par0 + k) & 1, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : IGA_ERR_HIP;
    }
    out.graph = g;
    if (hipGraphInstantiate(&out.exec, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      out.graph = nullptr;
      out.exec = nullptr;
      return IGA_ERR_HIP;
    }
    return IGA_OK;
  }

  int capture(hipStream_t st) {
    if (G <= 0) return IGA_OK;
    int rc = capture_one(g1[0], 1, 0, st);
    if (!rc) rc = capture_one(g1[1], 1, 1, st);
    if (!rc) rc = capture_one(g2, 2, 0, st);
    if (!rc && G >= 4) rc = capture_one(g4, 4, 0, st);
    if (!rc && G >= 8) rc = capture_one(g8, 8, 0, st);
    return rc;
  }

  // k more iterations from the `launched`-th, as few graph launches as the
  // parity allows
  int replay(int k, int& launched, hipStream_t st) {
    while (k > 0) {
      const FcgGraph* g;
      int m;
      if (launched & 1) {
        g = &g1[1];
        m = 1;
      } else if (k >= 8 && g8.exec) {
        g = &g8;
        m = 8;
      } else if (k >= 4 && g4.exec) {
        g = &g4;
        m = 4;
      } else if (k >= 2) {
        g = &g2;
        m = 2;
      } else {
        g = &g1[0];
        m = 1;
      }
      if (hipGraphLaunch(g->exec, st) != hipSuccess) return IGA_ERR_HIP;
      k -= m;
      launched += m;
    }
    return IGA_OK;
  }

  int run(double tol, int maxiter, int32_t* out_iters, double* out_rr, double* out_bb,
          hipStream_t st) {
    hipLaunchKernelGGL(k_fcg_init, dim3(vb), dim3(kVecThreads), 0, st, n, b, x, r, r32, P[1], Q[1],
                       F);
    hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kVecThreads), 0, st, vb, F, 1, tol * tol,
                       maxiter);
    IGA_CHECK_LAUNCH();
    FState host{};
    int launched = 0, status = IGA_OK;
    const bool graphs = g1[0].exec != nullptr;
    // iterations before the first look: the last run's count less one (four
    // on the first run), then one per look
    int next = last_iters > 1 ? last_iters - 1 : 4;
    for (;;) {
      if (hipMemcpyAsync(&host, F.st, sizeof(FState), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        status = IGA_ERR_HIP;
        break;
      }
      if (host.done) break;
      if (graphs) {
        status = replay(next, launched, st);
        if (status) break;
        next = last_iters > 1 ? 1 : 2;
      } else {
        for (int k = 0; k < 2 && !status; ++k, ++launched) status = iteration(launched & 1, st);
        if (status) break;
      }
    }
    if (status == IGA_OK) last_iters = host.iter;
    if (out_iters) *out_iters = host.iter;
    if (out_rr) *out_rr = host.rr;
    if (out_bb) *out_bb = host.bb;
    return status;
  }

  ~iga_fcg_plan() {
    for (FcgGraph* g : {&g8, &g4, &g2, &g1[0], &g1[1]}) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
    }
  }
};

extern "C" int iga_fcg_mg_plan(const iga_mg_t* mg, int32_t n, int32_t nnz, const int32_t* indptr,
                               const int32_t* cols, const double* vals, const double* b, double* x,
                               double* work, float* work32, int32_t graph_iters, void* stream,
                               iga_fcg_plan_t** out) {
  if (!out) return IGA_ERR_ARG;
  *out = nullptr;
  int rc = check_mg(mg);
  if (rc) return rc;
  if (!indptr || !cols || !vals || !b || !x || !work || !work32 || n <= 0) return IGA_ERR_ARG;
  if (mg->level[0].n != n) return IGA_ERR_SHAPE;
  iga_fcg_plan* P = new iga_fcg_plan();
  P->mg = *mg;
  P->n = n;
  P->nnz = nnz;
  P->indptr = indptr;
  P->cols = cols;
  P->vals = vals;
  P->b = b;
  P->x = x;
  P->F.st = reinterpret_cast<FState*>(work);
  P->r = work + 8;
  P->z = P->r + n;
  P->P[0] = P->z + n;
  P->P[1] = P->z + 2 * n;
  P->Q[0] = P->z + 3 * n;
  P->Q[1] = P->z + 4 * n;
  double* parts = P->z + 5 * n;
  P->F.P_a = parts;
  P->F.P_b = parts + kRedBlocks;
  P->F.P_c = parts + 2 * kRedBlocks;
  P->F.P_d = parts + 3 * kRedBlocks;
  P->r32 = work32;
  P->z32 = work32 + n;
  P->lpr = pick_lpr(n, nnz);
  P->vb = vec_blocks(n);
  P->rb = row_blocks(n, P->lpr);
  P->eb = blocks_for(n);
  P->G = graph_iters > 0 ? graph_iters : 0;
  rc = P->capture((hipStream_t)stream);
  if (rc) {
    delete P;
    return rc;
  }
  *out = P;
  return IGA_OK;
}

extern "C" int iga_fcg_mg_run(iga_fcg_plan_t* plan, double tol, int32_t maxiter,
                              int32_t* out_iters, double* out_rr, double* out_bb, void* stream) {
  if (!plan) return IGA_ERR_ARG;
  return plan->run(tol, maxiter, out_iters, out_rr, out_bb, (hipStream_t)stream);
}

extern "C" void iga_fcg_mg_destroy(iga_fcg_plan_t* plan) { delete plan; }

extern "C" int iga_fcg_mg(const iga_mg_t* mg, int32_t n, const int32_t* indptr,
                          const int32_t* cols, const double* vals, const double* b, double* x,
                          double tol, int32_t maxiter, int32_t graph_iters, double* work,
                          float* work32, int32_t* out_iters, double* out_rr, double* out_bb,
                          void* stream) {
  if (n == 0) {
    if (out_iters) *out_iters = 0;
    return IGA_OK;
  }
  if (!indptr) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int32_t nnz = 0;
  if (hipMemcpyAsync(&nnz, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return IGA_ERR_HIP;
  iga_fcg_plan_t* plan = nullptr;
  int rc = iga_fcg_mg_plan(mg, n, nnz, indptr, cols, vals, b, x, work, work32, graph_iters, stream,
                           &plan);
  if (rc) return rc;
  rc = plan->run(tol, maxiter, out_iters, out_rr, out_bb, st);
  delete plan;
  return rc;
}
