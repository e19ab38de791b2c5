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

constexpr int kGalRows = 36;  // fine rows of a P^T row (and P^T entries of a coarse column) held in registers

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
  const int na = a1 - a0;
  // the fine rows of P^T row I and their K-row bounds, one per lane
  int my_a = 0, my_q0 = 0, my_q1 = 0;
  double my_pa = 0.0;
  if (lane < na) {
    my_a = pt_cols[a0 + lane];
    my_pa = pt_vals[a0 + lane];
    my_q0 = k_ptr[my_a];
    my_q1 = k_ptr[my_a + 1];
  }
  const int r0 = wave_min(lane < na ? my_a / wf : 1 << 30) - R;
  const int c0 = wave_min(lane < na ? my_a % wf : 1 << 30) - R;
  // every K entry of the first kGalRows fine rows, one per lane, loaded at once
  int cb[kGalRows];
  double cv[kGalRows];
#pragma unroll
  for (int k = 0; k < kGalRows; ++k) {
    const int q0 = __shfl(my_q0, k, kWave), q1 = __shfl(my_q1, k, kWave);
    const int q = (k < na && lane < q1 - q0) ? q0 + lane : -1;
    cb[k] = k_cols[q >= 0 ? q : 0];
    cv[k] = q >= 0 ? k_vals[q] : 0.0;
    cb[k] = q >= 0 ? cb[k] : -1;
  }
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_wave_barrier();
  // T = P^T K row I in the box, fine rows in order (lanes never collide
  // within one row)
#pragma unroll
  for (int k = 0; k < kGalRows; ++k) {
    if (k < na) {
      const double pa = __shfl(my_pa, k, kWave);
      if (cb[k] >= 0) {
        const int br = cb[k] / wf - r0, bc = cb[k] % wf - c0;
        box[br * kGalBox + bc] += pa * cv[k];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int k = kGalRows; k < na; ++k) {
    int q0, q1, a;
    double pa;
    if (k < kWave) {
      q0 = __shfl(my_q0, k, kWave);
      q1 = __shfl(my_q1, k, kWave);
      pa = __shfl(my_pa, k, kWave);
    } else {
      a = pt_cols[a0 + k];
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
  // diagonal is zero or absent); the P^T entries of each coarse column J are
  // loaded at once
  bool diag = false;
  for (int k = c_ptr[I] + lane; k < c_ptr[I + 1]; k += kWave) {
    const int J = c_cols[k];
    const int s0 = pt_ptr[J], s1 = pt_ptr[J + 1];
    int fb[kGalRows];
    double fw[kGalRows];
#pragma unroll
    for (int m = 0; m < kGalRows; ++m) {
      const int q = s0 + m < s1 ? s0 + m : -1;
      fb[m] = pt_cols[q >= 0 ? q : s0];
      fw[m] = q >= 0 ? pt_vals[q] : 0.0;
      fb[m] = q >= 0 ? fb[m] : -1;
    }
    __builtin_amdgcn_sched_barrier(0);
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < kGalRows; ++m) {
      if (fb[m] < 0) continue;
      const int br = fb[m] / wf - r0, bc = fb[m] % wf - c0;
      if (br >= 0 && br < kGalBox && bc >= 0 && bc < kGalBox) acc += fw[m] * box[br * kGalBox + bc];
    }
    for (int q = s0 + kGalRows; q < s1; ++q) {
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

// level 0 of a new operator: vals64 = vals (when a copy is wanted), vals32 =
// (float) vals, and the row scalings (read from vals, so no pass waits on
// another)
__global__ void k_level0_prep(int n, int64_t nnz, const int32_t* __restrict__ indptr,
                              const int32_t* __restrict__ cols, const double* __restrict__ vals,
                              double* __restrict__ vals64, float* __restrict__ vals32,
                              double* __restrict__ dinv, float* __restrict__ dsq) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double v = vals[k];
    if (vals64) vals64[k] = v;
    vals32[k] = (float)v;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int lo = indptr[i], hi = indptr[i + 1];
    const int k = lower_bound(cols, lo, hi, (int)i);
    const double d = (k < hi && cols[k] == i) ? vals[k] : 0.0;
    const double di = d != 0.0 ? 1.0 / d : 1.0;
    dinv[i] = di;
    dsq[i] = (float)sqrt(fabs(di));
  }
}

// level 0's CSR copies (FP64 when csr64, FP32) from vals, or from the
// address in *slot when given
__global__ void k_csr0_copies(int64_t nnz, const double* __restrict__ vals,
                              const double* const* __restrict__ slot, double* __restrict__ csr64,
                              float* __restrict__ csr32) {
  const double* v = slot ? *slot : vals;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double x = v[k];
    if (csr64) csr64[k] = x;
    csr32[k] = (float)x;
  }
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

// Gauss-Jordan elimination of [A | I] without pivoting (A is SPD), one block,
// with the augmented matrix held in LDS (n <= kGjLds); written back at the end.
constexpr int kGjLds = 48;

// The FP32 inverse of a small SPD CSR matrix in one block: [A | I] is formed,
// eliminated and read out in LDS.
__global__ void __launch_bounds__(1024)
    k_dense_inverse_lds(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
                        const double* __restrict__ vals, float* __restrict__ inv) {
  __shared__ double M[kGjLds * 2 * kGjLds];
  __shared__ double fac[kGjLds];
  const int n2 = 2 * n;
  for (int t = threadIdx.x; t < n * n2; t += blockDim.x) {
    const int i = t / n2, j = t - i * n2;
    M[t] = j == n + i ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    for (int k = indptr[i]; k < indptr[i + 1]; ++k) M[i * n2 + cols[k]] = vals[k];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double piv = M[k * n2 + k];
    __syncthreads();
    for (int j = threadIdx.x; j < n2; j += blockDim.x) M[k * n2 + j] /= piv;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) fac[i] = M[i * n2 + k];
    __syncthreads();
    for (int t = threadIdx.x; t < n * n2; t += blockDim.x) {
      const int i = t / n2, j = t - i * n2;
      if (i != k) M[t] -= fac[i] * M[k * n2 + j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t - i * n;
    inv[t] = (float)M[i * n2 + n + j];
  }
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

// Device state of a flexible-CG solve (scalars never leave the device).
struct FState {
  double rr, bb, tol2, pq_prev;
  int iter, maxiter, done;
  int pending;  // an update x += alpha p, r -= alpha q waits (grid cycle)
  double res2;  // ||b - A x||^2 recomputed from x after the solve
  double* x_out;  // the caller's copy of x, written by the flush (grid cycle)
};

// Gate of the first kernel of a V-cycle. Inside flexible CG it is also the
// convergence test of the previous iteration: every block reduces the r.r
// partials of the last update and returns when ||r||^2 <= tol^2 ||b||^2;
// block 0 records rr and the flag, and counts the iteration it lets through.
// Every later kernel of the iteration reads that flag (`flag`).
struct Gate {
  const int* flag;   // return when *flag != 0 (null: never)
  const double* rr;  // r.r partials (null: only `flag`)
  int n_rr;
  FState* st;
  int clear_pending;  // block 0 also clears st->pending (the update was applied)
};

// A negative st->bb (set by an init that skips its own reduction) means the
// partials the first gate reduces are those of r = b: that gate records them
// as ||b||^2. Blocks that read st->bb while block 0 writes it see either the
// sentinel or the same value, so every block makes the same test.
__device__ __forceinline__ bool gate_decide(const Gate& g, double rr, double tol2, double bb) {
  const bool first = bb < 0.0;
  bb = first ? rr : bb;
  const bool done = !(rr > tol2 * bb);
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
    if (first) g.st->bb = bb;
    g.st->rr = rr;
    g.st->done = done ? 1 : 0;
    if (!done) g.st->iter += 1;
    if (g.clear_pending) g.st->pending = 0;
  }
  return done;
}

__device__ __forceinline__ bool gate_closed(const Gate& g) {
  if (g.rr) return gate_decide(g, reduce_partials(g.rr, g.n_rr), g.st->tol2, g.st->bb);
  return g.flag && *g.flag;
}

// gate_closed with its loads issued early (gate_load, beside the kernel's
// other loads) and the test made later (gate_test): one trip to memory less
struct GateVals {
  double share, tol2, bb;
  int flag;
};

// (blocks of 256 threads: at most kRedBlocks / 256 partials per thread)
__device__ __forceinline__ GateVals gate_load(const Gate& g) {
  GateVals v{0.0, 0.0, 0.0, 0};
  if (g.rr) {
    v.share = partials_share<kRedBlocks / 256>(g.rr, g.n_rr);
    v.tol2 = g.st->tol2;
    v.bb = g.st->bb;
  } else if (g.flag) {
    v.flag = *g.flag;
  }
  return v;
}

__device__ __forceinline__ bool gate_test(const Gate& g, const GateVals& v) {
  if (g.rr) return gate_decide(g, partials_total(v.share), v.tol2, v.bb);
  return v.flag != 0;
}

// Optional dot product fused into the kernel that writes a vector z:
// parts[block] = sum over the block's entries of z_i y_i (FP64).
struct DotOut {
  const double* y;
  double* parts;
};

__device__ __forceinline__ void store_partial_at(double* parts, int idx, double v) {
  __shared__ double scratch[16];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) parts[idx] = v;
}

// out (=|+=) D^-1/2 T (D^-1/2 v), T = T_W or its transpose; batch over
// blockIdx.z. A block computes a 16 x 16 tile of the grid: u = D^-1/2 v over
// the tile and its halo is staged in LDS (zero outside the grid), so the tap
// loop has no bounds tests on u; the weights stream from HBM, coalesced
// along the tile rows.
constexpr int kStTile = 16;

template <int KS, bool TR>
__global__ void __launch_bounds__(256)
    k_stencil(int h, int w, const float* __restrict__ W, const float* __restrict__ dsq,
              const float* __restrict__ v, float* __restrict__ out, int accumulate, Gate gate,
              DotOut dot) {
  if (gate_closed(gate)) return;
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
  const bool inside = r < h && c < w;
  const int i = inside ? r * w + c : 0;
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
  float z = 0.f;
  if (inside) {
    const float val = dsq[i] * acc;
    z = accumulate ? ob[i] + val : val;
    ob[i] = z;
  }
  if (dot.parts)  // one grid of tiles (blockIdx.z == 0): partial per tile
    store_partial_at(dot.parts, blockIdx.y * gridDim.x + blockIdx.x,
                     inside ? (double)z * dot.y[i] : 0.0);
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
                 Gate gate, int64_t vs) {
  if (gate_closed(gate)) return;
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
                 float* __restrict__ x, Gate gate, int64_t vs) {
  if (gate_closed(gate)) return;
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

// b = a, with the partials of b . dot.y when requested (grid-stride, at most
// kRedBlocks blocks)
__global__ void __launch_bounds__(256)
    k_copy_fs(int n, const float* __restrict__ a, float* __restrict__ b,
              const int* __restrict__ skip, DotOut dot) {
  if (skipped(skip)) return;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float v = a[i];
    b[i] = v;
    if (dot.parts) s += (double)v * dot.y[i];
  }
  if (dot.parts) store_partial_at(dot.parts, blockIdx.x, s);
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
                  int accumulate, const Gate& gate, const DotOut& dot, hipStream_t st) {
  const iga_mg_level_t& L = mg.level[l];
  if (!L.stencil) {
    if (gate.rr || dot.parts) return IGA_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(k_jacobi_f, dim3(blocks_for(L.n)), dim3(256), 0, st, L.n, L.omega, L.dsq, v,
                       out, accumulate, gate.flag);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }
  const dim3 grid((L.w + kStTile - 1) / kStTile, (L.h + kStTile - 1) / kStTile, 1);
  return by_ks(mg.ks, [&](auto K) {
    constexpr int KS = decltype(K)::value;
    if (transpose)
      hipLaunchKernelGGL((k_stencil<KS, true>), grid, dim3(256), 0, st, L.h, L.w, L.stencil, L.dsq,
                         v, out, accumulate, gate, dot);
    else
      hipLaunchKernelGGL((k_stencil<KS, false>), grid, dim3(256), 0, st, L.h, L.w, L.stencil,
                         L.dsq, v, out, accumulate, gate, dot);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

// tiles of the stencil kernel on level l (its partial count when it carries a dot)
static int stencil_tiles(const iga_mg_level_t& L) {
  return ((L.w + kStTile - 1) / kStTile) * ((L.h + kStTile - 1) / kStTile);
}

// Leading dimension of the dense tail matrix: rows padded with zeros to a
// multiple of 256 floats, so a wavefront reads a row as float4s unconditionally.
static inline int dense_ld(int n) { return (n + 255) / 256 * 256; }
constexpr int kDenseMaxRows = 2048;

// x = M b for the dense tail (n x n, leading dimension ld = dense_ld(n)): one
// wavefront per row, every lane issues its NV float4 loads at once.
template <int NV>
__global__ void __launch_bounds__(256)
    k_dense_gemv_ld(int n, int ld, const float* __restrict__ M, const float* __restrict__ b,
                    float* __restrict__ x, Gate gate, int64_t vs) {
  if (gate_closed(gate)) return;
  b += blockIdx.y * vs;
  x += blockIdx.y * vs;
  const int i = blockIdx.x * 4 + threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (i >= n) return;
  const float4* row = reinterpret_cast<const float4*>(M + (size_t)i * ld);
  float4 m[NV];
  float bv[NV][4];
#pragma unroll
  for (int v = 0; v < NV; ++v) m[v] = row[l + kWave * v];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int k = 0; k < 4; ++k) bv[v][k] = b[min(4 * (l + kWave * v) + k, n - 1)];
  float s = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v) s += m[v].x * bv[v][0] + m[v].y * bv[v][1] + m[v].z * bv[v][2] + m[v].w * bv[v][3];
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (l == 0) x[i] = s;
}

static int dense_tail_gemv(int n, const float* M, const float* b, float* x, const Gate& gate,
                           int batch, hipStream_t st) {
  const int ld = dense_ld(n);
  const int nv = ld / 256;
  auto go = [&](auto V) {
    constexpr int NV = decltype(V)::value;
    hipLaunchKernelGGL(k_dense_gemv_ld<NV>, dim3((n + 3) / 4, batch), dim3(256), 0, st, n, ld, M, b,
                       x, gate, (int64_t)n);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  };
  switch (nv) {
    case 1: return go(std::integral_constant<int, 1>{});
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 5: return go(std::integral_constant<int, 5>{});
    case 6: return go(std::integral_constant<int, 6>{});
    case 7: return go(std::integral_constant<int, 7>{});
    case 8: return go(std::integral_constant<int, 8>{});
    default: return IGA_ERR_UNSUPPORTED;
  }
}

// four vectors per level (x, b, r and the post-smoothed x of the grid cycle)
static int64_t level_offset(const iga_mg_t& mg, int l) {
  int64_t o = 0;
  for (int k = 0; k < l; ++k) o += 4 * (int64_t)mg.level[k].n;
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

// partials of x . y (x FP32, y FP64), grid-stride over at most kRedBlocks blocks
__global__ void __launch_bounds__(256)
    k_dot_fd(int n, const float* __restrict__ x, const int* __restrict__ skip, DotOut dot) {
  if (skipped(skip)) return;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += (double)x[i] * dot.y[i];
  store_partial_at(dot.parts, blockIdx.x, s);
}

// ---------------------------------------------------------------------------
// Grid V-cycle: tile kernels over the level grids (operators in stencil
// layout, transfers as 1-D bands). On the coarse levels each kernel computes
// a 16 x 16 tile of its output grid and recomputes what it needs of the
// previous step on the tile's halo in LDS, so every transfer runs fused with
// the smoother beside it; on level 0 (row tiles, below) each big array is
// read once per point:
//   k_g0_smooth     level 0: x0 = S b0 (learned stencil or damped Jacobi)
//   k_g0_resid      level 0: r0 = b0 - A0 x0
//   k_g_restrict    level l+1: b = P^T r_l, x = w D^-1 b, r = b - A x
//   k_g_tail        level t: x_t = M P^T r_{t-1} (dense tail)
//   k_g_prolong     level l: x += P x_{l+1}, x += w D^-1 (b - A x)
//   k_g_prolong_p   level t-2: the same from level t-1, whose values it forms
//                   from the tail first (the prolongation into t-1 folded in)
//   k_g0_prolong    level 0: x = x0 + P x1, u = D^-1/2 (b0 - A0 x)
//   k_g0_smooth_t   level 0: z = x + D^-1/2 T_W^T u (+ z . q)
// Values off the grid are zero. A level's radius-3 operator reaches a 22 x 22
// halo of a 16 x 16 tile and a 70 x 10 box around a 64 x 4 row tile.
// ---------------------------------------------------------------------------

constexpr int kGR = 3;                     // operator and smoother radius
constexpr int kGS = 2 * kGR + 1;           // 7
constexpr int kGN = kGS * kGS;             // 49 stencil offsets
constexpr int kGT = 16;                    // tile
constexpr int kGH1 = kGT + 2 * kGR;        // 22
constexpr int kGH2 = kGT + 4 * kGR;        // 28
constexpr int kGTh = 256;                  // threads per tile block
constexpr int kKPF = 4, kKFC = 6;          // band widths (coarse per fine, fine per coarse)
constexpr int kFineSpan = 56;              // fine rows/cols staged by k_g_restrict
constexpr int kCoarseSpan = 24;            // coarse rows/cols staged by the prolongations
constexpr int kTailFine = 8192, kTailCoarse = 2048;

__host__ __device__ constexpr int passes(int n) { return (n + kGTh - 1) / kGTh; }

struct GLevel {
  int h, w;
  const float* A;    // (49, n) stencil layout
  const float* dsq;  // D^-1/2
  float omega;
};

struct GBand {
  const int32_t* c_lo;
  const float* pw;  // [nf][4]
  const int32_t* f_lo;
  const float* rw;  // [nc][6]
};

__device__ __forceinline__ bool on_grid(int r, int c, int h, int w) {
  return r >= 0 && r < h && c >= 0 && c < w;
}

// the 49 stencil values of grid point i, all loads issued before any use
template <typename T>
__device__ __forceinline__ void load_taps(const T* __restrict__ A, int n, int i, T (&a)[kGN]) {
#pragma unroll
  for (int o = 0; o < kGN; ++o) a[o] = A[(size_t)o * n + i];
  __builtin_amdgcn_sched_barrier(0);
}

// sum_o a[o] v[ly + dy][lx + dx], v an LDS array of row length LD holding the
// neighbourhood of the point at (ly, lx)
template <int LD, typename T, typename V>
__device__ __forceinline__ T tap_sum(const T (&a)[kGN], const V (*v)[LD], int ly, int lx) {
  T acc = 0;
#pragma unroll
  for (int dy = -kGR; dy <= kGR; ++dy)
#pragma unroll
    for (int dx = -kGR; dx <= kGR; ++dx)
      acc += a[(dy + kGR) * kGS + dx + kGR] * (T)v[ly + dy][lx + dx];
  return acc;
}

// stage x_c over coarse rows [cr0, cr0 + 24) x cols [cc0, cc0 + 24), all
// loads in flight together
__device__ __forceinline__ void stage_coarse(float (*cs)[kCoarseSpan], const float* __restrict__ xc,
                                             int hc, int wc, int cr0, int cc0) {
  constexpr int N = kCoarseSpan * kCoarseSpan, NP = passes(N);
  float v[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int t = threadIdx.x + k * kGTh;
    const int rr = cr0 + t / kCoarseSpan, cc = cc0 + t % kCoarseSpan;
    const bool in = t < N && on_grid(rr, cc, hc, wc);
    v[k] = xc[in ? rr * wc + cc : 0];
    v[k] = in ? v[k] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int t = threadIdx.x + k * kGTh;
    if (t < N) cs[t / kCoarseSpan][t % kCoarseSpan] = v[k];
  }
}

// coarse tile of level l+1 from the fine residual r_f of level l. Every load
// that depends on no other (the skip flag, the staging origin, the band rows
// and scalings of the halo points, the tile point's taps) is issued first, in
// one group, and the fine residual is staged behind them: two trips to memory.
__global__ void __launch_bounds__(kGTh)
    k_g_restrict(int hf, int wf, GLevel gc, GBand br, GBand bc, const float* __restrict__ rf,
                 float* __restrict__ bo, float* __restrict__ xo, float* __restrict__ ro,
                 const int* __restrict__ skip) {
  __shared__ float fs[kFineSpan][kFineSpan];
  __shared__ float bs[kGH1][kGH1];
  __shared__ float xs[kGH1][kGH1];
  constexpr int NH = passes(kGH1 * kGH1);
  const int nc = gc.h * gc.w;
  const int R0 = blockIdx.y * kGT, C0 = blockIdx.x * kGT;
  const int skipv = skip ? *skip : 0;
  const int fr0 = br.f_lo[max(R0 - kGR, 0)], fc0 = bc.f_lo[max(C0 - kGR, 0)];
  int fa[NH], fb[NH];
  float wr[NH][kKFC], wc[NH][kKFC], dh[NH];
  bool hin[NH];
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int t = threadIdx.x + k * kGTh;
    const int ly = min(t / kGH1, kGH1 - 1), lx = t % kGH1;
    const int R = R0 - kGR + ly, C = C0 - kGR + lx;
    hin[k] = t < kGH1 * kGH1 && on_grid(R, C, gc.h, gc.w);
    const int Rc = hin[k] ? R : 0, Cc = hin[k] ? C : 0;
    fa[k] = br.f_lo[Rc];
    fb[k] = bc.f_lo[Cc];
#pragma unroll
    for (int a = 0; a < kKFC; ++a) {
      wr[k][a] = br.rw[Rc * kKFC + a];
      wc[k][a] = bc.rw[Cc * kKFC + a];
    }
    dh[k] = gc.dsq[Rc * gc.w + Cc];
  }
  const int ty = threadIdx.x / kGT, tx = threadIdx.x % kGT;
  const int R = R0 + ty, C = C0 + tx;
  const bool in = on_grid(R, C, gc.h, gc.w);
  const int i = in ? R * gc.w + C : 0;
  float a[kGN];
#pragma unroll
  for (int o = 0; o < kGN; ++o) a[o] = gc.A[(size_t)o * nc + i];
  __builtin_amdgcn_sched_barrier(0);
  if (skipv) return;
  {
    constexpr int N = kFineSpan * kFineSpan, NP = passes(N);
    float v[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      const int rr = fr0 + t / kFineSpan, cc = fc0 + t % kFineSpan;
      const bool fin = t < N && on_grid(rr, cc, hf, wf);
      v[k] = rf[fin ? rr * wf + cc : 0];
      v[k] = fin ? v[k] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      if (t < N) fs[t / kFineSpan][t % kFineSpan] = v[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    const int t = threadIdx.x + k * kGTh;
    const int ly = min(t / kGH1, kGH1 - 1), lx = t % kGH1;
    const int ka = fa[k] - fr0, kb = fb[k] - fc0;
    float bv = 0.f;
#pragma unroll
    for (int q = 0; q < kKFC; ++q) {
      const int ia = min(max(ka + q, 0), kFineSpan - 1);
      float s = 0.f;
#pragma unroll
      for (int u = 0; u < kKFC; ++u) s += wc[k][u] * fs[ia][min(max(kb + u, 0), kFineSpan - 1)];
      bv += wr[k][q] * s;
    }
    if (t < kGH1 * kGH1) {
      bs[ly][lx] = hin[k] ? bv : 0.f;
      xs[ly][lx] = hin[k] ? gc.omega * dh[k] * dh[k] * bv : 0.f;
    }
  }
  __syncthreads();
  const float ax = tap_sum<kGH1>(a, xs, ty + kGR, tx + kGR);
  if (!in) return;
  const float bv = bs[ty + kGR][tx + kGR];
  bo[i] = bv;
  xo[i] = xs[ty + kGR][tx + kGR];
  ro[i] = bv - ax;
}

// x_t = M (P^T r_f) on the tail level: every block restricts the whole
// (small) fine residual into LDS, then one wavefront per row of M. The row of
// M, the first fine values and the band rows of the first coarse points are
// loaded in one group with the skip flag.
template <int NV>
__global__ void __launch_bounds__(kGTh)
    k_g_tail(int hf, int wf, int hc, int wc, GBand br, GBand bc, const float* __restrict__ rf,
             const float* __restrict__ M, int ld, float* __restrict__ xo,
             const int* __restrict__ skip) {
  __shared__ float fs[kTailFine];
  __shared__ float bs[kTailCoarse];
  constexpr int NF = 8, NJ = 2;  // fine values and coarse points per thread in the first group
  const int nf = hf * wf, nc = hc * wc;
  const int i = blockIdx.x * 4 + threadIdx.x / kWave, l = threadIdx.x % kWave;
  const int skipv = skip ? *skip : 0;
  const float4* row = reinterpret_cast<const float4*>(M + (size_t)min(i, nc - 1) * ld);
  float4 m[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) m[v] = row[l + kWave * v];
  float fv[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const int t = threadIdx.x + k * kGTh;
    fv[k] = rf[min(t, nf - 1)];
  }
  int fa[NJ], fb[NJ];
  float wr[NJ][kKFC], wcl[NJ][kKFC];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int J = min((int)threadIdx.x + j * kGTh, nc - 1);
    const int R = J / wc, C = J - R * wc;
    fa[j] = br.f_lo[R];
    fb[j] = bc.f_lo[C];
#pragma unroll
    for (int a = 0; a < kKFC; ++a) {
      wr[j][a] = br.rw[R * kKFC + a];
      wcl[j][a] = bc.rw[C * kKFC + a];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  if (skipv) return;
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const int t = threadIdx.x + k * kGTh;
    if (t < nf) fs[t] = fv[k];
  }
  for (int t = threadIdx.x + NF * kGTh; t < nf; t += kGTh) fs[t] = rf[t];
  __syncthreads();
  // P^T r_f at coarse point J from the band rows (ja, war) x (jb, wac)
  auto restrict_at = [&](int ja, int jb, const float(&war)[kKFC], const float(&wac)[kKFC]) {
    float bv = 0.f;
#pragma unroll
    for (int a = 0; a < kKFC; ++a) {
      const int ra = min(ja + a, hf - 1);
      float s = 0.f;
#pragma unroll
      for (int q = 0; q < kKFC; ++q) s += wac[q] * fs[ra * wf + min(jb + q, wf - 1)];
      bv += war[a] * s;
    }
    return bv;
  };
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int J = threadIdx.x + j * kGTh;
    if (J < nc) bs[J] = restrict_at(fa[j], fb[j], wr[j], wcl[j]);
  }
  for (int J = threadIdx.x + NJ * kGTh; J < nc; J += kGTh) {
    const int R = J / wc, C = J - R * wc;
    float war[kKFC], wac[kKFC];
#pragma unroll
    for (int a = 0; a < kKFC; ++a) {
      war[a] = br.rw[R * kKFC + a];
      wac[a] = bc.rw[C * kKFC + a];
    }
    bs[J] = restrict_at(br.f_lo[R], bc.f_lo[C], war, wac);
  }
  __syncthreads();
  if (i >= nc) return;
  float s = 0.f;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int j = 4 * (l + kWave * v);
    s += m[v].x * bs[min(j, nc - 1)] + m[v].y * bs[min(j + 1, nc - 1)] +
         m[v].z * bs[min(j + 2, nc - 1)] + m[v].w * bs[min(j + 3, nc - 1)];
  }
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (l == 0) xo[i] = s;
}

// The first group of loads of a fine tile of level l >= 1 (the halo's
// x_pre and band rows, the tile point's taps, scaling and b) and, once the
// coarse values are staged in cs, the rest: x = x_pre + P x_c on the halo,
// then one damped Jacobi step x + w D^-1 (b - A x) on the tile.
struct ProlongTile {
  static constexpr int NH = passes(kGH1 * kGH1);
  float xp[NH], wr[NH][kKPF], wcl[NH][kKPF];
  int ka[NH], kb[NH];
  bool hin[NH];
  float a[kGN];
  float d, bi;
  int i;
  bool in;

  __device__ __forceinline__ void load(const GLevel& g, const GBand& br, const GBand& bc,
                                       const float* __restrict__ xpre,
                                       const float* __restrict__ bf, int r0, int c0) {
    const int n = g.h * g.w;
#pragma unroll
    for (int k = 0; k < NH; ++k) {
      const int t = threadIdx.x + k * kGTh;
      const int ly = min(t / kGH1, kGH1 - 1), lx = t % kGH1;
      const int rr = r0 - kGR + ly, cc = c0 - kGR + lx;
      hin[k] = t < kGH1 * kGH1 && on_grid(rr, cc, g.h, g.w);
      const int rc = hin[k] ? rr : 0, ccc = hin[k] ? cc : 0;
      xp[k] = xpre[rc * g.w + ccc];
      ka[k] = br.c_lo[rc];
      kb[k] = bc.c_lo[ccc];
#pragma unroll
      for (int q = 0; q < kKPF; ++q) {
        wr[k][q] = br.pw[rc * kKPF + q];
        wcl[k][q] = bc.pw[ccc * kKPF + q];
      }
    }
    const int ty = threadIdx.x / kGT, tx = threadIdx.x % kGT;
    const int rr = r0 + ty, cc = c0 + tx;
    in = on_grid(rr, cc, g.h, g.w);
    i = in ? rr * g.w + cc : 0;
#pragma unroll
    for (int o = 0; o < kGN; ++o) a[o] = g.A[(size_t)o * n + i];
    d = g.dsq[i];
    bi = bf[i];
  }

  // cs: the coarse values from (cr0, cc0), staged and synchronised
  __device__ __forceinline__ void finish(const GLevel& g, const float (*cs)[kCoarseSpan],
                                         float (*xs)[kGH1], int cr0, int cc0,
                                         float* __restrict__ xo) const {
#pragma unroll
    for (int k = 0; k < NH; ++k) {
      const int t = threadIdx.x + k * kGTh;
      const int ly = min(t / kGH1, kGH1 - 1), lx = t % kGH1;
      const int ia0 = ka[k] - cr0, ib0 = kb[k] - cc0;
      float v = 0.f;
#pragma unroll
      for (int p = 0; p < kKPF; ++p) {
        const int ia = min(max(ia0 + p, 0), kCoarseSpan - 1);
        float s2 = 0.f;
#pragma unroll
        for (int q = 0; q < kKPF; ++q)
          s2 += wcl[k][q] * cs[ia][min(max(ib0 + q, 0), kCoarseSpan - 1)];
        v += wr[k][p] * s2;
      }
      if (t < kGH1 * kGH1) xs[ly][lx] = hin[k] ? xp[k] + v : 0.f;
    }
    __syncthreads();
    const int ty = threadIdx.x / kGT, tx = threadIdx.x % kGT;
    const float ax = tap_sum<kGH1>(a, xs, ty + kGR, tx + kGR);
    if (!in) return;
    xo[i] = xs[ty + kGR][tx + kGR] + g.omega * d * d * (bi - ax);
  }
};

// fine tile of level l >= 1 from the coarse values of level l + 1. As in
// k_g_restrict, every independent load goes out in the first group; the
// coarse values follow.
__global__ void __launch_bounds__(kGTh)
    k_g_prolong(GLevel g, int hc, int wc, GBand br, GBand bc, const float* __restrict__ xc,
                const float* __restrict__ xpre, const float* __restrict__ bf,
                float* __restrict__ xo, const int* __restrict__ skip) {
  __shared__ float cs[kCoarseSpan][kCoarseSpan];
  __shared__ float xs[kGH1][kGH1];
  const int r0 = blockIdx.y * kGT, c0 = blockIdx.x * kGT;
  const int skipv = skip ? *skip : 0;
  const int cr0 = br.c_lo[max(r0 - kGR, 0)], cc0 = bc.c_lo[max(c0 - kGR, 0)];
  ProlongTile T;
  T.load(g, br, bc, xpre, bf, r0, c0);
  __builtin_amdgcn_sched_barrier(0);
  if (skipv) return;
  stage_coarse(cs, xc, hc, wc, cr0, cc0);
  __syncthreads();
  T.finish(g, cs, xs, cr0, cc0, xo);
}

// The prolongation from the tail into level p = t - 1 and its damped Jacobi
// post-smoothing, folded into the prolongation into level g = p - 1: a tile
// block of level g forms the level-p values its coarse window needs,
//   y = x_p + P_p x_t                    (the window and a border of kGR),
//   x_p' = y + w D_p^-1 (b_p - A_p y)    (the window),
// in LDS, then finishes as k_g_prolong with x_p' as the coarse values. The
// window is the coarse rows and columns the tile's halo reaches (at most
// kCoarseSpan, as for k_g_prolong's staging).
constexpr int kPBor = kCoarseSpan + 2 * kGR;  // the window with its border
constexpr int kPTail = 512;                   // tail values staged

struct ProlongP {
  GLevel g, p;     // levels g and p
  int ht, wt;      // level t
  GBand brg, bcg;  // g -> p (rows, columns)
  GBand brp, bcp;  // p -> t
};

__global__ void __launch_bounds__(kGTh)
    k_g_prolong_p(ProlongP a, const float* __restrict__ xt, const float* __restrict__ xp_pre,
                  const float* __restrict__ bp, const float* __restrict__ xpre,
                  const float* __restrict__ bf, float* __restrict__ xo,
                  const int* __restrict__ skip) {
  __shared__ float cs[kCoarseSpan][kCoarseSpan];  // x_p' (zero off the window and the grid)
  __shared__ float ys[kPBor][kPBor];              // y
  __shared__ float xs[kGH1][kGH1];
  __shared__ float ts[kPTail];                    // x_t
  __shared__ int ylo[2][kPBor];                   // tail band rows of the bordered window
  __shared__ float ypw[2][kPBor][kKPF];
  constexpr int NT = kPTail / kGTh;
  constexpr int NB = passes(kPBor * kPBor);
  const GLevel& P = a.p;
  const int np = P.h * P.w, nt = a.ht * a.wt;
  const int r0 = blockIdx.y * kGT, c0 = blockIdx.x * kGT;
  const int skipv = skip ? *skip : 0;
  // the window: coarse rows cr0 .. cr0 + wh - 1 and columns cc0 .. cc0 + ww - 1
  const int cr0 = a.brg.c_lo[max(r0 - kGR, 0)], cc0 = a.bcg.c_lo[max(c0 - kGR, 0)];
  const int cr1 = a.brg.c_lo[min(r0 + kGT + kGR - 1, a.g.h - 1)] + kKPF - 1;
  const int cc1 = a.bcg.c_lo[min(c0 + kGT + kGR - 1, a.g.w - 1)] + kKPF - 1;
  ProlongTile T;
  T.load(a.g, a.brg, a.bcg, xpre, bf, r0, c0);
  float tv[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) tv[k] = xt[min((int)threadIdx.x + k * kGTh, nt - 1)];
  __builtin_amdgcn_sched_barrier(0);
  if (skipv) return;
  const int wh = min(min(cr1, P.h - 1) - cr0 + 1, kCoarseSpan);
  const int ww = min(min(cc1, P.w - 1) - cc0 + 1, kCoarseSpan);
  const int bh = wh + 2 * kGR, bw = ww + 2 * kGR;
  // second group: the tail band rows of the bordered window's rows (threads
  // 0 .. bh-1) and columns (threads 64 .. 64+bw-1), x_p on it, and the
  // scaling, b_p and taps of this thread's first window point
  int yl = 0;
  float yw[kKPF] = {0.f, 0.f, 0.f, 0.f};
  const bool is_row = (int)threadIdx.x < bh;
  const bool is_col = (int)threadIdx.x >= 64 && (int)threadIdx.x < 64 + bw;
  if (is_row || is_col) {
    const int q = is_row ? min(max(cr0 - kGR + (int)threadIdx.x, 0), P.h - 1)
                         : min(max(cc0 - kGR + (int)threadIdx.x - 64, 0), P.w - 1);
    const GBand& B = is_row ? a.brp : a.bcp;
    yl = B.c_lo[q];
#pragma unroll
    for (int k = 0; k < kKPF; ++k) yw[k] = B.pw[q * kKPF + k];
  }
  float xv[NB];
  bool xin[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int t = threadIdx.x + k * kGTh;
    const int rr = cr0 - kGR + t / max(bw, 1), cc = cc0 - kGR + t % max(bw, 1);
    xin[k] = t < bh * bw && on_grid(rr, cc, P.h, P.w);
    xv[k] = xp_pre[xin[k] ? rr * P.w + cc : 0];
  }
  const int nw = wh * ww;
  auto win_point = [&](int t, int& K) {
    const int rr = cr0 + t / max(ww, 1), cc = cc0 + t % max(ww, 1);
    const bool on = t < nw && on_grid(rr, cc, P.h, P.w);
    K = on ? rr * P.w + cc : 0;
    return on;
  };
  int K0;
  const bool on0 = win_point(threadIdx.x, K0);
  float pa[kGN];
#pragma unroll
  for (int o = 0; o < kGN; ++o) pa[o] = P.A[(size_t)o * np + K0];
  const float pd0 = P.dsq[K0], pb0 = bp[K0];
  __builtin_amdgcn_sched_barrier(0);
  if (is_row) {
    ylo[0][threadIdx.x] = yl;
#pragma unroll
    for (int k = 0; k < kKPF; ++k) ypw[0][threadIdx.x][k] = yw[k];
  }
  if (is_col) {
    ylo[1][threadIdx.x - 64] = yl;
#pragma unroll
    for (int k = 0; k < kKPF; ++k) ypw[1][threadIdx.x - 64][k] = yw[k];
  }
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const int t = threadIdx.x + k * kGTh;
    if (t < nt) ts[t] = tv[k];
  }
  __syncthreads();
  // y on the bordered window
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int t = threadIdx.x + k * kGTh;
    if (t < bh * bw) {
      const int br = t / bw, bc = t - br * bw;
      float v = 0.f;
#pragma unroll
      for (int p = 0; p < kKPF; ++p) {
        const int ra = min(ylo[0][br] + p, a.ht - 1);
        float s2 = 0.f;
#pragma unroll
        for (int q = 0; q < kKPF; ++q) s2 += ypw[1][bc][q] * ts[ra * a.wt + min(ylo[1][bc] + q, a.wt - 1)];
        v += ypw[0][br][p] * s2;
      }
      ys[br][bc] = xin[k] ? xv[k] + v : 0.f;
    }
  }
  __syncthreads();
  // x_p' on the window (zero elsewhere in cs)
  auto smooth_at = [&](int t, int K, float d, float b, const float(&av)[kGN]) {
    const int wr = t / ww, wc = t - wr * ww;
    float s = 0.f;
#pragma unroll
    for (int dy = -kGR; dy <= kGR; ++dy)
#pragma unroll
      for (int dx = -kGR; dx <= kGR; ++dx)
        s += av[(dy + kGR) * kGS + dx + kGR] * ys[wr + kGR + dy][wc + kGR + dx];
    return ys[wr + kGR][wc + kGR] + P.omega * d * d * (b - s);
  };
  for (int t = threadIdx.x; t < kCoarseSpan * kCoarseSpan; t += kGTh)
    cs[t / kCoarseSpan][t % kCoarseSpan] = 0.f;
  __syncthreads();
  if (on0) {
    const int t = threadIdx.x;
    cs[t / ww][t % ww] = smooth_at(t, K0, pd0, pb0, pa);
  }
  for (int t = threadIdx.x + kGTh; t < nw; t += kGTh) {
    int K;
    if (!win_point(t, K)) continue;
#pragma unroll
    for (int o = 0; o < kGN; ++o) pa[o] = P.A[(size_t)o * np + K];
    cs[t / ww][t % ww] = smooth_at(t, K, P.dsq[K], bp[K], pa);
  }
  __syncthreads();
  T.finish(a.g, cs, xs, cr0, cc0, xo);
}

// ---------------------------------------------------------------------------
// Level-0 kernels of the grid cycle: row tiles of 64 x 4 points (one
// wavefront per 64-point row segment, so every load of a stencil array is
// 256 contiguous bytes), each big array (learned weights W, operator A) read
// once per point; only vectors are staged on the 70 x 10 halo box.
// ---------------------------------------------------------------------------

constexpr int kL0W = 64, kL0H = 4;
constexpr int kL0HW = kL0W + 2 * kGR, kL0HH = kL0H + 2 * kGR;  // 70 x 10
constexpr int kL0Box = kL0HW * kL0HH;
constexpr int kL0CW = 44, kL0CH = 12;  // coarse points staged by the level-0 prolongation

struct L0Tile {
  int r0, c0, ty, tx, rt, ct, it;
  bool in;
};

__device__ __forceinline__ L0Tile l0_tile(int h, int w) {
  L0Tile t;
  t.r0 = blockIdx.y * kL0H;
  t.c0 = blockIdx.x * kL0W;
  t.ty = threadIdx.x / kL0W;
  t.tx = threadIdx.x % kL0W;
  t.rt = t.r0 + t.ty;
  t.ct = t.c0 + t.tx;
  t.in = on_grid(t.rt, t.ct, h, w);
  t.it = t.in ? t.rt * w + t.ct : 0;
  return t;
}

// halo-box loads of v (times dsq when SCALE) into registers; stored by box_store
template <bool SCALE, typename T>
struct BoxLoad {
  static constexpr int NP = passes(kL0Box);
  T v[NP];
  float d[NP];
  bool in[NP];
  __device__ __forceinline__ void load(const T* __restrict__ src, const float* __restrict__ dsq,
                                       int h, int w, int r0, int c0) {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      const int rr = r0 - kGR + t / kL0HW, cc = c0 - kGR + t % kL0HW;
      in[k] = t < kL0Box && on_grid(rr, cc, h, w);
      const int i = in[k] ? rr * w + cc : 0;
      v[k] = src[i];
      if (SCALE) d[k] = dsq[i];
    }
  }
  __device__ __forceinline__ void store(T (*box)[kL0HW]) const {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      if (t < kL0Box) box[t / kL0HW][t % kL0HW] = in[k] ? (SCALE ? (T)d[k] * v[k] : v[k]) : (T)0;
    }
  }
};

// x0 = S b0: D^-1/2 T_W (D^-1/2 b0), or w D^-1 b0 without a stencil; the
// first kernel of a cycle (gated)
template <bool STW, typename WT = float>
__global__ void __launch_bounds__(kGTh)
    k_g0_smooth(GLevel g, const WT* __restrict__ W, const float* __restrict__ b,
                float* __restrict__ x, Gate gate) {
  __shared__ float ub[kL0HH][kL0HW];
  const int n = g.h * g.w;
  const L0Tile T = l0_tile(g.h, g.w);
  float a[kGN];
  if (STW)
#pragma unroll
    for (int o = 0; o < kGN; ++o) a[o] = W[(size_t)o * n + T.it];
  const float dt = g.dsq[T.it], bt = b[T.it];
  BoxLoad<true, float> bl;
  if (STW) bl.load(b, g.dsq, g.h, g.w, T.r0, T.c0);
  const GateVals gv = gate_load(gate);
  __builtin_amdgcn_sched_barrier(0);
  if (gate_test(gate, gv)) return;
  float v;
  if (STW) {
    bl.store(ub);
    __syncthreads();
    v = dt * tap_sum<kL0HW>(a, ub, T.ty + kGR, T.tx + kGR);
  } else {
    v = g.omega * dt * dt * bt;
  }
  if (T.in) x[T.it] = v;
}

// Flexible-CG state for the level-0 kernel that applies the waiting update
// of the previous iteration: alpha = (p.r)/(p.q) from the product's
// partials, r_new = r_old - alpha q (ping-pong buffers: the halo reads r_old),
// x += alpha p, r32 = (float) r_new and the partials of r_new . r_new.
struct FcgFuse {
  FState* st;
  const double* r_old;
  double* r_new;
  const double* p;
  const double* q;
  double* x;
  float* r32;
  const double* P_b;  // p.r partials of the product
  const double* P_c;  // p.q partials
  double* P_d;        // r.r partials out (one per tile)
  int nb;             // partials of the product
};

// the update (when one waits), then x0 = S b0 with b0 = r32 on the box
template <bool STW, typename WT = float>
__global__ void __launch_bounds__(kGTh)
    k_g0_smooth_upd(GLevel g, const WT* __restrict__ W, FcgFuse f, float* __restrict__ x0) {
  __shared__ float ub[kL0HH][kL0HW];
  const int n = g.h * g.w;
  const L0Tile T = l0_tile(g.h, g.w);
  float a[kGN];
  if (STW)
#pragma unroll
    for (int o = 0; o < kGN; ++o) a[o] = W[(size_t)o * n + T.it];
  const float dt = g.dsq[T.it];
  const double pt = f.p[T.it], xt = f.x[T.it];
  const double ro = f.r_old[T.it], qo = f.q[T.it];
  BoxLoad<false, double> rl, ql;
  BoxLoad<false, float> dl;
  rl.load(f.r_old, nullptr, g.h, g.w, T.r0, T.c0);
  ql.load(f.q, nullptr, g.h, g.w, T.r0, T.c0);
  dl.load(g.dsq, nullptr, g.h, g.w, T.r0, T.c0);
  // the state and the product's partials in the same group
  const int done = f.st->done, pending = f.st->pending;
  const double sc = partials_share<kRedBlocks / kGTh>(f.P_c, f.nb);
  const double sb = partials_share<kRedBlocks / kGTh>(f.P_b, f.nb);
  __builtin_amdgcn_sched_barrier(0);
  if (done) return;
  const bool pend = pending != 0;
  double alpha = 0.0, pq = 0.0;
  if (pend) {
    pq = partials_total(sc);
    alpha = partials_total(sb) / pq;
  }
#pragma unroll
  for (int k = 0; k < BoxLoad<false, float>::NP; ++k) {
    const int t = threadIdx.x + k * kGTh;
    const double rn = rl.v[k] - alpha * ql.v[k];
    if (t < kL0Box) ub[t / kL0HW][t % kL0HW] = rl.in[k] ? dl.v[k] * (float)rn : 0.f;
  }
  __syncthreads();
  float v;
  if (STW)
    v = dt * tap_sum<kL0HW>(a, ub, T.ty + kGR, T.tx + kGR);
  else
    v = g.omega * dt * ub[T.ty + kGR][T.tx + kGR];
  double rr = 0.0;
  if (T.in) {
    x0[T.it] = v;
    if (pend) f.x[T.it] = xt + alpha * pt;
  }
  // the FP64 new residual of the tile point
  {
    const double rn = ro - alpha * qo;
    if (T.in) {
      f.r_new[T.it] = rn;
      f.r32[T.it] = (float)rn;
      rr = rn * rn;
    }
  }
  store_partial_at(f.P_d, blockIdx.y * gridDim.x + blockIdx.x, rr);
  if (pend && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) f.st->pq_prev = pq;
}

// the waiting update in place at the end of a replay (each point its own),
// with the r.r partials for the check that follows
__global__ void __launch_bounds__(kGTh) k_fcg_flush(int h, int w, FcgFuse f) {
  const L0Tile T = l0_tile(h, w);
  // the caller's copy of x: every replay ends here, also one whose gate
  // closed inside it (x is then final already)
  double* xo = f.st->x_out;
  if (f.st->done) {
    if (xo && T.in) xo[T.it] = f.x[T.it];
    return;
  }
  const bool pend = f.st->pending != 0;
  double alpha = 0.0, pq = 0.0;
  if (pend) {
    pq = reduce_partials(f.P_c, f.nb);
    alpha = reduce_partials(f.P_b, f.nb) / pq;
  }
  double rr = 0.0;
  if (T.in) {
    const double rn = f.r_old[T.it] - alpha * f.q[T.it];
    double xn = f.x[T.it];
    if (pend) {
      f.r_new[T.it] = rn;
      f.r32[T.it] = (float)rn;
      xn += alpha * f.p[T.it];
      f.x[T.it] = xn;
    }
    if (xo) xo[T.it] = xn;
    rr = rn * rn;
  }
  store_partial_at(f.P_d, blockIdx.y * gridDim.x + blockIdx.x, rr);
  if (pend && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) f.st->pq_prev = pq;
}

// r0 = b0 - A0 x0
__global__ void __launch_bounds__(kGTh)
    k_g0_resid(GLevel g, const float* __restrict__ x, const float* __restrict__ b,
               float* __restrict__ r, Gate gate) {
  __shared__ float xb[kL0HH][kL0HW];
  const int n = g.h * g.w;
  const L0Tile T = l0_tile(g.h, g.w);
  float a[kGN];
#pragma unroll
  for (int o = 0; o < kGN; ++o) a[o] = g.A[(size_t)o * n + T.it];
  const float bt = b[T.it];
  BoxLoad<false, float> bl;
  bl.load(x, nullptr, g.h, g.w, T.r0, T.c0);
  const GateVals gv = gate_load(gate);
  __builtin_amdgcn_sched_barrier(0);
  if (gate_test(gate, gv)) return;
  bl.store(xb);
  __syncthreads();
  const float ax = tap_sum<kL0HW>(a, xb, T.ty + kGR, T.tx + kGR);
  if (T.in) r[T.it] = bt - ax;
}

// x = x0 + P x1 on the halo box (prolongation from the staged coarse values
// and the band rows of the box), then u = D^-1/2 (b0 - A0 x) and x on the tile
__global__ void __launch_bounds__(kGTh)
    k_g0_prolong(GLevel g, int hc, int wc, GBand br, GBand bc, const float* __restrict__ xc,
                 const float* __restrict__ x0, const float* __restrict__ b0,
                 float* __restrict__ xo, float* __restrict__ uo, const int* __restrict__ skip) {
  __shared__ float cs[kL0CH][kL0CW];
  __shared__ float xb[kL0HH][kL0HW];
  __shared__ int rlo[kL0HH], clo[kL0HW];
  __shared__ float rpw[kL0HH][kKPF], cpw[kL0HW][kKPF];
  const int n = g.h * g.w;
  const L0Tile T = l0_tile(g.h, g.w);
  // band rows of the box's rows (threads 0..9) and columns (threads 64..133)
  int bl = 0;
  float bw[kKPF] = {0.f, 0.f, 0.f, 0.f};
  const bool is_row = threadIdx.x < kL0HH;
  const bool is_col = threadIdx.x >= 64 && threadIdx.x < 64 + kL0HW;
  if (is_row || is_col) {
    const int q = is_row ? min(max(T.r0 - kGR + (int)threadIdx.x, 0), g.h - 1)
                         : min(max(T.c0 - kGR + (int)threadIdx.x - 64, 0), g.w - 1);
    const GBand& B = is_row ? br : bc;
    bl = B.c_lo[q];
#pragma unroll
    for (int k = 0; k < kKPF; ++k) bw[k] = B.pw[q * kKPF + k];
  }
  float a[kGN];
#pragma unroll
  for (int o = 0; o < kGN; ++o) a[o] = g.A[(size_t)o * n + T.it];
  const float bt = b0[T.it], dt = g.dsq[T.it];
  BoxLoad<false, float> xl;
  xl.load(x0, nullptr, g.h, g.w, T.r0, T.c0);
  const int cr0 = br.c_lo[max(T.r0 - kGR, 0)], cc0 = bc.c_lo[max(T.c0 - kGR, 0)];
  const int skipv = skip ? *skip : 0;
  __builtin_amdgcn_sched_barrier(0);
  if (skipv) return;
  {
    constexpr int N = kL0CH * kL0CW, NP = passes(N);
    float v[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      const int rr = cr0 + t / kL0CW, cc = cc0 + t % kL0CW;
      const bool in = t < N && on_grid(rr, cc, hc, wc);
      v[k] = xc[in ? rr * wc + cc : 0];
      v[k] = in ? v[k] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      if (t < N) cs[t / kL0CW][t % kL0CW] = v[k];
    }
  }
  if (is_row) {
    rlo[threadIdx.x] = bl;
#pragma unroll
    for (int k = 0; k < kKPF; ++k) rpw[threadIdx.x][k] = bw[k];
  }
  if (is_col) {
    clo[threadIdx.x - 64] = bl;
#pragma unroll
    for (int k = 0; k < kKPF; ++k) cpw[threadIdx.x - 64][k] = bw[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < BoxLoad<false, float>::NP; ++k) {
    const int t = threadIdx.x + k * kGTh;
    const int ly = min(t / kL0HW, kL0HH - 1), lx = t % kL0HW;
    const int ka = rlo[ly] - cr0, kb = clo[lx] - cc0;
    float v = 0.f;
#pragma unroll
    for (int p = 0; p < kKPF; ++p) {
      const int ia = min(max(ka + p, 0), kL0CH - 1);
      float s2 = 0.f;
#pragma unroll
      for (int q = 0; q < kKPF; ++q) s2 += cpw[lx][q] * cs[ia][min(max(kb + q, 0), kL0CW - 1)];
      v += rpw[ly][p] * s2;
    }
    if (t < kL0Box) xb[ly][lx] = xl.in[k] ? xl.v[k] + v : 0.f;
  }
  __syncthreads();
  const float ax = tap_sum<kL0HW>(a, xb, T.ty + kGR, T.tx + kGR);
  if (!T.in) return;
  xo[T.it] = xb[T.ty + kGR][T.tx + kGR];
  uo[T.it] = dt * (bt - ax);
}

// z = x + D^-1/2 T_W^T u (or x + w D^-1/2 u), u = D^-1/2 r, with the partial
// of z . dot.y per tile; the transposed stencil reads the neighbours' weights
template <bool STW, typename WT = float>
__global__ void __launch_bounds__(kGTh)
    k_g0_smooth_t(GLevel g, const WT* __restrict__ W, const float* __restrict__ x,
                  const float* __restrict__ u, float* __restrict__ z,
                  const int* __restrict__ skip, DotOut dot) {
  __shared__ float ubx[kL0HH][kL0HW];
  const int n = g.h * g.w;
  const L0Tile T = l0_tile(g.h, g.w);
  float a[kGN];
  if (STW) {
#pragma unroll
    for (int dy = -kGR; dy <= kGR; ++dy)
#pragma unroll
      for (int dx = -kGR; dx <= kGR; ++dx) {
        const int o = (dy + kGR) * kGS + dx + kGR;
        const bool in = on_grid(T.rt - dy, T.ct - dx, g.h, g.w);
        a[o] = W[(size_t)o * n + (in ? (T.rt - dy) * g.w + T.ct - dx : T.it)];
        a[o] = in ? a[o] : 0.f;
      }
  }
  const float xt = x[T.it], dt = g.dsq[T.it], ut = u[T.it];
  const double yt = dot.parts ? dot.y[T.it] : 0.0;
  BoxLoad<false, float> bl;
  if (STW) bl.load(u, nullptr, g.h, g.w, T.r0, T.c0);
  const int skipv = skip ? *skip : 0;
  __builtin_amdgcn_sched_barrier(0);
  if (skipv) return;
  float acc;
  if (STW) {
    bl.store(ubx);
    __syncthreads();
    acc = 0.f;
#pragma unroll
    for (int dy = -kGR; dy <= kGR; ++dy)
#pragma unroll
      for (int dx = -kGR; dx <= kGR; ++dx)
        acc += a[(dy + kGR) * kGS + dx + kGR] * ubx[T.ty + kGR - dy][T.tx + kGR - dx];
  } else {
    acc = g.omega * ut;
  }
  const float zi = xt + dt * acc;
  if (T.in) z[T.it] = zi;
  if (dot.parts)
    store_partial_at(dot.parts, blockIdx.y * gridDim.x + blockIdx.x, T.in ? (double)zi * yt : 0.0);
}

// FP64 flexible-CG product on level 0 in stencil layout: beta from the
// z . q_prev partials, p = z + beta p_prev on the halo box (LDS), q = A p on
// the tile; partials of p . q and p . r per tile
__global__ void __launch_bounds__(kGTh)
    k_fcg_spmv_grid(int h, int w, int nblk_a, const double* __restrict__ A,
                    const float* __restrict__ z32, const double* __restrict__ pp,
                    const double* __restrict__ r, double* __restrict__ p, double* __restrict__ q,
                    FState* st, const double* __restrict__ P_a, double* __restrict__ P_b,
                    double* __restrict__ P_c) {
  __shared__ double pb[kL0HH][kL0HW];
  const int n = h * w;
  const L0Tile T = l0_tile(h, w);
  double a[kGN];
#pragma unroll
  for (int o = 0; o < kGN; ++o) a[o] = A[(size_t)o * n + T.it];
  const double ri = r[T.it];
  BoxLoad<false, float> zl;
  BoxLoad<false, double> pl;
  zl.load(z32, nullptr, h, w, T.r0, T.c0);
  pl.load(pp, nullptr, h, w, T.r0, T.c0);
  const int done = st->done;
  const double pq_prev = st->pq_prev;
  const double sa = partials_share<kRedBlocks / kGTh>(P_a, nblk_a);
  __builtin_amdgcn_sched_barrier(0);
  if (done) return;
  const double beta = -partials_total(sa) / pq_prev;
#pragma unroll
  for (int k = 0; k < BoxLoad<false, float>::NP; ++k) {
    const int t = threadIdx.x + k * kGTh;
    if (t < kL0Box) pb[t / kL0HW][t % kL0HW] = zl.in[k] ? (double)zl.v[k] + beta * pl.v[k] : 0.0;
  }
  __syncthreads();
  const double qi = tap_sum<kL0HW>(a, pb, T.ty + kGR, T.tx + kGR);
  const double pi = pb[T.ty + kGR][T.tx + kGR];
  if (T.in) {
    p[T.it] = pi;
    q[T.it] = qi;
  }
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  store_partial_at(P_c, blk, T.in ? pi * qi : 0.0);
  store_partial_at(P_b, blk, T.in ? pi * ri : 0.0);
  // the update x += alpha p, r -= alpha q waits for the next cycle (or the flush)
  if (blk == 0 && threadIdx.x == 0) st->pending = 1;
}

// stencil layout of the level operators (blockIdx.y = level; the FP64 copy of
// level 0 too when op64 is set). A block converts 64 consecutive rows: their
// CSR segment is staged in LDS with coalesced reads, then every (offset, row)
// pair is written coalesced along the rows. A row's CSR entries are its
// in-grid neighbours in row-major order, so entry (dy, dx) sits at an offset
// computed from the clipped window; an entry found elsewhere (a pattern with
// holes, or a segment larger than the staging) is searched in the CSR row.
struct GridOpsArgs {
  int n_levels;
  int n[16], w[16];
  const int32_t* indptr[16];
  const int32_t* cols[16];
  const float* vals[16];
  float* op[16];
  const double* vals64;
  const double* const* vals64_slot;  // optional: device memory holding vals64
  double* op64;
  float* feats;  // level 0: K[i, i + off] / K[i, i] (FP32), with op64
  // level 0 from vals64 alone (vals[0] null): its CSR copies and scalings
  double* csr64;  // optional copy of vals64
  float* csr32;
  double* dinv;
  float* dsq;
};

constexpr int kOpsRows = 64;
constexpr int kOpsMax = kOpsRows * kGN;

__global__ void __launch_bounds__(kGTh) k_grid_ops(GridOpsArgs a) {
  // the staged segment: columns and values (FP32 levels' values widened
  // exactly, so one array serves both)
  __shared__ int sc[kOpsMax];
  __shared__ double sv[kOpsMax];
  __shared__ int sp[kOpsRows + 1];
  const int l = blockIdx.y;
  if (l >= a.n_levels) return;
  const int n = a.n[l], w = a.w[l], h = n / w;
  const int i0 = blockIdx.x * kOpsRows;
  if (i0 >= n) return;
  const int i1 = min(i0 + kOpsRows, n);
  const double* vals64 = a.vals64_slot ? *a.vals64_slot : a.vals64;
  const int32_t* cols = a.cols[l];
  const float* vals = a.vals[l];  // null: the FP32 values from vals64 (level 0)
  const int lo = a.indptr[l][i0], cnt = a.indptr[l][i1] - lo;
  const bool f64 = l == 0 && a.op64;
  const bool prep = l == 0 && !vals;
  const bool fits = cnt <= kOpsMax;
  if (threadIdx.x <= i1 - i0) sp[threadIdx.x] = a.indptr[l][i0 + threadIdx.x] - lo;
  if (fits && cnt > 0) {
    constexpr int NP = passes(kOpsMax);
    int c[NP];
    double v[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = min((int)threadIdx.x + k * kGTh, max(cnt - 1, 0));
      c[k] = cols[lo + t];
      v[k] = prep || f64 ? vals64[lo + t] : (double)vals[lo + t];
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int t = threadIdx.x + k * kGTh;
      if (t < cnt) {
        sc[t] = c[k];
        sv[t] = v[k];
        if (prep && a.csr32) {
          if (a.csr64) a.csr64[lo + t] = v[k];
          a.csr32[lo + t] = (float)v[k];
        }
      }
    }
  } else if (prep && a.csr32) {
    for (int t = threadIdx.x; t < cnt; t += kGTh) {
      const double v64 = vals64[lo + t];
      if (a.csr64) a.csr64[lo + t] = v64;
      a.csr32[lo + t] = (float)v64;
    }
  }
  __syncthreads();
  // row i's entry at grid offset (ey, ex): at its computed slot of the
  // staged segment (the in-grid neighbours in row-major order), else
  // searched in the CSR row (0 off the grid or absent); FP32 in v32
  auto entry = [&](int i, int ey, int ex, float& v32) -> double {
    v32 = 0.f;
    const int il = i - i0, r = i / w, c = i - r * w;
    if (!on_grid(r + ey, c + ex, h, w)) return 0.0;
    const int s0 = sp[il], e0 = sp[il + 1];
    const int dy0 = max(-kGR, -r), dx0 = max(-kGR, -c);
    const int ncol = min(kGR, w - 1 - c) - dx0 + 1;
    const int j = i + ey * w + ex;
    const int k = s0 + (ey - dy0) * ncol + (ex - dx0);
    if (fits && k < e0 && sc[k] == j) {
      v32 = (float)sv[k];
      return f64 || prep ? sv[k] : 0.0;
    }
    double v64 = 0.0;
    for (int q = lo + s0; q < lo + e0; ++q)
      if (cols[q] == j) {
        v64 = f64 || prep ? vals64[q] : 0.0;
        v32 = prep ? (float)v64 : vals[q];
      }
    return v64;
  };
  if (prep && (int)threadIdx.x < i1 - i0) {
    // the row's scalings: 1 / K_ii (1 where zero or absent), D^-1/2
    const int i = i0 + threadIdx.x;
    float dummy;
    const double d = entry(i, 0, 0, dummy);
    const double di = d != 0.0 ? 1.0 / d : 1.0;
    a.dinv[i] = di;
    a.dsq[i] = (float)sqrt(fabs(di));
  }
  for (int t = threadIdx.x; t < kGN * kOpsRows; t += kGTh) {
    const int il = t % kOpsRows, o = t / kOpsRows;
    const int i = i0 + il;
    if (i >= i1) continue;
    const int dy = o / kGS - kGR, dx = o % kGS - kGR;
    float v;
    const double v64 = entry(i, dy, dx, v);
    a.op[l][(size_t)o * n + i] = v;
    if (f64) {
      a.op64[(size_t)o * n + i] = v64;
      if (a.feats) {
        float dummy;
        double d = entry(i, 0, 0, dummy);
        if (d == 0.0) d = 1.0;
        a.feats[(size_t)o * n + i] = (float)(v64 / d);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Tensor Galerkin passes: K_c = P^T K P with P = P1 (x) P0 in two 1-D passes
// over FP64 stencil layouts (radius 3 in and out):
//   pass 0 (along grid rows): K'[(r, C), (dr, dC)] = sum_{c, c'} P0[c, C]
//          K[(r, c), (r + dr, c')] P0[c', C + dC]
//   pass 1 (across rows):     K_c[(R, C), (dR, dC)] = sum_{r, r'} P1[r, R]
//          K'[(r, C), (r', C + dC)] P1[r', R + dR]
// Each thread keeps u[j] = sum_c P[c, C] K[c, c'] for the KF + 6 points c'
// of its window (c' = f_lo[C] - 3 + j), then forms the 7 outputs of its
// offset line from u and the coarse-column bands of its neighbours. KF is
// the band width (fine values per coarse column, the row stride of rw): 6
// for one level's transfer; a coarse level two or three levels down is
// formed straight from a finer one through the product of the transfers in
// between, whose bands are 14 and 30 wide. The passes are the same: P^T K P
// with P = P_s ... P_{m-1} equals the level-by-level products.
// ---------------------------------------------------------------------------

// out[d] = sum_j u[j] P[f0 - 3 + j, C + d - 3] for d < 7 (zero for a coarse
// index off [0, nc)); P[f, C2] = rw[C2][f - f_lo[C2]] inside the band
template <int KF>
__device__ __forceinline__ void gal_close(const double (&u)[KF + 2 * kGR], int f0, int C, int nc,
                                          const int32_t* __restrict__ f_lo,
                                          const double* __restrict__ rw, double (&out)[kGS]) {
#pragma unroll
  for (int d = 0; d < kGS; ++d) {
    const int C2 = C + d - kGR;
    const bool in = C2 >= 0 && C2 < nc;
    const int Cc = in ? C2 : 0;
    const int off = f0 - kGR - f_lo[Cc];
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < KF + 2 * kGR; ++j) {
      const int k = j + off;
      const double wv = rw[Cc * KF + min(max(k, 0), KF - 1)];
      s += (k >= 0 && k < KF) ? u[j] * wv : 0.0;
    }
    out[d] = in ? s : 0.0;
  }
}

// Wide bands (two or three levels): the band of the close step is staged in
// LDS with gal_pad zeros before and after every column's KF values, so the
// KF + 6 values a neighbour contributes are read at consecutive addresses
// from one base, with no test per value (the host checks that the offsets
// stay inside the padding); the window's operator reads are unclamped (the
// band is zero past a column's last fine point, and the fine stencils carry
// kGalTail zero values past their end), consecutive per stencil offset.
constexpr int kGalTail = 32;
constexpr int kBufWord3 = 0x00020000;  // raw buffer descriptor, dword 3 (32-bit data)
constexpr int kGalLdsBand = 4096;  // doubles of LDS for the padded band

template <int KF>
constexpr int gal_pad() {
  return KF <= kKFC ? 9 : KF <= 14 ? 16 : 28;
}

template <int KF>
constexpr int gal_row() {
  return KF + 2 * gal_pad<KF>();
}

template <int KF>
__device__ __forceinline__ void gal_close_lds(const double (&u)[KF + 2 * kGR], int f0, int C, int nc,
                                              const int32_t* __restrict__ f_lo,
                                              const double* __restrict__ s_rw, double (&out)[kGS]) {
  int base[kGS];
#pragma unroll
  for (int d = 0; d < kGS; ++d) {
    const int C2 = C + d - kGR;
    const int Cc = C2 >= 0 && C2 < nc ? C2 : 0;
    // inside the staged band (off-grid neighbours read any values, dropped)
    base[d] = min(max(Cc * gal_row<KF>() + gal_pad<KF>() + (f0 - kGR - f_lo[Cc]), 0),
                  nc * gal_row<KF>() - (KF + 2 * kGR));
    out[d] = 0.0;
  }
#pragma unroll
  for (int j = 0; j < KF + 2 * kGR; ++j) {
#pragma unroll
    for (int d = 0; d < kGS; ++d) out[d] += u[j] * s_rw[base[d] + j];
    if (KF > kKFC && j % 6 == 5) __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int d = 0; d < kGS; ++d)
    if (C + d - kGR < 0 || C + d - kGR >= nc) out[d] = 0.0;
}

// the padded band of n coarse columns into LDS (whole workgroup)
template <int KF>
__device__ __forceinline__ void gal_stage_band(const double* __restrict__ rw, int n,
                                               double* __restrict__ s_rw) {
  constexpr int RS = gal_row<KF>();
  for (int i = threadIdx.x; i < n * RS; i += blockDim.x) {
    const int c = i / RS, k = i - c * RS - gal_pad<KF>();
    s_rw[i] = k >= 0 && k < KF ? rw[c * KF + k] : 0.0;
  }
  __syncthreads();
}

// One 1-D Galerkin step of a level (see above): the arguments of both passes.
struct GalLevel {
  int hf, wf, hc, wc;
  const int32_t* f_lo[2];
  const double* rw[2];  // rows of KF values
  const double* K;  // fine FP64 stencil
  double* Kp;       // intermediate
  double* S64;      // coarse FP64 stencil
  float* S32;       // coarse FP32 stencil (optional)
  const int32_t* indptr;
  const int32_t* cols;
  double* vals64;
  float* vals32;
  double* dinv;
  float* dsq;
  int regular;  // every coarse CSR row is exactly its clipped 7 x 7 window
};

// pass 0 entry t = (dr, r, C), C fastest; WIDE: the band of the close step
// is the padded LDS copy s_rw
template <int KF, bool WIDE>
__device__ __forceinline__ void gal_pass0_entry(const GalLevel& g, int64_t t, const double* s_rw) {
  const int64_t np = (int64_t)g.hf * g.wc;
  const int C = (int)(t % g.wc), r = (int)((t / g.wc) % g.hf), dr = (int)(t / np) - kGR;
  const int64_t nf = (int64_t)g.hf * g.wf;
  const int f0 = g.f_lo[0][C];
  double u[KF + 2 * kGR];
#pragma unroll
  for (int j = 0; j < KF + 2 * kGR; ++j) u[j] = 0.0;
  if constexpr (WIDE) {
    const double* row = g.K + (int64_t)(dr + kGR) * kGS * nf + (int64_t)r * g.wf + f0;
#pragma unroll
    for (int dc = -kGR; dc <= kGR; ++dc) {
      const double* kr = row + (int64_t)(dc + kGR) * nf;
#pragma unroll
      for (int a = 0; a < KF; ++a) u[a + dc + kGR] += g.rw[0][C * KF + a] * kr[a];
      // wide bands: one offset's reads in flight at a time (all of them
      // would not fit the registers)
      if (KF > kKFC) __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int a = 0; a < KF; ++a) {
      const double wa = g.rw[0][C * KF + a];
      const int64_t idx = (int64_t)r * g.wf + min(f0 + a, g.wf - 1);
#pragma unroll
      for (int dc = -kGR; dc <= kGR; ++dc)
        u[a + dc + kGR] += wa * g.K[((dr + kGR) * kGS + dc + kGR) * nf + idx];
    }
  }
  double out[kGS];
  if constexpr (WIDE)
    gal_close_lds<KF>(u, f0, C, g.wc, g.f_lo[0], s_rw, out);
  else
    gal_close<KF>(u, f0, C, g.wc, g.f_lo[0], g.rw[0], out);
#pragma unroll
  for (int d = 0; d < kGS; ++d) g.Kp[((dr + kGR) * kGS + d) * np + (int64_t)r * g.wc + C] = out[d];
}

// pass 1 entry t = (dC, R, C), C fastest: the coarse stencil line (., dC) of
// point (R, C) in FP64 (and FP32) stencil layout, its CSR values and, for
// dC = 0, the diagonal scalings (1 where the diagonal is zero)
template <int KF, bool WIDE>
__device__ __forceinline__ void gal_pass1_entry(const GalLevel& g, int64_t t, const double* s_rw) {
  const int64_t nc = (int64_t)g.hc * g.wc, np = (int64_t)g.hf * g.wc;
  const int C = (int)(t % g.wc), R = (int)((t / g.wc) % g.hc), dC = (int)(t / nc) - kGR;
  const int f0 = g.f_lo[1][R];
  double u[KF + 2 * kGR];
#pragma unroll
  for (int j = 0; j < KF + 2 * kGR; ++j) u[j] = 0.0;
  if constexpr (WIDE) {
    // buffer reads: the lane's offset in a register, each value's distance
    // from it (uniform) in a scalar register (the host keeps the
    // intermediate below 2 GiB)
    const __amdgpu_buffer_rsrc_t kp = __builtin_amdgcn_make_buffer_rsrc(g.Kp, (short)0, 0x7fffffff, kBufWord3);
    const int voff = (int)(((dC + kGR) * np + (int64_t)f0 * g.wc + C) * 8);
#pragma unroll
    for (int dr = -kGR; dr <= kGR; ++dr) {
#pragma unroll
      for (int a = 0; a < KF; ++a) {
        const int soff = (int)((((int64_t)(dr + kGR) * kGS) * np + (int64_t)a * g.wc) * 8);
        const double v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(kp, voff, soff, 0));
        u[a + dr + kGR] += g.rw[1][R * KF + a] * v;
      }
      if (KF > kKFC) __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int a = 0; a < KF; ++a) {
      const double wa = g.rw[1][R * KF + a];
      const int64_t idx = (int64_t)min(f0 + a, g.hf - 1) * g.wc + C;
#pragma unroll
      for (int dr = -kGR; dr <= kGR; ++dr)
        u[a + dr + kGR] += wa * g.Kp[((dr + kGR) * kGS + dC + kGR) * np + idx];
    }
  }
  double out[kGS];
  if constexpr (WIDE)
    gal_close_lds<KF>(u, f0, R, g.hc, g.f_lo[1], s_rw, out);
  else
    gal_close<KF>(u, f0, R, g.hc, g.f_lo[1], g.rw[1], out);
  const int I = R * g.wc + C;
  const bool cin = C + dC >= 0 && C + dC < g.wc;
  const int lo = g.indptr[I], hi = g.indptr[I + 1];
  const int dR0 = max(-kGR, -R), dC0 = max(-kGR, -C);
  const int ncol = min(kGR, g.wc - 1 - C) - dC0 + 1;
#pragma unroll
  for (int d = 0; d < kGS; ++d) {
    const int dR = d - kGR;
    const double v = cin ? out[d] : 0.0;
    const int o = d * kGS + dC + kGR;
    g.S64[(int64_t)o * nc + I] = v;
    if (g.S32) g.S32[(int64_t)o * nc + I] = (float)v;
    if (cin && R + dR >= 0 && R + dR < g.hc) {
      const int J = I + dR * g.wc + dC;
      int k = lo + (dR - dR0) * ncol + (dC - dC0);
      if (!g.regular && (k >= hi || g.cols[k] != J)) {
        k = -1;
        for (int q = lo; q < hi; ++q)
          if (g.cols[q] == J) k = q;
      }
      if (k >= 0) {
        g.vals64[k] = v;
        g.vals32[k] = (float)v;
      }
    }
  }
  if (dC == 0) {
    const double dv = out[kGR];
    const double di = dv != 0.0 ? 1.0 / dv : 1.0;
    g.dinv[I] = di;
    g.dsq[I] = (float)sqrt(fabs(di));
  }
}

__global__ void __launch_bounds__(256) k_gal_pass0(GalLevel g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)kGS * g.hf * g.wc) gal_pass0_entry<kKFC, false>(g, t, nullptr);
}

__global__ void __launch_bounds__(256) k_gal_pass1(GalLevel g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)kGS * g.hc * g.wc) gal_pass1_entry<kKFC, false>(g, t, nullptr);
}

// Band widths of the transfer from a level to the levels one, two and three
// below it, and the passes that form up to three coarse levels from one fine
// level in one launch each (g[k] the level k + 1 below; blockIdx.y = NT-1-k,
// so the blocks of the widest bands are dispatched first)
constexpr int kGalKF[3] = {kKFC, 14, 30};

struct GalMulti {
  GalLevel g[3];
};

template <int NT>
__global__ void __launch_bounds__(256) k_gal_multi0(GalMulti m) {
  __shared__ double s_rw[kGalLdsBand];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = NT - 1 - (int)blockIdx.y;
  if (k == 0) {
    const GalLevel& g = m.g[0];
    if (t < (int64_t)kGS * g.hf * g.wc) gal_pass0_entry<kGalKF[0], false>(g, t, nullptr);
  } else if (NT > 1 && k == 1) {
    const GalLevel& g = m.g[1];
    gal_stage_band<kGalKF[1]>(g.rw[0], g.wc, s_rw);
    if (t < (int64_t)kGS * g.hf * g.wc) gal_pass0_entry<kGalKF[1], true>(g, t, s_rw);
  } else if (NT > 2 && k == 2) {
    const GalLevel& g = m.g[2];
    gal_stage_band<kGalKF[2]>(g.rw[0], g.wc, s_rw);
    if (t < (int64_t)kGS * g.hf * g.wc) gal_pass0_entry<kGalKF[2], true>(g, t, s_rw);
  }
}

template <int NT>
__global__ void __launch_bounds__(256) k_gal_multi1(GalMulti m) {
  __shared__ double s_rw[kGalLdsBand];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = NT - 1 - (int)blockIdx.y;
  if (k == 0) {
    const GalLevel& g = m.g[0];
    if (t < (int64_t)kGS * g.hc * g.wc) gal_pass1_entry<kGalKF[0], false>(g, t, nullptr);
  } else if (NT > 1 && k == 1) {
    const GalLevel& g = m.g[1];
    gal_stage_band<kGalKF[1]>(g.rw[1], g.hc, s_rw);
    if (t < (int64_t)kGS * g.hc * g.wc) gal_pass1_entry<kGalKF[1], true>(g, t, s_rw);
  } else if (NT > 2 && k == 2) {
    const GalLevel& g = m.g[2];
    gal_stage_band<kGalKF[2]>(g.rw[1], g.hc, s_rw);
    if (t < (int64_t)kGS * g.hc * g.wc) gal_pass1_entry<kGalKF[2], true>(g, t, s_rw);
  }
}

// ---------------------------------------------------------------------------
// Dense tail from the tensor set-up: the V-cycle started at level t (damped
// Jacobi pre- and post-smoothing with M = w D^-1, exact coarsest solve) is
//   V_Lc = A_Lc^-1,   V_l = 2 M_l - M_l A_l M_l + G_l V_{l+1} G_l^T,
//   G_l = (I - M_l A_l) P_l,
// formed in FP64 from the levels' stencil layouts and transfer bands:
// k_tail_small (one block) the coarse inverse and V_{t+1}, G_t into scratch;
// k_tail_rows one wavefront per row of V_t (FP32, padded rows).
// ---------------------------------------------------------------------------

struct TailLevel {
  int n, h, w;
  const double* op;    // (49, n) FP64 stencil
  const double* dinv;  // D^-1
  const int32_t* f_lo[2];
  const double* rw[2];  // coarse-column bands of P_l (l -> l+1), d = 0 along rows
};

struct TailArgs {
  int t, lc;    // tail level, coarsest level
  int gj_in_g;  // the coarse inverse is formed in k_tail_g (n_lc <= kGjWave)
  double omega;
  TailLevel lv[16];
  double* work;        // scratch (see tail_work_doubles)
  float* coarse_inv;   // FP32 inverse of the coarsest level
};

// A_l[i][j] from the stencil (0 off the 7 x 7 neighbourhood), i, j < n. The
// read is unconditional, from a clamped offset: a read under a condition is
// waited for before the next one is issued.
__device__ __forceinline__ double a_entry(const TailLevel& L, int i, int j) {
  const int ir = i / L.w, ic = i - ir * L.w, jr = j / L.w, jc = j - jr * L.w;
  const int dr = jr - ir, dc = jc - ic;
  const bool in = dr >= -kGR && dr <= kGR && dc >= -kGR && dc <= kGR;
  const int o = (min(max(dr, -kGR), kGR) + kGR) * kGS + min(max(dc, -kGR), kGR) + kGR;
  const double v = L.op[(size_t)o * L.n + i];
  return in ? v : 0.0;
}

// Scratch of the dense tail (doubles): V ping-pong (vmax each), T (vmax), W
// (n_{t+1} x n_t), and G_l of every level l = t .. lc-1 (n_l x n_{l+1}) at
// goff[l].
struct TailWork {
  int64_t vmax, w_off, goff[16];
};

// G_l[i][c] = P[i][c] - m_i sum_o A[i][i+o] P[i+o][c] for every level, one
// thread per entry (blockIdx.y = level - t). The bands of coarse point c are
// read once; P[f][c] = w1[fr - f_lo1] w0[fc - f_lo0] inside them.
__device__ __forceinline__ double band_w(const double (&w)[kKFC], int k) {
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < kKFC; ++q) v = k == q ? w[q] : v;
  return v;
}

__device__ __forceinline__ void tail_g_entry(const TailArgs& a, const TailWork& tw, int l, int e) {
  const TailLevel& L = a.lv[l];
  const TailLevel& C = a.lv[l + 1];
  const int i = e / C.n, c = e - i * C.n;
  const int ir = i / L.w, ic = i - ir * L.w, cr = c / C.w, cc = c - cr * C.w;
  const int flo1 = L.f_lo[1][cr], flo0 = L.f_lo[0][cc];
  double w1[kKFC], w0[kKFC], av[kGN];
#pragma unroll
  for (int q = 0; q < kKFC; ++q) {
    w1[q] = L.rw[1][cr * kKFC + q];
    w0[q] = L.rw[0][cc * kKFC + q];
  }
#pragma unroll
  for (int o = 0; o < kGN; ++o) av[o] = L.op[(size_t)o * L.n + i];
  const double mi = a.omega * L.dinv[i];
  __builtin_amdgcn_sched_barrier(0);
  double s = 0.0;
#pragma unroll
  for (int o = 0; o < kGN; ++o) {
    const int rr = ir + o / kGS - kGR, cc2 = ic + o % kGS - kGR;
    s += av[o] * band_w(w1, rr - flo1) * band_w(w0, cc2 - flo0);
  }
  a.work[tw.goff[l] + e] = band_w(w1, ir - flo1) * band_w(w0, ic - flo0) - mi * s;
}

// Gauss-Jordan on [A | I] for n <= kGjWave in one wavefront: lane j holds
// column j in registers; in step k every lane reads lane k's column (the
// multipliers M[i][k]) as scalars (readlane), scales its entry k by the
// pivot's reciprocal and subtracts M[i][k] times it from every other entry
// i -- the operations of the elimination in LDS below, in the same order,
// with no LDS traffic or barrier per step. Rows i >= n are zero and stay
// zero.
// Writes the inverse (FP64 into inv64, FP32 into inv32), row-major n x n.
constexpr int kGjWave = kWave / 2;

__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int NR>
__device__ __forceinline__ void gj_wave_n(const TailLevel& C, int n, double* __restrict__ inv64,
                                          float* __restrict__ inv32) {
  const int j = threadIdx.x % kWave;
  double col[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const double av = a_entry(C, min(i, n - 1), min(j, n - 1));
    col[i] = (i < n && j < 2 * n) ? (j < n ? av : (j == n + i ? 1.0 : 0.0)) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (k < n) {
      double m[NR];
#pragma unroll
      for (int i = 0; i < NR; ++i) m[i] = lane_value(col[i], k);
      // the pivot's reciprocal (two Newton steps on the hardware estimate)
      double r = __builtin_amdgcn_rcp(m[k]);
      r = fma(fma(-m[k], r, 1.0), r, r);
      r = fma(fma(-m[k], r, 1.0), r, r);
      col[k] *= r;
#pragma unroll
      for (int i = 0; i < NR; ++i)
        if (i != k) col[i] -= m[i] * col[k];
    }
  }
  if (j >= n && j < 2 * n) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      if (i < n) {
        inv64[i * n + j - n] = col[i];
        inv32[i * n + j - n] = (float)col[i];
      }
    }
  }
}

// the register arrays sized to the level (the step loop is unrolled: rows
// beyond n would cost instructions in every step)
__device__ __forceinline__ void gj_wave(const TailLevel& C, int n, double* __restrict__ inv64,
                                        float* __restrict__ inv32) {
  if (n <= 9)
    gj_wave_n<9>(C, n, inv64, inv32);
  else if (n <= 16)
    gj_wave_n<16>(C, n, inv64, inv32);
  else if (n <= 25)
    gj_wave_n<25>(C, n, inv64, inv32);
  else
    gj_wave_n<kGjWave>(C, n, inv64, inv32);
}

// G_l for the levels l0 + blockIdx.y (below lc); with a.gj_in_g, the slice
// blockIdx.y = lc - l0 is the coarse inverse (V_lc at the front of the work
// and the FP32 copy), formed by one wavefront beside the products
__global__ void __launch_bounds__(256) k_tail_g(TailArgs a, TailWork tw, int l0) {
  __shared__ double s_st[kGN * kGjWave];
  const int l = l0 + blockIdx.y;
  if (l == a.lc && a.gj_in_g) {
    if (blockIdx.x != 0) return;
    const TailLevel& C = a.lv[a.lc];
    // every load of the staging in flight together (n_lc <= kGjWave)
    constexpr int NS = (kGN * kGjWave + 255) / 256;
    const int ns = kGN * C.n;
    double v[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) v[k] = C.op[min((int)threadIdx.x + k * 256, ns - 1)];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int e = threadIdx.x + k * 256;
      if (e < ns) s_st[e] = v[k];
    }
    __syncthreads();
    TailLevel Cs = C;
    Cs.op = s_st;
    if (threadIdx.x < kWave) gj_wave(Cs, C.n, a.work, a.coarse_inv);
    return;
  }
  if (l >= a.lc) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.lv[l].n * a.lv[l + 1].n) return;
  tail_g_entry(a, tw, l, e);
}

// one block: the coarse inverse (Gauss-Jordan, SPD without pivoting: in one
// wavefront up to 32 rows, in LDS above) and V_l = G_l V_{l+1} G_l^T + 2M -
// M A M for l = lc-1 .. t+1 (on the matrix cores when the operands fit the
// LDS); V_{t+1} ends at the front of the work
constexpr int kTailLds = 4800;  // doubles of LDS: the elimination, then V_{l+1}, G_l, T

// C(i, j) = sum_k A(i, k) B(k, j) for i < m, j < n, k < kk inside one
// workgroup on the matrix cores: each wave takes 16 x 16 tiles in turn; in
// step k0 lane l holds A(i0 + l % 16, k0 + l / 16) and B(k0 + l / 16,
// j0 + l % 16), and its results are the rows i0 + l / 16 + 4 r (r < 4) of
// column j0 + l % 16 (v_mfma_f64_16x16x4f64). Operands outside the ranges
// are zero; every result in range is handed to out(i, j, value).
using d4 = double __attribute__((ext_vector_type(4)));

template <class FA, class FB, class FO>
__device__ __forceinline__ void tile_dgemm(int m, int n, int kk, FA A, FB B, FO out) {
  const int nw = blockDim.x / kWave, l = threadIdx.x % kWave;
  const int tm = (m + 15) / 16, tn = (n + 15) / 16;
  for (int t = threadIdx.x / kWave; t < tm * tn; t += nw) {
    const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < kk; k0 += 4) {
      const int k = k0 + l / 16, i = i0 + l % 16, j = j0 + l % 16;
      const double av = i < m && k < kk ? A(i, k) : 0.0;
      const double bv = j < n && k < kk ? B(k, j) : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + l / 16 + 4 * r, j = j0 + l % 16;
      if (i < m && j < n) out(i, j, acc[r]);
    }
  }
}

// tile_dgemm with every k-step's operands of a tile read (from clamped
// indices) before its chain of matrix-core products: the reads overlap
// instead of each product waiting for its own. kk <= 4 KS.
template <int KS, class FA, class FB, class FO>
__device__ __forceinline__ void tile_dgemm_pre(int m, int n, int kk, FA A, FB B, FO out) {
  const int nw = blockDim.x / kWave, l = threadIdx.x % kWave;
  const int tm = (m + 15) / 16, tn = (n + 15) / 16;
  for (int t = threadIdx.x / kWave; t < tm * tn; t += nw) {
    const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
    const int i = i0 + l % 16, j = j0 + l % 16;
    const int ic = min(i, m - 1), jc = min(j, n - 1);
    double av[KS], bv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + l / 16, kc = min(k, kk - 1);
      av[s] = A(ic, kc);
      bv[s] = B(kc, jc);
      av[s] = i < m && k < kk ? av[s] : 0.0;
      bv[s] = j < n && k < kk ? bv[s] : 0.0;
    }
    d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int io = i0 + l / 16 + 4 * r;
      if (io < m && j < n) out(io, j, acc[r]);
    }
  }
}

// tiles per wave whose operator entries are read at once in the V_l
// product (16 waves: two cover V_l up to 80 x 80, more take further rounds)
constexpr int kTailTilesPerWave = 2;

__device__ __forceinline__ void tail_small_body(const TailArgs& a, const TailWork& tw) {
  __shared__ double M[kTailLds];
  __shared__ double fac[kGjLds];
  const TailLevel& C = a.lv[a.lc];
  const int n = C.n, n2 = 2 * n;
  double* Va = a.work;
  double* Vb = a.work + tw.vmax;
  double* T = a.work + 2 * tw.vmax;
  if (a.gj_in_g) {
    // formed by k_tail_g
  } else {
    for (int e = threadIdx.x; e < n * n2; e += blockDim.x) {
      const int i = e / n2, j = e - i * n2;
      M[e] = j < n ? a_entry(C, i, j) : (j == n + i ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
      const double piv = M[k * n2 + k];
      __syncthreads();
      for (int j = threadIdx.x; j < n2; j += blockDim.x) M[k * n2 + j] /= piv;
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += blockDim.x) fac[i] = M[i * n2 + k];
      __syncthreads();
      for (int e = threadIdx.x; e < n * n2; e += blockDim.x) {
        const int i = e / n2, j = e - i * n2;
        if (i != k) M[e] -= fac[i] * M[k * n2 + j];
      }
      __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int i = e / n, j = e - i * n;
      Va[e] = M[i * n2 + n + j];
      a.coarse_inv[e] = (float)M[i * n2 + n + j];
    }
    __syncthreads();
  }
  double* Vc = Va;
  double* Vn = Vb;
  for (int l = a.lc - 1; l > a.t; --l) {
    const TailLevel& L = a.lv[l];
    const int nl = L.n, nc = a.lv[l + 1].n;
    const double* g = a.work + tw.goff[l];
    if (nc * nc + 2 * nl * nc <= kTailLds) {
      // V_{l+1}, G_l and T in LDS
      double* sv = M;
      double* sg = M + nc * nc;
      double* st2 = sg + nl * nc;
      for (int e = threadIdx.x; e < nc * nc; e += blockDim.x) sv[e] = Vc[e];
      for (int e = threadIdx.x; e < nl * nc; e += blockDim.x) sg[e] = g[e];
      __syncthreads();
      // T = G V, then V_l = T G^T + 2M - MAM, both on the matrix cores
      // (16 x 16 tiles, v_mfma_f64_16x16x4f64, operands from LDS)
      tile_dgemm(nl, nc, nc, [&](int i, int m) { return sg[i * nc + m]; },
                 [&](int m, int k) { return sv[m * nc + k]; },
                 [&](int i, int k, double v) { st2[i * nc + k] = v; });
      __syncthreads();
      // the global reads of a lane's entries (operator entry, diagonal
      // scalings) are issued for all its tiles before the products
      const int tn = (nl + 15) / 16, nw = blockDim.x / kWave, wv = threadIdx.x / kWave;
      const int ln = threadIdx.x % kWave;
      for (int u0 = 0; wv + u0 * nw < tn * tn; u0 += kTailTilesPerWave) {
        double lin[kTailTilesPerWave][4];
#pragma unroll
        for (int u = 0; u < kTailTilesPerWave; ++u) {
          const int t = min(wv + (u0 + u) * nw, tn * tn - 1);
          const int i0 = (t / tn) * 16, j = min((t % tn) * 16 + ln % 16, nl - 1);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = min(i0 + ln / 16 + 4 * r, nl - 1);
            const double mi = a.omega * L.dinv[i], mj = a.omega * L.dinv[j];
            lin[u][r] = (i == j ? 2.0 * mi : 0.0) - mi * a_entry(L, i, j) * mj;
          }
        }
#pragma unroll
        for (int u = 0; u < kTailTilesPerWave; ++u) {
          const int t = wv + (u0 + u) * nw;
          if (t >= tn * tn) break;
          const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
          d4 acc = {0.0, 0.0, 0.0, 0.0};
          for (int k0 = 0; k0 < nc; k0 += 4) {
            const int k = k0 + ln / 16, i = i0 + ln % 16, j = j0 + ln % 16;
            const double av = i < nl && k < nc ? st2[i * nc + k] : 0.0;
            const double bv = j < nl && k < nc ? sg[j * nc + k] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = i0 + ln / 16 + 4 * r, j = j0 + ln % 16;
            if (i < nl && j < nl) Vn[i * nl + j] = acc[r] + lin[u][r];
          }
        }
      }
      __syncthreads();
      double* tmp = Vc;
      Vc = Vn;
      Vn = tmp;
      continue;
    }
    for (int e = threadIdx.x; e < nl * nc; e += blockDim.x) {
      const int i = e / nc, k = e - i * nc;
      double s = 0.0;
#pragma unroll 5
      for (int m = 0; m < nc; ++m) s += g[i * nc + m] * Vc[m * nc + k];
      T[e] = s;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nl * nl; e += blockDim.x) {
      const int i = e / nl, j = e - i * nl;
      double s = 0.0;
#pragma unroll 5
      for (int k = 0; k < nc; ++k) s += T[i * nc + k] * g[j * nc + k];
      const double mi = a.omega * L.dinv[i], mj = a.omega * L.dinv[j];
      s += (i == j ? 2.0 * mi : 0.0) - mi * a_entry(L, i, j) * mj;
      Vn[e] = s;
    }
    __syncthreads();
    double* tmp = Vc;
    Vc = Vn;
    Vn = tmp;
  }
  if (Vc != Va) {
    const int nc = a.lv[a.t + 1].n;
    for (int e = threadIdx.x; e < nc * nc; e += blockDim.x) Va[e] = Vc[e];
  }
}

__global__ void __launch_bounds__(1024) k_tail_small(TailArgs a, TailWork tw) { tail_small_body(a, tw); }

// W = V_{t+1} G_t^T (n_{t+1} x n_t), one thread per entry
__global__ void __launch_bounds__(256) k_tail_w(TailArgs a, TailWork tw) {
  const int nt = a.lv[a.t].n, nc = a.lv[a.t + 1].n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc * nt) return;
  const int k = e / nt, j = e - k * nt;
  const double* V = a.work;
  const double* g = a.work + tw.goff[a.t];
  double s = 0.0;
#pragma unroll 9
  for (int m = 0; m < nc; ++m) s += V[k * nc + m] * g[j * nc + m];
  a.work[tw.w_off + e] = s;
}

// V_t[i][j] = sum_k G_t[i][k] W[k][j] + 2M - M A M into the tail (FP32, ld
// floats per row, zero beyond n_t), one thread per entry
__global__ void __launch_bounds__(256) k_tail_v(TailArgs a, TailWork tw, float* __restrict__ tail,
                                                int ld) {
  const TailLevel& L = a.lv[a.t];
  const int nt = L.n, nc = a.lv[a.t + 1].n;
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ld) return;
  float out = 0.f;
  if (j < nt) {
    const double* g = a.work + tw.goff[a.t] + (size_t)i * nc;
    const double* W = a.work + tw.w_off;
    double s = 0.0;
#pragma unroll 9
    for (int k = 0; k < nc; ++k) s += g[k] * W[(size_t)k * nt + j];
    const double mi = a.omega * L.dinv[i];
    s += (i == j ? 2.0 * mi : 0.0) - mi * a_entry(L, i, j) * a.omega * L.dinv[j];
    out = (float)s;
  }
  tail[(size_t)i * ld + j] = out;
}

// V_t in one launch when the coarse inverse sits two levels below t (n_{t+1}
// small): every block forms V_{t+1} = G_{t+1} V_{t+2} G_{t+1}^T + 2M - MAM
// in LDS, then the columns of W = V_{t+1} G_t^T its tile needs, then its
// kTFT x kTFT tile of V_t = G_t W + 2M - MAM, all on the FP64 matrix cores
// (the products of k_tail_small, k_tail_w and k_tail_v, repeated per block
// where they are small). Blocks past n_t write the zero columns up to ld.
// LDS: V_{t+1}, then one region holding V_{t+2}, G_{t+1} and G_{t+1}V_{t+2}
// while V_{t+1} is formed and the G_t rows of the tile, W and the tile's
// 2M - MAM afterwards (those are read into registers first).
constexpr int kTFT = 32;           // tile of V_t
constexpr int kTFTh = 512;         // threads per block
constexpr int kTFLds = 16384;      // doubles of LDS
constexpr int kTFMaxN1 = 96;       // n_{t+1}
constexpr int kTFRowVals = (kTFT * kTFMaxN1 + kTFTh - 1) / kTFTh;  // G_t values per thread
constexpr int kTFLin = kTFT * kTFT / kTFTh;                         // tile entries per thread

__global__ void __launch_bounds__(kTFTh)
    k_tail_fused(TailArgs a, TailWork tw, float* __restrict__ tail, int ld) {
  __shared__ double lds[kTFLds];
  const int t = a.t;
  const TailLevel& Lt = a.lv[t];
  const TailLevel& L1 = a.lv[t + 1];
  const int nt = Lt.n, n1 = L1.n, n2 = a.lv[t + 2].n;
  const int i0 = blockIdx.y * kTFT, j0 = blockIdx.x * kTFT;
  if (j0 >= nt) {
    for (int e = threadIdx.x; e < kTFT * kTFT; e += kTFTh) {
      const int i = i0 + e / kTFT, j = j0 + e % kTFT;
      if (i < nt && j < ld) tail[(size_t)i * ld + j] = 0.f;
    }
    return;
  }
  double* sV1 = lds;             // n1 x n1
  double* sV2 = lds + n1 * n1;   // n2 x n2
  double* sG1 = sV2 + n2 * n2;   // n1 x n2
  double* sT1 = sG1 + n1 * n2;   // n1 x n2
  double* sGi = lds + n1 * n1;   // kTFT x n1 (after V_{t+1}): G_t rows i0 ..
  double* sGj = sGi + kTFT * n1; // kTFT x n1: G_t rows j0 ..
  double* sW = sGj + kTFT * n1;  // n1 x kTFT
  double* sLin = sW + n1 * kTFT; // kTFT x kTFT
  const double* V2 = a.work;     // the coarse inverse, at the front
  const double* G1 = a.work + tw.goff[t + 1];
  const double* Gt = a.work + tw.goff[t];
  // the tile's G_t rows and 2M - MAM into registers, beside the loads of
  // the first products' operands into LDS
  double gi[kTFRowVals], gj[kTFRowVals], lin[kTFLin];
#pragma unroll
  for (int k = 0; k < kTFRowVals; ++k) {
    const int e = threadIdx.x + k * kTFTh;
    const int r = e / n1, m = e - r * n1;
    const bool ok = e < kTFT * n1;
    gi[k] = Gt[(size_t)(ok && i0 + r < nt ? i0 + r : 0) * n1 + (ok ? m : 0)];
    gj[k] = Gt[(size_t)(ok && j0 + r < nt ? j0 + r : 0) * n1 + (ok ? m : 0)];
    gi[k] = ok && i0 + r < nt ? gi[k] : 0.0;
    gj[k] = ok && j0 + r < nt ? gj[k] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < kTFLin; ++k) {
    const int e = threadIdx.x + k * kTFTh;
    const int i = min(i0 + e / kTFT, nt - 1), j = min(j0 + e % kTFT, nt - 1);
    const double mi = a.omega * Lt.dinv[i], mj = a.omega * Lt.dinv[j];
    lin[k] = (i == j ? 2.0 * mi : 0.0) - mi * a_entry(Lt, i, j) * mj;
  }
  // V_{t+2}, G_{t+1} and level t+1's stencil and scalings: one group of
  // loads, then LDS (the stencil and scalings behind G_{t+1} V_{t+2})
  double* sOp = sT1 + n1 * n2;   // kGN x n1
  double* sDi = sOp + kGN * n1;  // n1
  {
    constexpr int N2 = (kGjWave * kGjWave + kTFTh - 1) / kTFTh;
    constexpr int N12 = (kTFMaxN1 * kGjWave + kTFTh - 1) / kTFTh;
    constexpr int NO = (kGN * kTFMaxN1 + kTFTh - 1) / kTFTh;
    double v2[N2], g1[N12], op[NO], di;
#pragma unroll
    for (int k = 0; k < N2; ++k) v2[k] = V2[min((int)threadIdx.x + k * kTFTh, n2 * n2 - 1)];
#pragma unroll
    for (int k = 0; k < N12; ++k) g1[k] = G1[min((int)threadIdx.x + k * kTFTh, n1 * n2 - 1)];
#pragma unroll
    for (int k = 0; k < NO; ++k) op[k] = L1.op[min((int)threadIdx.x + k * kTFTh, kGN * n1 - 1)];
    di = L1.dinv[min((int)threadIdx.x, n1 - 1)];
#pragma unroll
    for (int k = 0; k < N2; ++k) {
      const int e = threadIdx.x + k * kTFTh;
      if (e < n2 * n2) sV2[e] = v2[k];
    }
#pragma unroll
    for (int k = 0; k < N12; ++k) {
      const int e = threadIdx.x + k * kTFTh;
      if (e < n1 * n2) sG1[e] = g1[k];
    }
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      const int e = threadIdx.x + k * kTFTh;
      if (e < kGN * n1) sOp[e] = op[k];
    }
    if ((int)threadIdx.x < n1) sDi[threadIdx.x] = di;
  }
  __syncthreads();
  {
    TailLevel S1 = L1;
    S1.op = sOp;
    for (int e = threadIdx.x; e < n1 * n1; e += kTFTh) {
      const int i = e / n1, j = e - i * n1;
      const double mi = a.omega * sDi[i], mj = a.omega * sDi[j];
      sV1[e] = (i == j ? 2.0 * mi : 0.0) - mi * a_entry(S1, i, j) * mj;
    }
  }
  tile_dgemm_pre<kGjWave / 4>(n1, n2, n2, [&](int i, int k) { return sG1[i * n2 + k]; },
                              [&](int k, int j) { return sV2[k * n2 + j]; },
                              [&](int i, int j, double v) { sT1[i * n2 + j] = v; });
  __syncthreads();
  tile_dgemm_pre<kGjWave / 4>(n1, n1, n2, [&](int i, int k) { return sT1[i * n2 + k]; },
             [&](int k, int j) { return sG1[j * n2 + k]; },
             [&](int i, int j, double v) { sV1[i * n1 + j] += v; });
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kTFRowVals; ++k) {
    const int e = threadIdx.x + k * kTFTh;
    if (e < kTFT * n1) {
      sGi[e] = gi[k];
      sGj[e] = gj[k];
    }
  }
#pragma unroll
  for (int k = 0; k < kTFLin; ++k) sLin[threadIdx.x + k * kTFTh] = lin[k];
  __syncthreads();
  // W[m][c] = sum_k V_{t+1}[m][k] G_t[j0 + c][k]
  tile_dgemm_pre<kTFMaxN1 / 4>(n1, kTFT, n1, [&](int i, int k) { return sV1[i * n1 + k]; },
             [&](int k, int j) { return sGj[j * n1 + k]; },
             [&](int i, int j, double v) { sW[i * kTFT + j] = v; });
  __syncthreads();
  tile_dgemm_pre<kTFMaxN1 / 4>(kTFT, kTFT, n1, [&](int i, int k) { return sGi[i * n1 + k]; },
             [&](int k, int j) { return sW[k * kTFT + j]; },
             [&](int r, int c, double v) {
               const int i = i0 + r, j = j0 + c;
               if (i < nt) tail[(size_t)i * ld + j] = j < nt ? (float)(v + sLin[r * kTFT + c]) : 0.f;
             });
}

// whether k_tail_fused forms V_t (see there)
static bool tail_fused_fits(const iga_mg_setup_t& s, int t, int lc, bool gj_in_g) {
  if (!gj_in_g || lc != t + 2) return false;
  const int64_t n1 = s.level[t + 1].n, n2 = s.level[t + 2].n;
  return n1 <= kTFMaxN1 &&
         n1 * n1 + std::max(n2 * n2 + 2 * n1 * n2 + kGN * n1 + n1, 3 * kTFT * n1 + kTFT * kTFT) <= kTFLds;
}

static inline GLevel glevel(const iga_mg_t& mg, int l) {
  const iga_mg_level_t& L = mg.level[l];
  return GLevel{L.h, L.w, mg.grid->op[l], L.dsq, L.omega};
}

static inline GBand gband(const iga_mg_t& mg, int l, int d) {
  const iga_mg_band_t& B = mg.grid->band[l][d];
  return GBand{B.c_lo, B.pw, B.f_lo, B.rw};
}

// whether the grid V-cycle applies (see iga_mg_grid_t)
static bool grid_cycle_ok(const iga_mg_t& mg) {
  const iga_mg_grid_t* G = mg.grid;
  if (!G || !mg.tail || G->radius != kGR) return false;
  const int t = mg.tail_level;
  if (t < 2 || t >= mg.n_levels - 1 || G->n_levels < t) return false;
  if (mg.level[t - 1].n > kTailFine || mg.level[t].n > kTailCoarse) return false;
  if (mg.level[0].stencil && mg.ks != kGS) return false;
  for (int l = 1; l < t; ++l)
    if (mg.level[l].stencil) return false;
  for (int l = 0; l < t; ++l) {
    if (!G->op[l]) return false;
    for (int d = 0; d < 2; ++d) {
      const iga_mg_band_t& B = G->band[l][d];
      if (!B.c_lo || !B.pw || !B.f_lo || !B.rw) return false;
    }
  }
  return true;
}

static inline dim3 tiles_of(const iga_mg_level_t& L) {
  return dim3((L.w + kGT - 1) / kGT, (L.h + kGT - 1) / kGT);
}

static inline dim3 row_tiles(const iga_mg_level_t& L) {
  return dim3((L.w + kL0W - 1) / kL0W, (L.h + kL0H - 1) / kL0H);
}

// Per-level buffers of the grid cycle: x (pre-smoothed), b, r, x (post)
static int vcycle_grid(const iga_mg_t& mg, const float* b0, float* z, const Gate& gate,
                       const DotOut& dot, int* dot_blocks, hipStream_t st,
                       const FcgFuse* fz = nullptr) {
  const int t = mg.tail_level;
  float* X[16];
  float* B[16];
  float* Rr[16];
  float* XO[16];
  for (int l = 0; l <= t; ++l) {
    float* base = mg.work + level_offset(mg, l);
    const int n = mg.level[l].n;
    X[l] = base;
    B[l] = base + n;
    Rr[l] = base + 2 * n;
    XO[l] = base + 3 * n;
  }
  const int* skip = gate.flag;
  const iga_mg_level_t& L0 = mg.level[0];
  // the FP16 copy of the level-0 smoother, when given, in place of the stencil
  const _Float16* W16 = L0.stencil ? static_cast<const _Float16*>(L0.stencil16) : nullptr;
  const GLevel g0 = glevel(mg, 0);
  const dim3 t0 = row_tiles(L0);
  if (fz) {
    // flexible CG: the waiting update first, its convergence test (the
    // gate) in the residual kernel
    if (W16)
      hipLaunchKernelGGL((k_g0_smooth_upd<true, _Float16>), t0, dim3(kGTh), 0, st, g0, W16, *fz,
                         X[0]);
    else if (L0.stencil)
      hipLaunchKernelGGL(k_g0_smooth_upd<true>, t0, dim3(kGTh), 0, st, g0, L0.stencil, *fz, X[0]);
    else
      hipLaunchKernelGGL((k_g0_smooth_upd<false, float>), t0, dim3(kGTh), 0, st, g0, nullptr, *fz,
                         X[0]);
    hipLaunchKernelGGL(k_g0_resid, t0, dim3(kGTh), 0, st, g0, X[0], b0, Rr[0], gate);
  } else {
    if (W16)
      hipLaunchKernelGGL((k_g0_smooth<true, _Float16>), t0, dim3(kGTh), 0, st, g0, W16, b0, X[0],
                         gate);
    else if (L0.stencil)
      hipLaunchKernelGGL(k_g0_smooth<true>, t0, dim3(kGTh), 0, st, g0, L0.stencil, b0, X[0], gate);
    else
      hipLaunchKernelGGL((k_g0_smooth<false, float>), t0, dim3(kGTh), 0, st, g0, nullptr, b0, X[0],
                         gate);
    hipLaunchKernelGGL(k_g0_resid, t0, dim3(kGTh), 0, st, g0, X[0], b0, Rr[0],
                       Gate{skip, nullptr, 0, nullptr, 0});
  }
  IGA_CHECK_LAUNCH();
  for (int l = 0; l + 1 < t; ++l) {
    const iga_mg_level_t& F = mg.level[l];
    hipLaunchKernelGGL(k_g_restrict, tiles_of(mg.level[l + 1]), dim3(kGTh), 0, st, F.h, F.w,
                       glevel(mg, l + 1), gband(mg, l, 1), gband(mg, l, 0), Rr[l], B[l + 1],
                       X[l + 1], Rr[l + 1], skip);
    IGA_CHECK_LAUNCH();
  }
  {
    const iga_mg_level_t& F = mg.level[t - 1];
    const iga_mg_level_t& C = mg.level[t];
    const int ld = dense_ld(C.n);
    auto go = [&](auto V) {
      constexpr int NV = decltype(V)::value;
      hipLaunchKernelGGL(k_g_tail<NV>, dim3((C.n + 3) / 4), dim3(kGTh), 0, st, F.h, F.w, C.h, C.w,
                         gband(mg, t - 1, 1), gband(mg, t - 1, 0), Rr[t - 1], mg.tail, ld, XO[t],
                         skip);
      IGA_CHECK_LAUNCH();
      return (int)IGA_OK;
    };
    int rc;
    switch (ld / 256) {
      case 1: rc = go(std::integral_constant<int, 1>{}); break;
      case 2: rc = go(std::integral_constant<int, 2>{}); break;
      case 3: rc = go(std::integral_constant<int, 3>{}); break;
      case 4: rc = go(std::integral_constant<int, 4>{}); break;
      case 5: rc = go(std::integral_constant<int, 5>{}); break;
      case 6: rc = go(std::integral_constant<int, 6>{}); break;
      case 7: rc = go(std::integral_constant<int, 7>{}); break;
      case 8: rc = go(std::integral_constant<int, 8>{}); break;
      default: rc = IGA_ERR_UNSUPPORTED;
    }
    if (rc) return rc;
  }
  // the prolongation into level t - 1 joins the one into level t - 2 when
  // that is a tile level (t >= 3) and the tail fits the staging
  const bool fuse_prolong = t >= 3 && mg.level[t].n <= kPTail;
  for (int l = t - 1; l >= 1; --l) {
    if (fuse_prolong && l == t - 1) continue;
    if (fuse_prolong && l == t - 2) {
      const ProlongP pp{glevel(mg, l),       glevel(mg, t - 1),  mg.level[t].h,
                        mg.level[t].w,       gband(mg, l, 1),    gband(mg, l, 0),
                        gband(mg, t - 1, 1), gband(mg, t - 1, 0)};
      hipLaunchKernelGGL(k_g_prolong_p, tiles_of(mg.level[l]), dim3(kGTh), 0, st, pp, XO[t],
                         X[t - 1], B[t - 1], X[l], B[l], XO[l], skip);
    } else {
      const iga_mg_level_t& C = mg.level[l + 1];
      hipLaunchKernelGGL(k_g_prolong, tiles_of(mg.level[l]), dim3(kGTh), 0, st, glevel(mg, l),
                         C.h, C.w, gband(mg, l, 1), gband(mg, l, 0), XO[l + 1], X[l], B[l], XO[l],
                         skip);
    }
    IGA_CHECK_LAUNCH();
  }
  const iga_mg_level_t& L1 = mg.level[1];
  hipLaunchKernelGGL(k_g0_prolong, t0, dim3(kGTh), 0, st, g0, L1.h, L1.w, gband(mg, 0, 1),
                     gband(mg, 0, 0), XO[1], X[0], b0, XO[0], B[0], skip);
  const dim3 tg = t0;
  const bool fuse = dot.parts && (int)(tg.x * tg.y) <= kRedBlocks;
  const DotOut d = fuse ? dot : DotOut{nullptr, nullptr};
  if (W16)
    hipLaunchKernelGGL((k_g0_smooth_t<true, _Float16>), tg, dim3(kGTh), 0, st, g0, W16, XO[0],
                       B[0], z, skip, d);
  else if (L0.stencil)
    hipLaunchKernelGGL(k_g0_smooth_t<true>, tg, dim3(kGTh), 0, st, g0, L0.stencil, XO[0], B[0], z,
                       skip, d);
  else
    hipLaunchKernelGGL((k_g0_smooth_t<false, float>), tg, dim3(kGTh), 0, st, g0, nullptr, XO[0],
                       B[0], z, skip, d);
  IGA_CHECK_LAUNCH();
  if (fuse) {
    *dot_blocks = (int)(tg.x * tg.y);
  } else if (dot.parts) {
    const int nb = std::min(blocks_for(L0.n), kRedBlocks);
    hipLaunchKernelGGL(k_dot_fd, dim3(nb), dim3(256), 0, st, L0.n, z, skip, dot);
    IGA_CHECK_LAUNCH();
    *dot_blocks = nb;
  }
  return IGA_OK;
}

// x = V(b) starting at level l0 (b, x of level l0's size). Levels with a
// stencil smoother run pre-smooth / residual / restrict and prolong /
// residual / post-smooth as separate kernels; damped-Jacobi levels fuse the
// smoother into the residual kernels; below the tail level the cycle is one
// dense product. `gate` gates the first kernel, its flag every later one.
// With `dot.parts` set, the partials of x . dot.y come with the last kernel
// (or one more kernel) and their count is returned in *dot_blocks.
static int vcycle(const iga_mg_t& mg, int l0, const float* b0, float* x0, const Gate& gate,
                  const DotOut& dot, int* dot_blocks, hipStream_t st,
                  const FcgFuse* fz = nullptr) {
  if (l0 == 0 && grid_cycle_ok(mg)) return vcycle_grid(mg, b0, x0, gate, dot, dot_blocks, st, fz);
  if (fz) return IGA_ERR_UNSUPPORTED;
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
  const int* skip = gate.flag;
  const Gate later{gate.flag, nullptr, 0, nullptr, 0};
  const DotOut none{nullptr, nullptr};
  bool first = true;
  auto next_gate = [&]() -> Gate {
    if (!first) return later;
    first = false;
    return gate;
  };
  bool dot_done = dot.parts == nullptr;
  int rc;
  const int t = (mg.tail && mg.tail_level >= l0 && mg.tail_level < Lc) ? mg.tail_level : Lc;
  for (int l = l0; l < t; ++l) {
    const iga_mg_level_t& L = mg.level[l];
    if (L.stencil) {
      if ((rc = smooth(mg, l, B[l], X[l], false, 0, next_gate(), none, st))) return rc;
      if ((rc = csr_f(L.n, L.nnz, L.indptr, L.cols, L.vals, X[l], B[l], Rr[l], 1, 1, 0, 0, skip, st)))
        return rc;
    } else {
      const Gate g = next_gate();
      rc = launch_rows(L.n, L.nnz, [&](auto K, int rb) {
        constexpr int LPR = decltype(K)::value;
        hipLaunchKernelGGL(k_jacobi_pre<LPR>, dim3(rb), dim3(256), 0, st, L.n, L.indptr, L.cols,
                           L.vals, L.omega, L.dsq, B[l], X[l], Rr[l], g, (int64_t)0);
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
    const Gate g = next_gate();
    if (t < Lc) {
      if ((rc = dense_tail_gemv(C.n, mg.tail, B[t], X[t], g, 1, st))) return rc;
    } else {
      hipLaunchKernelGGL(k_dense_gemv, dim3((C.n + 3) / 4), dim3(256), 0, st, C.n, mg.coarse_inv,
                         B[t], X[t], g, (int64_t)0);
      IGA_CHECK_LAUNCH();
    }
  }
  for (int l = t - 1; l >= l0; --l) {
    const iga_mg_level_t& L = mg.level[l];
    if ((rc = csr_f(L.n, L.p_nnz, L.p_ptr, L.p_cols, L.p_vals, X[l + 1], nullptr, X[l], 2, 1, 0, 0,
                    skip, st)))
      return rc;
    if (L.stencil) {
      if ((rc = csr_f(L.n, L.nnz, L.indptr, L.cols, L.vals, X[l], B[l], Rr[l], 1, 1, 0, 0, skip, st)))
        return rc;
      const bool fuse = l == l0 && !dot_done && stencil_tiles(L) <= kRedBlocks;
      if ((rc = smooth(mg, l, Rr[l], X[l], true, 1, later, fuse ? dot : none, st))) return rc;
      if (fuse) {
        *dot_blocks = stencil_tiles(L);
        dot_done = true;
      }
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
        const int nb = std::min(blocks_for(L.n), kRedBlocks);
        hipLaunchKernelGGL(k_copy_fs, dim3(nb), dim3(256), 0, st, L.n, X[l], x0, skip,
                           dot_done ? none : dot);
        IGA_CHECK_LAUNCH();
        if (!dot_done) {
          *dot_blocks = nb;
          dot_done = true;
        }
      }
    }
  }
  if (!dot_done) {
    const int n = mg.level[l0].n;
    const int nb = std::min(blocks_for(n), kRedBlocks);
    hipLaunchKernelGGL(k_dot_fd, dim3(nb), dim3(256), 0, st, n, x0, skip, dot);
    IGA_CHECK_LAUNCH();
    *dot_blocks = nb;
  }
  return IGA_OK;
}

static int check_mg(const iga_mg_t* mg) {
  if (!mg || mg->n_levels < 1 || mg->n_levels > 16 || !mg->work || !mg->coarse_inv)
    return IGA_ERR_ARG;
  if (mg->tail) {
    if (mg->tail_level < 0 || mg->tail_level >= mg->n_levels - 1) return IGA_ERR_ARG;
    if (mg->level[mg->tail_level].n > kDenseMaxRows) return IGA_ERR_UNSUPPORTED;
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
// host reads the convergence flag once per replay. One iteration is
//   z = V(r)                       (V-cycle; its last kernel adds z . q_prev)
//   beta = -(z . q_prev) / (p_prev . q_prev),  p = z + beta p_prev,  q = A p
//                                  (one kernel: p formed row by row)
//   alpha = (p . r) / (p . q),  x += alpha p,  r -= alpha q
// and the test ||r|| <= tol ||b|| is made by the first kernel of the next
// V-cycle (the gate).
// ---------------------------------------------------------------------------

struct FBufs {
  FState* st;
  double* P_a;  // [kRedBlocks] z.q_prev
  double* P_b;  // [kRedBlocks] p.r
  double* P_c;  // [kRedBlocks] p.q
  double* P_d;  // [kRedBlocks] r.r
};

// the solver state at the start of a run, written by k_fcg_init when `fold`
// (then the first gate takes ||b||^2 from its r.r partials: no check kernel)
struct FInit {
  int fold;
  double tol2;
  int maxiter;
};

// x = 0, r = b, the first "previous" direction zero, the b.b partials; with
// b_in, b is read from there and kept in b (the run's copy of the caller's
// right-hand side)
__global__ void __launch_bounds__(kVecThreads)
    k_fcg_init(int n, const double* __restrict__ b_in, double* __restrict__ b, double* x,
               double* r, float* r32, double* p1, double* q1, FBufs F, double* x_out, FInit ini) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    FState* st = F.st;
    st->x_out = x_out;
    if (ini.fold) {
      st->bb = -1.0;
      st->rr = 0.0;
      st->tol2 = ini.tol2;
      st->maxiter = ini.maxiter;
      st->iter = 0;
      st->pq_prev = 1.0;
      st->pending = 0;
      st->done = 0;
    }
  }
  const double* src = b_in ? b_in : b;
  double bb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double bi = src[i];
    if (b_in) b[i] = bi;
    x[i] = 0.0;
    r[i] = bi;
    r32[i] = (float)bi;
    p1[i] = 0.0;  // the "previous" direction of the first iteration (beta = 0)
    q1[i] = 0.0;
    bb += bi * bi;
  }
  store_partial(F.P_d, bb);
}

// init: record ||b||^2 and the tolerance; otherwise record ||r||^2 and the
// flag after the last iteration of a graph (the iteration count is kept by
// the gates)
__global__ void k_fcg_check(int nblk, FBufs F, int init, double tol2, int maxiter) {
  FState* st = F.st;
  const double rr = reduce_partials(F.P_d, nblk);
  if (threadIdx.x == 0) {
    if (init) {
      st->bb = rr;
      st->tol2 = tol2;
      st->maxiter = maxiter;
      st->iter = 0;
      st->pq_prev = 1.0;
    }
    st->pending = 0;
    st->rr = rr;
    st->done = !(rr > st->tol2 * st->bb) ? 1 : 0;
  }
}

// beta from the z . q_prev partials; p = z + beta p_prev (z = the FP32
// V-cycle output, widened), q = A p with p formed at every column read; the
// partials of p . q and p . r
template <int LPR>
__global__ void __launch_bounds__(kVecThreads)
    k_fcg_spmv(int n, int nblk_a, const int32_t* __restrict__ indptr,
               const int32_t* __restrict__ cols, const double* __restrict__ vals,
               const float* __restrict__ z32, const double* __restrict__ pp,
               const double* __restrict__ r, double* __restrict__ p, double* __restrict__ q,
               FBufs F) {
  const FState* st = F.st;
  if (st->done) return;
  const double beta = -reduce_partials(F.P_a, nblk_a) / st->pq_prev;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double pq = 0.0, pr = 0.0;
  for (int row = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; row < n; row += groups) {
    const int lo = indptr[row], hi = indptr[row + 1];
    double s = 0.0;
    for (int k = lo + l; k < hi; k += LPR) {
      const int j = cols[k];
      s += vals[k] * ((double)z32[j] + beta * pp[j]);
    }
#pragma unroll
    for (int off = LPR / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, LPR);
    if (l == 0) {
      const double pi = (double)z32[row] + beta * pp[row];
      p[row] = pi;
      q[row] = s;
      pq += pi * s;
      pr += pi * r[row];
    }
  }
  __shared__ double scratch[16];
  pq = block_sum(pq, scratch);
  if (threadIdx.x == 0) F.P_c[blockIdx.x] = pq;
  pr = block_sum(pr, scratch);
  if (threadIdx.x == 0) F.P_b[blockIdx.x] = pr;
}

// independent FP64 residual of the result: partials of ||b - A x||^2 (P_b),
// then their sum into the state
template <int LPR>
__global__ void __launch_bounds__(kVecThreads)
    k_fcg_resid(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
                const double* __restrict__ vals, const double* __restrict__ x,
                const double* __restrict__ b, FBufs F) {
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double ss = 0.0;
  for (int row = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; row < n; row += groups) {
    const double s = row_dot<LPR>(row, indptr, cols, vals, x, l);
    if (l == 0) {
      const double ri = b[row] - s;
      ss += ri * ri;
    }
  }
  store_partial(F.P_b, ss);
}

__global__ void k_fcg_resid_fin(int nblk, FBufs F) {
  const double s = reduce_partials(F.P_b, nblk);
  if (threadIdx.x == 0) F.st->res2 = s;
}

// the end-of-batch test of k_fcg_check (not at init) and the residual sum
// of k_fcg_resid_fin in one kernel
__global__ void k_fcg_check_res(int nblk_d, int nblk_b, FBufs F) {
  FState* st = F.st;
  const double rr = reduce_partials(F.P_d, nblk_d);
  const double s = reduce_partials(F.P_b, nblk_b);
  if (threadIdx.x == 0) {
    st->pending = 0;
    st->rr = rr;
    st->done = !(rr > st->tol2 * st->bb) ? 1 : 0;
    st->res2 = s;
  }
}

__global__ void __launch_bounds__(kVecThreads)
    k_fcg_update(int n, int nblk, const double* __restrict__ p, const double* __restrict__ q,
                 double* x, double* r, float* r32, FBufs F) {
  FState* st = F.st;
  if (st->done) return;
  const double pq = reduce_partials(F.P_c, nblk);
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

// scratch layout of the dense tail (TailWork)
static TailWork tail_layout(const iga_mg_setup_t& s, int t, int64_t* total) {
  TailWork w{};
  w.vmax = 1;
  for (int l = t + 1; l < s.n_levels; ++l)
    w.vmax = std::max<int64_t>(w.vmax, (int64_t)s.level[l].n * s.level[l].n);
  int64_t o = 3 * w.vmax;
  w.w_off = o;
  o += (int64_t)s.level[t + 1].n * s.level[t].n;
  for (int l = t; l + 1 < s.n_levels; ++l) {
    w.goff[l] = o;
    o += (int64_t)s.level[l].n * s.level[l + 1].n;
  }
  if (total) *total = o;
  return w;
}

static int64_t tail_dense_doubles(const iga_mg_setup_t& s, int t) {
  int64_t total = 0;
  tail_layout(s, t, &total);
  return total;
}

extern "C" int64_t iga_mg_tail_dense_doubles(const iga_mg_setup_t* s, int32_t level) {
  if (!s || level < 1 || level >= s->n_levels - 1) return -1;
  return tail_dense_doubles(*s, level);
}

static int tail_dense_check(const iga_mg_setup_t& s) {
  const iga_mg_setup_grid_t& G = *s.grid;
  const int t = G.tail_level, lc = s.n_levels - 1;
  if (t < 1 || t >= lc || !G.tail || !G.tail_work || !s.coarse_inv) return IGA_ERR_ARG;
  if (s.level[lc].n > kGjLds || s.level[t + 1].n > 128 || s.level[t].n > kDenseMaxRows)
    return IGA_ERR_UNSUPPORTED;
  return IGA_OK;
}

// the dense tail of level t from the tensor set-up (after the Galerkin passes)
static int tail_dense(const iga_mg_setup_t& s, hipStream_t st) {
  const iga_mg_setup_grid_t& G = *s.grid;
  const int t = G.tail_level, lc = s.n_levels - 1;
  const int rc = tail_dense_check(s);
  if (rc) return rc;
  TailArgs a{};
  a.t = t;
  a.lc = lc;
  a.omega = G.omega;
  for (int l = t; l <= lc; ++l) {
    TailLevel& L = a.lv[l];
    L.n = s.level[l].n;
    L.h = s.level[l].h;
    L.w = s.level[l].w;
    L.op = G.op64[l];
    L.dinv = s.level[l].dinv;
    for (int d = 0; d < 2; ++d) {
      L.f_lo[d] = l < lc ? G.f_lo[l][d] : nullptr;
      L.rw[d] = l < lc ? G.rw[l][d] : nullptr;
    }
  }
  a.work = G.tail_work;
  a.coarse_inv = s.coarse_inv;
  const TailWork tw = tail_layout(s, t, nullptr);
  int gmax = 1;
  for (int l = t; l < lc; ++l) gmax = std::max(gmax, s.level[l].n * s.level[l + 1].n);
  const int nt = s.level[t].n, nc = s.level[t + 1].n, ld = dense_ld(nt);
  // a small coarse inverse is formed beside the G products
  a.gj_in_g = s.level[lc].n <= kGjWave;
  hipLaunchKernelGGL(k_tail_g, dim3((gmax + 255) / 256, lc - t + a.gj_in_g), dim3(256), 0, st, a, tw, t);
  if (tail_fused_fits(s, t, lc, a.gj_in_g)) {
    hipLaunchKernelGGL(k_tail_fused, dim3(ld / kTFT, (nt + kTFT - 1) / kTFT), dim3(kTFTh), 0, st, a,
                       tw, G.tail, ld);
  } else {
    hipLaunchKernelGGL(k_tail_small, dim3(1), dim3(1024), 0, st, a, tw);
    hipLaunchKernelGGL(k_tail_w, dim3((nc * nt + 255) / 256), dim3(256), 0, st, a, tw);
    hipLaunchKernelGGL(k_tail_v, dim3((ld + 255) / 256, nt), dim3(256), 0, st, a, tw, G.tail, ld);
  }
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
  const iga_mg_setup_grid_t* G = s->grid;
  if (!G) {
    // the finest level's copy, FP32 copy and scalings in one pass; the
    // Galerkin kernel writes those of every coarser level with its values
    // (with the tensor set-up the stencil conversion writes them)
    const iga_mg_setup_level_t& V = s->level[0];
    const int64_t m = std::max<int64_t>(V.nnz, V.n);
    hipLaunchKernelGGL(k_level0_prep, dim3((int)std::min<int64_t>(blocks_for(m), 8192)), dim3(256), 0,
                       st, V.n, (int64_t)V.nnz, V.indptr, V.cols, vals0,
                       V.vals64 != vals0 ? V.vals64 : nullptr, V.vals32, V.dinv, V.dsq);
  }
  if (G) {
    for (int l = 0; l < L; ++l) {
      const iga_mg_setup_level_t& V = s->level[l];
      if (!G->op64[l] || (int64_t)V.h * V.w != V.n) return IGA_ERR_ARG;
      if (l + 1 < L && (!G->f_lo[l][0] || !G->f_lo[l][1] || !G->rw[l][0] || !G->rw[l][1] || !G->work))
        return IGA_ERR_ARG;
    }
    if (s->radius != kGR) return IGA_ERR_UNSUPPORTED;
    if (G->stages < 0 || G->stages > 63) return IGA_ERR_ARG;
    if (G->gal_multi) {
      if (!G->gal_work) return IGA_ERR_ARG;
      for (int m = 1; m < L; ++m) {
        const int src = G->gal_src[m];
        if (src < 0 || src >= m || m - src > 3 || !G->gal_f_lo[m][0] || !G->gal_f_lo[m][1] ||
            !G->gal_rw[m][0] || !G->gal_rw[m][1])
          return IGA_ERR_ARG;
        // the levels of one source are consecutive below it
        if (m - src > 1 && G->gal_src[m - 1] != src) return IGA_ERR_ARG;
        // wide bands are staged whole in LDS; the intermediate of a launch
        // is addressed with 32-bit byte offsets
        const int row = m - src == 1 ? 0 : m - src == 2 ? gal_row<kGalKF[1]>() : gal_row<kGalKF[2]>();
        if ((int64_t)std::max(s->level[m].h, s->level[m].w) * row > kGalLdsBand ||
            (int64_t)kGN * s->level[src].h * s->level[m].w * 8 * 3 >= INT32_MAX)
          return IGA_ERR_UNSUPPORTED;
      }
    }
  }
  const int stages = G && G->stages ? G->stages : 3;
  // level 0 with its CSR copies (bit 0), without them (bit 4), or the
  // copies alone (bit 5)
  if (G && (stages & 32)) {
    const iga_mg_setup_level_t& V = s->level[0];
    const int64_t nnz = V.nnz;
    hipLaunchKernelGGL(k_csr0_copies, dim3((unsigned)std::min<int64_t>(blocks_for(nnz), 8192)),
                       dim3(256), 0, st, nnz, vals0, G->vals0_slot,
                       G->vals0_slot || V.vals64 != vals0 ? V.vals64 : nullptr, V.vals32);
  }
  if (G && (stages & (1 | 16))) {
    // level 0 in stencil layout (FP32, FP64, features) from its CSR
    const iga_mg_setup_level_t& V = s->level[0];
    GridOpsArgs a{};
    a.n_levels = 1;
    a.n[0] = V.n;
    a.w[0] = V.w;
    a.indptr[0] = V.indptr;
    a.cols[0] = V.cols;
    a.vals[0] = nullptr;  // from vals0: the CSR copies and scalings too
    // read through the slot when given (the values' address is then known
    // only to the device: always copy them)
    a.vals64_slot = G->vals0_slot;
    const bool copies = (stages & 1) != 0;
    a.csr64 = copies && (G->vals0_slot || V.vals64 != vals0) ? V.vals64 : nullptr;
    a.csr32 = copies ? V.vals32 : nullptr;
    a.dinv = V.dinv;
    a.dsq = V.dsq;
    float* op32_0 = G->op32[0];
    if (!op32_0) return IGA_ERR_ARG;
    a.op[0] = op32_0;
    a.vals64 = vals0;
    a.op64 = G->op64[0];
    a.feats = G->feats;
    hipLaunchKernelGGL(k_grid_ops, dim3((V.n + kOpsRows - 1) / kOpsRows, 1), dim3(kGTh), 0, st, a);
  }
  // stage 2 whole (bit 1), or its first launch (bit 2) and the rest (bit 3)
  const bool gal_first = (stages & 2) || (stages & 4);
  const bool gal_rest = (stages & 2) || (stages & 8);
  bool first_launch = true;
  if (G && (gal_first || gal_rest)) {
    auto gal = [&](int l) {
      const iga_mg_setup_level_t& F = s->level[l];
      const iga_mg_setup_level_t& C = s->level[l + 1];
      GalLevel g{};
      g.hf = F.h;
      g.wf = F.w;
      g.hc = C.h;
      g.wc = C.w;
      for (int d = 0; d < 2; ++d) {
        g.f_lo[d] = G->f_lo[l][d];
        g.rw[d] = G->rw[l][d];
      }
      g.K = G->op64[l];
      g.Kp = G->work;
      g.S64 = G->op64[l + 1];
      g.S32 = G->op32[l + 1];
      g.indptr = C.indptr;
      g.cols = C.cols;
      g.vals64 = C.vals64;
      g.vals32 = C.vals32;
      g.dinv = C.dinv;
      g.dsq = C.dsq;
      g.regular = G->regular[l + 1];
      return g;
    };
    if (G->gal_multi) {
      // the levels below each source level, up to three per launch pair
      for (int src = 0; src + 1 < L;) {
        GalMulti m{};
        int nt = 0;
        int64_t b0 = 0, b1 = 0, woff = 0;
        while (nt < 3 && src + nt + 1 < L && G->gal_src[src + nt + 1] == src) {
          const int tl = src + nt + 1;
          const iga_mg_setup_level_t& F = s->level[src];
          const iga_mg_setup_level_t& C = s->level[tl];
          GalLevel g = gal(tl - 1);
          for (int d = 0; d < 2; ++d) {
            g.f_lo[d] = G->gal_f_lo[tl][d];
            g.rw[d] = G->gal_rw[tl][d];
          }
          g.hf = F.h;
          g.wf = F.w;
          g.K = G->op64[src];
          g.Kp = G->gal_work + woff;
          woff += (int64_t)kGN * F.h * C.w;
          b0 = std::max(b0, ((int64_t)kGS * F.h * C.w + 255) / 256);
          b1 = std::max(b1, ((int64_t)kGS * C.n + 255) / 256);
          m.g[nt++] = g;
        }
        if (nt == 0) return IGA_ERR_ARG;
        const dim3 g0((unsigned)b0, nt), g1((unsigned)b1, nt);
        const bool p0 = first_launch ? gal_first : gal_rest;
        first_launch = false;
        if (nt == 1) {
          if (p0) hipLaunchKernelGGL(k_gal_multi0<1>, g0, dim3(256), 0, st, m);
          if (gal_rest) hipLaunchKernelGGL(k_gal_multi1<1>, g1, dim3(256), 0, st, m);
        } else if (nt == 2) {
          if (p0) hipLaunchKernelGGL(k_gal_multi0<2>, g0, dim3(256), 0, st, m);
          if (gal_rest) hipLaunchKernelGGL(k_gal_multi1<2>, g1, dim3(256), 0, st, m);
        } else {
          if (p0) hipLaunchKernelGGL(k_gal_multi0<3>, g0, dim3(256), 0, st, m);
          if (gal_rest) hipLaunchKernelGGL(k_gal_multi1<3>, g1, dim3(256), 0, st, m);
        }
        src += nt;
      }
    } else {
      for (int l = 0; l + 1 < L; ++l) {
        const iga_mg_setup_level_t& F = s->level[l];
        const iga_mg_setup_level_t& C = s->level[l + 1];
        const int64_t n0 = (int64_t)kGS * F.h * C.w, n1 = (int64_t)kGS * C.n;
        const GalLevel g = gal(l);
        if (first_launch ? gal_first : gal_rest)
          hipLaunchKernelGGL(k_gal_pass0, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, st, g);
        first_launch = false;
        if (gal_rest)
          hipLaunchKernelGGL(k_gal_pass1, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, g);
      }
    }
  } else if (!G) {
    for (int l = 0; l + 1 < L; ++l) {
      const iga_mg_setup_level_t& V = s->level[l];
      const iga_mg_setup_level_t& C = s->level[l + 1];
      hipLaunchKernelGGL(k_galerkin_rows, dim3((C.n + 3) / 4), dim3(256), 0, st, C.n, V.w,
                         s->radius, V.indptr, V.cols, V.vals64, V.pt_ptr, V.pt_cols, V.pt_vals,
                         C.indptr, C.cols, C.vals64, C.vals32, C.dinv, C.dsq);
    }
  }
  IGA_CHECK_LAUNCH();
  if (!gal_rest) return IGA_OK;
  // with the tensor set-up and a tail level, the dense tail (and with it the
  // coarse inverse) comes straight from the stencils
  if (G && G->tail_level > 0) return tail_dense(*s, st);
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
  if (n <= kGjLds) {
    hipLaunchKernelGGL(k_dense_inverse_lds, dim3(1), dim3(1024), 0, st, n, indptr, cols, vals, inv);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }
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

// out (n x ld, zero beyond column n) = a^T (a n x n)
__global__ void k_transpose_f(int n, int ld, const float* __restrict__ a, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * ld) return;
  const int i = (int)(t / ld), j = (int)(t % ld);
  out[t] = j < n ? a[(int64_t)j * n + i] : 0.f;
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
                         L.vals, L.omega, L.dsq, Bv[l], X[l], Rr[l], Gate{}, (int64_t)L.n);
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
                       mg.coarse_inv, Bv[Lc], X[Lc], Gate{}, (int64_t)C.n);
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
  const int64_t nl = (int64_t)Bn * dense_ld(Bn);
  hipLaunchKernelGGL(k_transpose_f, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, Bn,
                     dense_ld(Bn), X[t], out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int64_t iga_mg_tail_floats(const iga_mg_t* mg, int32_t level) {
  if (!mg || mg->n_levels < 2 || mg->n_levels > 16 || level < 0 || level >= mg->n_levels - 1)
    return -1;
  return tail_floats(*mg, level);
}

extern "C" int iga_mg_dense_ld(int32_t n) { return n > 0 ? dense_ld(n) : 0; }

extern "C" int iga_mg_grid_ops(const iga_mg_t* mg, const double* vals64, void* stream) {
  if (!mg || !mg->grid || mg->n_levels < 1 || mg->n_levels > 16) return IGA_ERR_ARG;
  const iga_mg_grid_t& G = *mg->grid;
  if (G.radius != kGR) return IGA_ERR_UNSUPPORTED;
  if (G.n_levels < 1 || G.n_levels > mg->n_levels) return IGA_ERR_ARG;
  GridOpsArgs a{};
  a.n_levels = G.n_levels;
  int nmax = 0;
  for (int l = 0; l < G.n_levels; ++l) {
    const iga_mg_level_t& L = mg->level[l];
    if (!L.indptr || !L.cols || !L.vals || !G.op[l]) return IGA_ERR_ARG;
    if (L.n <= 0 || L.w <= 0 || (int64_t)L.h * L.w != L.n) return IGA_ERR_SHAPE;
    a.n[l] = L.n;
    a.w[l] = L.w;
    a.indptr[l] = L.indptr;
    a.cols[l] = L.cols;
    a.vals[l] = L.vals;
    a.op[l] = G.op[l];
    nmax = std::max(nmax, L.n);
  }
  a.vals64 = vals64;
  a.op64 = vals64 ? G.op64 : nullptr;
  a.feats = a.op64 ? G.feats : nullptr;
  hipLaunchKernelGGL(k_grid_ops, dim3((nmax + kOpsRows - 1) / kOpsRows, G.n_levels), dim3(kGTh),
                     0, (hipStream_t)stream, a);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_mg_tail(const iga_mg_t* mg, int32_t level, float* work, float* out,
                           void* stream) {
  if (!mg || !work || !out || !mg->coarse_inv) return IGA_ERR_ARG;
  if (mg->n_levels < 2 || mg->n_levels > 16 || level < 0 || level >= mg->n_levels - 1)
    return IGA_ERR_ARG;
  if (mg->level[level].n > kDenseMaxRows) return IGA_ERR_UNSUPPORTED;
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
  int dot_blocks = 0;
  return vcycle(*mg, level, b, x, Gate{}, DotOut{}, &dot_blocks, (hipStream_t)stream);
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
                         accumulate, Gate{}, DotOut{});
    else
      hipLaunchKernelGGL((k_stencil<KS, false>), grid, dim3(256), 0, st, h, w, stencil, dsq, v,
                         out, accumulate, Gate{}, DotOut{});
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
// operator whose values are rewritten in place. One graph per iteration
// count is kept; a run first replays as many iterations as the last run
// needed without looking (one launch), then one iteration per look, so few
// iterations run past convergence (their kernels return at once but still
// cost a launch each) and a run like the last one reads the flag once.
struct FcgGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct iga_fcg_plan {
  iga_mg_t mg;
  int n, nnz, lpr, vb, rb, eb, G;
  bool grid64;  // level-0 product in FP64 stencil layout (grid cycle)
  int nb_spmv;  // blocks (partials) of the product kernel
  const int32_t* indptr;
  const int32_t* cols;
  const double* vals;
  double* b;  // written by a run given b_in
  double* x;
  FBufs F;
  double *r, *P[2], *Q[2];
  double* R[2];  // grid cycle: residual ping-pong (an iteration of parity k writes R[k])
  int nt0;       // row tiles of level 0 (grid cycle: partials of r.r, p.r, p.q)
  float *r32, *z32;
  FState* host_pinned = nullptr;  // the state as the host reads it
  static constexpr int kMaxGraph = 16;
  FcgGraph gk[kMaxGraph + 1];  // gk[k]: k iterations from parity 0
  FcgGraph g1odd;              // one iteration from parity 1
  // iteration counts of the last kHist runs (newest at hist[0]); the first
  // replay of a run takes the most frequent one (ties: the newer)
  static constexpr int kHist = 8;
  int hist[kHist] = {};
  int n_hist = 0;

  int predicted_iters() const {
    int best = 0, best_count = 0;
    for (int a = 0; a < n_hist; ++a) {
      int c = 0;
      for (int b = 0; b < n_hist; ++b) c += hist[b] == hist[a];
      if (c > best_count) {
        best = hist[a];
        best_count = c;
      }
    }
    return best;
  }

  // the waiting update of the iteration before one of parity par, applied
  // from R[1 - par] into r_new (R[par] inside a cycle, in place at a flush)
  FcgFuse fuse(int par, double* r_new) const {
    return FcgFuse{F.st, R[1 - par], r_new, P[1 - par], Q[1 - par], x, r32, F.P_b, F.P_c, F.P_d,
                   nt0};
  }

  int iteration(int par, hipStream_t st) {
    double* p = P[par];
    double* pp = P[1 - par];
    double* q = Q[par];
    double* qp = Q[1 - par];
    const DotOut dot{qp, F.P_a};
    int na = 0;
    if (grid64) {
      // the previous update runs in the cycle's first kernel, the gate in
      // its second; the product leaves its own update waiting
      const Gate gate{&F.st->done, F.P_d, nt0, F.st, 1};
      const FcgFuse fz = fuse(par, R[par]);
      int rc = vcycle(mg, 0, r32, z32, gate, dot, &na, st, &fz);
      if (rc) return rc;
      const iga_mg_level_t& L0 = mg.level[0];
      hipLaunchKernelGGL(k_fcg_spmv_grid, row_tiles(L0), dim3(kGTh), 0, st, L0.h, L0.w, na,
                         mg.grid->op64, z32, pp, R[par], p, q, F.st, F.P_a, F.P_b, F.P_c);
      IGA_CHECK_LAUNCH();
      return (int)IGA_OK;
    }
    const Gate gate{&F.st->done, F.P_d, vb, F.st, 0};
    int rc = vcycle(mg, 0, r32, z32, gate, dot, &na, st);
    if (rc) return rc;
    {
      by_lpr(lpr, [&](auto L) {
        constexpr int LPR = decltype(L)::value;
        hipLaunchKernelGGL(k_fcg_spmv<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, na, indptr, cols,
                           vals, z32, pp, r, p, q, F);
        return 0;
      });
    }
    hipLaunchKernelGGL(k_fcg_update, dim3(vb), dim3(kVecThreads), 0, st, n, nb_spmv, p, q, x, r,
                       r32, F);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  }

  // end of a batch whose last iteration had parity par: the waiting update
  // (grid cycle) and the convergence test
  int check(int par, hipStream_t st) {
    if (grid64) {
      const FcgFuse fz = fuse(1 - par, R[par]);  // R[par] in place
      hipLaunchKernelGGL(k_fcg_flush, row_tiles(mg.level[0]), dim3(kGTh), 0, st, mg.level[0].h,
                         mg.level[0].w, fz);
    }
    hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kVecThreads), 0, st, grid64 ? nt0 : vb, F, 0,
                       0.0, 0);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }

  // grid cycle, end of a replay: the waiting update, the FP64 residual of x
  // and the test (one kernel for the test and the residual's sum)
  int check_residual(int par, hipStream_t st) {
    const FcgFuse fz = fuse(1 - par, R[par]);  // R[par] in place
    hipLaunchKernelGGL(k_fcg_flush, row_tiles(mg.level[0]), dim3(kGTh), 0, st, mg.level[0].h,
                       mg.level[0].w, fz);
    by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL(k_fcg_resid<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols,
                         vals, x, b, F);
      return 0;
    });
    hipLaunchKernelGGL(k_fcg_check_res, dim3(1), dim3(kVecThreads), 0, st, nt0, rb, F);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }

  int capture_one(FcgGraph& out, int iters, int par0, hipStream_t st) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return IGA_ERR_HIP;
    int rc = 0;
    for (int k = 0; k < iters && !rc; ++k) rc = iteration((par0 + k) & 1, st);
    // grid cycle: the FP64 residual of x is part of every replay
    if (!rc) rc = grid64 ? check_residual((par0 + iters - 1) & 1, st) : check((par0 + iters - 1) & 1, st);
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

  // one graph per iteration count up to G (a batch is one launch: its end,
  // the waiting update, the test and the residual, runs once) and one
  // single iteration from the odd parity
  int capture(hipStream_t st) {
    if (G <= 0) return IGA_OK;
    int rc = capture_one(g1odd, 1, 1, st);
    for (int k = 1; k <= std::min(G, kMaxGraph) && !rc; ++k) rc = capture_one(gk[k], k, 0, st);
    return rc;
  }

  // k more iterations from the `launched`-th
  int replay(int k, int& launched, hipStream_t st) {
    const int gmax = std::min(G, kMaxGraph);
    while (k > 0) {
      const FcgGraph* g;
      int m;
      if (launched & 1) {
        g = &g1odd;
        m = 1;
      } else {
        m = std::min(k, gmax);
        g = &gk[m];
      }
      if (hipGraphLaunch(g->exec, st) != hipSuccess) return IGA_ERR_HIP;
      k -= m;
      launched += m;
    }
    return IGA_OK;
  }

  int true_residual(hipStream_t st) {
    by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL(k_fcg_resid<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols,
                         vals, x, b, F);
      return 0;
    });
    hipLaunchKernelGGL(k_fcg_resid_fin, dim3(1), dim3(kVecThreads), 0, st, rb, F);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }

  int run(double tol, int maxiter, int32_t* out_iters, double* out_rr, double* out_bb,
          double* out_res, double* x_out, const double* b_in, hipStream_t st) {
    const bool graphs = gk[1].exec != nullptr;
    // with graphs the init writes the state itself and the first gate
    // records ||b||^2 (one kernel less); it runs at least one iteration
    const bool fold = graphs && maxiter > 0;
    // the first iteration (parity 0) reads R[1] on the grid cycle
    // on the grid cycle the flush at the end of every replay writes x_out
    hipLaunchKernelGGL(k_fcg_init, dim3(vb), dim3(kVecThreads), 0, st, n, b_in, b, x,
                       grid64 ? R[1] : r, r32, P[1], Q[1], F, grid64 ? x_out : nullptr,
                       FInit{fold ? 1 : 0, tol * tol, maxiter});
    if (!fold)
      hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kVecThreads), 0, st, vb, F, 1, tol * tol,
                         maxiter);
    IGA_CHECK_LAUNCH();
    FState host{};
    int launched = 0, status = IGA_OK;
    // the FP64 residual is recomputed after every batch when asked for, so
    // the look that finds convergence also reads it
    bool res_fresh = false;
    // iterations before the first look: the most frequent count of the
    // recent runs (four on the first run), then one per look; never more
    // than maxiter in all
    int next = n_hist > 0 && predicted_iters() > 0 ? predicted_iters() : 4;
    for (bool look = false;; look = true) {
      // the state is read into pinned memory (an asynchronous copy; a
      // pageable destination takes a staged, blocking one)
      if (look && (hipMemcpyAsync(host_pinned, F.st, sizeof(FState), hipMemcpyDeviceToHost, st) !=
                       hipSuccess ||
                   hipStreamSynchronize(st) != hipSuccess)) {
        status = IGA_ERR_HIP;
        break;
      }
      if (look) host = *host_pinned;
      if ((look && host.done) || launched >= maxiter) {
        if (!look) continue;
        if (!out_res || res_fresh) break;
        if ((status = true_residual(st))) break;
        res_fresh = true;
        continue;
      }
      const int k = std::min(next, maxiter - launched);
      if (graphs) {
        status = replay(k, launched, st);
        if (status) break;
        next = 1;
      } else {
        for (int j = 0; j < std::min(k, 2) && !status; ++j, ++launched)
          status = iteration(launched & 1, st);
        if (!status) status = check((launched - 1) & 1, st);
        if (status) break;
      }
      res_fresh = graphs && grid64;  // the replayed graphs end with the residual
      if (out_res && !res_fresh) {
        if ((status = true_residual(st))) break;
        res_fresh = true;
      }
      // the caller's copy of x rides behind the batch
      if (x_out && !grid64 && hipMemcpyAsync(x_out, x, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice,
                                  st) != hipSuccess) {
        status = IGA_ERR_HIP;
        break;
      }
    }
    if (status == IGA_OK) {
      for (int k = kHist - 1; k > 0; --k) hist[k] = hist[k - 1];
      hist[0] = host.iter;
      n_hist = std::min(n_hist + 1, kHist);
    }
    if (out_iters) *out_iters = host.iter;
    if (out_rr) *out_rr = host.rr;
    if (out_bb) *out_bb = host.bb;
    if (out_res) *out_res = res_fresh && status == IGA_OK ? sqrt(host.res2) : -1.0;
    return status;
  }

  ~iga_fcg_plan() {
    if (host_pinned) (void)hipHostFree(host_pinned);
    for (int k = 0; k <= kMaxGraph + 1; ++k) {
      FcgGraph* g = k <= kMaxGraph ? &gk[k] : &g1odd;
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
  if (hipHostMalloc(reinterpret_cast<void**>(&P->host_pinned), sizeof(FState)) != hipSuccess) {
    P->host_pinned = nullptr;
    delete P;
    return IGA_ERR_HIP;
  }
  P->mg = *mg;
  P->n = n;
  P->nnz = nnz;
  P->indptr = indptr;
  P->cols = cols;
  P->vals = vals;
  // a run writes b only when given b_in (iga_fcg_mg passes none)
  P->b = const_cast<double*>(b);
  P->x = x;
  P->F.st = reinterpret_cast<FState*>(work);
  P->r = work + 8;
  P->P[0] = P->r + n;
  P->P[1] = P->r + 2 * n;
  P->Q[0] = P->r + 3 * n;
  P->Q[1] = P->r + 4 * n;
  P->R[0] = P->r;
  P->R[1] = P->r + 5 * n;
  double* parts = P->r + 6 * n;
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
  {
    const dim3 tg = row_tiles(mg->level[0]);
    P->grid64 = grid_cycle_ok(*mg) && mg->grid->op64 && (int)(tg.x * tg.y) <= kRedBlocks;
    P->nb_spmv = P->grid64 ? (int)(tg.x * tg.y) : P->rb;
    P->nt0 = (int)(tg.x * tg.y);
  }
  rc = P->capture((hipStream_t)stream);
  if (rc) {
    delete P;
    return rc;
  }
  *out = P;
  return IGA_OK;
}

extern "C" int iga_fcg_mg_run(iga_fcg_plan_t* plan, double tol, int32_t maxiter,
                              int32_t* out_iters, double* out_rr, double* out_bb,
                              double* out_res, double* x_out, const double* b_in, void* stream) {
  if (!plan) return IGA_ERR_ARG;
  return plan->run(tol, maxiter, out_iters, out_rr, out_bb, out_res, x_out, b_in,
                   (hipStream_t)stream);
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
  rc = plan->run(tol, maxiter, out_iters, out_rr, out_bb, nullptr, nullptr, nullptr, st);
  delete plan;
  return rc;
}
