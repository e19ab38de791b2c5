// CSR structure, sparse matrix-vector products and the device-resident
// Jacobi-preconditioned conjugate-gradient solver.
//
// CG keeps every scalar on the device (no host round trip per iteration):
// each iteration is three kernels -- SpMV fused with p.Ap, the x/r/z update
// fused with r.z and r.r, and the direction update. Each kernel writes one
// partial sum per block; every block of the consuming kernel re-reduces the
// partials in the same fixed order, so the scalars are deterministic and no
// block waits for another. The single-reduction variant (Chronopoulos-Gear)
// needs two kernels per iteration. Kernels read a device `done` flag, so the
// host enqueues iterations in batches and synchronises only once per batch.
#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "matfree.hpp"

namespace iga {

// ---------------------------------------------------------------------------
// CSR pattern of the free-free block on a tensor patch
// ---------------------------------------------------------------------------

struct NbInfo {
  int dim, vec;
  int ncp[3];
  const int32_t* lo[3];
  const int32_t* hi[3];
};

template <bool FILL>
__global__ void k_csr_rows(NbInfo nb, int n_free, const int32_t* __restrict__ freed,
                           const int32_t* __restrict__ f2f, int32_t* counts_or_indptr,
                           int32_t* cols, int32_t* rows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_free) return;
  const int g = freed[r];
  const int s = g / nb.vec;
  int ii[3] = {0, 0, 0};
  {
    int rem = s;
    for (int d = 0; d < nb.dim; ++d) {
      ii[d] = rem % nb.ncp[d];
      rem /= nb.ncp[d];
    }
  }
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < nb.dim; ++d) {
    lo[d] = nb.lo[d][ii[d]];
    hi[d] = nb.hi[d][ii[d]];
  }
  const int s0 = nb.ncp[0], s1 = nb.ncp[0] * (nb.dim > 1 ? nb.ncp[1] : 1);
  int pos = FILL ? counts_or_indptr[r] : 0;
  for (int j2 = lo[2]; j2 <= hi[2]; ++j2)
    for (int j1 = lo[1]; j1 <= hi[1]; ++j1)
      for (int j0 = lo[0]; j0 <= hi[0]; ++j0) {
        const int sj = j0 + j1 * s0 + j2 * s1;
        for (int c = 0; c < nb.vec; ++c) {
          const int f = f2f[(size_t)nb.vec * sj + c];
          if (f < 0) continue;
          if (FILL) {
            cols[pos] = f;
            if (rows) rows[pos] = r;
          }
          ++pos;
        }
      }
  if (!FILL) counts_or_indptr[r] = pos;
}

static int make_nb(int dim, const int32_t* n_cp, const int32_t* const* lo,
                   const int32_t* const* hi, int vec, NbInfo& nb) {
  if (!n_cp || !lo || !hi || dim < 1 || dim > 3 || vec < 1) return IGA_ERR_ARG;
  nb.dim = dim;
  nb.vec = vec;
  for (int d = 0; d < 3; ++d) {
    nb.ncp[d] = d < dim ? n_cp[d] : 1;
    nb.lo[d] = d < dim ? lo[d] : nullptr;
    nb.hi[d] = d < dim ? hi[d] : nullptr;
  }
  return IGA_OK;
}

// ---------------------------------------------------------------------------
// exclusive scan (int32), recursive over 1024-element blocks
// ---------------------------------------------------------------------------

constexpr int kScanBlock = 1024;

__global__ void k_scan_block(int n, const int32_t* __restrict__ in, int32_t* out_plus1,
                             int32_t* block_sums) {
  __shared__ int32_t buf[kScanBlock];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  buf[threadIdx.x] = i < n ? in[i] : 0;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int v = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += v;
    __syncthreads();
  }
  if (i < n) out_plus1[i] = buf[threadIdx.x];
  if (threadIdx.x == kScanBlock - 1 && block_sums) block_sums[blockIdx.x] = buf[threadIdx.x];
}

__global__ void k_scan_add(int n, int32_t* out_plus1, const int32_t* __restrict__ offs) {
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  if (i < n && blockIdx.x > 0) out_plus1[i] += offs[blockIdx.x - 1];
}

// inclusive scan of in[0..n) into out[0..n)
static int scan_inclusive(int n, const int32_t* in, int32_t* out, int32_t* ws, hipStream_t st) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kScanBlock), 0, st, n, in, out, nullptr);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }
  int32_t* sums = ws;
  int32_t* sums_scan = ws + nb;
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kScanBlock), 0, st, n, in, out, sums);
  IGA_CHECK_LAUNCH();
  int rc = scan_inclusive(nb, sums, sums_scan, ws + 2 * nb, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, st, n, out, sums_scan);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

static int64_t scan_ws(int64_t n) {
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb <= 1) return 0;
  return 2 * nb + scan_ws(nb);
}

// ---------------------------------------------------------------------------
// SpMV and Jacobi
// ---------------------------------------------------------------------------

__global__ void k_diag_inv(int n, const int32_t* __restrict__ indptr,
                           const int32_t* __restrict__ cols, const double* __restrict__ vals,
                           double* inv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int lo = indptr[r], hi = indptr[r + 1];
  const int k = lower_bound(cols, lo, hi, r);
  const double d = (k < hi && cols[k] == r) ? vals[k] : 0.0;
  inv[r] = d != 0.0 ? 1.0 / d : 1.0;
}

template <int LPR>
__global__ void __launch_bounds__(256)
    k_spmv(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
           const double* __restrict__ vals, const double* __restrict__ x, double* y) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = gid / LPR, l = gid % LPR;
  if (r >= n) return;
  const double s = row_dot<LPR>(r, indptr, cols, vals, x, l);
  if (l == 0) y[r] = s;
}

// ---------------------------------------------------------------------------
// Jacobi-PCG with device-resident scalars
// ---------------------------------------------------------------------------
//
// Every reduction writes one partial per block; the kernel that consumes the
// scalar re-reduces the partials in every block, in the same fixed order, so
// all blocks agree bit for bit and no atomics or grid barriers are needed.

__global__ void __launch_bounds__(kVecThreads)
    k_cg_init(int n, const double* __restrict__ b, const double* __restrict__ dinv, double* x,
              double* r, double* z, double* p, CGBufs B) {
  double rz = 0.0, bb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double bi = b[i];
    const double zi = dinv[i] * bi;
    x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    p[i] = zi;
    rz += bi * zi;
    bb += bi * bi;
  }
  store_partial(B.P_rz, rz);
  store_partial(B.P_bb, bb);
}

__global__ void k_cg_init_fin(int nblk, CGBufs B, double tol2, int maxiter) {
  const double bb = reduce_partials(B.P_bb, nblk);
  if (threadIdx.x == 0) {
    CGState* st = B.st;
    st->bb = bb;
    st->rr = bb;
    st->tol2 = tol2;
    st->iter = 0;
    st->maxiter = maxiter;
    st->done = (!(bb > tol2 * bb) || maxiter <= 0) ? 1 : 0;
  }
}

// q = A p, one row per group of LPR lanes, and the p.q partials (one per
// block). Blocks of kSpmvThreads (16 waves) keep up to kRedBlocks x 16 x
// 64 / LPR rows in flight, so the grid-stride loop makes few passes and each
// pass's dependent loads (columns, then x) overlap across the whole grid.
constexpr int kSpmvThreads = 1024;

template <int LPR, typename TV>
__global__ void __launch_bounds__(kSpmvThreads)
    k_cg_spmv(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
              const TV* __restrict__ vals, const double* __restrict__ p, double* q, CGBufs B) {
  if (B.st->done) return;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double pq = 0.0;
  for (int r = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; r < n; r += groups) {
    const double s = row_dot<LPR, TV>(r, indptr, cols, vals, p, l);
    if (l == 0) {
      q[r] = s;
      pq += s * p[r];
    }
  }
  store_partial(B.P_pq, pq);
}

__global__ void __launch_bounds__(kVecThreads)
    k_cg_update(int n, int nblk_pq, int nblk, int par, const double* __restrict__ dinv,
                const double* __restrict__ p, const double* __restrict__ q, double* x, double* r,
                double* z, CGBufs B) {
  if (B.st->done) return;
  const double pq = reduce_partials(B.P_pq, nblk_pq);
  const double rz_old = reduce_partials(B.P_rz + par * kRedBlocks, nblk);
  const double alpha = rz_old / pq;
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    const double zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  store_partial(B.P_rz + (1 - par) * kRedBlocks, rz);
  store_partial(B.P_rr, rr);
}

__global__ void __launch_bounds__(kVecThreads)
    k_cg_dir(int n, int nblk, int par, const double* __restrict__ z, double* p, CGBufs B) {
  CGState* st = B.st;
  if (st->done) return;
  const double rz_new = reduce_partials(B.P_rz + (1 - par) * kRedBlocks, nblk);
  const double rz_old = reduce_partials(B.P_rz + par * kRedBlocks, nblk);
  const double beta = rz_new / rz_old;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
  if (blockIdx.x == 0) {
    const double rr = reduce_partials(B.P_rr, nblk);
    if (threadIdx.x == 0) {
      st->rr = rr;
      st->iter += 1;
      st->done = (!(rr > st->tol2 * st->bb) || st->iter >= st->maxiter) ? 1 : 0;
    }
  }
}

template <int LPR>
__global__ void __launch_bounds__(kVecThreads)
    k_residual(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
               const double* __restrict__ vals, const double* __restrict__ u,
               const double* __restrict__ b, double* P_rr, double* P_bb) {
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double rr = 0.0, bb = 0.0;
  for (int r = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; r < n; r += groups) {
    const double s = row_dot<LPR>(r, indptr, cols, vals, u, l);
    if (l == 0) {
      const double d = s - b[r];
      rr += d * d;
      bb += b[r] * b[r];
    }
  }
  store_partial(P_rr, rr);
  store_partial(P_bb, bb);
}

__global__ void k_residual_fin(int nblk, const double* P_rr, const double* P_bb, double* out) {
  const double rr = reduce_partials(P_rr, nblk);
  const double bb = reduce_partials(P_bb, nblk);
  if (threadIdx.x == 0) {
    out[0] = rr;
    out[1] = bb;
  }
}

__global__ void k_zero_u32(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// random Fourier features
// ---------------------------------------------------------------------------

constexpr int kRffTile = 256;

__global__ void __launch_bounds__(256)
    k_rff(int n, int dim, const double* __restrict__ pts, int m, const double* __restrict__ omega,
          const double* __restrict__ phase, double amp, double* out) {
  __shared__ double om[kRffTile * 3];
  __shared__ double ph[kRffTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double xv[3] = {0.0, 0.0, 0.0};
  if (i < n)
    for (int d = 0; d < dim; ++d) xv[d] = pts[(size_t)i * dim + d];
  double s = 0.0;
  for (int m0 = 0; m0 < m; m0 += kRffTile) {
    const int cnt = min(kRffTile, m - m0);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      ph[t] = phase[m0 + t];
      for (int d = 0; d < dim; ++d) om[t * 3 + d] = omega[(size_t)(m0 + t) * dim + d];
    }
    __syncthreads();
    for (int t = 0; t < cnt; ++t) {
      double arg = ph[t];
      for (int d = 0; d < dim; ++d) arg += om[t * 3 + d] * xv[d];
      s += cos(arg);
    }
  }
  if (i < n) out[i] = amp * s;
}

}  // namespace iga

using namespace iga;

extern "C" int iga_csr_count(int32_t dim, const int32_t* n_cp, const int32_t* const* nb_lo,
                             const int32_t* const* nb_hi, int32_t vec, int32_t n_free,
                             const int32_t* freed, const int32_t* full_to_free, int32_t* counts,
                             void* stream) {
  NbInfo nb;
  if (int rc = make_nb(dim, n_cp, nb_lo, nb_hi, vec, nb)) return rc;
  if (n_free == 0) return IGA_OK;
  hipLaunchKernelGGL(k_csr_rows<false>, dim3((n_free + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, nb, n_free, freed, full_to_free, counts, nullptr,
                     nullptr);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_csr_fill(int32_t dim, const int32_t* n_cp, const int32_t* const* nb_lo,
                            const int32_t* const* nb_hi, int32_t vec, int32_t n_free,
                            const int32_t* freed, const int32_t* full_to_free,
                            const int32_t* indptr, int32_t* cols, int32_t* rows, void* stream) {
  NbInfo nb;
  if (int rc = make_nb(dim, n_cp, nb_lo, nb_hi, vec, nb)) return rc;
  if (n_free == 0) return IGA_OK;
  hipLaunchKernelGGL(k_csr_rows<true>, dim3((n_free + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, nb, n_free, freed, full_to_free,
                     const_cast<int32_t*>(indptr), cols, rows);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int64_t iga_scan_workspace(int32_t n) { return scan_ws(n) + 1; }

extern "C" int iga_scan_exclusive(int32_t n, const int32_t* in, int32_t* out, int32_t* ws,
                                  void* stream) {
  if (!in || !out) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned int*>(out), 1);
  IGA_CHECK_LAUNCH();
  if (n == 0) return IGA_OK;
  return scan_inclusive(n, in, out + 1, ws, st);
}

extern "C" int iga_csr_diag_inv(int32_t n, const int32_t* indptr, const int32_t* cols,
                                const double* vals, double* inv, void* stream) {
  if (n == 0) return IGA_OK;
  hipLaunchKernelGGL(k_diag_inv, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                     indptr, cols, vals, inv);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

static int nnz_of(int n, const int32_t* indptr, hipStream_t st, int* nnz) {
  int32_t v = 0;
  if (hipMemcpyAsync(&v, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess)
    return IGA_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
  *nnz = v;
  return IGA_OK;
}

// Lanes per row of a pattern, read from its row pointers once per (indptr,
// n) and remembered, so a repeated product issues no host round trip (and
// can be captured in a graph). The choice only tunes speed: a stale entry
// for a reused address gives the same product.
static int spmv_lpr(int n, const int32_t* indptr, hipStream_t st, int* lpr) {
  static std::mutex mu;
  static std::unordered_map<const int32_t*, std::pair<int, int>> seen;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = seen.find(indptr);
    if (it != seen.end() && it->second.first == n) {
      *lpr = it->second.second;
      return IGA_OK;
    }
  }
  int nnz = 0;
  if (int rc = nnz_of(n, indptr, st, &nnz)) return rc;
  *lpr = pick_lpr(n, nnz);
  std::lock_guard<std::mutex> lock(mu);
  seen[indptr] = {n, *lpr};
  return IGA_OK;
}

extern "C" int iga_spmv(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                        const double* x, double* y, void* stream) {
  if (n == 0) return IGA_OK;
  hipStream_t st = (hipStream_t)stream;
  int lpr = 8;
  if (int rc = spmv_lpr(n, indptr, st, &lpr)) return rc;
  return by_lpr(lpr, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    const long long threads = (long long)n * LPR;
    hipLaunchKernelGGL(k_spmv<LPR>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, n,
                       indptr, cols, vals, x, y);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}

// workspace (doubles), large enough for both loops:
//   reliable-update loop: state (8), r, z, p, q (4n), partials (5 kRedBlocks)
//   single-reduction loop: state (16), r, u, p, s, w (5n), partials (6 kRedBlocks)
extern "C" int64_t iga_cg_workspace(int32_t n) { return 16 + 5 * (int64_t)n + 6 * kRedBlocks; }

// ---------------------------------------------------------------------------
// Single-reduction Jacobi-PCG (Chronopoulos & Gear, 1989): the same iterates
// as the textbook recurrence in exact arithmetic, with one reduction point
// per iteration, so an iteration is the operator (w = A u with the w.u
// partials fused in) plus one vector kernel:
//   beta_i = gamma_i / gamma_{i-1},  alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1})
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,  u = D^-1 r
// with gamma = r.u, delta = w.u; the stopping test ||r|| <= tol ||b|| is made
// on r_i before iteration i, as in the textbook loop.
// ---------------------------------------------------------------------------

struct CG1Bufs {
  CGState* st;
  double* gam;   // [2] gamma by iteration parity
  double* alp;   // [2] alpha by iteration parity
  double* P_ru;  // [2][kRedBlocks]
  double* P_rr;  // [2][kRedBlocks]
  double* P_wu;  // [kRedBlocks]
  double* P_bb;  // [kRedBlocks]
};

static CG1Bufs cg1_bufs(double* work, int32_t n) {
  CG1Bufs B;
  B.st = reinterpret_cast<CGState*>(work);
  B.gam = work + 8;
  B.alp = work + 10;
  double* parts = work + 16 + 5 * (int64_t)n;
  B.P_ru = parts;
  B.P_rr = parts + 2 * kRedBlocks;
  B.P_wu = parts + 4 * kRedBlocks;
  B.P_bb = parts + 5 * kRedBlocks;
  return B;
}

__global__ void __launch_bounds__(kVecThreads)
    k_cg1_init(int n, const double* __restrict__ b, const double* __restrict__ dinv, double* x,
               double* r, double* u, double* p, double* s, CG1Bufs B) {
  double ru = 0.0, bb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double bi = b[i];
    const double ui = dinv[i] * bi;
    x[i] = 0.0;
    r[i] = bi;
    u[i] = ui;
    p[i] = 0.0;
    s[i] = 0.0;
    ru += bi * ui;
    bb += bi * bi;
  }
  store_partial(B.P_ru, ru);
  store_partial(B.P_rr, bb);
  store_partial(B.P_bb, bb);
}

__global__ void k_cg1_init_fin(int nblk, CG1Bufs B, double tol2, int maxiter) {
  const double bb = reduce_partials(B.P_bb, nblk);
  if (threadIdx.x == 0) {
    CGState* st = B.st;
    st->bb = bb;
    st->rr = bb;
    st->tol2 = tol2;
    st->iter = 0;
    st->maxiter = maxiter;
    st->done = (!(bb > tol2 * bb) || maxiter <= 0) ? 1 : 0;
  }
}

// iteration `it` (host counter): test r_it, then the update. Every block
// reduces the same partials in the same order, so every block takes the same
// branch; block 0 records the state. gamma/alpha of iteration it are written
// to slot it&1 and read by iteration it+1 from the other slot.
__global__ void __launch_bounds__(kVecThreads)
    k_cg1_update(int n, int it, int nblk_wu, int nblk, const double* __restrict__ dinv,
                 const double* __restrict__ w, double* u, double* x, double* r, double* p,
                 double* s, CG1Bufs B) {
  CGState* st = B.st;
  if (st->done) return;
  const int par = it & 1;
  const double gamma = reduce_partials(B.P_ru + par * kRedBlocks, nblk);
  const double rho = reduce_partials(B.P_rr + par * kRedBlocks, nblk);
  const double delta = reduce_partials(B.P_wu, nblk_wu);
  if (!(rho > st->tol2 * st->bb) || it >= st->maxiter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rho;
      st->iter = it;
      st->done = 1;
    }
    return;
  }
  const double beta = it == 0 ? 0.0 : gamma / B.gam[1 - par];
  const double alpha = it == 0 ? gamma / delta : gamma / (delta - beta * gamma / B.alp[1 - par]);
  double ru = 0.0, rr = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double pi = u[i] + beta * p[i];
    const double si = w[i] + beta * s[i];
    p[i] = pi;
    s[i] = si;
    x[i] += alpha * pi;
    const double ri = r[i] - alpha * si;
    const double ui = dinv[i] * ri;
    r[i] = ri;
    u[i] = ui;
    ru += ri * ui;
    rr += ri * ri;
  }
  store_partial(B.P_ru + (1 - par) * kRedBlocks, ru);
  store_partial(B.P_rr + (1 - par) * kRedBlocks, rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    B.gam[par] = gamma;
    B.alp[par] = alpha;
    st->rr = rho;
    st->iter = it + 1;
  }
}

// Two pinned host copies of the solver state and their events, per host
// thread and device (created at the first solve there, kept for the
// process), so that reading the state of one batch overlaps the next.
struct StatePipe {
  CGState* host = nullptr;  // [2], pinned
  hipEvent_t ev[2] = {nullptr, nullptr};
};

static int state_pipe(StatePipe** out) {
  constexpr int kMaxDev = 64;
  thread_local StatePipe pipes[kMaxDev];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return IGA_ERR_HIP;
  StatePipe& P = pipes[dev];
  if (!P.host) {
    void* h = nullptr;
    if (hipHostMalloc(&h, 2 * sizeof(CGState), hipHostMallocDefault) != hipSuccess) return IGA_ERR_HIP;
    for (int i = 0; i < 2; ++i)
      if (hipEventCreateWithFlags(&P.ev[i], hipEventDisableTiming) != hipSuccess) return IGA_ERR_HIP;
    P.host = static_cast<CGState*>(h);
  }
  *out = &P;
  return IGA_OK;
}

// `op(u, w, P)` must produce w = A u and the w.u partials into P (returns the
// number of partial blocks, or a negative status); a set done flag must make
// it a no-op.
template <typename Op>
static int cg1_loop(int32_t n, const double* dinv, const double* b, double* x, double tol,
                    int32_t maxiter, double* work, int32_t* out_iters, double* out_rr,
                    double* out_bb, hipStream_t st, Op op) {
  if (n == 0) {
    if (out_iters) *out_iters = 0;
    if (out_rr) *out_rr = 0.0;
    if (out_bb) *out_bb = 0.0;
    return IGA_OK;
  }
  CG1Bufs B = cg1_bufs(work, n);
  double* r = work + 16;
  double* u = r + n;
  double* p = u + n;
  double* s = p + n;
  double* w = s + n;
  const int vb = vec_blocks(n);
  hipLaunchKernelGGL(k_cg1_init, dim3(vb), dim3(kVecThreads), 0, st, n, b, dinv, x, r, u, p, s, B);
  hipLaunchKernelGGL(k_cg1_init_fin, dim3(1), dim3(kVecThreads), 0, st, vb, B, tol * tol, maxiter);
  IGA_CHECK_LAUNCH();
  // The host keeps one batch queued ahead of the state it reads: batch k+1
  // is enqueued before the state after batch k is awaited, so the GPU never
  // drains between batches (after convergence the batch in flight is
  // early-exit launches). The state copies go to pinned host slots.
  StatePipe* pipe = nullptr;
  if (int rc = state_pipe(&pipe)) return rc;
  const int batch = 8;
  int launched = 0;
  auto enqueue = [&](int slot) -> int {
    for (int k = 0; k < batch; ++k, ++launched) {
      const int nb = op(u, w, B.P_wu, B.st);
      if (nb < 0) return nb;
      hipLaunchKernelGGL(k_cg1_update, dim3(vb), dim3(kVecThreads), 0, st, n, launched, nb, vb,
                         dinv, w, u, x, r, p, s, B);
      IGA_CHECK_LAUNCH();
    }
    if (hipMemcpyAsync(&pipe->host[slot], B.st, sizeof(CGState), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipEventRecord(pipe->ev[slot], st) != hipSuccess)
      return IGA_ERR_HIP;
    return IGA_OK;
  };
  // the state after the initialisation
  if (hipMemcpyAsync(&pipe->host[0], B.st, sizeof(CGState), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipEventRecord(pipe->ev[0], st) != hipSuccess)
    return IGA_ERR_HIP;
  CGState host{};
  for (int k = 0;; ++k) {
    // batch k+1 (slot (k+1)&1) queued behind the state of slot k&1
    if (int rc = enqueue((k + 1) & 1)) return rc;
    if (hipEventSynchronize(pipe->ev[k & 1]) != hipSuccess) return IGA_ERR_HIP;
    host = pipe->host[k & 1];
    if (host.done) break;
  }
  if (out_iters) *out_iters = host.iter;
  if (out_rr) *out_rr = host.rr;
  if (out_bb) *out_bb = host.bb;
  return IGA_OK;
}


static CGBufs cg_bufs(double* work, int32_t n) {
  CGBufs B;
  B.st = reinterpret_cast<CGState*>(work);
  double* parts = work + 8 + 4 * (int64_t)n;
  B.P_pq = parts;
  B.P_rz = parts + kRedBlocks;
  B.P_rr = parts + 3 * kRedBlocks;
  B.P_bb = parts + 4 * kRedBlocks;
  return B;
}

// blocks of the SpMV (kSpmvThreads each, at most kRedBlocks partials)
static int spmv_blocks(int32_t n, int lpr) {
  const int64_t b = ((int64_t)n * lpr + kSpmvThreads - 1) / kSpmvThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kRedBlocks));
}

// Residual replacement for CG on an inexact operator: r = b - A x with the
// exact (FP64) matrix, z = D^-1 r, and the r.z and r.r partials into the
// given r.z slot (one partial per block; launched with the vector-kernel
// block count, rows strided over the grid). `force` runs it after the loop
// has stopped.
template <int LPR>
__global__ void __launch_bounds__(kVecThreads)
    k_cg_replace(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
                 const double* __restrict__ vals, const double* __restrict__ b,
                 const double* __restrict__ dinv, const double* __restrict__ x, double* r,
                 double* z, CGBufs B, int slot, int force) {
  if (B.st->done && !force) return;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; i < n; i += groups) {
    const double s = row_dot<LPR>(i, indptr, cols, vals, x, l);
    if (l == 0) {
      const double ri = b[i] - s;
      const double zi = dinv[i] * ri;
      r[i] = ri;
      z[i] = zi;
      rz += ri * zi;
      rr += ri * ri;
    }
  }
  store_partial(B.P_rz + slot * kRedBlocks, rz);
  store_partial(B.P_rr, rr);
}

// After a forced replacement: the true residual decides whether to stop.
__global__ void k_cg_recheck(int nblk, CGBufs B) {
  const double rr = reduce_partials(B.P_rr, nblk);
  if (threadIdx.x == 0) {
    CGState* st = B.st;
    st->rr = rr;
    st->done = (!(rr > st->tol2 * st->bb) || st->iter >= st->maxiter) ? 1 : 0;
  }
}

// When the check fails, the direction restarts from the replaced residual:
// the recursion had run past the true residual, so the old direction is
// orders of magnitude smaller than the new r and no longer pairs with it.
__global__ void __launch_bounds__(kVecThreads)
    k_cg_restart(int n, const double* __restrict__ z, double* p, CGBufs B) {
  if (B.st->done) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = z[i];
}

// Jacobi-PCG loop; `apply(p, q, B)` must produce q = A p and the p.q
// partials (returns the number of partial blocks, or a negative status).
struct NoReplace {
  int every = 0;
  int operator()(const double*, double*, double*, CGBufs, int, int) const { return IGA_OK; }
};

// `replace(x, r, z, B, slot, force)`, when `replace.every` > 0, swaps the
// recursive residual for the true one every `every` iterations and once more
// when the recursion reports convergence (reliable updates: the search
// direction is kept, only r and z are refreshed).
template <typename Apply, typename Replace = NoReplace>
static int cg_loop(int32_t n, const double* dinv, const double* b, double* x, double tol,
                   int32_t maxiter, double* work, int32_t* out_iters, double* out_rr,
                   double* out_bb, hipStream_t st, Apply apply, Replace replace = Replace{},
                   bool queued = false) {
  if (n == 0) {
    if (out_iters) *out_iters = 0;
    if (out_rr) *out_rr = 0.0;
    if (out_bb) *out_bb = 0.0;
    return IGA_OK;
  }
  CGBufs B = cg_bufs(work, n);
  double* r = work + 8;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  const int vb = vec_blocks(n);
  hipLaunchKernelGGL(k_cg_init, dim3(vb), dim3(kVecThreads), 0, st, n, b, dinv, x, r, z, p, B);
  hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(kVecThreads), 0, st, vb, B, tol * tol, maxiter);
  IGA_CHECK_LAUNCH();
  if (queued && replace.every <= 0) {
    // one batch queued ahead of the state the host reads (see cg1_loop);
    // after convergence the batch in flight is early-exit launches
    StatePipe* pipe = nullptr;
    if (int rc = state_pipe(&pipe)) return rc;
    int launched = 0;
    auto enqueue = [&](int slot) -> int {
      for (int k = 0; k < 8; ++k, ++launched) {
        const int par = launched & 1;
        const int pb = apply(p, q, B);
        if (pb < 0) return pb;
        hipLaunchKernelGGL(k_cg_update, dim3(vb), dim3(kVecThreads), 0, st, n, pb, vb, par, dinv,
                           p, q, x, r, z, B);
        hipLaunchKernelGGL(k_cg_dir, dim3(vb), dim3(kVecThreads), 0, st, n, vb, par, z, p, B);
        IGA_CHECK_LAUNCH();
      }
      if (hipMemcpyAsync(&pipe->host[slot], B.st, sizeof(CGState), hipMemcpyDeviceToHost, st) !=
              hipSuccess ||
          hipEventRecord(pipe->ev[slot], st) != hipSuccess)
        return IGA_ERR_HIP;
      return IGA_OK;
    };
    if (hipMemcpyAsync(&pipe->host[0], B.st, sizeof(CGState), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipEventRecord(pipe->ev[0], st) != hipSuccess)
      return IGA_ERR_HIP;
    CGState host{};
    for (int k = 0;; ++k) {
      if (int rc = enqueue((k + 1) & 1)) return rc;
      if (hipEventSynchronize(pipe->ev[k & 1]) != hipSuccess) return IGA_ERR_HIP;
      host = pipe->host[k & 1];
      if (host.done) break;
    }
    if (out_iters) *out_iters = host.iter;
    if (out_rr) *out_rr = host.rr;
    if (out_bb) *out_bb = host.bb;
    return IGA_OK;
  }
  CGState host{};
  const int batch = 16;
  int launched = 0;
  bool verified = false;
  for (;;) {
    if (hipMemcpyAsync(&host, B.st, sizeof(CGState), hipMemcpyDeviceToHost, st) != hipSuccess)
      return IGA_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
    if (host.done) {
      if (replace.every <= 0 || verified || host.iter >= maxiter) break;
      // the recursion says converged: check with the true residual
      if (int rc = replace(x, r, z, B, launched & 1, 1)) return rc;
      hipLaunchKernelGGL(k_cg_recheck, dim3(1), dim3(kVecThreads), 0, st, vb, B);
      hipLaunchKernelGGL(k_cg_restart, dim3(vb), dim3(kVecThreads), 0, st, n, z, p, B);
      IGA_CHECK_LAUNCH();
      verified = true;
      continue;
    }
    verified = false;
    for (int k = 0; k < batch; ++k, ++launched) {
      const int par = launched & 1;
      const int pb = apply(p, q, B);
      if (pb < 0) return pb;
      hipLaunchKernelGGL(k_cg_update, dim3(vb), dim3(kVecThreads), 0, st, n, pb, vb, par, dinv, p,
                         q, x, r, z, B);
      if (replace.every > 0 && (launched + 1) % replace.every == 0)
        if (int rc = replace(x, r, z, B, 1 - par, 0)) return rc;
      hipLaunchKernelGGL(k_cg_dir, dim3(vb), dim3(kVecThreads), 0, st, n, vb, par, z, p, B);
      IGA_CHECK_LAUNCH();
    }
  }
  if (out_iters) *out_iters = host.iter;
  if (out_rr) *out_rr = host.rr;
  if (out_bb) *out_bb = host.bb;
  return IGA_OK;
}

// one CSR operator application on the device-resident state: q = A p and
// the p.q partials into P (returns the number of partial blocks)
template <typename TV>
static int csr_apply(int32_t n, int lpr, int rb, const int32_t* indptr, const int32_t* cols,
                     const TV* vals, const double* p, double* q, double* P, CGState* state,
                     hipStream_t st) {
  CGBufs V{};
  V.st = state;
  V.P_pq = P;
  int rc = by_lpr(lpr, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    hipLaunchKernelGGL((k_cg_spmv<LPR, TV>), dim3(rb), dim3(kSpmvThreads), 0, st, n, indptr, cols,
                       vals, p, q, V);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
  return rc ? rc : rb;
}

// `single`: the single-reduction loop (two kernels per iteration), else the
// textbook three-kernel loop
static int cg_solve_csr(int32_t n, const int32_t* indptr, const int32_t* cols, const void* vals,
                        int32_t precision, const double* dinv, const double* b, double* x,
                        double tol, int32_t maxiter, double* work, int32_t* out_iters,
                        double* out_rr, double* out_bb, hipStream_t st, bool single = false,
                        bool queued = false) {
  int nnz = 0;
  if (n > 0)
    if (int rc = nnz_of(n, indptr, st, &nnz)) return rc;
  const int lpr = pick_lpr(n, nnz);
  const int rb = spmv_blocks(n, lpr);
  auto op = [&](const double* p, double* q, double* P, CGState* state) -> int {
    if (precision == IGA_FP32)
      return csr_apply(n, lpr, rb, indptr, cols, reinterpret_cast<const float*>(vals), p, q, P,
                       state, st);
    return csr_apply(n, lpr, rb, indptr, cols, reinterpret_cast<const double*>(vals), p, q, P, state,
                     st);
  };
  if (single)
    return cg1_loop(n, dinv, b, x, tol, maxiter, work, out_iters, out_rr, out_bb, st, op);
  return cg_loop(n, dinv, b, x, tol, maxiter, work, out_iters, out_rr, out_bb, st,
                 [&](const double* p, double* q, CGBufs B) { return op(p, q, B.P_pq, B.st); },
                 NoReplace{}, queued);
}

extern "C" int iga_cg_solve(int32_t n, const int32_t* indptr, const int32_t* cols,
                            const void* vals, int32_t precision, const double* dinv,
                            const double* b, double* x, double tol, int32_t maxiter, double* work,
                            int32_t* out_iters, double* out_rr, double* out_bb, void* stream) {
  if (!indptr || !cols || !vals || !dinv || !b || !x || !work) return IGA_ERR_ARG;
  if (precision != IGA_FP64 && precision != IGA_FP32) return IGA_ERR_UNSUPPORTED;
  return cg_solve_csr(n, indptr, cols, vals, precision, dinv, b, x, tol, maxiter, work, out_iters,
                      out_rr, out_bb, (hipStream_t)stream, false, true);
}

extern "C" int iga_cg1_solve(int32_t n, const int32_t* indptr, const int32_t* cols,
                             const void* vals, int32_t precision, const double* dinv,
                             const double* b, double* x, double tol, int32_t maxiter, double* work,
                             int32_t* out_iters, double* out_rr, double* out_bb, void* stream) {
  if (!indptr || !cols || !vals || !dinv || !b || !x || !work) return IGA_ERR_ARG;
  if (precision != IGA_FP64 && precision != IGA_FP32) return IGA_ERR_UNSUPPORTED;
  return cg_solve_csr(n, indptr, cols, vals, precision, dinv, b, x, tol, maxiter, work, out_iters,
                      out_rr, out_bb, (hipStream_t)stream, true);
}

// ---------------------------------------------------------------------------
// The same Jacobi-PCG for callers that cannot allocate: its workspace (the
// solver state, r, z, p, q and the reduction partials) is the file-scope
// device array `scratch`, one copy per GPU, held until the process ends.
// Host code never passes the array's name (a host-side stand-in, not a
// device address) to a kernel: the first call on each device looks up the
// device address with hipGetSymbolAddress and keeps it per device.
// ---------------------------------------------------------------------------

constexpr int64_t kNoAllocDoubles = int64_t(1) << 21;  // 16 MiB per GPU
static_assert(kNoAllocDoubles >= 16 + 5 * int64_t(IGA_CG_NOALLOC_MAX_N) + 6 * kRedBlocks,
              "scratch too small for IGA_CG_NOALLOC_MAX_N (see iga_cg_workspace)");

static __device__ __align__(256) double scratch[kNoAllocDoubles];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[256] = {};

static int scratch_for_current_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return IGA_ERR_HIP;
  if (dev < 0 || dev >= 256) return IGA_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return IGA_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return IGA_OK;
}

extern "C" int iga_cg_solve_noalloc(int32_t n, const int32_t* indptr, const int32_t* cols,
                                    const void* vals, int32_t precision, const double* dinv,
                                    const double* b, double* x, double tol, int32_t maxiter,
                                    int32_t* out_iters, double* out_rr, double* out_bb,
                                    void* stream) {
  if (!indptr || !cols || !vals || !dinv || !b || !x) return IGA_ERR_ARG;
  if (precision != IGA_FP64 && precision != IGA_FP32) return IGA_ERR_UNSUPPORTED;
  if (n < 0 || n > IGA_CG_NOALLOC_MAX_N) return IGA_ERR_SHAPE;
  double* work = nullptr;
  if (int rc = scratch_for_current_device(&work)) return rc;
  // every field the loop reads (state, r, z, p, x, partials) is written by
  // its initialisation kernels, so the previous call's contents need no
  // clearing
  return cg_solve_csr(n, indptr, cols, vals, precision, dinv, b, x, tol, maxiter, work, out_iters,
                      out_rr, out_bb, (hipStream_t)stream);
}

// CG on an FP32 matrix with FP64 vectors and recurrences, and reliable
// updates: every `replace_every` iterations, and when the recursion reports
// convergence, the residual is recomputed with the FP64 matrix. The stopping
// test therefore holds for the FP64 operator.
struct CsrReplace {
  int every;
  int32_t n, lpr, vb;
  const int32_t* indptr;
  const int32_t* cols;
  const double* vals64;
  const double* b;
  const double* dinv;
  hipStream_t st;
  int operator()(const double* x, double* r, double* z, CGBufs B, int slot, int force) const {
    return by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL((k_cg_replace<LPR>), dim3(vb), dim3(kVecThreads), 0, st, n, indptr, cols,
                         vals64, b, dinv, x, r, z, B, slot, force);
      IGA_CHECK_LAUNCH();
      return (int)IGA_OK;
    });
  }
};

extern "C" int iga_cg_solve_mixed(int32_t n, const int32_t* indptr, const int32_t* cols,
                                  const float* vals32, const double* vals64, const double* dinv,
                                  const double* b, double* x, double tol, int32_t maxiter,
                                  int32_t replace_every, double* work, int32_t* out_iters,
                                  double* out_rr, double* out_bb, void* stream) {
  if (!indptr || !cols || !vals32 || !vals64 || !dinv || !b || !x || !work) return IGA_ERR_ARG;
  if (replace_every < 1) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int nnz = 0;
  if (n > 0)
    if (int rc = nnz_of(n, indptr, st, &nnz)) return rc;
  const int lpr = pick_lpr(n, nnz);
  const int rb = spmv_blocks(n, lpr);
  CsrReplace rep{replace_every, n, lpr, vec_blocks(n), indptr, cols, vals64, b, dinv, st};
  return cg_loop(
      n, dinv, b, x, tol, maxiter, work, out_iters, out_rr, out_bb, st,
      [&](const double* p, double* q, CGBufs B) -> int {
        int rc = by_lpr(lpr, [&](auto L) {
          constexpr int LPR = decltype(L)::value;
          hipLaunchKernelGGL((k_cg_spmv<LPR, float>), dim3(rb), dim3(kSpmvThreads), 0, st, n,
                             indptr, cols, vals32, p, q, B);
          IGA_CHECK_LAUNCH();
          return (int)IGA_OK;
        });
        return rc ? rc : rb;
      },
      rep);
}

extern "C" int iga_cg_solve_mf(const iga_patch_t* patch, const void* mf_work,
                               int64_t mf_work_bytes, int32_t n_free,
                               const int32_t* freed, const int32_t* full_to_free,
                               const double* dinv, const double* b, double* x, double tol,
                               int32_t maxiter, double* work, int32_t* out_iters, double* out_rr,
                               double* out_bb, void* stream) {
  if (!patch || !mf_work || !dinv || !b || !x || !work) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  // textbook loop; the p.q partials come out of the operator's gather kernel
  return cg_loop(n_free, dinv, b, x, tol, maxiter, work, out_iters, out_rr, out_bb, st,
                 [&](const double* p, double* q, CGBufs B) -> int {
                   int nb = 0;
                   int rc = mf_apply_launch(patch, mf_work, mf_work_bytes, n_free, freed,
                                            full_to_free, p, q, &B.st->done, B.P_pq, &nb, st);
                   return rc ? rc : nb;
                 },
                 NoReplace{}, true);
}

extern "C" int iga_residual_norms(int32_t n, const int32_t* indptr, const int32_t* cols,
                                  const double* vals, const double* u, const double* b,
                                  double* out, double* work, void* stream) {
  if (!indptr || !cols || !vals || !u || !b || !out || !work) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned int*>(out), 4);
    IGA_CHECK_LAUNCH();
    return IGA_OK;
  }
  int nnz = 0;
  if (int rc = nnz_of(n, indptr, st, &nnz)) return rc;
  double* P_rr = work;
  double* P_bb = work + kRedBlocks;
  const int lpr = pick_lpr(n, nnz);
  const int rb = row_blocks(n, lpr);
  int rc = by_lpr(lpr, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    hipLaunchKernelGGL(k_residual<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols, vals,
                       u, b, P_rr, P_bb);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
  if (rc) return rc;
  hipLaunchKernelGGL(k_residual_fin, dim3(1), dim3(kVecThreads), 0, st, rb, P_rr, P_bb, out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

__global__ void __launch_bounds__(kVecThreads)
    k_diff_norms(int n, const double* __restrict__ y, const double* __restrict__ b, double* P_rr,
                 double* P_bb) {
  double rr = 0.0, bb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double d = y[i] - b[i];
    rr += d * d;
    bb += b[i] * b[i];
  }
  store_partial(P_rr, rr);
  store_partial(P_bb, bb);
}

__global__ void k_inv_guard(int n, const double* __restrict__ d, double* inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[i] = d[i] != 0.0 ? 1.0 / d[i] : 1.0;
}

extern "C" int iga_diff_norms(int32_t n, const double* y, const double* b, double* out,
                              double* work, void* stream) {
  if (!y || !b || !out || !work) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int vb = vec_blocks(n);
  hipLaunchKernelGGL(k_diff_norms, dim3(vb), dim3(kVecThreads), 0, st, n, y, b, work,
                     work + kRedBlocks);
  hipLaunchKernelGGL(k_residual_fin, dim3(1), dim3(kVecThreads), 0, st, vb, work,
                     work + kRedBlocks, out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_inv_guard(int32_t n, const double* d, double* inv, void* stream) {
  if (!d || !inv) return IGA_ERR_ARG;
  if (n == 0) return IGA_OK;
  hipLaunchKernelGGL(k_inv_guard, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, d,
                     inv);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

template <int LPR>
__global__ void __launch_bounds__(kVecThreads)
    k_residual_vec(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
                   const double* __restrict__ vals, const double* __restrict__ u,
                   const double* __restrict__ b, double* rvec, double* P_rr, double* P_bb) {
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  double rr = 0.0, bb = 0.0;
  for (int r = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; r < n; r += groups) {
    const double s = row_dot<LPR>(r, indptr, cols, vals, u, l);
    if (l == 0) {
      const double d = b[r] - s;
      rvec[r] = d;
      rr += d * d;
      bb += b[r] * b[r];
    }
  }
  store_partial(P_rr, rr);
  store_partial(P_bb, bb);
}

__global__ void k_axpy(int n, double a, const double* __restrict__ x, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

extern "C" int iga_residual_vec(int32_t n, const int32_t* indptr, const int32_t* cols,
                                const double* vals, const double* u, const double* b, double* r,
                                double* out, double* work, void* stream) {
  if (!indptr || !cols || !vals || !u || !b || !r || !out || !work) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) return IGA_OK;
  int nnz = 0;
  if (int rc = nnz_of(n, indptr, st, &nnz)) return rc;
  const int lpr = pick_lpr(n, nnz);
  const int rb = row_blocks(n, lpr);
  int rc = by_lpr(lpr, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    hipLaunchKernelGGL(k_residual_vec<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols,
                       vals, u, b, r, work, work + kRedBlocks);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
  if (rc) return rc;
  hipLaunchKernelGGL(k_residual_fin, dim3(1), dim3(kVecThreads), 0, st, rb, work, work + kRedBlocks,
                     out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_axpy(int32_t n, double a, const double* x, double* y, void* stream) {
  if (!x || !y) return IGA_ERR_ARG;
  if (n == 0) return IGA_OK;
  hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, a, x, y);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_rff_eval(int32_t n, int32_t dim, const double* pts, int32_t m,
                            const double* omega, const double* phase, double amp, double* out,
                            void* stream) {
  if (!pts || !omega || !phase || !out || dim < 1 || dim > 3) return IGA_ERR_ARG;
  if (n == 0) return IGA_OK;
  hipLaunchKernelGGL(k_rff, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, dim, pts,
                     m, omega, phase, amp, out);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}
