// Batched Jacobi-PCG over many independent operators in one persistent launch.
//
// One workgroup owns one instance for the whole solve: its vectors live in
// HBM/L2, every dot product is a workgroup reduction, and the iteration loop
// runs inside the kernel until the instance converges (||r|| <= tol ||b||,
// FP64 recurrences, the accelerator semantics of jax cg) or reaches maxiter.
// When the recursion stops, the true residual b - A x is recomputed in FP64;
// above tol the instance restarts from it (at most three times, sharing
// maxiter), so the reported relative residual is the true one.
// No grid barrier and no host round trip are needed, so a whole batch of
// uneven instances is one launch that a hipGraph can replay. Instances are
// issued largest first (the order array) so the longest solves start early.
#include <mutex>

#include "common.hpp"

namespace iga {

struct BatchInst {
  int32_t n;
  int32_t pad;
  const int32_t* indptr;
  const int32_t* cols;
  const double* vals;
  const double* dinv;
  const double* b;
  double* x;
  double* work;  // r, z, p, q (4n)
};

__device__ __forceinline__ double block_allsum(double v, double* scratch, double* bcast) {
  const double s = block_sum(v, scratch);
  if (threadIdx.x == 0) *bcast = s;
  __syncthreads();
  const double r = *bcast;
  __syncthreads();
  return r;
}

template <int LPR>
__global__ void __launch_bounds__(1024)
    k_batched_cg(const BatchInst* __restrict__ insts, const int32_t* __restrict__ order,
                 double tol, int maxiter, int32_t* out_iters, double* out_rel) {
  __shared__ double scratch[16];
  __shared__ double bcast;
  const int id = order[blockIdx.x];
  const BatchInst I = insts[id];
  const int n = I.n;
  double* r = I.work;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  double rz = 0.0, bb = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double bi = I.b[i];
    const double zi = I.dinv[i] * bi;
    I.x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    p[i] = zi;
    rz += bi * zi;
    bb += bi * bi;
  }
  rz = block_allsum(rz, scratch, &bcast);
  bb = block_allsum(bb, scratch, &bcast);
  const double tol2bb = tol * tol * bb;
  double rr = bb;
  int it = 0;
  const int lane = threadIdx.x % LPR;
  const int groups = blockDim.x / LPR;
  double rr_true = bb;
  for (int restarts = 0;; ++restarts) {
    while (rr > tol2bb && it < maxiter) {
      double pq = 0.0;
      for (int row = threadIdx.x / LPR; row < n; row += groups) {
        const double s = row_dot<LPR>(row, I.indptr, I.cols, I.vals, p, lane);
        if (lane == 0) {
          q[row] = s;
          pq += s * p[row];
        }
      }
      pq = block_allsum(pq, scratch, &bcast);  // also orders the q writes before their reads
      const double alpha = rz / pq;
      double rzn = 0.0, rrn = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        I.x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        const double zi = I.dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        rzn += ri * zi;
        rrn += ri * ri;
      }
      rzn = block_allsum(rzn, scratch, &bcast);
      rr = block_allsum(rrn, scratch, &bcast);
      const double beta = rzn / rz;
      rz = rzn;
      for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = z[i] + beta * p[i];
      __syncthreads();
      ++it;
    }
    // the recursion has stopped: recompute r = b - A x in FP64 and restart
    // from it (r, z = D^-1 r, p = z) while the true residual is above tol,
    // at most three times within maxiter
    double rrt = 0.0, rzt = 0.0;
    for (int row = threadIdx.x / LPR; row < n; row += groups) {
      const double s = row_dot<LPR>(row, I.indptr, I.cols, I.vals, I.x, lane);
      if (lane == 0) {
        const double ri = I.b[row] - s;
        const double zi = I.dinv[row] * ri;
        r[row] = ri;
        z[row] = zi;
        p[row] = zi;
        rrt += ri * ri;
        rzt += ri * zi;
      }
    }
    rr_true = block_allsum(rrt, scratch, &bcast);
    rzt = block_allsum(rzt, scratch, &bcast);
    if (rr_true <= tol2bb || it >= maxiter || restarts >= 3) break;
    rr = rr_true;
    rz = rzt;
  }
  if (threadIdx.x == 0) {
    out_iters[id] = it;
    out_rel[id] = bb > 0.0 ? sqrt(rr_true / bb) : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Segmented batched Jacobi-PCG. The rows of every instance are cut into
// blocks of at most `rpb` rows (one workgroup each); every kernel runs over
// the blocks of all instances at once, so a large instance is spread over
// the whole GPU instead of one compute unit, and each iteration is three
// launches for the whole batch. Per-instance scalars are reduced from the
// per-block partials of the instance in a fixed order by every block that
// needs them (bitwise reproducible, no inter-block wait). A converged
// instance's blocks return at once. rz is double-buffered by the parity of
// the global iteration count k: the direction kernel reads slot k & 1 and
// the instance's first block writes slot (k + 1) & 1.

struct SegInst {
  double bb, tol2, rz[2], rr, pad0;
  int32_t it, done, restarts, pad1;
};
static_assert(sizeof(SegInst) == 64, "per-instance state is 64 bytes");

struct SegArgs {
  const BatchInst* insts;
  const int32_t* blk_inst;  // instance of each block
  const int32_t* blk_row0;  // first row of each block (within its instance)
  const int32_t* inst_blk0;  // first block of each instance, n_inst + 1 entries
  SegInst* st;
  double* part;  // [3][n_blocks]: p.q, r.z, r.r partials
  int32_t* active;  // [0] instances still iterating, [1] their blocks
  int32_t* blist;   // the blocks of the instances still iterating
  int32_t n_blocks, rpb;
};

// Sum of the instance's block partials part[b0 .. b1) in a fixed order,
// valid in every thread of the block (one wave reduces, LDS broadcasts).
__device__ __forceinline__ double seg_sum(const double* part, int b0, int b1, double* bc) {
  if (threadIdx.x < kWave) {
    double v = 0.0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += kWave) v += part[b];
    v = group_sum<kWave>(v);
    if (threadIdx.x == 0) *bc = v;
  }
  __syncthreads();
  const double r = *bc;
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_seg_init(SegArgs A) {
  __shared__ double scratch[16];
  const int b = blockIdx.x;
  const BatchInst I = A.insts[A.blk_inst[b]];
  const int r0 = A.blk_row0[b], r1 = min(I.n, r0 + A.rpb);
  double* r = I.work;
  double* z = r + I.n;
  double* p = z + I.n;
  double rz = 0.0, bb = 0.0;
  for (int i = r0 + (int)threadIdx.x; i < r1; i += blockDim.x) {
    const double bi = I.b[i], zi = I.dinv[i] * bi;
    I.x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    p[i] = zi;
    rz += bi * zi;
    bb += bi * bi;
  }
  rz = block_sum(rz, scratch);
  bb = block_sum(bb, scratch);
  if (threadIdx.x == 0) {
    A.part[A.n_blocks + b] = rz;
    A.part[2 * A.n_blocks + b] = bb;
  }
}

// One thread per instance: the initial state from the partials.
__global__ void __launch_bounds__(256)
    k_seg_state(SegArgs A, int n_inst, double tol, int maxiter) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  const int b0 = A.inst_blk0[i], b1 = A.inst_blk0[i + 1];
  double rz = 0.0, bb = 0.0;
  for (int b = b0; b < b1; ++b) {
    rz += A.part[A.n_blocks + b];
    bb += A.part[2 * A.n_blocks + b];
  }
  SegInst& S = A.st[i];
  S.bb = bb;
  S.tol2 = tol * tol * bb;
  S.rz[0] = rz;
  S.rz[1] = rz;
  S.rr = bb;
  S.it = 0;
  S.restarts = 0;
  S.done = (bb <= S.tol2 || maxiter <= 0) ? 1 : 0;
}

template <int LPR>
__global__ void __launch_bounds__(256) k_seg_spmv(SegArgs A) {
  __shared__ double scratch[16];
  const int b = A.blist[blockIdx.x], inst = A.blk_inst[b];
  if (A.st[inst].done) return;
  const BatchInst I = A.insts[inst];
  const int r0 = A.blk_row0[b], r1 = min(I.n, r0 + A.rpb);
  const double* p = I.work + 2 * I.n;
  double* q = I.work + 3 * I.n;
  const int lane = threadIdx.x % LPR, groups = blockDim.x / LPR;
  double pq = 0.0;
  for (int row = r0 + (int)threadIdx.x / LPR; row < r1; row += groups) {
    const double s = row_dot<LPR>(row, I.indptr, I.cols, I.vals, p, lane);
    if (lane == 0) {
      q[row] = s;
      pq += s * p[row];
    }
  }
  pq = block_sum(pq, scratch);
  if (threadIdx.x == 0) A.part[b] = pq;
}

__global__ void __launch_bounds__(256) k_seg_update(SegArgs A, int k) {
  __shared__ double scratch[16];
  __shared__ double bc;
  const int b = A.blist[blockIdx.x], inst = A.blk_inst[b];
  const SegInst& S = A.st[inst];
  if (S.done) return;
  const double pq = seg_sum(A.part, A.inst_blk0[inst], A.inst_blk0[inst + 1], &bc);
  const double alpha = S.rz[k & 1] / pq;
  const BatchInst I = A.insts[inst];
  const int r0 = A.blk_row0[b], r1 = min(I.n, r0 + A.rpb);
  double* r = I.work;
  double* z = r + I.n;
  const double* p = z + I.n;
  const double* q = p + I.n;
  double rz = 0.0, rr = 0.0;
  for (int i = r0 + (int)threadIdx.x; i < r1; i += blockDim.x) {
    I.x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    const double zi = I.dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = block_sum(rz, scratch);
  rr = block_sum(rr, scratch);
  if (threadIdx.x == 0) {
    A.part[A.n_blocks + b] = rz;
    A.part[2 * A.n_blocks + b] = rr;
  }
}

__global__ void __launch_bounds__(256) k_seg_dir(SegArgs A, int k, int maxiter) {
  __shared__ double bc;
  const int b = A.blist[blockIdx.x], inst = A.blk_inst[b];
  SegInst& S = A.st[inst];
  if (S.done) return;
  const int b0 = A.inst_blk0[inst], b1 = A.inst_blk0[inst + 1];
  const double rzn = seg_sum(A.part + A.n_blocks, b0, b1, &bc);
  const double rr = seg_sum(A.part + 2 * A.n_blocks, b0, b1, &bc);
  const double beta = rzn / S.rz[k & 1];
  const BatchInst I = A.insts[inst];
  const int r0 = A.blk_row0[b], r1 = min(I.n, r0 + A.rpb);
  const double* z = I.work + I.n;
  double* p = I.work + 2 * I.n;
  for (int i = r0 + (int)threadIdx.x; i < r1; i += blockDim.x) p[i] = z[i] + beta * p[i];
  // the instance's state, once: every block of the instance has read
  // rz[k & 1] and done above (this kernel writes only the other slot; done
  // is read again only by the next launches)
  if (b == b0 && threadIdx.x == 0) {
    S.rz[(k + 1) & 1] = rzn;
    S.rr = rr;
    S.it += 1;
    if (rr <= S.tol2 || S.it >= maxiter) S.done = 1;
  }
}

// The instances still iterating and the list of their blocks, in instance
// order (one workgroup of 1024 threads, an exclusive scan per 1024
// instances): active[0] = instances, active[1] = blocks.
__global__ void __launch_bounds__(1024) k_seg_compact(SegArgs A, int n_inst) {
  __shared__ int32_t cnt[1024];
  __shared__ int32_t base, ninst;
  if (threadIdx.x == 0) {
    base = 0;
    ninst = 0;
  }
  __syncthreads();
  for (int i0 = 0; i0 < n_inst; i0 += 1024) {
    const int i = i0 + threadIdx.x;
    const bool live = i < n_inst && !A.st[i].done;
    const int b0 = live ? A.inst_blk0[i] : 0, nb = live ? A.inst_blk0[i + 1] - b0 : 0;
    cnt[threadIdx.x] = nb;
    __syncthreads();
    // inclusive scan (Hillis-Steele) of the block counts
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = threadIdx.x >= off ? cnt[threadIdx.x - off] : 0;
      __syncthreads();
      cnt[threadIdx.x] += v;
      __syncthreads();
    }
    const int start = base + cnt[threadIdx.x] - nb;
    for (int c = 0; c < nb; ++c) A.blist[start + c] = b0 + c;
    if (live) atomicAdd(&ninst, 1);
    __syncthreads();
    if (threadIdx.x == 0) base += cnt[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    A.active[0] = ninst;
    A.active[1] = base;
  }
}

// True residual r = b - A x (FP64) with z = D^-1 r, p = z, and its partials.
template <int LPR>
__global__ void __launch_bounds__(256) k_seg_resid(SegArgs A) {
  __shared__ double scratch[16];
  const int b = blockIdx.x;
  const BatchInst I = A.insts[A.blk_inst[b]];
  const int r0 = A.blk_row0[b], r1 = min(I.n, r0 + A.rpb);
  double* r = I.work;
  double* z = r + I.n;
  double* p = z + I.n;
  const int lane = threadIdx.x % LPR, groups = blockDim.x / LPR;
  double rr = 0.0, rz = 0.0;
  for (int row = r0 + (int)threadIdx.x / LPR; row < r1; row += groups) {
    const double s = row_dot<LPR>(row, I.indptr, I.cols, I.vals, I.x, lane);
    if (lane == 0) {
      const double ri = I.b[row] - s, zi = I.dinv[row] * ri;
      r[row] = ri;
      z[row] = zi;
      p[row] = zi;
      rr += ri * ri;
      rz += ri * zi;
    }
  }
  rr = block_sum(rr, scratch);
  rz = block_sum(rz, scratch);
  if (threadIdx.x == 0) {
    A.part[A.n_blocks + b] = rz;
    A.part[2 * A.n_blocks + b] = rr;
  }
}

// One thread per instance after a true-residual pass: restart the instance
// from it while it is above tol (at most three restarts, within maxiter);
// otherwise report. Converged instances report again (same values).
__global__ void __launch_bounds__(256)
    k_seg_restart(SegArgs A, int n_inst, int k, int maxiter, int32_t* out_iters,
                  double* out_rel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  SegInst& S = A.st[i];
  const int b0 = A.inst_blk0[i], b1 = A.inst_blk0[i + 1];
  double rz = 0.0, rr = 0.0;
  for (int b = b0; b < b1; ++b) {
    rz += A.part[A.n_blocks + b];
    rr += A.part[2 * A.n_blocks + b];
  }
  if (S.done == 1 && rr > S.tol2 && S.it < maxiter && S.restarts < 3) {
    S.restarts += 1;
    S.rz[k & 1] = rz;
    S.rr = rr;
    S.done = 0;
  } else {
    S.done = 1;
  }
  out_iters[i] = S.it;
  out_rel[i] = S.bb > 0.0 ? sqrt(rr / S.bb) : 0.0;
}

}  // namespace iga

using namespace iga;

extern "C" int32_t iga_batched_cg_seg_plan(int32_t n_inst, const int32_t* n_rows, int32_t rpb,
                                           int32_t* blk_inst, int32_t* blk_row0,
                                           int32_t* inst_blk0) {
  if (n_inst < 0 || !n_rows || rpb <= 0) return IGA_ERR_ARG;
  int64_t nb = 0;
  for (int32_t i = 0; i < n_inst; ++i) {
    if (n_rows[i] < 0) return IGA_ERR_ARG;
    const int32_t cnt = n_rows[i] > 0 ? (n_rows[i] + rpb - 1) / rpb : 1;
    if (inst_blk0) inst_blk0[i] = (int32_t)nb;
    for (int32_t c = 0; c < cnt; ++c) {
      if (blk_inst) blk_inst[nb] = i;
      if (blk_row0) blk_row0[nb] = c * rpb;
      ++nb;
    }
    if (nb > INT32_MAX / 4) return IGA_ERR_SHAPE;
  }
  if (inst_blk0) inst_blk0[n_inst] = (int32_t)nb;
  return (int32_t)nb;
}

extern "C" int64_t iga_batched_cg_seg_workspace(int32_t n_inst, int32_t n_blocks) {
  if (n_inst < 0 || n_blocks < 0) return -1;
  return (int64_t)n_inst * (int64_t)sizeof(SegInst) + 3 * (int64_t)n_blocks * 8 + 64 +
         (int64_t)n_blocks * 4;
}

extern "C" int iga_batched_cg_seg(int32_t n_inst, const void* insts, int32_t n_blocks,
                                  const int32_t* blk_inst, const int32_t* blk_row0,
                                  const int32_t* inst_blk0, int32_t rows_per_block,
                                  int32_t lanes_per_row, double tol, int32_t maxiter, void* work,
                                  int64_t work_bytes, int32_t* out_iters, double* out_rel,
                                  void* stream) {
  if (n_inst == 0) return IGA_OK;
  if (n_inst < 0 || !insts || !blk_inst || !blk_row0 || !inst_blk0 || !work || !out_iters ||
      !out_rel || n_blocks < n_inst || rows_per_block <= 0)
    return IGA_ERR_ARG;
  if (work_bytes < iga_batched_cg_seg_workspace(n_inst, n_blocks)) return IGA_ERR_SHAPE;
  if (reinterpret_cast<uintptr_t>(work) % 64) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  SegArgs A;
  A.insts = reinterpret_cast<const BatchInst*>(insts);
  A.blk_inst = blk_inst;
  A.blk_row0 = blk_row0;
  A.inst_blk0 = inst_blk0;
  A.st = reinterpret_cast<SegInst*>(work);
  A.part = reinterpret_cast<double*>(reinterpret_cast<char*>(work) + (size_t)n_inst * sizeof(SegInst));
  A.active = reinterpret_cast<int32_t*>(A.part + 3 * (size_t)n_blocks);
  A.blist = A.active + 16;
  A.n_blocks = n_blocks;
  A.rpb = rows_per_block;
  // host copy of the active count, pinned once per process; concurrent
  // calls are serialised on it
  static int32_t* host_active = nullptr;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!host_active && hipHostMalloc((void**)&host_active, 64, hipHostMallocDefault) != hipSuccess)
    return IGA_ERR_HIP;
  const dim3 gb((unsigned)n_blocks), gi((unsigned)((n_inst + 255) / 256));
  return by_lpr(lanes_per_row, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    // the active instances and their blocks, read back by the host
    auto compact = [&]() -> int {
      hipLaunchKernelGGL(k_seg_compact, dim3(1), dim3(1024), 0, st, A, n_inst);
      IGA_CHECK_LAUNCH();
      if (hipMemcpyAsync(host_active, A.active, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st) !=
              hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return (int)IGA_ERR_HIP;
      return (int)IGA_OK;
    };
    hipLaunchKernelGGL(k_seg_init, gb, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_seg_state, gi, dim3(256), 0, st, A, n_inst, tol, maxiter);
    IGA_CHECK_LAUNCH();
    constexpr int kChunk = 16;  // iterations enqueued between host checks
    int k = 0;
    for (;;) {
      // iterate until every instance has met tol on its recursion or maxiter;
      // each group of iterations launches only the blocks of the instances
      // that were still iterating at its start
      for (;;) {
        if (int rc = compact()) return rc;
        if (host_active[0] == 0) break;
        const dim3 ga((unsigned)host_active[1]);
        for (int c = 0; c < kChunk; ++c, ++k) {
          hipLaunchKernelGGL(k_seg_spmv<LPR>, ga, dim3(256), 0, st, A);
          hipLaunchKernelGGL(k_seg_update, ga, dim3(256), 0, st, A, k);
          hipLaunchKernelGGL(k_seg_dir, ga, dim3(256), 0, st, A, k, maxiter);
        }
        IGA_CHECK_LAUNCH();
      }
      // true residuals; restart the instances above tol
      hipLaunchKernelGGL(k_seg_resid<LPR>, gb, dim3(256), 0, st, A);
      hipLaunchKernelGGL(k_seg_restart, gi, dim3(256), 0, st, A, n_inst, k, maxiter, out_iters,
                         out_rel);
      IGA_CHECK_LAUNCH();
      if (int rc = compact()) return rc;
      if (host_active[0] == 0) return (int)IGA_OK;
    }
  });
}

extern "C" int iga_batched_cg(int32_t n_inst, const void* insts, const int32_t* order,
                              int32_t lanes_per_row, int32_t threads, double tol, int32_t maxiter,
                              int32_t* out_iters, double* out_rel, void* stream) {
  if (!insts || !order || !out_iters || !out_rel) return IGA_ERR_ARG;
  if (n_inst == 0) return IGA_OK;
  if (threads < 64 || threads > 1024 || threads % 64) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const BatchInst* I = reinterpret_cast<const BatchInst*>(insts);
  return by_lpr(lanes_per_row, [&](auto L) {
    constexpr int LPR = decltype(L)::value;
    hipLaunchKernelGGL(k_batched_cg<LPR>, dim3(n_inst), dim3(threads), 0, st, I, order, tol,
                       maxiter, out_iters, out_rel);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  });
}
