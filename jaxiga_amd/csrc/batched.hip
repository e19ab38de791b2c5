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

}  // namespace iga

using namespace iga;

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
