// Learned preconditioner: a bias-free, activation-free convolutional U-Net on
// the control-point grid, with hand-written FP32 forward, backward and Adam
// kernels for gfx950, and the flexible-CG solve that applies it.
//
// Tensors are NCHW float32. A stride-s 3x3 convolution with zero padding 1
// maps H -> (H-1)/s + 1 and has weights (C_out, C_in, 3, 3); the transposed
// ("up") convolution has weights (C_in, C_out, 3, 3), an explicit output size,
// and is the exact adjoint of the forward convolution with the same weights.
// An up layer adds its skip tensor: out = conv_t(in) + skip.
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace iga {

constexpr int kCoTile = 8;  // output channels per thread

// out[n][co][y][x] (+)= sum_ci sum_k w[co][ci][k] in[n][ci][s*y+ky-1][s*x+kx-1]
template <int COT>
__global__ void __launch_bounds__(256)
    k_conv(int B, int Cin, int H, int W, int Cout, int Ho, int Wo, int s,
           const float* __restrict__ in, const float* __restrict__ w, float* __restrict__ out,
           int accumulate) {
  const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int co0 = blockIdx.y * COT;
  const int64_t npix = (int64_t)B * Ho * Wo;
  if (pix >= npix) return;
  const int x = pix % Wo, y = (pix / Wo) % Ho, n = pix / ((int64_t)Wo * Ho);
  float acc[COT];
#pragma unroll
  for (int t = 0; t < COT; ++t) acc[t] = 0.f;
  const float* inb = in + (size_t)n * Cin * H * W;
  for (int ci = 0; ci < Cin; ++ci) {
    const float* ip = inb + (size_t)ci * H * W;
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
      const int iy = s * y + ky - 1;
      if (iy < 0 || iy >= H) continue;
#pragma unroll
      for (int kx = 0; kx < 3; ++kx) {
        const int ix = s * x + kx - 1;
        if (ix < 0 || ix >= W) continue;
        const float v = ip[(size_t)iy * W + ix];
#pragma unroll
        for (int t = 0; t < COT; ++t)
          if (co0 + t < Cout) acc[t] += w[(((size_t)(co0 + t) * Cin + ci) * 3 + ky) * 3 + kx] * v;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < COT; ++t) {
    if (co0 + t >= Cout) break;
    float* o = out + (((size_t)n * Cout + co0 + t) * Ho + y) * Wo + x;
    *o = accumulate ? *o + acc[t] : acc[t];
  }
}

// Transposed convolution, weights (Cin, Cout, 3, 3) (Cin = channels of `in`):
// out[n][co][Y][X] (+)= sum_ci sum_k w[ci][co][k] in[n][ci][y][x], Y = s*y+ky-1.
template <int COT>
__global__ void __launch_bounds__(256)
    k_conv_t(int B, int Cin, int h, int wd, int Cout, int H, int W, int s,
             const float* __restrict__ in, const float* __restrict__ w, float* __restrict__ out,
             int accumulate) {
  const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int co0 = blockIdx.y * COT;
  const int64_t npix = (int64_t)B * H * W;
  if (pix >= npix) return;
  const int X = pix % W, Y = (pix / W) % H, n = pix / ((int64_t)W * H);
  float acc[COT];
#pragma unroll
  for (int t = 0; t < COT; ++t) acc[t] = 0.f;
  const float* inb = in + (size_t)n * Cin * h * wd;
#pragma unroll
  for (int ky = 0; ky < 3; ++ky) {
    const int ny = Y + 1 - ky;
    if (ny < 0 || ny % s) continue;
    const int y = ny / s;
    if (y >= h) continue;
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int nx = X + 1 - kx;
      if (nx < 0 || nx % s) continue;
      const int x = nx / s;
      if (x >= wd) continue;
      for (int ci = 0; ci < Cin; ++ci) {
        const float v = inb[((size_t)ci * h + y) * wd + x];
#pragma unroll
        for (int t = 0; t < COT; ++t)
          if (co0 + t < Cout) acc[t] += w[(((size_t)ci * Cout + co0 + t) * 3 + ky) * 3 + kx] * v;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < COT; ++t) {
    if (co0 + t >= Cout) break;
    float* o = out + (((size_t)n * Cout + co0 + t) * H + Y) * W + X;
    *o = accumulate ? *o + acc[t] : acc[t];
  }
}

// Weight gradient partials: G[a][b][k] = sum_{n,y,x} gy[n][a][y][x] * x[n][b][s*y+ky-1][s*x+kx-1]
// over a slice of the (n, y, x) positions; grid (A*Bc, nsplit).
__global__ void __launch_bounds__(256)
    k_wgrad(int N, int Bc, int H, int W, int A, int h, int wd, int s, const float* __restrict__ xin,
            const float* __restrict__ gy, float* __restrict__ part, int nsplit) {
  const int ab = blockIdx.x;
  const int a = ab / Bc, b = ab % Bc;
  const int64_t npos = (int64_t)N * h * wd;
  const int64_t chunk = (npos + nsplit - 1) / nsplit;
  const int64_t p0 = blockIdx.y * chunk, p1 = min(npos, p0 + chunk);
  float acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = 0.f;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int x = p % wd, y = (p / wd) % h, n = p / ((int64_t)wd * h);
    const float g = gy[(((size_t)n * A + a) * h + y) * wd + x];
    const float* xp = xin + ((size_t)n * Bc + b) * H * W;
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
      const int iy = s * y + ky - 1;
      if (iy < 0 || iy >= H) continue;
#pragma unroll
      for (int kx = 0; kx < 3; ++kx) {
        const int ix = s * x + kx - 1;
        if (ix < 0 || ix >= W) continue;
        acc[ky * 3 + kx] += g * xp[(size_t)iy * W + ix];
      }
    }
  }
  __shared__ float red[9][256];
#pragma unroll
  for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
#pragma unroll
      for (int k = 0; k < 9; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 9) part[((size_t)ab * 9 + threadIdx.x) * nsplit + blockIdx.y] = red[threadIdx.x][0];
}

// g[(a,b,k)] = sum over splits (fixed order); `transpose_ab` stores at [b][a][k]
__global__ void k_wgrad_sum(int A, int Bc, int nsplit, const float* __restrict__ part,
                            float* __restrict__ g, int transpose_ab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A * Bc * 9) return;
  const int k = i % 9, ab = i / 9;
  const int a = ab / Bc, b = ab % Bc;
  float s = 0.f;
  for (int j = 0; j < nsplit; ++j) s += part[(size_t)i * nsplit + j];
  const int o = transpose_ab ? ((b * A + a) * 9 + k) : i;
  g[o] = s;
}

// per-sample ||p - t||^2 and ||t||^2 (one block per sample), then the loss
// mean_n num_n/den_n and the gradient 2 (p - t) / (den_n N)
__global__ void k_nmse_sums(int64_t per, const float* __restrict__ p, const float* __restrict__ t,
                            double* sums) {
  __shared__ double scratch[16];
  const int n = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < per; i += blockDim.x) {
    const double d = (double)p[n * per + i] - t[n * per + i];
    a += d * d;
    b += (double)t[n * per + i] * t[n * per + i];
  }
  a = block_sum(a, scratch);
  b = block_sum(b, scratch);
  if (threadIdx.x == 0) {
    sums[2 * n] = a;
    sums[2 * n + 1] = b;
  }
}

__global__ void k_nmse_grad(int N, int64_t per, const float* __restrict__ p,
                            const float* __restrict__ t, const double* __restrict__ sums,
                            float* __restrict__ g, double* loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    double l = 0.0;
    for (int n = 0; n < N; ++n) l += sums[2 * n] / sums[2 * n + 1];
    *loss = l / N;
  }
  if (i >= (int64_t)N * per) return;
  const int n = i / per;
  g[i] = (float)(2.0 * ((double)p[i] - t[i]) / (sums[2 * n + 1] * N));
}

__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_copy_f(int64_t n, const float* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_zero_f(int64_t n, float* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.f;
}

__global__ void k_add_f(int64_t n, const float* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] += a[i];
}

// ---------------------------------------------------------------------------
// network driver
// ---------------------------------------------------------------------------

struct NetPlan {
  int64_t in_off;        // input activation offset (floats)
  int64_t off[16];       // layer output offsets
  int64_t total;
};

static int64_t act_size(const iga_net_t& net, int l, int batch) {
  // l = -1: network input (1 channel at H[0] x W[0]); else layer l output
  if (l < 0) return (int64_t)batch * net.H[0] * net.W[0];
  return (int64_t)batch * net.cout[l] * net.H[l + 1] * net.W[l + 1];
}

static NetPlan plan(const iga_net_t& net, int batch) {
  NetPlan P{};
  int64_t o = 0;
  P.in_off = o;
  o += (act_size(net, -1, batch) + 63) & ~(int64_t)63;
  for (int l = 0; l < net.n_layers; ++l) {
    P.off[l] = o;
    o += (act_size(net, l, batch) + 63) & ~(int64_t)63;
  }
  P.total = o;
  return P;
}

static int in_H(const iga_net_t& net, int src) { return net.H[src + 1]; }
static int in_W(const iga_net_t& net, int src) { return net.W[src + 1]; }
static int in_C(const iga_net_t& net, int src) { return src < 0 ? 1 : net.cout[src]; }

static void launch_conv(int B, int Cin, int H, int W, int Cout, int Ho, int Wo, int s,
                        const float* in, const float* w, float* out, int acc, hipStream_t st) {
  const int64_t npix = (int64_t)B * Ho * Wo;
  dim3 grid((unsigned)((npix + 255) / 256), (Cout + kCoTile - 1) / kCoTile);
  hipLaunchKernelGGL(k_conv<kCoTile>, grid, dim3(256), 0, st, B, Cin, H, W, Cout, Ho, Wo, s, in, w,
                     out, acc);
}

static void launch_conv_t(int B, int Cin, int h, int wd, int Cout, int H, int W, int s,
                          const float* in, const float* w, float* out, int acc, hipStream_t st) {
  const int64_t npix = (int64_t)B * H * W;
  dim3 grid((unsigned)((npix + 255) / 256), (Cout + kCoTile - 1) / kCoTile);
  hipLaunchKernelGGL(k_conv_t<kCoTile>, grid, dim3(256), 0, st, B, Cin, h, wd, Cout, H, W, s, in,
                     w, out, acc);
}

static int net_forward(const iga_net_t& net, int batch, const float* params, float* act,
                       hipStream_t st) {
  const NetPlan P = plan(net, batch);
  auto A = [&](int l) { return act + (l < 0 ? P.in_off : P.off[l]); };
  for (int l = 0; l < net.n_layers; ++l) {
    const int src = net.src[l];
    const int Ci = in_C(net, src), Hi = in_H(net, src), Wi = in_W(net, src);
    const int Co = net.cout[l], Ho = net.H[l + 1], Wo = net.W[l + 1], s = net.stride[l];
    const float* w = params + net.woff[l];
    if (Ci != net.cin[l]) return IGA_ERR_SHAPE;
    if (net.kind[l] == 0) {
      launch_conv(batch, Ci, Hi, Wi, Co, Ho, Wo, s, A(src), w, A(l), 0, st);
    } else {
      const int sk = net.skip[l];
      const int64_t n = act_size(net, l, batch);
      hipLaunchKernelGGL(k_copy_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, A(sk), A(l));
      launch_conv_t(batch, Ci, Hi, Wi, Co, Ho, Wo, s, A(src), w, A(l), 1, st);
    }
    IGA_CHECK_LAUNCH();
  }
  return IGA_OK;
}

static int64_t wgrad_splits(int64_t npos, int pairs) {
  int64_t ns = std::max<int64_t>(1, 2048 / std::max(pairs, 1));
  ns = std::min<int64_t>(ns, std::max<int64_t>(1, npos / 256));
  return std::min<int64_t>(ns, 256);
}

// gact: a second activation-sized buffer for the activation gradients;
// part: scratch for the weight-gradient partials (iga_net_scratch floats).
static int net_backward(const iga_net_t& net, int batch, const float* params, const float* act,
                        float* gact, float* grads, float* part, hipStream_t st) {
  const NetPlan P = plan(net, batch);
  auto A = [&](int l) { return act + (l < 0 ? P.in_off : P.off[l]); };
  auto G = [&](int l) { return gact + (l < 0 ? P.in_off : P.off[l]); };
  // activation gradients accumulate over consumers: clear all but the output's
  {
    const int64_t n = P.off[net.n_layers - 1];
    if (n > 0)
      hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, gact);
  }
  for (int l = net.n_layers - 1; l >= 0; --l) {
    const int src = net.src[l];
    const int Ci = in_C(net, src), Hi = in_H(net, src), Wi = in_W(net, src);
    const int Co = net.cout[l], Ho = net.H[l + 1], Wo = net.W[l + 1], s = net.stride[l];
    const float* w = params + net.woff[l];
    float* gw = grads + net.woff[l];
    if (net.kind[l] == 0) {
      // weight grad: (Co, Ci) = corr(g_out (Co, small), in (Ci, big))
      const int pairs = Co * Ci;
      const int64_t npos = (int64_t)batch * Ho * Wo;
      const int ns = (int)wgrad_splits(npos, pairs);
      hipLaunchKernelGGL(k_wgrad, dim3(pairs, ns), dim3(256), 0, st, batch, Ci, Hi, Wi, Co, Ho, Wo,
                         s, A(src), G(l), part, ns);
      hipLaunchKernelGGL(k_wgrad_sum, dim3((pairs * 9 + 255) / 256), dim3(256), 0, st, Co, Ci, ns,
                         part, gw, 0);
      launch_conv_t(batch, Co, Ho, Wo, Ci, Hi, Wi, s, G(l), w, G(src), 1, st);
    } else {
      // out = conv_t(in, w) + skip; w (Ci, Co): dW[c][o] = sum in[c][y] g[o][s y + k - 1]
      const int pairs = Ci * Co;
      const int64_t npos = (int64_t)batch * Hi * Wi;
      const int ns = (int)wgrad_splits(npos, pairs);
      hipLaunchKernelGGL(k_wgrad, dim3(pairs, ns), dim3(256), 0, st, batch, Co, Ho, Wo, Ci, Hi, Wi,
                         s, G(l), A(src), part, ns);
      hipLaunchKernelGGL(k_wgrad_sum, dim3((pairs * 9 + 255) / 256), dim3(256), 0, st, Ci, Co, ns,
                         part, gw, 0);
      launch_conv(batch, Co, Ho, Wo, Ci, Hi, Wi, s, G(l), w, G(src), 1, st);
      const int64_t n = act_size(net, l, batch);
      hipLaunchKernelGGL(k_add_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, G(l),
                         G(net.skip[l]));
    }
    IGA_CHECK_LAUNCH();
  }
  return IGA_OK;
}

}  // namespace iga

using namespace iga;

extern "C" int64_t iga_net_act_floats(const iga_net_t* net, int32_t batch) {
  if (!net || net->n_layers < 1 || net->n_layers > 16) return -1;
  return plan(*net, batch).total;
}

extern "C" int64_t iga_net_act_offset(const iga_net_t* net, int32_t batch, int32_t layer) {
  if (!net || layer < -1 || layer >= net->n_layers) return -1;
  const NetPlan P = plan(*net, batch);
  return layer < 0 ? P.in_off : P.off[layer];
}

extern "C" int64_t iga_net_scratch_floats(const iga_net_t* net, int32_t batch) {
  if (!net) return -1;
  int64_t m = 0;
  for (int l = 0; l < net->n_layers; ++l) {
    const int pairs = net->cin[l] * net->cout[l];
    const int64_t npos = (int64_t)batch * (net->kind[l] == 0 ? net->H[l + 1] * net->W[l + 1]
                                                            : net->H[net->src[l] + 1] *
                                                                  net->W[net->src[l] + 1]);
    m = std::max<int64_t>(m, (int64_t)pairs * 9 * wgrad_splits(npos, pairs));
  }
  return m;
}

extern "C" int iga_net_forward(const iga_net_t* net, int32_t batch, const float* params,
                               float* act, void* stream) {
  if (!net || !params || !act) return IGA_ERR_ARG;
  return net_forward(*net, batch, params, act, (hipStream_t)stream);
}

extern "C" int iga_net_backward(const iga_net_t* net, int32_t batch, const float* params,
                                const float* act, float* gact, float* grads, float* scratch,
                                void* stream) {
  if (!net || !params || !act || !gact || !grads || !scratch) return IGA_ERR_ARG;
  return net_backward(*net, batch, params, act, gact, grads, scratch, (hipStream_t)stream);
}

extern "C" int iga_nmse(int32_t batch, int64_t per_sample, const float* pred, const float* target,
                        float* grad, double* sums, double* loss, void* stream) {
  if (!pred || !target || !grad || !sums || !loss) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_nmse_sums, dim3(batch), dim3(1024), 0, st, per_sample, pred, target, sums);
  const int64_t n = (int64_t)batch * per_sample;
  hipLaunchKernelGGL(k_nmse_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, batch,
                     per_sample, pred, target, sums, grad, loss);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1,
                        float b2, float eps, int32_t step, void* stream) {
  if (!w || !g || !m || !v || step < 1) return IGA_ERR_ARG;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                     w, g, m, v, lr, b1, b2, eps, c1, c2);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

// ---------------------------------------------------------------------------
// Flexible CG (Notay's FCG with one stored direction) with the learned
// two-level preconditioner
//   x1 = w D^-1 r ; r1 = r - K x1 ; x2 = x1 + D^-1/2 U(D^-1/2 r1) ; r2 = r - K x2 ;
//   z  = x2 + w D^-1 r2
// (U: the network on the control-point grid, K: FP32 CSR values). Every
// scalar stays on the device; one iteration is captured into a hipGraph and
// replayed, and the host reads the convergence flag once per replay batch.
// ---------------------------------------------------------------------------

namespace iga {

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
    k_fcg_init(int n, const double* __restrict__ b, double* x, double* r, double* p1, double* q1,
               FBufs F) {
  double bb = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] = 0.0;
    r[i] = b[i];
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

__global__ void k_pc_pre(int n, double omega, const double* __restrict__ dinv,
                         const double* __restrict__ r, double* x1, const FState* st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x1[i] = omega * dinv[i] * r[i];
}

template <int LPR, typename TV>
__global__ void __launch_bounds__(kVecThreads)
    k_resid(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ cols,
            const TV* __restrict__ vals, const double* __restrict__ r, const double* __restrict__ x,
            double* out, const FState* st) {
  if (st->done) return;
  const int l = threadIdx.x % LPR;
  const int groups = gridDim.x * (blockDim.x / LPR);
  for (int row = blockIdx.x * (blockDim.x / LPR) + threadIdx.x / LPR; row < n; row += groups) {
    const double s = row_dot<LPR, TV>(row, indptr, cols, vals, x, l);
    if (l == 0) out[row] = r[row] - s;
  }
}

__global__ void k_to_grid(int n, const int32_t* __restrict__ cell, const double* __restrict__ dsq,
                          const double* __restrict__ v, float* grid, const FState* st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) grid[cell[i]] = (float)(dsq[i] * v[i]);
}

__global__ void k_from_grid(int n, const int32_t* __restrict__ cell, const double* __restrict__ dsq,
                            const float* __restrict__ grid, const double* __restrict__ x1,
                            double* x2, const FState* st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x2[i] = x1[i] + dsq[i] * (double)grid[cell[i]];
}

__global__ void k_pc_post(int n, double omega, const double* __restrict__ dinv,
                          const double* __restrict__ x2, const double* __restrict__ r2, double* z,
                          const FState* st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x2[i] + omega * dinv[i] * r2[i];
}

__global__ void __launch_bounds__(kVecThreads)
    k_fcg_zq(int n, const double* __restrict__ z, const double* __restrict__ qp, FBufs F) {
  if (F.st->done) return;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += z[i] * qp[i];
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
                 const double* __restrict__ q, double* x, double* r, FBufs F) {
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
    rr += ri * ri;
  }
  store_partial(F.P_d, rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->pq_prev = pq;
}

}  // namespace iga

extern "C" int64_t iga_fcg_workspace(int32_t n) { return 8 + 10 * (int64_t)n + 4 * kRedBlocks; }

extern "C" int iga_fcg_learned(const iga_net_t* net, const float* params, float* act, int32_t n,
                               const int32_t* indptr, const int32_t* cols, const double* vals,
                               const float* vals_pc, const double* dinv, const double* dsq,
                               const int32_t* cell, double omega, const double* b, double* x,
                               double tol, int32_t maxiter, int32_t graph_iters, double* work,
                               int32_t* out_iters, double* out_rr, double* out_bb, void* stream) {
  if (!net || !params || !act || !indptr || !cols || !vals || !vals_pc || !dinv || !dsq || !cell ||
      !b || !x || !work)
    return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    if (out_iters) *out_iters = 0;
    return IGA_OK;
  }
  FBufs F;
  F.st = reinterpret_cast<FState*>(work);
  double* r = work + 8;
  double* z = r + n;
  double* P[2] = {z + n, z + 2 * n};
  double* Q[2] = {z + 3 * n, z + 4 * n};
  double* x1 = z + 5 * n;
  double* x2 = z + 6 * n;
  double* t1 = z + 7 * n;
  double* t2 = z + 8 * n;
  double* parts = z + 9 * n;
  F.P_a = parts;
  F.P_b = parts + kRedBlocks;
  F.P_c = parts + 2 * kRedBlocks;
  F.P_d = parts + 3 * kRedBlocks;
  int nnz = 0;
  {
    int32_t v = 0;
    if (hipMemcpyAsync(&v, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess)
      return IGA_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
    nnz = v;
  }
  const int lpr = pick_lpr(n, nnz);
  const int vb = vec_blocks(n);
  const int rb = row_blocks(n, lpr);
  const int eb = (n + 255) / 256;
  const NetPlan plan_ = plan(*net, 1);
  float* grid_in = act + plan_.in_off;
  const float* grid_out = act + plan_.off[net->n_layers - 1];
  {
    // boundary cells of the network input stay zero
    const int64_t cells = (int64_t)net->H[0] * net->W[0];
    hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, cells,
                       grid_in);
  }
  hipLaunchKernelGGL(k_fcg_init, dim3(vb), dim3(kVecThreads), 0, st, n, b, x, r, P[1], Q[1], F);
  hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kVecThreads), 0, st, vb, F, 1, tol * tol, maxiter);
  IGA_CHECK_LAUNCH();

  auto iteration = [&](int par) -> int {
    double* p = P[par];
    double* pp = P[1 - par];
    double* q = Q[par];
    double* qp = Q[1 - par];
    hipLaunchKernelGGL(k_pc_pre, dim3(eb), dim3(256), 0, st, n, omega, dinv, r, x1, F.st);
    int rc = by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL((k_resid<LPR, float>), dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols,
                         vals_pc, r, x1, t1, F.st);
      return 0;
    });
    hipLaunchKernelGGL(k_to_grid, dim3(eb), dim3(256), 0, st, n, cell, dsq, t1, grid_in, F.st);
    rc = net_forward(*net, 1, params, act, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_from_grid, dim3(eb), dim3(256), 0, st, n, cell, dsq, grid_out, x1, x2, F.st);
    by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL((k_resid<LPR, float>), dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols,
                         vals_pc, r, x2, t2, F.st);
      return 0;
    });
    hipLaunchKernelGGL(k_pc_post, dim3(eb), dim3(256), 0, st, n, omega, dinv, x2, t2, z, F.st);
    hipLaunchKernelGGL(k_fcg_zq, dim3(vb), dim3(kVecThreads), 0, st, n, z, qp, F);
    hipLaunchKernelGGL(k_fcg_dir, dim3(vb), dim3(kVecThreads), 0, st, n, vb, z, pp, p, r, F);
    by_lpr(lpr, [&](auto L) {
      constexpr int LPR = decltype(L)::value;
      hipLaunchKernelGGL(k_fcg_spmv<LPR>, dim3(rb), dim3(kVecThreads), 0, st, n, indptr, cols, vals,
                         p, q, F);
      return 0;
    });
    hipLaunchKernelGGL(k_fcg_update, dim3(vb), dim3(kVecThreads), 0, st, n, rb, vb, p, q, x, r, F);
    hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(kVecThreads), 0, st, vb, F, 0, 0.0, 0);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  };

  const int G = graph_iters > 0 ? (graph_iters + 1) & ~1 : 0;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  if (G > 0) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return IGA_ERR_HIP;
    int rc = 0;
    for (int k = 0; k < G && !rc; ++k) rc = iteration(k & 1);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      return rc ? rc : IGA_ERR_HIP;
    }
    graph = g;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      hipGraphDestroy(graph);
      return IGA_ERR_HIP;
    }
  }
  FState host{};
  int launched = 0;
  int status = IGA_OK;
  for (;;) {
    if (hipMemcpyAsync(&host, F.st, sizeof(FState), hipMemcpyDeviceToHost, st) != hipSuccess) {
      status = IGA_ERR_HIP;
      break;
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
      status = IGA_ERR_HIP;
      break;
    }
    if (host.done) break;
    if (exec) {
      if (hipGraphLaunch(exec, st) != hipSuccess) {
        status = IGA_ERR_HIP;
        break;
      }
      launched += G;
    } else {
      for (int k = 0; k < 2; ++k, ++launched) {
        const int rc = iteration(launched & 1);
        if (rc) {
          status = rc;
          break;
        }
      }
      if (status) break;
    }
  }
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  if (out_iters) *out_iters = host.iter;
  if (out_rr) *out_rr = host.rr;
  if (out_bb) *out_bb = host.bb;
  return status;
}
