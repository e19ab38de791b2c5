// Convolutional networks on the control-point grid with hand-written FP32
// forward, backward and Adam kernels for gfx950 (the smoother generator of the
// learned multigrid preconditioner is one).
//
// Tensors are NCHW float32. A stride-s 3x3 convolution with zero padding 1
// maps H -> (H-1)/s + 1 and has weights (C_out, C_in, 3, 3); the transposed
// ("up") convolution has weights (C_in, C_out, 3, 3), an explicit output size,
// and is the exact adjoint of the forward convolution with the same weights.
// An up layer adds its skip tensor: out = conv_t(in) + skip. Each layer may
// add a per-channel bias and apply a ReLU, in that order, after the skip add.
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

// Stride-1 3x3 convolution as an implicit GEMM on the FP32 matrix cores:
//   out[n][co][p] = act(sum_kk W[co][kk] X[kk][p] + bias[co]),
//   kk = ci*9 + k, X[kk][p] = in[n][ci][p + off(k)]  (zero outside the image).
// A block computes a row segment of 64 pixels for every output channel
// (Cout <= 64, Cin <= 64): the input rows y-1..y+1 of the segment are staged
// in LDS; each wave owns 16 pixels and four 16x16 output-channel tiles, and
// each step of 4 along kk issues four v_mfma_f32_16x16x4f32 sharing the
// pixel operand. Output mapping: lane l holds pixel l%16 and channels
// 4*(l/16)+r of each tile.
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kConvSeg = 64;
constexpr int kConvMaxC = 64;

constexpr int kConvCh = 8;                 // input channels per weight chunk
constexpr int kConvKc = kConvCh * 9;       // kk per chunk (a multiple of 4)
constexpr int kConvWs = kConvKc + 1;       // padded LDS row of the weight chunk

__global__ void __launch_bounds__(256)
    k_conv_mfma(int Cin, int H, int W, int Cout, const float* __restrict__ in,
                const float* __restrict__ w, const float* __restrict__ bias, int relu,
                float* __restrict__ out) {
  __shared__ float tile[kConvMaxC][3][kConvSeg + 2];
  __shared__ float wsm[kConvMaxC][kConvWs];
  const int x0 = blockIdx.x * kConvSeg, y = blockIdx.y, n = blockIdx.z;
  const float* inb = in + (size_t)n * Cin * H * W;
  for (int t = threadIdx.x; t < Cin * 3 * (kConvSeg + 2); t += blockDim.x) {
    const int xx = t % (kConvSeg + 2), dy = (t / (kConvSeg + 2)) % 3, ci = t / (3 * (kConvSeg + 2));
    const int gy = y + dy - 1, gx = x0 + xx - 1;
    tile[ci][dy][xx] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? inb[((size_t)ci * H + gy) * W + gx] : 0.f;
  }
  const int wave = threadIdx.x / kWave, l = threadIdx.x % kWave;
  const int px = wave * 16 + l % 16;  // pixel of the B operand and of the output
  const int kq = l / 16;              // row of the 4-deep k slice this lane feeds
  const int K = Cin * 9;
  f4v acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
  for (int ci0 = 0; ci0 < Cin; ci0 += kConvCh) {
    __syncthreads();  // the previous chunk's weights are consumed (and the tile is staged)
    for (int t = threadIdx.x; t < kConvMaxC * kConvKc; t += blockDim.x) {
      const int co = t / kConvKc, kl = t - co * kConvKc;
      const int kk = ci0 * 9 + kl;
      wsm[co][kl] = (co < Cout && kk < K) ? w[(size_t)co * K + kk] : 0.f;
    }
    __syncthreads();
#pragma unroll 2
    for (int kl0 = 0; kl0 < kConvKc; kl0 += 4) {
      const int kl = kl0 + kq;
      const int ci = ci0 + kl / 9, k = kl % 9;
      const float bv = ci < Cin ? tile[ci][k / 3][px + k % 3] : 0.f;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wsm[t * 16 + l % 16][kl], bv, acc[t], 0, 0, 0);
    }
  }
  const int x = x0 + px;
  if (x >= W) return;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = t * 16 + 4 * kq + r;
      if (co >= Cout) continue;
      float v = acc[t][r];
      if (bias) v += bias[co];
      if (relu) v = fmaxf(v, 0.f);
      out[(((size_t)n * Cout + co) * H + y) * W + x] = v;
    }
}

// FP16 operands of v_mfma_f32_16x16x16f16 (FP32 accumulation): lane l feeds
// A[l%16][4(l/16)..+3] and B[4(l/16)..+3][l%16] per 16-deep step.
typedef _Float16 h4v __attribute__((ext_vector_type(4)));
// FP16-operand 3x3 stride-1 convolution on v_mfma_f32_16x16x16f16, FP32
// accumulation. A block computes kCvRows image rows x kCvPx pixels (one
// wavefront per row, four 16-pixel MFMA tiles each) for all output channels,
// looping over 16-channel input chunks: per chunk the input rows (with a
// one-pixel halo) are staged channel-last and the packed FP16 weights
// w16[co][tap][ci] of the chunk as [co][tap][16], so a lane's 4-deep A and B
// operands are one 8-byte LDS read each (rows padded to 20 halves). Global
// reads are issued from clamped addresses, eight at a time, and zeroed off
// the image afterwards.
__global__ void k_w16_pack(int Cin, int Cp, int Cout, int Cop, const float* __restrict__ w,
                           _Float16* __restrict__ w16) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Cop * 9 * Cp) return;
  const int ci = t % Cp, tap = (t / Cp) % 9, co = t / (9 * Cp);
  w16[t] = (_Float16)((co < Cout && ci < Cin) ? w[((size_t)co * Cin + ci) * 9 + tap] : 0.f);
}

constexpr int kCvRows = 4;            // image rows per block (one wavefront each)
constexpr int kCvPx = 64;             // pixels per row
constexpr int kCvHx = kCvPx + 2;      // with the halo
constexpr int kCvCs = 20;             // halves per staged pixel / weight row (16 + pad)

static size_t conv3_lds_bytes(int Cp, int Cop) {
  (void)Cp;
  return sizeof(_Float16) * ((size_t)(kCvRows + 2) * kCvHx * kCvCs + (size_t)Cop * 9 * kCvCs);
}

template <int COT>
__global__ void __launch_bounds__(256)
    k_conv3_f16(int Cin, int Cp, int H, int W, int Cout, const float* __restrict__ in,
                const _Float16* __restrict__ w16, const float* __restrict__ bias, int relu,
                float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) _Float16 tin[(kCvRows + 2) * kCvHx * kCvCs];
  __shared__ __attribute__((aligned(16))) _Float16 twt[COT * 16 * 9 * kCvCs];
  const int x0 = blockIdx.x * kCvPx, y0 = blockIdx.y * kCvRows, n = blockIdx.z;
  const int wave = threadIdx.x / kWave, l = threadIdx.x % kWave;
  const int lr = l % 16, kq = l / 16;
  const float* inb = in + (size_t)n * Cin * H * W;
  f4v acc[4][COT];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int t = 0; t < COT; ++t) acc[p][t] = f4v{0.f, 0.f, 0.f, 0.f};
  constexpr int NIN = 16 * (kCvRows + 2) * kCvHx;  // input values per chunk
  constexpr int NW = COT * 16 * 9 * 4;             // 4-half weight pieces per chunk
  for (int ci0 = 0; ci0 < Cp; ci0 += 16) {
    for (int t0 = 0; t0 < NIN; t0 += 8 * 256) {
      float v[8];
      int dst[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = t0 + u * 256 + threadIdx.x;
        const int xx = t % kCvHx, rr = (t / kCvHx) % (kCvRows + 2), c = t / (kCvHx * (kCvRows + 2));
        const int gy = y0 + rr - 1, gx = x0 + xx - 1, ci = ci0 + c;
        const bool ok = t < NIN && ci < Cin && gy >= 0 && gy < H && gx >= 0 && gx < W;
        const int cy = min(max(gy, 0), H - 1), cx = min(max(gx, 0), W - 1), cc = min(ci, Cin - 1);
        v[u] = inb[((size_t)cc * H + cy) * W + cx];
        v[u] = ok ? v[u] : 0.f;
        dst[u] = t < NIN ? (rr * kCvHx + xx) * kCvCs + c : -1;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (dst[u] >= 0) tin[dst[u]] = (_Float16)v[u];
    }
    for (int t0 = 0; t0 < NW; t0 += 4 * 256) {
      h4v v[4];
      int dst[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = min(t0 + u * 256 + (int)threadIdx.x, NW - 1);
        const int q = t % 4, tap = (t / 4) % 9, co = t / 36;
        v[u] = *reinterpret_cast<const h4v*>(w16 + ((size_t)co * 9 + tap) * Cp + ci0 + 4 * q);
        dst[u] = t0 + u * 256 + (int)threadIdx.x < NW ? (co * 9 + tap) * kCvCs + 4 * q : -1;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (dst[u] >= 0) *reinterpret_cast<h4v*>(twt + dst[u]) = v[u];
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = tap / 3, kx = tap - 3 * ky;
      h4v av[COT];
#pragma unroll
      for (int t = 0; t < COT; ++t)
        av[t] = *reinterpret_cast<const h4v*>(twt + ((t * 16 + lr) * 9 + tap) * kCvCs + 4 * kq);
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const h4v bv = *reinterpret_cast<const h4v*>(
            tin + ((wave + ky) * kCvHx + p * 16 + lr + kx) * kCvCs + 4 * kq);
#pragma unroll
        for (int t = 0; t < COT; ++t)
          acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x16f16(av[t], bv, acc[p][t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int y = y0 + wave;
  if (y >= H) return;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int x = x0 + p * 16 + lr;
    if (x >= W) continue;
#pragma unroll
    for (int t = 0; t < COT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = t * 16 + 4 * kq + r;
        if (co >= Cout) continue;
        float v = acc[p][t][r];
        if (bias) v += bias[co];
        if (relu) v = fmaxf(v, 0.f);
        out[(((size_t)n * Cout + co) * H + y) * W + x] = v;
      }
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

// out[n][c][p] = act(out[n][c][p] + bias[c]) over C channels of HW pixels
__global__ void k_bias_act(int64_t total, int C, int64_t HW, const float* __restrict__ bias,
                           int relu, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  float v = out[i];
  if (bias) v += bias[(i / HW) % C];
  if (relu) v = fmaxf(v, 0.f);
  out[i] = v;
}

// g *= (a > 0): the ReLU derivative read from the stored post-activation output
__global__ void k_relu_mask(int64_t total, const float* __restrict__ a, float* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total && !(a[i] > 0.f)) g[i] = 0.f;
}

// gb[c] = sum over batch and pixels of g[n][c][p]; one block per channel
__global__ void k_bias_grad(int B, int C, int64_t HW, const float* __restrict__ g,
                            float* __restrict__ gb) {
  __shared__ double scratch[16];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < (int64_t)B * HW; t += blockDim.x) {
    const int64_t n = t / HW, p = t - n * HW;
    s += g[(n * C + c) * HW + p];
  }
  s = block_sum(s, scratch);
  if (threadIdx.x == 0) gb[c] = (float)s;
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
  int64_t w16_off[16];   // packed FP16 weights of FP16 matrix-core layers, or -1
  int64_t total;
};

// FP16 matrix-core convolution: stride-1 forward layers of at most
// kConvMaxC channels, when the descriptor asks for FP16 operands
static bool f16_layer(const iga_net_t& net, int l, int Ci) {
  return (net.precision == 1 || net.precision == 2) && net.kind[l] == 0 && net.stride[l] == 1 &&
         Ci <= kConvMaxC &&
         net.cout[l] <= kConvMaxC &&
         conv3_lds_bytes((Ci + 15) / 16 * 16, (net.cout[l] + 15) / 16 * 16) <= 64 * 1024;
}

static int in_C(const iga_net_t& net, int src) {
  return src < 0 ? (net.c_in > 0 ? net.c_in : 1) : net.cout[src];
}

static int64_t act_size(const iga_net_t& net, int l, int batch) {
  // l = -1: network input (c_in channels at H[0] x W[0]); else layer l output
  if (l < 0) return (int64_t)batch * in_C(net, -1) * net.H[0] * net.W[0];
  return (int64_t)batch * net.cout[l] * net.H[l + 1] * net.W[l + 1];
}

static int in_H(const iga_net_t& net, int src) { return net.H[src + 1]; }
static int in_W(const iga_net_t& net, int src) { return net.W[src + 1]; }

static NetPlan plan(const iga_net_t& net, int batch) {
  NetPlan P{};
  int64_t o = 0;
  P.in_off = o;
  o += (act_size(net, -1, batch) + 63) & ~(int64_t)63;
  for (int l = 0; l < net.n_layers; ++l) {
    P.off[l] = o;
    o += (act_size(net, l, batch) + 63) & ~(int64_t)63;
  }
  for (int l = 0; l < net.n_layers; ++l) {
    P.w16_off[l] = -1;
    const int Ci = in_C(net, net.src[l]);
    if (!f16_layer(net, l, Ci)) continue;
    const int64_t halves = (int64_t)((net.cout[l] + 15) / 16 * 16) * 9 * ((Ci + 15) / 16 * 16);
    P.w16_off[l] = o;
    o += ((halves + 1) / 2 + 63) & ~(int64_t)63;
  }
  P.total = o;
  return P;
}

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

// input / output (optional): the input slot and the last layer's slot read
// and written there instead of in act
static int net_forward(const iga_net_t& net, int batch, const float* params, float* act,
                       hipStream_t st, const float* input = nullptr, float* output = nullptr) {
  const NetPlan P = plan(net, batch);
  auto A = [&](int l) -> float* {
    if (l < 0 && input) return const_cast<float*>(input);
    if (l == net.n_layers - 1 && output) return output;
    return act + (l < 0 ? P.in_off : P.off[l]);
  };
  for (int l = 0; l < net.n_layers; ++l) {
    const int src = net.src[l];
    const int Ci = in_C(net, src), Hi = in_H(net, src), Wi = in_W(net, src);
    const int Co = net.cout[l], Ho = net.H[l + 1], Wo = net.W[l + 1], s = net.stride[l];
    const float* w = params + net.woff[l];
    if (Ci != net.cin[l]) return IGA_ERR_SHAPE;
    const int64_t n = act_size(net, l, batch);
    const float* bias = net.boff[l] >= 0 ? params + net.boff[l] : nullptr;
    if (net.kind[l] == 0 && s == 1 && Ci <= kConvMaxC && Co <= kConvMaxC) {
      // stride 1 on the matrix cores (FP32, or FP16 operands when the
      // descriptor asks for it), bias and ReLU in the epilogue
      dim3 grid((Wo + kConvSeg - 1) / kConvSeg, Ho, batch);
      if (P.w16_off[l] >= 0) {
        grid = dim3((Wo + kCvPx - 1) / kCvPx, (Ho + kCvRows - 1) / kCvRows, batch);
        const int Cp = (Ci + 15) / 16 * 16, Cop = (Co + 15) / 16 * 16;
        _Float16* w16 = reinterpret_cast<_Float16*>(act + P.w16_off[l]);
        const int nw = Cop * 9 * Cp;
        // precision 2: packed already (iga_net_pack16 into the same act)
        if (net.precision != 2)
          hipLaunchKernelGGL(k_w16_pack, dim3((nw + 255) / 256), dim3(256), 0, st, Ci, Cp, Co, Cop,
                             w, w16);
        const size_t lds = 0;  // static LDS (conv3_lds_bytes)
        switch (Cop / 16) {
          case 1:
            hipLaunchKernelGGL((k_conv3_f16<1>), grid, dim3(256), lds, st, Ci, Cp, Hi, Wi, Co,
                               A(src), w16, bias, net.relu[l], A(l));
            break;
          case 2:
            hipLaunchKernelGGL((k_conv3_f16<2>), grid, dim3(256), lds, st, Ci, Cp, Hi, Wi, Co,
                               A(src), w16, bias, net.relu[l], A(l));
            break;
          case 3:
            hipLaunchKernelGGL((k_conv3_f16<3>), grid, dim3(256), lds, st, Ci, Cp, Hi, Wi, Co,
                               A(src), w16, bias, net.relu[l], A(l));
            break;
          default:
            hipLaunchKernelGGL((k_conv3_f16<4>), grid, dim3(256), lds, st, Ci, Cp, Hi, Wi, Co,
                               A(src), w16, bias, net.relu[l], A(l));
        }
      } else {
        hipLaunchKernelGGL(k_conv_mfma, grid, dim3(256), 0, st, Ci, Hi, Wi, Co, A(src), w, bias,
                           net.relu[l], A(l));
      }
      IGA_CHECK_LAUNCH();
      continue;
    }
    if (net.kind[l] == 0) {
      launch_conv(batch, Ci, Hi, Wi, Co, Ho, Wo, s, A(src), w, A(l), 0, st);
    } else {
      const int sk = net.skip[l];
      hipLaunchKernelGGL(k_copy_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, A(sk), A(l));
      launch_conv_t(batch, Ci, Hi, Wi, Co, Ho, Wo, s, A(src), w, A(l), 1, st);
    }
    if (net.boff[l] >= 0 || net.relu[l])
      hipLaunchKernelGGL(k_bias_act, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, Co,
                         (int64_t)Ho * Wo, net.boff[l] >= 0 ? params + net.boff[l] : nullptr,
                         net.relu[l], A(l));
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
    const int64_t nout = act_size(net, l, batch);
    if (net.relu[l])
      hipLaunchKernelGGL(k_relu_mask, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, nout,
                         A(l), G(l));
    if (net.boff[l] >= 0)
      hipLaunchKernelGGL(k_bias_grad, dim3(Co), dim3(256), 0, st, batch, Co, (int64_t)Ho * Wo, G(l),
                         grads + net.boff[l]);
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
      hipLaunchKernelGGL(k_add_f, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, nout,
                         G(l), G(net.skip[l]));
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

extern "C" int iga_net_pack16(const iga_net_t* net, int32_t batch, const float* params,
                              float* act, void* stream) {
  if (!net || !params || !act || net->n_layers < 1) return IGA_ERR_ARG;
  const NetPlan P = plan(*net, batch);
  hipStream_t st = (hipStream_t)stream;
  for (int l = 0; l < net->n_layers; ++l) {
    if (P.w16_off[l] < 0) continue;
    const int Ci = in_C(*net, net->src[l]), Co = net->cout[l];
    const int Cp = (Ci + 15) / 16 * 16, Cop = (Co + 15) / 16 * 16;
    const int nw = Cop * 9 * Cp;
    hipLaunchKernelGGL(k_w16_pack, dim3((nw + 255) / 256), dim3(256), 0, st, Ci, Cp, Co, Cop,
                       params + net->woff[l], reinterpret_cast<_Float16*>(act + P.w16_off[l]));
  }
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

extern "C" int iga_net_infer(const iga_net_t* net, int32_t batch, const float* params, float* act,
                             const float* input, float* output, void* stream) {
  if (!net || !params || !act || !input || !output || net->n_layers < 1) return IGA_ERR_ARG;
  return net_forward(*net, batch, params, act, (hipStream_t)stream, input, output);
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
