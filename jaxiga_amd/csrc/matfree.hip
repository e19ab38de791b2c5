// Matrix-free, sum-factorised operator apply y = K v for the scalar Poisson
// form on a single tensor-product NURBS patch (gfx950).
//
// Setup (k_mf_geometry), once per operator: at every Gauss point of every
// element, the symmetric factor G = w a J^{-T} J^{-1} (w: quadrature weight
// times |det J|, a: coefficient) and, on rational patches, the NURBS weight
// function W and its parametric gradient -- all in FP64, stored in the apply
// precision, point-major within an element.
//
// Apply, two streaming passes and no atomics:
//  k_mf_apply  -- one group of lanes per element: gather the element's
//                 coefficients, contract them direction by direction with the
//                 element's 1-D extracted tables (values and derivatives at
//                 the Gauss points, staged in LDS) to get the parametric
//                 gradient, apply G (with the quotient-rule correction when
//                 the patch is rational), contract back with the transposed
//                 1-D tables, and store the element's local result;
//  k_mf_gather -- one lane per free dof: sum the local results of the
//                 elements supporting it (fixed order, bitwise reproducible).
//
// Rational patches: with u~_i = w_i u_i, U = sum u~_i N_i and W = sum w_i N_i
// are both separable sums, u_h = U / W and grad u_h = (grad U - u_h grad W)/W;
// the test side carries the transposed correction r_i = w_i [ grad N_i . f/W
// - N_i (grad W . f)/W^2 ] with f = G grad u_h.
#include <algorithm>

#include "common.hpp"
#include "matfree.hpp"

namespace iga {

template <int DIM, int P>
struct MfShape {
  static constexpr int N1 = P + 1;       // 1-D functions per element
  static constexpr int Q1 = P + 1;       // 1-D Gauss points per element
  static constexpr int NL = ipow(N1, DIM);
  static constexpr int NQ = ipow(Q1, DIM);
  static constexpr int NG = DIM == 2 ? 3 : 6;  // stored entries of symmetric G
  static constexpr int NW = 1 + DIM;           // W and grad W (rational only)
  static constexpr int LANES = pow2ceil(NQ > NL ? NQ : NL);  // lanes per element
  // one wavefront per block where an element fits in one: the apply has a
  // barrier between every pair of contractions, and a one-wave barrier costs
  // next to nothing, while four waves would wait on the slowest of them
  static constexpr int BLOCK = LANES > 64 ? LANES : 64;
};

// index of G[a][b] (a <= b) in the packed symmetric layout
template <int DIM>
__device__ __forceinline__ int gsym(int a, int b) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  if constexpr (DIM == 2) return a == 0 ? b : 2;
  else return a == 0 ? b : (a == 1 ? 2 + b : 5);
}

// ---------------------------------------------------------------------------
// setup: geometric factors at the Gauss points
// ---------------------------------------------------------------------------

template <int DIM, int P, typename T>
__global__ void __launch_bounds__(256)
    k_mf_geometry(iga_patch_t pt, const double* __restrict__ qp, const double* __restrict__ qw,
                  int coef_kind, double coef_const, const double* __restrict__ coef_points,
                  T* __restrict__ geo, T* __restrict__ wgeo, int64_t n_e) {
  using S = MfShape<DIM, P>;
  constexpr int N1 = S::N1, Q1 = S::Q1, NQ = S::NQ, NG = S::NG, NW = S::NW;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_e * NQ) return;
  const int64_t e = t / NQ;
  const int q = (int)(t % NQ);
  int ei[3] = {0, 0, 0}, qi[3] = {q % Q1, (q / Q1) % Q1, q / (Q1 * Q1)};
  {
    int64_t rem = e;
    for (int d = 0; d < DIM; ++d) {
      ei[d] = (int)(rem % pt.n_elem[d]);
      rem /= pt.n_elem[d];
    }
  }
  double v1[DIM][N1], d1[DIM][N1], jac = 1.0;
  for (int d = 0; d < DIM; ++d) {
    double B[N1], dB[N1];
    bernstein<P>(qp[qi[d]], B, dB);
    const double h = pt.hi[d][ei[d]] - pt.lo[d][ei[d]];
    jac *= 0.5 * h * qw[qi[d]];
    for (int a = 0; a < N1; ++a) {
      const double* C = pt.ext[d] + ((size_t)ei[d] * N1 + a) * N1;
      double s = 0.0, ds = 0.0;
      for (int b = 0; b < N1; ++b) {
        s += C[b] * B[b];
        ds += C[b] * dB[b];
      }
      v1[d][a] = s;
      d1[d][a] = ds * (2.0 / h);
    }
  }
  // W, grad W, sum M x, sum dM x over the element's functions
  double W = 0.0, dW[DIM] = {}, X[DIM] = {}, Y[DIM][DIM] = {};
  for (int i = 0; i < S::NL; ++i) {
    int ai[3] = {i % N1, (i / N1) % N1, i / (N1 * N1)};
    int sdof = 0, stride = 1;
    double N = 1.0, dN[DIM];
    for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
    for (int d = 0; d < DIM; ++d) {
      sdof += (pt.first[d][ei[d]] + ai[d]) * stride;
      stride *= pt.n_cp[d];
      N *= v1[d][ai[d]];
      for (int k = 0; k < DIM; ++k) dN[k] *= (k == d) ? d1[d][ai[d]] : v1[d][ai[d]];
    }
    const double w = pt.wgts[sdof];
    W += N * w;
    for (int k = 0; k < DIM; ++k) dW[k] += dN[k] * w;
    for (int c = 0; c < DIM; ++c) {
      const double xc = pt.cpts[(size_t)sdof * DIM + c];
      X[c] += N * w * xc;
      for (int k = 0; k < DIM; ++k) Y[k][c] += dN[k] * w * xc;
    }
  }
  double J[DIM][DIM], x[DIM];
  for (int c = 0; c < DIM; ++c) x[c] = X[c] / W;
  for (int k = 0; k < DIM; ++k)
    for (int c = 0; c < DIM; ++c) J[k][c] = (Y[k][c] - dW[k] * x[c]) / W;
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
  const double coef = coef_kind == IGA_COEF_POINTS ? coef_points[e * NQ + q] : coef_const;
  const double wa = jac * fabs(det) * coef;
  // G[a][b] = wa * sum_c inv[c][a] inv[c][b]; layout [e][g][q]
  for (int a = 0; a < DIM; ++a)
    for (int b = a; b < DIM; ++b) {
      double s = 0.0;
      for (int c = 0; c < DIM; ++c) s += inv[c][a] * inv[c][b];
      geo[(e * NG + gsym<DIM>(a, b)) * NQ + q] = (T)(wa * s);
    }
  if (wgeo) {
    wgeo[(e * NW + 0) * NQ + q] = (T)W;
    for (int k = 0; k < DIM; ++k) wgeo[(e * NW + 1 + k) * NQ + q] = (T)dW[k];
  }
}

// 1-D extracted tables of every span: tab[d][e_d][q][a] values, dtab derivatives
template <int P, typename T>
__global__ void k_mf_tables(int n_elem, const double* __restrict__ ext, const double* __restrict__ lo,
                            const double* __restrict__ hi, const double* __restrict__ qp,
                            T* __restrict__ tab, T* __restrict__ dtab) {
  constexpr int N1 = P + 1, Q1 = P + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * Q1 * N1) return;
  const int a = t % N1, q = (t / N1) % Q1, e = t / (N1 * Q1);
  double B[N1], dB[N1];
  bernstein<P>(qp[q], B, dB);
  const double* C = ext + ((size_t)e * N1 + a) * N1;
  double s = 0.0, ds = 0.0;
  for (int b = 0; b < N1; ++b) {
    s += C[b] * B[b];
    ds += C[b] * dB[b];
  }
  tab[t] = (T)s;
  dtab[t] = (T)(ds * 2.0 / (hi[e] - lo[e]));
}

// ---------------------------------------------------------------------------
// apply
// ---------------------------------------------------------------------------

struct MfArgs {
  int64_t n_e;
  const int* skip;     // optional device flag: nonzero -> the launch is a no-op
  const int32_t* f2f;  // (n_dofs) free index or -1
  const void* tab[3];
  const void* dtab[3];
};

// out[o] over a 3-index tensor laid out [z][y][x] (x fastest): contract the
// axis `ax` of `in` (extent nin) against M[qo][i] (M row-major, extent nout x nin).
template <typename T>
__device__ __forceinline__ T contract(const T* in, const T* M, int ax, int nin, int x, int y, int z,
                                      const int* ext_in) {
  // ext_in: extents of `in` along x, y, z; (x, y, z) is the output coordinate,
  // whose coordinate along `ax` indexes M's row.
  T s = 0;
  const int sx = 1, sy = ext_in[0], sz = ext_in[0] * ext_in[1];
  if (ax == 0) {
    const T* row = M + x * nin;
    const T* base = in + y * sy + z * sz;
    for (int i = 0; i < nin; ++i) s += row[i] * base[i * sx];
  } else if (ax == 1) {
    const T* row = M + y * nin;
    const T* base = in + x * sx + z * sz;
    for (int i = 0; i < nin; ++i) s += row[i] * base[i * sy];
  } else {
    const T* row = M + z * nin;
    const T* base = in + x * sx + y * sy;
    for (int i = 0; i < nin; ++i) s += row[i] * base[i * sz];
  }
  return s;
}

// transposed contraction: contract axis `ax` of `in` (extent nq) against
// M[q][o] (M row-major nq x nout), output coordinate along ax indexes M's column.
template <typename T>
__device__ __forceinline__ T contract_t(const T* in, const T* M, int ax, int nq, int nout, int x,
                                        int y, int z, const int* ext_in) {
  T s = 0;
  const int sx = 1, sy = ext_in[0], sz = ext_in[0] * ext_in[1];
  if (ax == 0) {
    const T* base = in + y * sy + z * sz;
    for (int q = 0; q < nq; ++q) s += M[q * nout + x] * base[q * sx];
  } else if (ax == 1) {
    const T* base = in + x * sx + z * sz;
    for (int q = 0; q < nq; ++q) s += M[q * nout + y] * base[q * sy];
  } else {
    const T* base = in + x * sx + y * sy;
    for (int q = 0; q < nq; ++q) s += M[q * nout + z] * base[q * sz];
  }
  return s;
}

template <int DIM, int P, typename T, bool RATIONAL>
__global__ void __launch_bounds__(256)
    k_mf_apply(iga_patch_t pt, MfArgs A, const T* __restrict__ geo, const T* __restrict__ wgeo,
               const double* __restrict__ v, T* __restrict__ rloc) {
  if (A.skip && *A.skip) return;
  using S = MfShape<DIM, P>;
  constexpr int N1 = S::N1, Q1 = S::Q1, NL = S::NL, NQ = S::NQ, NG = S::NG, NW = S::NW;
  constexpr int LANES = S::LANES;
  constexpr int EPB = S::BLOCK / LANES;  // elements per block
  constexpr int BUF = NQ > NL ? (NQ > N1 * N1 * Q1 ? NQ : N1 * N1 * Q1) : NL;
  constexpr int NBUF = DIM == 3 ? 8 : 6;
  __shared__ T tabs[EPB][DIM][2][Q1 * N1];
  __shared__ T buf[EPB][NBUF][BUF + 1];
  __shared__ T wloc[EPB][NL];

  const int le = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  const int64_t e = (int64_t)blockIdx.x * EPB + le;
  const bool live = e < A.n_e;
  int ei[3] = {0, 0, 0};
  if (live) {
    int64_t rem = e;
    for (int d = 0; d < DIM; ++d) {
      ei[d] = (int)(rem % pt.n_elem[d]);
      rem /= pt.n_elem[d];
    }
  }
  // stage the element's 1-D tables and gather its coefficients
  for (int t = lane; t < DIM * Q1 * N1; t += LANES) {
    const int d = t / (Q1 * N1), o = t % (Q1 * N1);
    T tv = 0, dv = 0;
    if (live) {
      tv = reinterpret_cast<const T*>(A.tab[d])[(size_t)ei[d] * Q1 * N1 + o];
      dv = reinterpret_cast<const T*>(A.dtab[d])[(size_t)ei[d] * Q1 * N1 + o];
    }
    tabs[le][d][0][o] = tv;
    tabs[le][d][1][o] = dv;
  }
  T* u = buf[le][0];
  if (lane < NL) {
    T val = 0, wv = 1;
    if (live) {
      int ai[3] = {lane % N1, (lane / N1) % N1, lane / (N1 * N1)};
      int sdof = 0, stride = 1;
      for (int d = 0; d < DIM; ++d) {
        sdof += (pt.first[d][ei[d]] + ai[d]) * stride;
        stride *= pt.n_cp[d];
      }
      const int f = A.f2f[sdof];
      val = f >= 0 ? (T)v[f] : (T)0;
      if (RATIONAL) wv = (T)pt.wgts[sdof];
    }
    wloc[le][lane] = wv;
    u[lane] = RATIONAL ? val * wv : val;
  }
  __syncthreads();
  const T* B0 = tabs[le][0][0];
  const T* D0 = tabs[le][0][1];
  const T* B1 = tabs[le][1][0];
  const T* D1 = tabs[le][1][1];
  const T* g = geo + (size_t)(live ? e : 0) * NG * NQ;

  if constexpr (DIM == 2) {
    // u[i1][i0] -> A0 = B0 u, A0d = D0 u : [i1][q0]
    T* A0 = buf[le][1];
    T* A0d = buf[le][2];
    {
      const int ext[3] = {N1, N1, 1};
      if (lane < Q1 * N1) {
        const int x = lane % Q1, y = lane / Q1;
        A0[lane] = contract(u, B0, 0, N1, x, y, 0, ext);
        A0d[lane] = contract(u, D0, 0, N1, x, y, 0, ext);
      }
    }
    __syncthreads();
    // gradient and value at [q1][q0]
    T* g0 = buf[le][3];
    T* g1 = buf[le][4];
    T* gv = buf[le][5];
    {
      const int ext[3] = {Q1, N1, 1};
      if (lane < NQ) {
        const int x = lane % Q1, y = lane / Q1;
        const T a = contract(A0d, B1, 1, N1, x, y, 0, ext);
        const T b = contract(A0, D1, 1, N1, x, y, 0, ext);
        T val = RATIONAL ? contract(A0, B1, 1, N1, x, y, 0, ext) : (T)0;
        T gr[2] = {a, b};
        // flux
        T f[2], sc = 0;
        if (RATIONAL) {
          const T* wg = wgeo + (size_t)(live ? e : 0) * NW * NQ;
          const T W = wg[lane], dW0 = wg[NQ + lane], dW1 = wg[2 * NQ + lane];
          const T uh = val / W;
          gr[0] = (gr[0] - uh * dW0) / W;
          gr[1] = (gr[1] - uh * dW1) / W;
          f[0] = g[0 * NQ + lane] * gr[0] + g[1 * NQ + lane] * gr[1];
          f[1] = g[1 * NQ + lane] * gr[0] + g[2 * NQ + lane] * gr[1];
          sc = -(dW0 * f[0] + dW1 * f[1]) / (W * W);
          f[0] /= W;
          f[1] /= W;
        } else {
          f[0] = g[0 * NQ + lane] * gr[0] + g[1 * NQ + lane] * gr[1];
          f[1] = g[1 * NQ + lane] * gr[0] + g[2 * NQ + lane] * gr[1];
        }
        g0[lane] = f[0];
        g1[lane] = f[1];
        gv[lane] = sc;
      }
    }
    __syncthreads();
    // transpose along dir 1: [q1][q0] -> [i1][q0]
    T* C0 = buf[le][1];  // -> followed by D0^T
    T* C1 = buf[le][2];  // -> followed by B0^T
    {
      const int ext[3] = {Q1, Q1, 1};
      if (lane < Q1 * N1) {
        const int x = lane % Q1, y = lane / Q1;
        C0[lane] = contract_t(g0, B1, 1, Q1, N1, x, y, 0, ext);
        T c1 = contract_t(g1, D1, 1, Q1, N1, x, y, 0, ext);
        if (RATIONAL) c1 += contract_t(gv, B1, 1, Q1, N1, x, y, 0, ext);
        C1[lane] = c1;
      }
    }
    __syncthreads();
    if (lane < NL && live) {
      const int ext[3] = {Q1, N1, 1};
      const int x = lane % N1, y = lane / N1;
      T r = contract_t(C0, D0, 0, Q1, N1, x, y, 0, ext) + contract_t(C1, B0, 0, Q1, N1, x, y, 0, ext);
      if (RATIONAL) r *= wloc[le][lane];
      rloc[(size_t)e * NL + lane] = r;
    }
  } else {
    const T* B2 = tabs[le][2][0];
    const T* D2 = tabs[le][2][1];
    // u[i2][i1][i0] -> A0 = B0 u, A0d = D0 u : [i2][i1][q0]
    T* A0 = buf[le][1];
    T* A0d = buf[le][2];
    {
      const int ext[3] = {N1, N1, N1};
      if (lane < Q1 * N1 * N1) {
        const int x = lane % Q1, y = (lane / Q1) % N1, z = lane / (Q1 * N1);
        A0[lane] = contract(u, B0, 0, N1, x, y, z, ext);
        A0d[lane] = contract(u, D0, 0, N1, x, y, z, ext);
      }
    }
    __syncthreads();
    // dir 1: [i2][q1][q0]: A1 = B1 A0, A1x = B1 A0d, A1y = D1 A0
    T* A1 = buf[le][3];
    T* A1x = buf[le][4];
    T* A1y = buf[le][5];
    {
      const int ext[3] = {Q1, N1, N1};
      if (lane < Q1 * Q1 * N1) {
        const int x = lane % Q1, y = (lane / Q1) % Q1, z = lane / (Q1 * Q1);
        A1[lane] = contract(A0, B1, 1, N1, x, y, z, ext);
        A1x[lane] = contract(A0d, B1, 1, N1, x, y, z, ext);
        A1y[lane] = contract(A0, D1, 1, N1, x, y, z, ext);
      }
    }
    __syncthreads();
    // dir 2: gradient and value at [q2][q1][q0], then the flux
    T* f0 = buf[le][0];
    T* f1 = buf[le][1];
    T* f2 = buf[le][2];
    T* fs = buf[le][6];
    {
      const int ext[3] = {Q1, Q1, N1};
      if (lane < NQ) {
        const int x = lane % Q1, y = (lane / Q1) % Q1, z = lane / (Q1 * Q1);
        T gr[3] = {contract(A1x, B2, 2, N1, x, y, z, ext), contract(A1y, B2, 2, N1, x, y, z, ext),
                   contract(A1, D2, 2, N1, x, y, z, ext)};
        T W = 1, dW[3] = {0, 0, 0};
        if (RATIONAL) {
          const T val = contract(A1, B2, 2, N1, x, y, z, ext);
          const T* wg = wgeo + (size_t)(live ? e : 0) * NW * NQ;
          W = wg[lane];
          dW[0] = wg[NQ + lane];
          dW[1] = wg[2 * NQ + lane];
          dW[2] = wg[3 * NQ + lane];
          const T uh = val / W;
          for (int k = 0; k < 3; ++k) gr[k] = (gr[k] - uh * dW[k]) / W;
        }
        T f[3];
        for (int a = 0; a < 3; ++a) {
          T s = 0;
          for (int b = 0; b < 3; ++b) s += g[gsym<3>(a, b) * NQ + lane] * gr[b];
          f[a] = s;
        }
        T sc = 0;
        if (RATIONAL) {
          sc = -(dW[0] * f[0] + dW[1] * f[1] + dW[2] * f[2]) / (W * W);
          for (int a = 0; a < 3; ++a) f[a] /= W;
        }
        // buffers 0..2 (u, A0, A0d) were last read before the previous barrier
        f0[lane] = f[0];
        f1[lane] = f[1];
        f2[lane] = f[2];
        fs[lane] = sc;
      }
    }
    __syncthreads();
    // transpose dir 2: [q2][q1][q0] -> [i2][q1][q0]
    T* C0 = buf[le][3];  // f0 -> then B1^T, D0^T
    T* C1 = buf[le][4];  // f1 -> then D1^T, B0^T
    T* C2 = buf[le][5];  // f2 (D2^T) + fs (B2^T) -> then B1^T, B0^T
    {
      const int ext[3] = {Q1, Q1, Q1};
      if (lane < Q1 * Q1 * N1) {
        const int x = lane % Q1, y = (lane / Q1) % Q1, z = lane / (Q1 * Q1);
        C0[lane] = contract_t(f0, B2, 2, Q1, N1, x, y, z, ext);
        C1[lane] = contract_t(f1, B2, 2, Q1, N1, x, y, z, ext);
        T c2 = contract_t(f2, D2, 2, Q1, N1, x, y, z, ext);
        if (RATIONAL) c2 += contract_t(fs, B2, 2, Q1, N1, x, y, z, ext);
        C2[lane] = c2;
      }
    }
    __syncthreads();
    // transpose dir 1: [i2][q1][q0] -> [i2][i1][q0]
    T* E0 = buf[le][0];  // C0 via B1^T, then D0^T
    T* E1 = buf[le][1];  // C1 via D1^T + C2 via B1^T, then B0^T
    {
      const int ext[3] = {Q1, Q1, N1};
      if (lane < Q1 * N1 * N1) {
        const int x = lane % Q1, y = (lane / Q1) % N1, z = lane / (Q1 * N1);
        E0[lane] = contract_t(C0, B1, 1, Q1, N1, x, y, z, ext);
        E1[lane] = contract_t(C1, D1, 1, Q1, N1, x, y, z, ext) +
                   contract_t(C2, B1, 1, Q1, N1, x, y, z, ext);
      }
    }
    __syncthreads();
    if (lane < NL && live) {
      const int ext[3] = {Q1, N1, N1};
      const int x = lane % N1, y = (lane / N1) % N1, z = lane / (N1 * N1);
      T r = contract_t(E0, D0, 0, Q1, N1, x, y, z, ext) + contract_t(E1, B0, 0, Q1, N1, x, y, z, ext);
      if (RATIONAL) r *= wloc[le][lane];
      rloc[(size_t)e * NL + lane] = r;
    }
  }
}

// y[r] = sum of the local results of the elements supporting free dof r
template <int DIM, int P, typename T, typename TO>
__global__ void __launch_bounds__(256)
    k_mf_gather(iga_patch_t pt, int n_free, const int32_t* __restrict__ freed,
                const T* __restrict__ rloc, TO* __restrict__ y, const int* skip) {
  if (skip && *skip) return;
  constexpr int N1 = P + 1, NL = ipow(N1, DIM);
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_free) return;
  int s = freed[r];
  int id[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < DIM; ++d) {
    id[d] = s % pt.n_cp[d];
    s /= pt.n_cp[d];
    lo[d] = pt.elo[d][id[d]];
    hi[d] = pt.ehi[d][id[d]];
  }
  double sum = 0.0;
  for (int e2 = lo[2]; e2 <= hi[2]; ++e2)
    for (int e1 = lo[1]; e1 <= hi[1]; ++e1)
      for (int e0 = lo[0]; e0 <= hi[0]; ++e0) {
        const int ee[3] = {e0, e1, e2};
        int li = 0, stride = 1, e = 0, estride = 1;
        for (int d = 0; d < DIM; ++d) {
          li += (id[d] - pt.first[d][ee[d]]) * stride;
          stride *= N1;
          e += ee[d] * estride;
          estride *= pt.n_elem[d];
        }
        sum += (double)rloc[(size_t)e * NL + li];
      }
  y[r] = (TO)sum;
}

}  // namespace iga

// ------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------

using namespace iga;

namespace {

struct MfLayout {
  int64_t n_e, geo, wgeo, tabs[3], dtabs[3], rloc, total;  // byte offsets
};

template <int DIM, int P>
MfLayout mf_layout(const iga_patch_t& pt, int tsize, int rational) {
  using S = MfShape<DIM, P>;
  MfLayout L{};
  L.n_e = 1;
  for (int d = 0; d < DIM; ++d) L.n_e *= pt.n_elem[d];
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off += (bytes + 255) & ~(int64_t)255;
    return o;
  };
  L.geo = take(L.n_e * S::NG * S::NQ * tsize);
  L.wgeo = rational ? take(L.n_e * S::NW * S::NQ * tsize) : -1;
  for (int d = 0; d < DIM; ++d) {
    L.tabs[d] = take((int64_t)pt.n_elem[d] * S::Q1 * S::N1 * tsize);
    L.dtabs[d] = take((int64_t)pt.n_elem[d] * S::Q1 * S::N1 * tsize);
  }
  L.rloc = take(L.n_e * S::NL * tsize);
  L.total = off;
  return L;
}

template <typename F>
int by_shape(int dim, int degree, F f) {
  switch (dim * 10 + degree) {
    case 21: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{});
    case 22: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 2>{});
    case 23: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 3>{});
    case 24: return f(std::integral_constant<int, 2>{}, std::integral_constant<int, 4>{});
    case 31: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 1>{});
    case 32: return f(std::integral_constant<int, 3>{}, std::integral_constant<int, 2>{});
    default: return IGA_ERR_UNSUPPORTED;
  }
}

}  // namespace

extern "C" int64_t iga_mf_workspace(const iga_patch_t* patch, int32_t precision, int32_t rational) {
  if (!patch) return -1;
  const int ts = precision == IGA_FP32 ? 4 : 8;
  int64_t out = -1;
  by_shape(patch->dim, patch->degree, [&](auto D, auto P) {
    out = mf_layout<decltype(D)::value, decltype(P)::value>(*patch, ts, rational).total;
    return 0;
  });
  return out;
}

extern "C" int iga_mf_setup(const iga_patch_t* patch, const iga_rule_t* rule, int32_t precision,
                            int32_t rational, int32_t coef_kind, double coef_const,
                            const double* coef_points, void* work, int64_t work_bytes,
                            void* stream) {
  if (!patch || !rule || !work) return IGA_ERR_ARG;
  if (rule->n_q1 != patch->degree + 1) return IGA_ERR_UNSUPPORTED;
  if (precision != IGA_FP64 && precision != IGA_FP32) return IGA_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  return by_shape(patch->dim, patch->degree, [&](auto D, auto Pc) -> int {
    constexpr int DIM = decltype(D)::value, P = decltype(Pc)::value;
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
      const MfLayout L = mf_layout<DIM, P>(*patch, sizeof(T), rational);
      if (work_bytes < L.total) return IGA_ERR_SHAPE;
      char* base = reinterpret_cast<char*>(work);
      using S = MfShape<DIM, P>;
      const int64_t pts = L.n_e * S::NQ;
      if (pts > 0) {
        hipLaunchKernelGGL((k_mf_geometry<DIM, P, T>), dim3((unsigned)((pts + 255) / 256)), dim3(256),
                           0, st, *patch, rule->pts, rule->wts, coef_kind, coef_const, coef_points,
                           reinterpret_cast<T*>(base + L.geo),
                           rational ? reinterpret_cast<T*>(base + L.wgeo) : nullptr, L.n_e);
        IGA_CHECK_LAUNCH();
      }
      for (int d = 0; d < DIM; ++d) {
        const int n = patch->n_elem[d] * S::Q1 * S::N1;
        hipLaunchKernelGGL((k_mf_tables<P, T>), dim3((n + 255) / 256), dim3(256), 0, st,
                           patch->n_elem[d], patch->ext[d], patch->lo[d], patch->hi[d], rule->pts,
                           reinterpret_cast<T*>(base + L.tabs[d]),
                           reinterpret_cast<T*>(base + L.dtabs[d]));
        IGA_CHECK_LAUNCH();
      }
      return IGA_OK;
    };
    return precision == IGA_FP32 ? run(float{}) : run(double{});
  });
}

namespace iga {

int mf_apply_launch(const iga_patch_t* patch, int precision, int rational, const void* work,
                    int64_t work_bytes, int n_free, const int32_t* freed,
                    const int32_t* full_to_free, const double* x, double* y, const int* skip,
                    hipStream_t st) {
  if (!patch || !work || !freed || !full_to_free || !x || !y) return IGA_ERR_ARG;
  if (precision != IGA_FP64 && precision != IGA_FP32) return IGA_ERR_UNSUPPORTED;
  return by_shape(patch->dim, patch->degree, [&](auto D, auto Pc) -> int {
    constexpr int DIM = decltype(D)::value, P = decltype(Pc)::value;
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
      using S = MfShape<DIM, P>;
      const MfLayout L = mf_layout<DIM, P>(*patch, sizeof(T), rational);
      if (work_bytes < L.total) return IGA_ERR_SHAPE;
      char* base = const_cast<char*>(reinterpret_cast<const char*>(work));
      MfArgs A;
      A.n_e = L.n_e;
      A.skip = skip;
      A.f2f = full_to_free;
      for (int d = 0; d < 3; ++d) {
        A.tab[d] = d < DIM ? base + L.tabs[d] : nullptr;
        A.dtab[d] = d < DIM ? base + L.dtabs[d] : nullptr;
      }
      T* rloc = reinterpret_cast<T*>(base + L.rloc);
      constexpr int EPB = S::BLOCK / S::LANES;
      const int64_t blocks = (L.n_e + EPB - 1) / EPB;
      if (blocks > 0) {
        if (rational)
          hipLaunchKernelGGL((k_mf_apply<DIM, P, T, true>), dim3((unsigned)blocks), dim3(S::BLOCK), 0, st,
                             *patch, A, reinterpret_cast<const T*>(base + L.geo),
                             reinterpret_cast<const T*>(base + L.wgeo), x, rloc);
        else
          hipLaunchKernelGGL((k_mf_apply<DIM, P, T, false>), dim3((unsigned)blocks), dim3(S::BLOCK), 0, st,
                             *patch, A, reinterpret_cast<const T*>(base + L.geo), nullptr, x, rloc);
        IGA_CHECK_LAUNCH();
      }
      if (n_free > 0) {
        hipLaunchKernelGGL((k_mf_gather<DIM, P, T, double>), dim3((n_free + 255) / 256), dim3(256), 0,
                           st, *patch, n_free, freed, rloc, y, skip);
        IGA_CHECK_LAUNCH();
      }
      return IGA_OK;
    };
    return precision == IGA_FP32 ? run(float{}) : run(double{});
  });
}

}  // namespace iga

extern "C" int iga_mf_apply(const iga_patch_t* patch, int32_t precision, int32_t rational,
                            const void* work, int64_t work_bytes, int32_t n_free,
                            const int32_t* freed, const int32_t* full_to_free, const void* x,
                            void* y, void* stream) {
  return iga::mf_apply_launch(patch, precision, rational, work, work_bytes, n_free, freed,
                              full_to_free, reinterpret_cast<const double*>(x),
                              reinterpret_cast<double*>(y), nullptr, (hipStream_t)stream);
}
