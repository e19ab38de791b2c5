// Matrix-free, sum-factorised operator apply y = K v for the scalar Poisson
// form and for linear elasticity on a single tensor-product NURBS patch
// (gfx950).
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
//
// Linear elasticity (k_mf_apply_el): the same contractions per displacement
// component; set-up stores M = sqrt(w |det J|) J^-1 per point instead of G,
// and the flux of component c is T_ck = sum_x S_cx M_xk with S the stress
// tensor of sigma = C eps(grad u) (Voigt, engineering shear).
#include <algorithm>
#include <mutex>
#include <unordered_map>

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

template <int DIM, int P, typename T, bool ELAS>
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
  if constexpr (ELAS) {
    // M[x][k] = sqrt(w |det J|) (J^-1)[x][k]; layout [e][x * DIM + k][q]
    const double sc = sqrt(jac * fabs(det));
    for (int xx = 0; xx < DIM; ++xx)
      for (int k = 0; k < DIM; ++k)
        geo[(e * DIM * DIM + xx * DIM + k) * NQ + q] = (T)(sc * inv[xx][k]);
  } else {
    const double coef = coef_kind == IGA_COEF_POINTS ? coef_points[e * NQ + q] : coef_const;
    const double wa = jac * fabs(det) * coef;
    // G[a][b] = wa * sum_c inv[c][a] inv[c][b]; layout [e][g][q]
    for (int a = 0; a < DIM; ++a)
      for (int b = a; b < DIM; ++b) {
        double s = 0.0;
        for (int c = 0; c < DIM; ++c) s += inv[c][a] * inv[c][b];
        geo[(e * NG + gsym<DIM>(a, b)) * NQ + q] = (T)(wa * s);
      }
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

// Voigt matrix of the elastic material (row-major NS x NS, NS = 3 or 6)
struct MfElastic {
  double c[36];
};

// Linear elasticity, vec = DIM: one group of lanes per element as in
// k_mf_apply. The forward contractions run component by component into
// per-component gradient (and, rational, value) buffers; at each Gauss point
// one lane forms H = P M^T (P_ck = d u_c / d xi_k, M = sqrt(w |det J|) J^-1),
// the engineering strain, sigma = C eps and the flux T_ck = sum_x S_cx M_xk;
// the transposed contractions then run component by component. Local
// results are stored [e][i][component], the dof order.
template <int DIM, int P, typename T, bool RATIONAL>
__global__ void __launch_bounds__(256)
    k_mf_apply_el(iga_patch_t pt, MfArgs A, const MfElastic cm, const T* __restrict__ geo,
                  const T* __restrict__ wgeo, const double* __restrict__ v, T* __restrict__ rloc) {
  if (A.skip && *A.skip) return;
  using S = MfShape<DIM, P>;
  constexpr int N1 = S::N1, Q1 = S::Q1, NL = S::NL, NQ = S::NQ, NW = S::NW;
  constexpr int VEC = DIM, NM = DIM * DIM, NS = DIM == 2 ? 3 : 6;
  constexpr int LANES = S::LANES;
  constexpr int EPB = S::BLOCK / LANES;  // elements per block
  constexpr int BUF = NQ > NL ? (NQ > N1 * N1 * Q1 ? NQ : N1 * N1 * Q1) : NL;
  constexpr int NWK = DIM == 2 ? 2 : 5;      // work buffers of the contractions
  constexpr int GB = VEC + NWK;              // first per-component point buffer
  constexpr int NBUF = GB + VEC * (DIM + 1);  // u[c], work, then [c][k] (k = DIM: value)
  __shared__ T tabs[EPB][DIM][2][Q1 * N1];
  __shared__ T buf[EPB][NBUF][BUF + 1];
  __shared__ T wloc[EPB][NL];
  __shared__ double cmat[NS * NS];

  if (threadIdx.x < NS * NS) cmat[threadIdx.x] = cm.c[threadIdx.x];
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
  if (lane < NL) {
    T wv = 1;
    T val[VEC];
#pragma unroll
    for (int c = 0; c < VEC; ++c) val[c] = 0;
    if (live) {
      const int ai[3] = {lane % N1, (lane / N1) % N1, lane / (N1 * N1)};
      int sdof = 0, stride = 1;
      for (int d = 0; d < DIM; ++d) {
        sdof += (pt.first[d][ei[d]] + ai[d]) * stride;
        stride *= pt.n_cp[d];
      }
#pragma unroll
      for (int c = 0; c < VEC; ++c) {
        const int f = A.f2f[(size_t)VEC * sdof + c];
        val[c] = f >= 0 ? (T)v[f] : (T)0;
      }
      if (RATIONAL) wv = (T)pt.wgts[sdof];
    }
    wloc[le][lane] = wv;
#pragma unroll
    for (int c = 0; c < VEC; ++c) buf[le][c][lane] = RATIONAL ? val[c] * wv : val[c];
  }
  __syncthreads();
  const T* B0 = tabs[le][0][0];
  const T* D0 = tabs[le][0][1];
  const T* B1 = tabs[le][1][0];
  const T* D1 = tabs[le][1][1];
  auto G = [&](int c, int k) { return buf[le][GB + c * (DIM + 1) + k]; };

  // ---- forward: parametric gradient (and value) of every component ---------
  for (int c = 0; c < VEC; ++c) {
    const T* u = buf[le][c];
    if constexpr (DIM == 2) {
      T* A0 = buf[le][VEC];
      T* A0d = buf[le][VEC + 1];
      {
        const int ext[3] = {N1, N1, 1};
        if (lane < Q1 * N1) {
          const int x = lane % Q1, y = lane / Q1;
          A0[lane] = contract(u, B0, 0, N1, x, y, 0, ext);
          A0d[lane] = contract(u, D0, 0, N1, x, y, 0, ext);
        }
      }
      __syncthreads();
      {
        const int ext[3] = {Q1, N1, 1};
        if (lane < NQ) {
          const int x = lane % Q1, y = lane / Q1;
          G(c, 0)[lane] = contract(A0d, B1, 1, N1, x, y, 0, ext);
          G(c, 1)[lane] = contract(A0, D1, 1, N1, x, y, 0, ext);
          G(c, 2)[lane] = RATIONAL ? contract(A0, B1, 1, N1, x, y, 0, ext) : (T)0;
        }
      }
      __syncthreads();
    } else {
      const T* B2 = tabs[le][2][0];
      const T* D2 = tabs[le][2][1];
      T* A0 = buf[le][VEC];
      T* A0d = buf[le][VEC + 1];
      {
        const int ext[3] = {N1, N1, N1};
        if (lane < Q1 * N1 * N1) {
          const int x = lane % Q1, y = (lane / Q1) % N1, z = lane / (Q1 * N1);
          A0[lane] = contract(u, B0, 0, N1, x, y, z, ext);
          A0d[lane] = contract(u, D0, 0, N1, x, y, z, ext);
        }
      }
      __syncthreads();
      T* A1 = buf[le][VEC + 2];
      T* A1x = buf[le][VEC + 3];
      T* A1y = buf[le][VEC + 4];
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
      {
        const int ext[3] = {Q1, Q1, N1};
        if (lane < NQ) {
          const int x = lane % Q1, y = (lane / Q1) % Q1, z = lane / (Q1 * Q1);
          G(c, 0)[lane] = contract(A1x, B2, 2, N1, x, y, z, ext);
          G(c, 1)[lane] = contract(A1y, B2, 2, N1, x, y, z, ext);
          G(c, 2)[lane] = contract(A1, D2, 2, N1, x, y, z, ext);
          G(c, 3)[lane] = RATIONAL ? contract(A1, B2, 2, N1, x, y, z, ext) : (T)0;
        }
      }
      __syncthreads();
    }
  }

  // ---- at the Gauss points: strain, stress, parametric flux ------------------
  if (lane < NQ) {
    const T* m = geo + (size_t)(live ? e : 0) * NM * NQ;
    T M[DIM][DIM];
#pragma unroll
    for (int xx = 0; xx < DIM; ++xx)
#pragma unroll
      for (int k = 0; k < DIM; ++k) M[xx][k] = m[(xx * DIM + k) * NQ + lane];
    T Pg[VEC][DIM];
#pragma unroll
    for (int c = 0; c < VEC; ++c)
#pragma unroll
      for (int k = 0; k < DIM; ++k) Pg[c][k] = G(c, k)[lane];
    T W = 1, dW[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dW[k] = 0;
    if (RATIONAL) {
      const T* wg = wgeo + (size_t)(live ? e : 0) * NW * NQ;
      W = wg[lane];
#pragma unroll
      for (int k = 0; k < DIM; ++k) dW[k] = wg[(1 + k) * NQ + lane];
#pragma unroll
      for (int c = 0; c < VEC; ++c) {
        const T uh = G(c, DIM)[lane] / W;
#pragma unroll
        for (int k = 0; k < DIM; ++k) Pg[c][k] = (Pg[c][k] - uh * dW[k]) / W;
      }
    }
    // H[c][x] = sum_k M[x][k] P[c][k]
    T H[VEC][DIM];
#pragma unroll
    for (int c = 0; c < VEC; ++c)
#pragma unroll
      for (int xx = 0; xx < DIM; ++xx) {
        T s = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += M[xx][k] * Pg[c][k];
        H[c][xx] = s;
      }
    // engineering strain, stress (Voigt), symmetric stress tensor
    T eps[NS], sig[NS], St[DIM][DIM];
    if constexpr (DIM == 2) {
      eps[0] = H[0][0];
      eps[1] = H[1][1];
      eps[2] = H[0][1] + H[1][0];
    } else {
      eps[0] = H[0][0];
      eps[1] = H[1][1];
      eps[2] = H[2][2];
      eps[3] = H[1][2] + H[2][1];
      eps[4] = H[0][2] + H[2][0];
      eps[5] = H[0][1] + H[1][0];
    }
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      T s = 0;
#pragma unroll
      for (int b = 0; b < NS; ++b) s += (T)cmat[a * NS + b] * eps[b];
      sig[a] = s;
    }
    if constexpr (DIM == 2) {
      St[0][0] = sig[0];
      St[1][1] = sig[1];
      St[0][1] = St[1][0] = sig[2];
    } else {
      St[0][0] = sig[0];
      St[1][1] = sig[1];
      St[2][2] = sig[2];
      St[1][2] = St[2][1] = sig[3];
      St[0][2] = St[2][0] = sig[4];
      St[0][1] = St[1][0] = sig[5];
    }
#pragma unroll
    for (int c = 0; c < VEC; ++c) {
      T Tf[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        T s = 0;
#pragma unroll
        for (int xx = 0; xx < DIM; ++xx) s += St[c][xx] * M[xx][k];
        Tf[k] = s;
      }
      T sc = 0;
      if (RATIONAL) {
        T dot = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) dot += dW[k] * Tf[k];
        sc = -dot / (W * W);
#pragma unroll
        for (int k = 0; k < DIM; ++k) Tf[k] /= W;
      }
#pragma unroll
      for (int k = 0; k < DIM; ++k) G(c, k)[lane] = Tf[k];
      G(c, DIM)[lane] = sc;
    }
  }
  __syncthreads();

  // ---- transposed contractions, component by component ----------------------
  for (int c = 0; c < VEC; ++c) {
    T r = 0;
    if constexpr (DIM == 2) {
      T* C0 = buf[le][VEC];
      T* C1 = buf[le][VEC + 1];
      {
        const int ext[3] = {Q1, Q1, 1};
        if (lane < Q1 * N1) {
          const int x = lane % Q1, y = lane / Q1;
          C0[lane] = contract_t(G(c, 0), B1, 1, Q1, N1, x, y, 0, ext);
          T c1 = contract_t(G(c, 1), D1, 1, Q1, N1, x, y, 0, ext);
          if (RATIONAL) c1 += contract_t(G(c, 2), B1, 1, Q1, N1, x, y, 0, ext);
          C1[lane] = c1;
        }
      }
      __syncthreads();
      if (lane < NL) {
        const int ext[3] = {Q1, N1, 1};
        const int x = lane % N1, y = lane / N1;
        r = contract_t(C0, D0, 0, Q1, N1, x, y, 0, ext) + contract_t(C1, B0, 0, Q1, N1, x, y, 0, ext);
      }
    } else {
      const T* B2 = tabs[le][2][0];
      const T* D2 = tabs[le][2][1];
      T* C0 = buf[le][VEC];
      T* C1 = buf[le][VEC + 1];
      T* C2 = buf[le][VEC + 2];
      {
        const int ext[3] = {Q1, Q1, Q1};
        if (lane < Q1 * Q1 * N1) {
          const int x = lane % Q1, y = (lane / Q1) % Q1, z = lane / (Q1 * Q1);
          C0[lane] = contract_t(G(c, 0), B2, 2, Q1, N1, x, y, z, ext);
          C1[lane] = contract_t(G(c, 1), B2, 2, Q1, N1, x, y, z, ext);
          T c2 = contract_t(G(c, 2), D2, 2, Q1, N1, x, y, z, ext);
          if (RATIONAL) c2 += contract_t(G(c, 3), B2, 2, Q1, N1, x, y, z, ext);
          C2[lane] = c2;
        }
      }
      __syncthreads();
      T* E0 = buf[le][VEC + 3];
      T* E1 = buf[le][VEC + 4];
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
      if (lane < NL) {
        const int ext[3] = {Q1, N1, N1};
        const int x = lane % N1, y = (lane / N1) % N1, z = lane / (N1 * N1);
        r = contract_t(E0, D0, 0, Q1, N1, x, y, z, ext) + contract_t(E1, B0, 0, Q1, N1, x, y, z, ext);
      }
    }
    if (lane < NL && live) {
      if (RATIONAL) r *= wloc[le][lane];
      rloc[((size_t)e * NL + lane) * VEC + c] = r;
    }
    // the work buffers are written again by the next component
    __syncthreads();
  }
}

// y[r] = sum of the local results of the elements supporting free dof r;
// grid-stride over at most kRedBlocks blocks, with the x.y partials (one per
// block) when `parts` is set
template <int DIM, int P, typename T, typename TO, int VEC>
__global__ void __launch_bounds__(256)
    k_mf_gather(iga_patch_t pt, int n_free, const int32_t* __restrict__ freed,
                const T* __restrict__ rloc, TO* __restrict__ y, const int* skip,
                const double* __restrict__ x, double* parts) {
  if (skip && *skip) return;
  constexpr int N1 = P + 1, NL = ipow(N1, DIM);
  double dot = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n_free; r += gridDim.x * blockDim.x) {
    // free dof -> (control point, component); rloc is [e][i][component]
    const int full = freed[r];
    int s = full / VEC;
    const int comp = full - s * VEC;
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
          sum += (double)rloc[((size_t)e * NL + li) * VEC + comp];
        }
    y[r] = (TO)sum;
    if (parts) dot += sum * x[r];
  }
  if (parts) store_partial(parts, dot);
}

// ---------------------------------------------------------------------------
// separable patches: one lane per element, geometry from 1-D tables
// ---------------------------------------------------------------------------
//
// On a polynomial patch whose map is a tensor product of 1-D maps (control
// point coordinate c depends on the c-th index only, e.g. every refined box),
// J = diag(x_d'(xi_d)), so at Gauss point q = (q_0, .., q_{D-1})
//   G_q = a_q prod_k jw_k(q_k) diag(s_d(q_d)),
//   jw_d = wref(q_d) (h_d / 2) |x_d'|,  s_d = 1 / x_d'^2,
// and the square roots of these 1-D factors are folded into the 1-D tables
// of each span. The only stream per point is the coefficient a_q (none when
// it is constant), stored point-major, coef[q][e], in FP64, FP32 or BF16. A lane owns one element: it gathers the
// element's coefficients, runs the sum-factorised gradient, flux and
// transposed contractions in registers (no LDS, no barrier) and stores its
// local results local-function-major, rloc[i][e], so both the stores and the
// gather's reads of one local index are contiguous across lanes.

// BF16 storage of the coefficient: FP32 rounded to nearest even
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <typename TC>
__device__ __forceinline__ TC store_coef(double a) {
  if constexpr (std::is_same<TC, uint16_t>::value) return f32_to_bf16((float)a);
  else return (TC)a;
}

// Consecutive blocks go to the eight XCDs in turn; renumber them so that
// each XCD runs one contiguous range of elements (and keeps the coefficient
// vector entries of that range in its own L2).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t per = nb / 8;
  if (per == 0 || b >= per * 8) return b;
  return (b % 8) * per + b / 8;
}

struct MfsArgs {
  int64_t n_e;
  const int* skip;
  const int32_t* f2f;
  const void* tab[3];   // (n_elem_d, Q1, N1) scaled values
  const void* dtab[3];  // (n_elem_d, Q1, N1) scaled derivatives
  const void* coef;     // (NQ, n_e) in TC (BF16: point pairs), or null (constant)
  double coef_const;
};

// 1-D tables of direction d with the separable geometry folded in (FP64,
// stored in T): bt[q][a] = sqrt(jw(q)) N_a(q) and dt[q][a] = sqrt(jw(q) s(q))
// N_a'(q), jw = wref(q) (h/2) |x_d'|, s = 1/x_d'^2. The direction-d term of
// the element operator at point q is then a_q (dt_d (x) bt_others)^T (dt_d (x)
// bt_others): every geometric factor is in the tables, none per point.
template <int P, typename T>
__global__ void k_mfs_tables(int n_elem, int d, int dim, int64_t stride, const double* __restrict__ ext,
                             const int32_t* __restrict__ first, const double* __restrict__ lo,
                             const double* __restrict__ hi, const double* __restrict__ qp,
                             const double* __restrict__ qw, const double* __restrict__ cpts,
                             T* __restrict__ tab, T* __restrict__ dtab) {
  constexpr int N1 = P + 1, Q1 = P + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * Q1) return;
  const int q = t % Q1, e = t / Q1;
  double B[N1], dB[N1], v[N1], dv[N1];
  bernstein<P>(qp[q], B, dB);
  const double h = hi[e] - lo[e];
  double xp = 0.0;
#pragma unroll
  for (int a = 0; a < N1; ++a) {
    const double* C = ext + ((size_t)e * N1 + a) * N1;
    double s = 0.0, ds = 0.0;
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      s += C[b] * B[b];
      ds += C[b] * dB[b];
    }
    v[a] = s;
    dv[a] = ds * (2.0 / h);
    xp += dv[a] * cpts[(size_t)(first[e] + a) * stride * dim + d];
  }
  const double rj = sqrt(qw[q] * 0.5 * h * fabs(xp));
  const double rd = rj / fabs(xp);
#pragma unroll
  for (int a = 0; a < N1; ++a) {
    tab[((size_t)e * Q1 + q) * N1 + a] = (T)(rj * v[a]);
    dtab[((size_t)e * Q1 + q) * N1 + a] = (T)(rd * dv[a]);
  }
}

// coefficient at the Gauss points, (n_e, NQ) FP64 -> (NQ, n_e) TC; BF16 in
// pairs of points, word [(q/2) n_e + e] = (point q even: low half, odd: high)
template <typename TC>
__global__ void k_mfs_coef(int64_t n_e, int nq, const double* __restrict__ in, TC* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_e * nq) return;
  const int64_t e = t / nq;
  const int q = (int)(t % nq);
  if constexpr (std::is_same<TC, uint16_t>::value)
    out[((int64_t)(q >> 1) * n_e + e) * 2 + (q & 1)] = store_coef<TC>(in[t]);
  else
    out[(int64_t)q * n_e + e] = store_coef<TC>(in[t]);
}

// scalar dof of local function (a_0, .., a_{D-1}) of the element at ei
template <int DIM>
__device__ __forceinline__ int mfs_dof(const iga_patch_t& pt, const int* ei, const int* ai) {
  int s = 0, stride = 1;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int f = pt.regular ? ei[d] : pt.first[d][ei[d]];
    s += (f + ai[d]) * stride;
    stride *= pt.n_cp[d];
  }
  return s;
}

// MODE 0: y_loc = K_e u_loc; MODE 1: diag(K_e) (squared tables, no input)
template <int DIM, int P, typename T, typename TC, int MODE>
__global__ void __launch_bounds__(256)
    k_mfs_apply(iga_patch_t pt, MfsArgs A, const double* __restrict__ v, T* __restrict__ rloc) {
  if (A.skip && *A.skip) return;
  constexpr int N1 = P + 1, Q1 = P + 1, NL = ipow(N1, DIM);
  // a block covers a run of elements along direction 0 at fixed (e_1, e_2):
  // the direction-1 and -2 spans, and so their tables, are uniform across
  // the block (scalar loads); only direction 0 varies per lane
  const int n0 = pt.n_elem[0];
  const int nb0 = (n0 + blockDim.x - 1) / blockDim.x;
  const int64_t b = xcd_block(blockIdx.x, gridDim.x);
  const int64_t row = b / nb0;
  const int e0 = (int)(b % nb0) * blockDim.x + threadIdx.x;
  if (e0 >= n0) return;
  const int64_t e = row * n0 + e0;
  int ei[3] = {e0, 0, 0};
  if constexpr (DIM == 2) {
    ei[1] = (int)row;
  } else {
    ei[1] = (int)(row % pt.n_elem[1]);
    ei[2] = (int)(row / pt.n_elem[1]);
  }
  // scaled 1-D tables of the element's spans (squared for the diagonal)
  T B[DIM][Q1][N1], D[DIM][Q1][N1];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const T* tb = reinterpret_cast<const T*>(A.tab[d]) + (size_t)ei[d] * Q1 * N1;
    const T* td = reinterpret_cast<const T*>(A.dtab[d]) + (size_t)ei[d] * Q1 * N1;
#pragma unroll
    for (int q = 0; q < Q1; ++q)
#pragma unroll
      for (int a = 0; a < N1; ++a) {
        T bv = tb[q * N1 + a], dv = td[q * N1 + a];
        if (MODE == 1) {
          bv *= bv;
          dv *= dv;
        }
        B[d][q][a] = bv;
        D[d][q][a] = dv;
      }
  }
  auto coef_at = [&](int q) -> T {
    if (!A.coef) return (T)A.coef_const;
    if constexpr (std::is_same<TC, uint16_t>::value) {
      const uint32_t w = reinterpret_cast<const uint32_t*>(A.coef)[(int64_t)(q >> 1) * A.n_e + e];
      return (T)bf16_to_f32((uint16_t)((q & 1) ? (w >> 16) : (w & 0xFFFFu)));
    } else {
      return (T) reinterpret_cast<const TC*>(A.coef)[(int64_t)q * A.n_e + e];
    }
  };
  T r[NL];
  if constexpr (DIM == 2) {
    T A0[N1][Q1], A0d[N1][Q1];  // [i1][q0]
    if constexpr (MODE == 0) {
      T u[N1][N1];
#pragma unroll
      for (int i1 = 0; i1 < N1; ++i1)
#pragma unroll
        for (int i0 = 0; i0 < N1; ++i0) {
          const int ai[2] = {i0, i1};
          const int f = A.f2f[mfs_dof<2>(pt, ei, ai)];
          // unconditional load from a clamped index: all loads issue at once
          const double val = v[f >= 0 ? f : 0];
          u[i1][i0] = f >= 0 ? (T)val : (T)0;
        }
#pragma unroll
      for (int i1 = 0; i1 < N1; ++i1)
#pragma unroll
        for (int q0 = 0; q0 < Q1; ++q0) {
          T s = 0, sd = 0;
#pragma unroll
          for (int i0 = 0; i0 < N1; ++i0) {
            s += B[0][q0][i0] * u[i1][i0];
            sd += D[0][q0][i0] * u[i1][i0];
          }
          A0[i1][q0] = s;
          A0d[i1][q0] = sd;
        }
    }
    T C0[N1][Q1], C1[N1][Q1];
#pragma unroll
    for (int i1 = 0; i1 < N1; ++i1)
#pragma unroll
      for (int q0 = 0; q0 < Q1; ++q0) C0[i1][q0] = C1[i1][q0] = 0;
#pragma unroll
    for (int q1 = 0; q1 < Q1; ++q1)
#pragma unroll
      for (int q0 = 0; q0 < Q1; ++q0) {
        const T c = coef_at(q1 * Q1 + q0);
        T f0, f1;
        if constexpr (MODE == 0) {
          T g0 = 0, g1 = 0;
#pragma unroll
          for (int i1 = 0; i1 < N1; ++i1) {
            g0 += B[1][q1][i1] * A0d[i1][q0];
            g1 += D[1][q1][i1] * A0[i1][q0];
          }
          f0 = c * g0;
          f1 = c * g1;
        } else {
          f0 = f1 = c;
        }
#pragma unroll
        for (int i1 = 0; i1 < N1; ++i1) {
          C0[i1][q0] += B[1][q1][i1] * f0;
          C1[i1][q0] += D[1][q1][i1] * f1;
        }
      }
#pragma unroll
    for (int i1 = 0; i1 < N1; ++i1)
#pragma unroll
      for (int i0 = 0; i0 < N1; ++i0) {
        T s = 0;
#pragma unroll
        for (int q0 = 0; q0 < Q1; ++q0) s += D[0][q0][i0] * C0[i1][q0] + B[0][q0][i0] * C1[i1][q0];
        r[i1 * N1 + i0] = s;
      }
  } else {
    // one 1-D Gauss point q0 at a time: direction 0 at q0, then directions 1
    // and 2, the flux and the transposed directions 2 and 1 on that slice,
    // accumulated into r through the transposed direction 0 (fewer values
    // live at once than whole-element stages)
    T u[N1][N1][N1];
    if constexpr (MODE == 0) {
#pragma unroll
      for (int i2 = 0; i2 < N1; ++i2)
#pragma unroll
        for (int i1 = 0; i1 < N1; ++i1)
#pragma unroll
          for (int i0 = 0; i0 < N1; ++i0) {
            const int ai[3] = {i0, i1, i2};
            const int f = A.f2f[mfs_dof<3>(pt, ei, ai)];
            // unconditional load from a clamped index: all loads issue at once
            const double val = v[f >= 0 ? f : 0];
            u[i2][i1][i0] = f >= 0 ? (T)val : (T)0;
          }
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) r[i] = 0;
#pragma unroll
    for (int q0 = 0; q0 < Q1; ++q0) {
      T A1[N1][Q1], A1x[N1][Q1], A1y[N1][Q1];  // [i2][q1] at this q0
      if constexpr (MODE == 0) {
#pragma unroll
        for (int i2 = 0; i2 < N1; ++i2) {
          T a0[N1], ad[N1];
#pragma unroll
          for (int i1 = 0; i1 < N1; ++i1) {
            T s = 0, sd = 0;
#pragma unroll
            for (int i0 = 0; i0 < N1; ++i0) {
              s += B[0][q0][i0] * u[i2][i1][i0];
              sd += D[0][q0][i0] * u[i2][i1][i0];
            }
            a0[i1] = s;
            ad[i1] = sd;
          }
#pragma unroll
          for (int q1 = 0; q1 < Q1; ++q1) {
            T a = 0, ax = 0, ay = 0;
#pragma unroll
            for (int i1 = 0; i1 < N1; ++i1) {
              a += B[1][q1][i1] * a0[i1];
              ax += B[1][q1][i1] * ad[i1];
              ay += D[1][q1][i1] * a0[i1];
            }
            A1[i2][q1] = a;
            A1x[i2][q1] = ax;
            A1y[i2][q1] = ay;
          }
        }
      }
      T C0[N1][Q1], C1[N1][Q1], C2[N1][Q1];  // [i2][q1]
#pragma unroll
      for (int q1 = 0; q1 < Q1; ++q1) {
        T f0[Q1], f1[Q1], f2[Q1];
#pragma unroll
        for (int q2 = 0; q2 < Q1; ++q2) {
          const T c = coef_at((q2 * Q1 + q1) * Q1 + q0);
          if constexpr (MODE == 0) {
            T g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
            for (int i2 = 0; i2 < N1; ++i2) {
              g0 += B[2][q2][i2] * A1x[i2][q1];
              g1 += B[2][q2][i2] * A1y[i2][q1];
              g2 += D[2][q2][i2] * A1[i2][q1];
            }
            f0[q2] = c * g0;
            f1[q2] = c * g1;
            f2[q2] = c * g2;
          } else {
            f0[q2] = f1[q2] = f2[q2] = c;
          }
        }
#pragma unroll
        for (int i2 = 0; i2 < N1; ++i2) {
          T c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
          for (int q2 = 0; q2 < Q1; ++q2) {
            c0 += B[2][q2][i2] * f0[q2];
            c1 += B[2][q2][i2] * f1[q2];
            c2 += D[2][q2][i2] * f2[q2];
          }
          C0[i2][q1] = c0;
          C1[i2][q1] = c1;
          C2[i2][q1] = c2;
        }
      }
#pragma unroll
      for (int i2 = 0; i2 < N1; ++i2)
#pragma unroll
        for (int i1 = 0; i1 < N1; ++i1) {
          T e0 = 0, e1 = 0;
#pragma unroll
          for (int q1 = 0; q1 < Q1; ++q1) {
            e0 += B[1][q1][i1] * C0[i2][q1];
            e1 += D[1][q1][i1] * C1[i2][q1] + B[1][q1][i1] * C2[i2][q1];
          }
#pragma unroll
          for (int i0 = 0; i0 < N1; ++i0)
            r[(i2 * N1 + i1) * N1 + i0] += D[0][q0][i0] * e0 + B[0][q0][i0] * e1;
        }
    }
  }
#pragma unroll
  for (int i = 0; i < NL; ++i) rloc[(int64_t)i * A.n_e + e] = r[i];
}

// y[r] = sum over the elements supporting free dof r of their local result
// (local-function-major rloc), in a fixed order
template <int DIM, int P, typename T>
__device__ __forceinline__ double mfs_gather_one(const iga_patch_t& pt, int64_t n_e, int s,
                                                 const T* __restrict__ rloc) {
  constexpr int N1 = P + 1;
  int id[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    id[d] = s % pt.n_cp[d];
    s /= pt.n_cp[d];
  }
  double sum = 0.0;
  if (pt.regular) {
    // function i is supported on spans i-P .. i with local index P-k on span
    // i-P+k: every (k_0, .., k_{D-1}) is read unconditionally from a clamped
    // address and the absent spans are multiplied out, so all loads are in
    // flight together; the order of the sum is fixed
    constexpr int NS = ipow(N1, DIM);
    T vals[NS];
    bool ok[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      const int k[3] = {t % N1, (t / N1) % N1, t / (N1 * N1)};
      int64_t e = 0, estride = 1;
      int li = 0, stride = 1;
      bool in = true;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int ed = id[d] - P + k[d];
        in = in && ed >= 0 && ed < pt.n_elem[d];
        e += (int64_t)(ed < 0 ? 0 : ed >= pt.n_elem[d] ? pt.n_elem[d] - 1 : ed) * estride;
        estride *= pt.n_elem[d];
        li += (P - k[d]) * stride;
        stride *= N1;
      }
      vals[t] = rloc[(int64_t)li * n_e + e];
      ok[t] = in;
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) sum += ok[t] ? (double)vals[t] : 0.0;
    return sum;
  }
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = pt.elo[d][id[d]];
    hi[d] = pt.ehi[d][id[d]];
  }
  for (int e2 = lo[2]; e2 <= hi[2]; ++e2)
    for (int e1 = lo[1]; e1 <= hi[1]; ++e1)
      for (int e0 = lo[0]; e0 <= hi[0]; ++e0) {
        const int ee[3] = {e0, e1, e2};
        int li = 0, stride = 1;
        int64_t e = 0, estride = 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          li += (id[d] - pt.first[d][ee[d]]) * stride;
          stride *= N1;
          e += ee[d] * estride;
          estride *= pt.n_elem[d];
        }
        sum += (double)rloc[(int64_t)li * n_e + e];
      }
  return sum;
}

// y[r] = sum over the elements supporting free dof r of their local result
// (local-function-major rloc), in a fixed order; grid-stride over at most
// kRedBlocks blocks, with the x.y partials (one per block) when `parts` is set
template <int DIM, int P, typename T>
__global__ void __launch_bounds__(256)
    k_mfs_gather(iga_patch_t pt, int n_free, int64_t n_e, const int32_t* __restrict__ freed,
                 const T* __restrict__ rloc, double* __restrict__ y, const int* skip,
                 const double* __restrict__ x, double* parts) {
  if (skip && *skip) return;
  double dot = 0.0;
  const int64_t b = xcd_block(blockIdx.x, gridDim.x);
  for (int64_t r = b * blockDim.x + threadIdx.x; r < n_free; r += (int64_t)gridDim.x * blockDim.x) {
    const double sum = mfs_gather_one<DIM, P, T>(pt, n_e, freed[r], rloc);
    y[r] = sum;
    if (parts) dot += sum * x[r];
  }
  if (parts) store_partial(parts, dot);
}

}  // namespace iga

// ------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------

using namespace iga;

namespace {

// Workspace header: what iga_mf_setup built. The library also keeps a copy
// keyed by the workspace address, so that an apply can check (without a
// device read) that the workspace was set up, and for which layout.
struct MfInfo {
  int32_t magic, dim, degree, precision, flags, coef_kind;
  int64_t n_e, bytes;
  double coef_const;
  double cmat[36];  // elasticity: the Voigt matrix
};
constexpr int32_t kMfMagic = 0x4d465731;  // "MFW1"
constexpr int64_t kMfHeader = 512;
static_assert(sizeof(MfInfo) <= kMfHeader, "the workspace header holds MfInfo");

std::mutex g_mf_mu;
std::unordered_map<const void*, MfInfo> g_mf;

struct MfLayout {
  int64_t n_e, geo, wgeo, tabs[3], dtabs[3], coef, rloc, total;  // byte offsets
};

// general layout: G (NG per point) and, when rational, W and grad W per point;
// separable layout: 1-D tables (values, derivatives, jw, s) and the
// coefficient per point in the storage type of `precision`
template <int DIM, int P>
MfLayout mf_layout(const iga_patch_t& pt, int precision, int flags) {
  using S = MfShape<DIM, P>;
  const int ts = precision == IGA_FP64 ? 8 : 4;  // arithmetic type
  const int cs = precision == IGA_FP64 ? 8 : precision == IGA_FP32 ? 4 : 2;  // coefficient storage
  MfLayout L{};
  for (int d = 0; d < 3; ++d) L.tabs[d] = L.dtabs[d] = -1;
  L.geo = L.wgeo = L.coef = -1;
  L.n_e = 1;
  for (int d = 0; d < DIM; ++d) L.n_e *= pt.n_elem[d];
  int64_t off = kMfHeader;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off += (bytes + 255) & ~(int64_t)255;
    return o;
  };
  if (flags & IGA_MF_SEPARABLE) {
    for (int d = 0; d < DIM; ++d) {
      L.tabs[d] = take((int64_t)pt.n_elem[d] * S::Q1 * S::N1 * ts);
      L.dtabs[d] = take((int64_t)pt.n_elem[d] * S::Q1 * S::N1 * ts);
    }
    L.coef = take(L.n_e * (cs == 2 ? (S::NQ + 1) / 2 * 4 : S::NQ * cs));
  } else {
    const int ng = (flags & IGA_MF_ELASTICITY) ? DIM * DIM : S::NG;
    L.geo = take(L.n_e * ng * S::NQ * ts);
    L.wgeo = (flags & IGA_MF_RATIONAL) ? take(L.n_e * S::NW * S::NQ * ts) : -1;
    for (int d = 0; d < DIM; ++d) {
      L.tabs[d] = take((int64_t)pt.n_elem[d] * S::Q1 * S::N1 * ts);
      L.dtabs[d] = take((int64_t)pt.n_elem[d] * S::Q1 * S::N1 * ts);
    }
  }
  L.rloc = take(L.n_e * S::NL * ((flags & IGA_MF_ELASTICITY) ? DIM : 1) * ts);
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

// valid combinations: FP64 / FP32 with either layout (rational only in the
// general one), BF16 (BF16 coefficient, FP32 arithmetic) separable only
int check_flags(int precision, int flags) {
  if (flags & ~(IGA_MF_RATIONAL | IGA_MF_SEPARABLE | IGA_MF_ELASTICITY)) return IGA_ERR_ARG;
  if ((flags & IGA_MF_RATIONAL) && (flags & IGA_MF_SEPARABLE)) return IGA_ERR_ARG;
  // elasticity: the general layout, FP64 or FP32
  if ((flags & IGA_MF_ELASTICITY) && (flags & IGA_MF_SEPARABLE)) return IGA_ERR_UNSUPPORTED;
  if (precision == IGA_FP64 || precision == IGA_FP32) return IGA_OK;
  if (precision == IGA_BF16 && (flags & IGA_MF_SEPARABLE)) return IGA_OK;
  return IGA_ERR_UNSUPPORTED;
}

int lookup(const iga_patch_t* patch, const void* work, int64_t work_bytes, MfInfo* out) {
  std::lock_guard<std::mutex> lock(g_mf_mu);
  auto it = g_mf.find(work);
  if (it == g_mf.end()) return IGA_ERR_ARG;  // not set up by iga_mf_setup
  const MfInfo& I = it->second;
  int64_t n_e = 1;
  for (int d = 0; d < patch->dim; ++d) n_e *= patch->n_elem[d];
  if (I.dim != patch->dim || I.degree != patch->degree || I.n_e != n_e || work_bytes < I.bytes)
    return IGA_ERR_SHAPE;
  *out = I;
  return IGA_OK;
}

}  // namespace

extern "C" int64_t iga_mf_workspace(const iga_patch_t* patch, int32_t precision, int32_t flags) {
  if (!patch || check_flags(precision, flags)) return -1;
  int64_t out = -1;
  by_shape(patch->dim, patch->degree, [&](auto D, auto P) {
    out = mf_layout<decltype(D)::value, decltype(P)::value>(*patch, precision, flags).total;
    return 0;
  });
  return out;
}

static int mf_setup(const iga_patch_t* patch, const iga_rule_t* rule, int32_t precision,
                    int32_t flags, int32_t coef_kind, double coef_const, const double* coef_points,
                    const double* cmat, void* work, int64_t work_bytes, void* stream) {
  if (!patch || !rule || !work) return IGA_ERR_ARG;
  if (rule->n_q1 != patch->degree + 1) return IGA_ERR_UNSUPPORTED;
  if (int rc = check_flags(precision, flags)) return rc;
  if (coef_kind == IGA_COEF_POINTS && !coef_points) return IGA_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int rational = flags & IGA_MF_RATIONAL;
  int rc = by_shape(patch->dim, patch->degree, [&](auto D, auto Pc) -> int {
    constexpr int DIM = decltype(D)::value, P = decltype(Pc)::value;
    using S = MfShape<DIM, P>;
    const MfLayout L = mf_layout<DIM, P>(*patch, precision, flags);
    if (work_bytes < L.total) return IGA_ERR_SHAPE;
    char* base = reinterpret_cast<char*>(work);
    const int64_t pts = L.n_e * S::NQ;
    if (flags & IGA_MF_SEPARABLE) {
      auto tables = [&](auto tag) -> int {
        using T = decltype(tag);
        int64_t stride = 1;
        for (int d = 0; d < DIM; ++d) {
          const int n = patch->n_elem[d] * S::Q1;
          hipLaunchKernelGGL((k_mfs_tables<P, T>), dim3((n + 255) / 256), dim3(256), 0, st,
                             patch->n_elem[d], d, DIM, stride, patch->ext[d], patch->first[d],
                             patch->lo[d], patch->hi[d], rule->pts, rule->wts, patch->cpts,
                             reinterpret_cast<T*>(base + L.tabs[d]),
                             reinterpret_cast<T*>(base + L.dtabs[d]));
          IGA_CHECK_LAUNCH();
          stride *= patch->n_cp[d];
        }
        return IGA_OK;
      };
      if (int rc2 = precision == IGA_FP64 ? tables(double{}) : tables(float{})) return rc2;
      if (coef_kind == IGA_COEF_POINTS && pts > 0) {
        const unsigned g = (unsigned)((pts + 255) / 256);
        if (precision == IGA_FP64)
          hipLaunchKernelGGL(k_mfs_coef<double>, dim3(g), dim3(256), 0, st, L.n_e, S::NQ, coef_points,
                             reinterpret_cast<double*>(base + L.coef));
        else if (precision == IGA_FP32)
          hipLaunchKernelGGL(k_mfs_coef<float>, dim3(g), dim3(256), 0, st, L.n_e, S::NQ, coef_points,
                             reinterpret_cast<float*>(base + L.coef));
        else
          hipLaunchKernelGGL(k_mfs_coef<uint16_t>, dim3(g), dim3(256), 0, st, L.n_e, S::NQ,
                             coef_points, reinterpret_cast<uint16_t*>(base + L.coef));
        IGA_CHECK_LAUNCH();
      }
      return IGA_OK;
    }
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
      if (pts > 0) {
        if (flags & IGA_MF_ELASTICITY)
          hipLaunchKernelGGL((k_mf_geometry<DIM, P, T, true>), dim3((unsigned)((pts + 255) / 256)),
                             dim3(256), 0, st, *patch, rule->pts, rule->wts, coef_kind, coef_const,
                             coef_points, reinterpret_cast<T*>(base + L.geo),
                             rational ? reinterpret_cast<T*>(base + L.wgeo) : nullptr, L.n_e);
        else
          hipLaunchKernelGGL((k_mf_geometry<DIM, P, T, false>), dim3((unsigned)((pts + 255) / 256)),
                             dim3(256), 0, st, *patch, rule->pts, rule->wts, coef_kind, coef_const,
                             coef_points, reinterpret_cast<T*>(base + L.geo),
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
  if (rc) return rc;
  MfInfo I{};
  I.magic = kMfMagic;
  I.dim = patch->dim;
  I.degree = patch->degree;
  I.precision = precision;
  I.flags = flags;
  I.coef_kind = coef_kind;
  I.coef_const = coef_const;
  if (cmat) {
    const int ns = patch->dim == 2 ? 3 : 6;
    for (int t = 0; t < ns * ns; ++t) I.cmat[t] = cmat[t];
  }
  I.n_e = 1;
  for (int d = 0; d < patch->dim; ++d) I.n_e *= patch->n_elem[d];
  I.bytes = iga_mf_workspace(patch, precision, flags);
  // the header travels with the workspace (stream-ordered copy from the
  // library's registry entry, which outlives the call)
  {
    std::lock_guard<std::mutex> lock(g_mf_mu);
    auto& slot = g_mf[work];
    slot = I;
    if (hipMemcpyAsync(work, &slot, sizeof(MfInfo), hipMemcpyHostToDevice, st) != hipSuccess)
      return IGA_ERR_HIP;
  }
  return IGA_OK;
}

extern "C" int iga_mf_setup(const iga_patch_t* patch, const iga_rule_t* rule, int32_t precision,
                            int32_t flags, int32_t coef_kind, double coef_const,
                            const double* coef_points, void* work, int64_t work_bytes,
                            void* stream) {
  if (flags & IGA_MF_ELASTICITY) return IGA_ERR_ARG;  // iga_mf_setup_elasticity
  return mf_setup(patch, rule, precision, flags, coef_kind, coef_const, coef_points, nullptr, work,
                  work_bytes, stream);
}

extern "C" int iga_mf_setup_elasticity(const iga_patch_t* patch, const iga_rule_t* rule,
                                       int32_t precision, int32_t flags, const double* cmat,
                                       void* work, int64_t work_bytes, void* stream) {
  if (!cmat) return IGA_ERR_ARG;
  return mf_setup(patch, rule, precision, flags | IGA_MF_ELASTICITY, IGA_COEF_CONST, 1.0, nullptr,
                  cmat, work, work_bytes, stream);
}

namespace iga {

static int mf_launch(const iga_patch_t* patch, const MfInfo& I, const void* work, int n_free,
                     const int32_t* freed, const int32_t* full_to_free, const double* x, double* y,
                     const int* skip, int mode, double* parts, int* nparts, hipStream_t st) {
  // gather blocks: one row per lane, at most kRedBlocks blocks (grid-stride)
  const int gb = (int)std::max<int64_t>(1, std::min<int64_t>((n_free + 255) / 256, kRedBlocks));
  if (nparts) *nparts = gb;
  return by_shape(patch->dim, patch->degree, [&](auto D, auto Pc) -> int {
    constexpr int DIM = decltype(D)::value, P = decltype(Pc)::value;
    using S = MfShape<DIM, P>;
    const MfLayout L = mf_layout<DIM, P>(*patch, I.precision, I.flags);
    char* base = const_cast<char*>(reinterpret_cast<const char*>(work));
    if (I.flags & IGA_MF_SEPARABLE) {
      MfsArgs A{};
      A.n_e = L.n_e;
      A.skip = skip;
      A.f2f = full_to_free;
      for (int d = 0; d < DIM; ++d) {
        A.tab[d] = base + L.tabs[d];
        A.dtab[d] = base + L.dtabs[d];
      }
      A.coef = I.coef_kind == IGA_COEF_POINTS ? base + L.coef : nullptr;
      A.coef_const = I.coef_const;
      auto run = [&](auto ttag, auto ctag) -> int {
        using T = decltype(ttag);
        using TC = decltype(ctag);
        T* rloc = reinterpret_cast<T*>(base + L.rloc);
        // blocks of 64..256 lanes along direction 0, one run of a row each
        const int n0 = patch->n_elem[0];
        const int bw = std::min(256, (n0 + 63) / 64 * 64);
        const int64_t blocks = L.n_e > 0 ? (L.n_e / n0) * ((n0 + bw - 1) / bw) : 0;
        if (blocks > 0) {
          if (mode == 0)
            hipLaunchKernelGGL((k_mfs_apply<DIM, P, T, TC, 0>), dim3((unsigned)blocks), dim3(bw), 0,
                               st, *patch, A, x, rloc);
          else
            hipLaunchKernelGGL((k_mfs_apply<DIM, P, T, TC, 1>), dim3((unsigned)blocks), dim3(bw), 0,
                               st, *patch, A, x, rloc);
          IGA_CHECK_LAUNCH();
        }
        if (n_free > 0) {
          hipLaunchKernelGGL((k_mfs_gather<DIM, P, T>), dim3(gb), dim3(256), 0, st, *patch, n_free,
                             L.n_e, freed, rloc, y, skip, x, parts);
          IGA_CHECK_LAUNCH();
        }
        return IGA_OK;
      };
      if (I.precision == IGA_FP64) return run(double{}, double{});
      if (I.precision == IGA_FP32) return run(float{}, float{});
      return run(float{}, uint16_t{});
    }
    if (mode != 0) return (int)IGA_ERR_UNSUPPORTED;  // the diagonal is sum-factorised on separable patches
    const bool rational = I.flags & IGA_MF_RATIONAL;
    if (I.flags & IGA_MF_ELASTICITY) {
      auto run_el = [&](auto tag) -> int {
        using T = decltype(tag);
        MfArgs A;
        A.n_e = L.n_e;
        A.skip = skip;
        A.f2f = full_to_free;
        for (int d = 0; d < 3; ++d) {
          A.tab[d] = d < DIM ? base + L.tabs[d] : nullptr;
          A.dtab[d] = d < DIM ? base + L.dtabs[d] : nullptr;
        }
        MfElastic cm;
        for (int t = 0; t < 36; ++t) cm.c[t] = I.cmat[t];
        T* rloc = reinterpret_cast<T*>(base + L.rloc);
        constexpr int EPB = S::BLOCK / S::LANES;
        const int64_t blocks = (L.n_e + EPB - 1) / EPB;
        if (blocks > 0) {
          if (rational)
            hipLaunchKernelGGL((k_mf_apply_el<DIM, P, T, true>), dim3((unsigned)blocks), dim3(S::BLOCK),
                               0, st, *patch, A, cm, reinterpret_cast<const T*>(base + L.geo),
                               reinterpret_cast<const T*>(base + L.wgeo), x, rloc);
          else
            hipLaunchKernelGGL((k_mf_apply_el<DIM, P, T, false>), dim3((unsigned)blocks), dim3(S::BLOCK),
                               0, st, *patch, A, cm, reinterpret_cast<const T*>(base + L.geo), nullptr,
                               x, rloc);
          IGA_CHECK_LAUNCH();
        }
        if (n_free > 0) {
          hipLaunchKernelGGL((k_mf_gather<DIM, P, T, double, DIM>), dim3(gb), dim3(256), 0, st,
                             *patch, n_free, freed, rloc, y, skip, x, parts);
          IGA_CHECK_LAUNCH();
        }
        return IGA_OK;
      };
      return I.precision == IGA_FP32 ? run_el(float{}) : run_el(double{});
    }
    auto run = [&](auto tag) -> int {
      using T = decltype(tag);
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
        hipLaunchKernelGGL((k_mf_gather<DIM, P, T, double, 1>), dim3(gb), dim3(256), 0, st, *patch,
                           n_free, freed, rloc, y, skip, x, parts);
        IGA_CHECK_LAUNCH();
      }
      return IGA_OK;
    };
    return I.precision == IGA_FP32 ? run(float{}) : run(double{});
  });
}

int mf_apply_launch(const iga_patch_t* patch, const void* work, int64_t work_bytes, int n_free,
                    const int32_t* freed, const int32_t* full_to_free, const double* x, double* y,
                    const int* skip, double* parts, int* nparts, hipStream_t st) {
  if (!patch || !work || !freed || !full_to_free || !x || !y) return IGA_ERR_ARG;
  MfInfo I;
  if (int rc = lookup(patch, work, work_bytes, &I)) return rc;
  return mf_launch(patch, I, work, n_free, freed, full_to_free, x, y, skip, 0, parts, nparts, st);
}

}  // namespace iga

extern "C" int iga_mf_apply(const iga_patch_t* patch, const void* work, int64_t work_bytes,
                            int32_t n_free, const int32_t* freed, const int32_t* full_to_free,
                            const void* x, void* y, void* stream) {
  return iga::mf_apply_launch(patch, work, work_bytes, n_free, freed, full_to_free,
                              reinterpret_cast<const double*>(x), reinterpret_cast<double*>(y),
                              nullptr, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int iga_mf_diag(const iga_patch_t* patch, const void* work, int64_t work_bytes,
                           int32_t n_free, const int32_t* freed, const int32_t* full_to_free,
                           double* diag, void* stream) {
  if (!patch || !work || !freed || !full_to_free || !diag) return IGA_ERR_ARG;
  MfInfo I;
  if (int rc = lookup(patch, work, work_bytes, &I)) return rc;
  if (!(I.flags & IGA_MF_SEPARABLE)) return IGA_ERR_UNSUPPORTED;
  // the input vector is not read in diagonal mode; pass diag as a valid address
  return iga::mf_launch(patch, I, work, n_free, freed, full_to_free, diag, diag, nullptr, 1,
                        nullptr, nullptr, (hipStream_t)stream);
}
