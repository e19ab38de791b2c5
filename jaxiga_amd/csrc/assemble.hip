// Galerkin assembly for a single tensor-product NURBS patch on gfx950, in two
// streaming passes with no atomics on the matrix values.
//
// Pass 1, element matrices. Planar patches with p <= 3 and at most four
// Gauss points per direction (k_element_w): one wavefront per element, lane
// l holding local function l % 16 at Gauss point 4k + l / 16 in step k --
// the operand layout of the 16x16x4 MFMA -- so the physical gradients feed
// v_mfma_f64_16x16x4 / v_mfma_f32_16x16x4 from registers; the FP64 rational
// geometry is sum-factorised once per Gauss point. Other shapes (k_element):
// one workgroup per element with the basis in LDS. Both write K_e in the
// gradient-product form K_e = sum_cd E^cd (x) G_c^T W G_d, and accumulate the
// load (source work) and the Dirichlet lifting -K_fp g.
//
// Pass 2, CSR values (k_gather_points): one wavefront per control point sums
// the element-matrix rows of its supporting elements into a stencil box in
// LDS in a fixed order and writes the box out in CSR order, with the slots
// from a ballot over the present columns. This is the reference's scatter-add
// into unique CSR slots, realised as a deterministic gather: the assembled
// values are bitwise reproducible run to run. k_gather_rows (one wavefront
// per row) and k_gather (one lane per nonzero) serve vector Poisson and
// callers without row pointers.
#include "common.hpp"

namespace iga {

struct AsmDev {
  int problem, vec, precision;
  int coef_kind;
  double coef_const;
  const double* coef_points;
  double cmat[36];
  int src_kind;
  double src_const[3];
  const double* src_points;
  const int32_t* f2f;
  const double* gfull;
  double* rhs;
  double* diag;
  int n_free;  // rhs/diag stride between batch instances
  // E^cd_ab = sum_st C_st Bs(s; a, c) Bs(t; b, d), index ((c*dim + d)*dim + a)*dim + b
  double ecd[81];
  // planar elasticity with no normal-shear coupling (C_02 = C_12 = 0): only
  // eight of the sixteen E^cd_ab are nonzero
  int ortho;
  float ecd32[16];  // planar E^cd_ab rounded to FP32 (FP32 combination)
};

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Output mapping of the 16x16x4 MFMA accumulators: lane l holds column l%16;
// register r holds row l/16 + 4r for the f64 instruction and 4*(l/16) + r for
// the f32 instruction. The A operand of lane l is A[l%16][l/16], the B
// operand B[l/16][l%16], for both.
template <>
struct Mfma<double> {
  using acc_t = d4;
  static __device__ __forceinline__ int row(int k4, int r) { return k4 + 4 * r; }
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  using acc_t = f4;
  static __device__ __forceinline__ int row(int k4, int r) { return 4 * k4 + r; }
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Voigt pair (a, b) of strain component s: 2-D (xx, yy, xy), 3-D
// (xx, yy, zz, yz, xz, xy) -- engineering shear.
template <int DIM>
__device__ __forceinline__ void voigt(int s, int& a, int& b) {
  if constexpr (DIM == 2) {
    a = (s == 1) ? 1 : 0;
    b = (s == 0) ? 0 : 1;
  } else {
    const int A[6] = {0, 1, 2, 1, 0, 0};
    const int Bv[6] = {0, 1, 2, 2, 2, 1};
    a = A[s];
    b = Bv[s];
  }
}

// Strain component s of basis function i moving in component c.
template <int DIM>
__device__ __forceinline__ double strain(const double* g, int c, int s) {
  int a, b;
  voigt<DIM>(s, a, b);
  double v = 0.0;
  if (c == a) v += g[b];
  if (a != b && c == b) v += g[a];
  return v;
}

template <int DIM, int P, int PROB>
struct Shape {
  static constexpr int N1 = P + 1;
  static constexpr int NL = ipow(N1, DIM);
  static constexpr int NLP = pow2ceil(NL);
  static constexpr int VEC = PROB == IGA_ELASTICITY ? DIM : 1;
  static constexpr int NS = PROB == IGA_ELASTICITY ? (DIM == 2 ? 3 : 6) : DIM;
  static constexpr int NLV = NL * VEC;
  static constexpr int TI = (NLV + 15) / 16;
};

// Element pass. Per element (one workgroup): extracted 1-D tables, the
// rational basis and physical gradients at every Gauss point (FP64), then the
// element matrix in gradient-product form on the matrix cores:
//   M_cd[i][j] = sum_q G_c[q][i] (w_q G_d[q][j])      (G_c = d R / d x_c)
//   Poisson:    K_e[i][j] = sum_c M_cc[i][j]  with a(x_q) folded into w_q
//   elasticity: K_e[(i,a),(j,b)] = sum_cd E^cd_ab M_cd[i][j],
//               E^cd_ab = sum_st C_st Bs(s; a, c) Bs(t; b, d)
// where Bs(s; a, c) = 1 when the engineering-Voigt strain s contains
// d u_a / d x_c. Each wave owns 16x16 tiles of (i, j) and keeps every M_cd of
// its tile in accumulators, so the constitutive combination, the lifting
// -K_fp g and the diagonal are done in registers; the element matrix is
// streamed to the element-matrix buffer.
template <int DIM, int P, int PROB, typename T>
__global__ void __launch_bounds__(256)
    k_element(iga_patch_t pt, int nq1, const double* __restrict__ qp,
              const double* __restrict__ qw, const AsmDev a, T* __restrict__ ke_all) {
  using S = Shape<DIM, P, PROB>;
  constexpr int N1 = S::N1, NL = S::NL, NLP = S::NLP, NS = S::NS, NLV = S::NLV, VEC = S::VEC;
  constexpr int NLT = (NL + 15) / 16 * 16;  // tile-padded basis count
  constexpr int TI = NLT / 16;
  constexpr int NCD = PROB == IGA_ELASTICITY ? DIM * DIM : 1;
  const int nq = DIM == 2 ? nq1 * nq1 : nq1 * nq1 * nq1;
  const int nqp = (nq + 3) & ~3;

  extern __shared__ double smem[];
  double* N1t = smem;                   // [DIM][nq1][N1]
  double* dN1t = N1t + DIM * nq1 * N1;  // [DIM][nq1][N1]
  double* Rs = dN1t + DIM * nq1 * N1;   // [nq][NL]
  double* Wf = Rs + nq * NL;            // [nq] weight
  T* GA = reinterpret_cast<T*>(Wf + nq);  // [DIM][nqp][NLT] gradients
  T* GW = GA + DIM * nqp * NLT;           // [DIM][nqp][NLT] weighted gradients
  int* fdof = reinterpret_cast<int*>(GW + DIM * nqp * NLT);  // [NL*3] free index or -1
  int* gdof = fdof + NL * 3;                                 // [NL] scalar dofs
  // small elasticity elements: each wave forms one M_cd tile into LDS and all
  // threads combine the entries; larger ones combine in registers
  constexpr bool kLdsCombine = PROB == IGA_ELASTICITY && NLT == 16 && DIM == 2;
  double* Mb = reinterpret_cast<double*>((reinterpret_cast<uintptr_t>(gdof + NL) + 7) &
                                         ~(uintptr_t)7);  // [DIM*DIM][16][16]
  // constitutive data and constant source staged in LDS: indexing the
  // kernel-argument copies with runtime indices would place the whole
  // argument block in per-lane scratch memory
  __shared__ double cm[36], srcc[3], Ecd[81];
  if (threadIdx.x < 36) cm[threadIdx.x] = a.cmat[threadIdx.x];
  if (threadIdx.x < 3) srcc[threadIdx.x] = a.src_const[threadIdx.x];

  const int e = blockIdx.x;
  // batch instance: per-instance inputs and outputs
  const int64_t bi = blockIdx.y, n_e = gridDim.x;
  const double* coefp = a.coef_points ? a.coef_points + bi * n_e * nq : nullptr;
  const double* srcp = a.src_points ? a.src_points + bi * n_e * nq * a.vec : nullptr;
  double* rhsp = a.rhs ? a.rhs + bi * a.n_free : nullptr;
  double* diagp = a.diag ? a.diag + bi * a.n_free : nullptr;
  T* ke = ke_all ? ke_all + bi * n_e * S::NLV * S::NLV : nullptr;
  int ei[3] = {0, 0, 0};
  {
    int rem = e;
    for (int d = 0; d < DIM; ++d) {
      ei[d] = rem % pt.n_elem[d];
      rem /= pt.n_elem[d];
    }
  }

  // ---- 1-D extracted tables, zeroed operand tiles --------------------------
  for (int t = threadIdx.x; t < DIM * nq1 * N1; t += blockDim.x) {
    const int ai = t % N1;
    const int q = (t / N1) % nq1;
    const int d = t / (N1 * nq1);
    double B[N1], dB[N1];
    bernstein<P>(qp[q], B, dB);
    const double* C = pt.ext[d] + ((size_t)ei[d] * N1 + ai) * N1;
    double s = 0.0, ds = 0.0;
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      s += C[b] * B[b];
      ds += C[b] * dB[b];
    }
    const double h = pt.hi[d][ei[d]] - pt.lo[d][ei[d]];
    N1t[t] = s;
    dN1t[t] = ds * (2.0 / h);
  }
  for (int t = threadIdx.x; t < 2 * DIM * nqp * NLT; t += blockDim.x) GA[t] = (T)0;
  if constexpr (PROB == IGA_ELASTICITY) {
    // E^cd_ab from the Voigt matrix
    if (threadIdx.x < DIM * DIM * DIM * DIM) {
      const int t = threadIdx.x;
      const int bb = t % DIM, aa = (t / DIM) % DIM, dd = (t / (DIM * DIM)) % DIM, cc = t / (DIM * DIM * DIM);
      double v = 0.0;
      for (int s1 = 0; s1 < NS; ++s1) {
        int p1, q1;
        voigt<DIM>(s1, p1, q1);
        const bool in1 = (aa == p1 && cc == q1) || (p1 != q1 && aa == q1 && cc == p1);
        if (!in1) continue;
        for (int s2 = 0; s2 < NS; ++s2) {
          int p2, q2;
          voigt<DIM>(s2, p2, q2);
          const bool in2 = (bb == p2 && dd == q2) || (p2 != q2 && bb == q2 && dd == p2);
          if (in2) v += cm[s1 * NS + s2];
        }
      }
      Ecd[t] = v;
    }
  }

  // ---- lane -> local basis function ------------------------------------
  const int lane = threadIdx.x % NLP;
  const int slot = threadIdx.x / NLP;
  const int nslots = blockDim.x / NLP;
  const bool active = lane < NL;
  int ai[3] = {0, 0, 0};
  {
    int rem = lane;
    for (int d = 0; d < DIM; ++d) {
      ai[d] = rem % N1;
      rem /= N1;
    }
  }
  double Pc[DIM];
  double wgt = 1.0;
  int sdof = 0;
  if (active) {
    int stride = 1;
    for (int d = 0; d < DIM; ++d) {
      sdof += (pt.first[d][ei[d]] + ai[d]) * stride;
      stride *= pt.n_cp[d];
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) Pc[c] = pt.cpts[(size_t)sdof * DIM + c];
    wgt = pt.wgts[sdof];
  } else {
#pragma unroll
    for (int c = 0; c < DIM; ++c) Pc[c] = 0.0;
  }
  if (slot == 0 && active) {
    gdof[lane] = sdof;
    for (int c = 0; c < a.vec; ++c) fdof[lane * a.vec + c] = a.f2f[(size_t)a.vec * sdof + c];
  }
  double jac_ref = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) jac_ref *= 0.5 * (pt.hi[d][ei[d]] - pt.lo[d][ei[d]]);
  // weights and weighted control points of the element's functions
  double* ctl = Mb + (kLdsCombine ? DIM * DIM * 256 : 0);  // [NL][DIM+1]
  double* frm = ctl + NL * (DIM + 1);                        // [nq][1 + DIM + DIM*DIM]
  double* wkq = frm + nq * (1 + DIM + DIM * DIM);            // [nq] w_q times the coefficient
  constexpr int FR = 1 + DIM + DIM * DIM;
  if (slot == 0 && active) {
    ctl[lane * (DIM + 1)] = wgt;
#pragma unroll
    for (int c = 0; c < DIM; ++c) ctl[lane * (DIM + 1) + 1 + c] = wgt * Pc[c];
  }
  __syncthreads();

  // ---- per Gauss point, one lane: weight function, geometry, J^-1 (FP64) ---
  for (int q = threadIdx.x; q < nq; q += blockDim.x) {
    const int qi[3] = {q % nq1, (q / nq1) % nq1, q / (nq1 * nq1)};
    double W = 0.0, dW[DIM] = {}, X[DIM] = {}, Y[DIM][DIM] = {};
#pragma unroll
    for (int f = 0; f < NL; ++f) {
      const int fi[3] = {f % N1, (f / N1) % N1, f / (N1 * N1)};
      double N = 1.0, dN[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int o = (d * nq1 + qi[d]) * N1 + fi[d];
        const double v = N1t[o], dv = dN1t[o];
        N *= v;
#pragma unroll
        for (int k = 0; k < DIM; ++k) dN[k] *= (k == d) ? dv : v;
      }
      const double* c = ctl + f * (DIM + 1);
      W += N * c[0];
#pragma unroll
      for (int k = 0; k < DIM; ++k) dW[k] += dN[k] * c[0];
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) {
        X[cc] += N * c[1 + cc];
#pragma unroll
        for (int k = 0; k < DIM; ++k) Y[k][cc] += dN[k] * c[1 + cc];
      }
    }
    const double iW = 1.0 / W;
    double xp[DIM], J[DIM][DIM], inv[DIM][DIM], det;
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) xp[cc] = X[cc] * iW;
#pragma unroll
    for (int k = 0; k < DIM; ++k)
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) J[k][cc] = (Y[k][cc] - dW[k] * xp[cc]) * iW;
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
    double* fr = frm + q * FR;
    fr[0] = W;
#pragma unroll
    for (int k = 0; k < DIM; ++k) fr[1 + k] = dW[k];
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc)
#pragma unroll
      for (int k = 0; k < DIM; ++k) fr[1 + DIM + cc * DIM + k] = inv[cc][k];
    double wq = jac_ref * fabs(det);
#pragma unroll
    for (int d = 0; d < DIM; ++d) wq *= qw[qi[d]];
    double wk = wq;
    if constexpr (PROB == IGA_POISSON)
      wk *= a.coef_kind == IGA_COEF_POINTS ? coefp[(size_t)e * nq + q] : a.coef_const;
    Wf[q] = wq;
    wkq[q] = wk;
  }
  __syncthreads();

  // ---- per (point, function): R and the physical gradients ---------------
  for (int t = threadIdx.x; t < nq * NL; t += blockDim.x) {
    const int q = t / NL, f = t - q * NL;
    const int qi[3] = {q % nq1, (q / nq1) % nq1, q / (nq1 * nq1)};
    const int fi[3] = {f % N1, (f / N1) % N1, f / (N1 * N1)};
    double N = 1.0, dN[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int o = (d * nq1 + qi[d]) * N1 + fi[d];
      const double v = N1t[o], dv = dN1t[o];
      N *= v;
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] *= (k == d) ? dv : v;
    }
    const double* fr = frm + q * FR;
    const double wa = ctl[f * (DIM + 1)];
    const double iW = 1.0 / fr[0];
    const double R = N * wa * iW;
    double gp[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) gp[k] = (dN[k] * wa - R * fr[1 + k]) * iW;
    const double wk = wkq[q];
#pragma unroll
    for (int cc = 0; cc < DIM; ++cc) {
      double g = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) g += fr[1 + DIM + cc * DIM + k] * gp[k];
      GA[(cc * nqp + q) * NLT + f] = (T)g;
      GW[(cc * nqp + q) * NLT + f] = (T)(wk * g);
    }
    Rs[q * NL + f] = R;
  }
  __syncthreads();

  // ---- M_cd on the matrix cores -------------------------------------------
  if constexpr (kLdsCombine) {
    const int wave = threadIdx.x / kWave, wl = threadIdx.x % kWave;
    const int r16 = wl % 16, k4 = wl / 16;
    for (int cd = wave; cd < DIM * DIM; cd += blockDim.x / kWave) {
      const int c = cd / DIM, d = cd % DIM;
      auto acc = Mfma<T>::zero();
      for (int q0 = 0; q0 < nqp; q0 += 4)
        acc = Mfma<T>::step(GA[(c * nqp + q0 + k4) * NLT + r16], GW[(d * nqp + q0 + k4) * NLT + r16], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) Mb[(cd * 16 + Mfma<T>::row(k4, r)) * 16 + r16] = (double)acc[r];
    }
    __syncthreads();
    const bool lift = rhsp != nullptr && a.gfull != nullptr;
    T* kout = ke ? ke + (size_t)e * NLV * NLV : nullptr;
    for (int t = threadIdx.x; t < NLV * NLV; t += blockDim.x) {
      const int ri = t / NLV, cj = t - ri * NLV;
      const int i = ri / VEC, aa = ri - i * VEC, j = cj / VEC, bb = cj - j * VEC;
      double v = 0.0;
#pragma unroll
      for (int cd = 0; cd < DIM * DIM; ++cd)
        v += Ecd[(cd * DIM + aa) * DIM + bb] * Mb[(cd * 16 + i) * 16 + j];
      if (kout) kout[t] = (T)v;
      const int rr = fdof[ri];
      if (rr < 0) continue;
      const int cc = fdof[cj];
      if (cc >= 0) {
        if (diagp && ri == cj) unsafeAtomicAdd(diagp + rr, v);
      } else if (lift) {
        const double gv = a.gfull[(size_t)DIM * gdof[j] + bb];
        if (gv != 0.0) unsafeAtomicAdd(rhsp + rr, -v * gv);
      }
    }
  } else {
    const bool lift = rhsp != nullptr && a.gfull != nullptr;
    T* kout = ke ? ke + (size_t)e * NLV * NLV : nullptr;
    const int wave = threadIdx.x / kWave, wl = threadIdx.x % kWave;
    const int nwaves = blockDim.x / kWave;
    const int r16 = wl % 16, k4 = wl / 16;
    for (int tile = wave; tile < TI * TI; tile += nwaves) {
      const int i0 = (tile / TI) * 16, j0 = (tile % TI) * 16;
      typename Mfma<T>::acc_t acc[NCD];
#pragma unroll
      for (int t = 0; t < NCD; ++t) acc[t] = Mfma<T>::zero();
      if constexpr (PROB == IGA_POISSON) {
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          for (int q0 = 0; q0 < nqp; q0 += 4) {
            const int o = (c * nqp + q0 + k4) * NLT;
            acc[0] = Mfma<T>::step(GA[o + i0 + r16], GW[o + j0 + r16], acc[0]);
          }
      } else {
        for (int q0 = 0; q0 < nqp; q0 += 4) {
          T av[DIM], bv[DIM];
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            const int o = (c * nqp + q0 + k4) * NLT;
            av[c] = GA[o + i0 + r16];
            bv[c] = GW[o + j0 + r16];
          }
#pragma unroll
          for (int c = 0; c < DIM; ++c)
#pragma unroll
            for (int d = 0; d < DIM; ++d) acc[c * DIM + d] = Mfma<T>::step(av[c], bv[d], acc[c * DIM + d]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + Mfma<T>::row(k4, r), j = j0 + r16;
        if (i >= NL || j >= NL) continue;
        if constexpr (PROB == IGA_POISSON) {
          const double v = (double)acc[0][r];
          if (kout) kout[i * NLV + j] = acc[0][r];
          for (int c = 0; c < a.vec; ++c) {
            const int rr = fdof[i * a.vec + c];
            if (rr < 0) continue;
            const int cc = fdof[j * a.vec + c];
            if (cc >= 0) {
              if (diagp && i == j) unsafeAtomicAdd(diagp + rr, v);
            } else if (lift) {
              const double gv = a.gfull[(size_t)a.vec * gdof[j] + c];
              if (gv != 0.0) unsafeAtomicAdd(rhsp + rr, -v * gv);
            }
          }
        } else {
#pragma unroll
          for (int aa = 0; aa < VEC; ++aa)
#pragma unroll
            for (int bb = 0; bb < VEC; ++bb) {
              double v = 0.0;
#pragma unroll
              for (int c = 0; c < DIM; ++c)
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  v += Ecd[((c * DIM + d) * DIM + aa) * DIM + bb] * (double)acc[c * DIM + d][r];
              const int ri = i * VEC + aa, cj = j * VEC + bb;
              if (kout) kout[ri * NLV + cj] = (T)v;
              const int rr = fdof[ri];
              if (rr < 0) continue;
              const int cc = fdof[cj];
              if (cc >= 0) {
                if (diagp && ri == cj) unsafeAtomicAdd(diagp + rr, v);
              } else if (lift) {
                const double gv = a.gfull[(size_t)DIM * gdof[j] + bb];
                if (gv != 0.0) unsafeAtomicAdd(rhsp + rr, -v * gv);
              }
            }
        }
      }
    }
  }

  // ---- source work ----------------------------------------------------------
  if (rhsp && a.src_kind != IGA_SRC_NONE) {
    for (int idx = threadIdx.x; idx < NL * a.vec; idx += blockDim.x) {
      const int i = idx / a.vec, c = idx % a.vec;
      const int r = fdof[i * a.vec + c];
      if (r < 0) continue;
      double f = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double fv = a.src_kind == IGA_SRC_CONST
                              ? srcc[c]
                              : srcp[((size_t)e * nq + q) * a.vec + c];
        f += Wf[q] * fv * Rs[q * NL + i];
      }
      unsafeAtomicAdd(rhsp + r, f);
    }
  }
}

// Packed symmetric element matrix: the blocks (i, j), i <= j, of the NL x NL
// block matrix (VEC x VEC values each, row-major), stored by block rows --
// row i holds blocks i .. NL-1 -- so block (i, j) is the
// i NL - i (i - 1) / 2 + (j - i)-th. Block (j, i) of the matrix is the
// transpose of the stored (i, j): the element matrices are symmetric, so
// the element pass writes about half of them and the gather reads half.
__host__ __device__ constexpr int sym_ke_values(int nl, int vec) { return nl * (nl + 1) / 2 * vec * vec; }

__device__ __forceinline__ int sym_block(int i, int j, int nl) {
  return i * nl - (i * (i - 1)) / 2 + (j - i);
}

// Element pass, one wavefront per element, for planar patches with at most
// four functions and four Gauss points per direction. Lane l owns the local
// function i = l % 16 (i0 = i % 4, i1 = i / 4, padded to 4 x 4) and, in step
// k = 0..3, the Gauss point q = 4k + l / 16 (q0 = l / 16, q1 = k). That is the
// operand layout of the 16x16x4 MFMA step k -- A[i][q] = G_c[q][i] and
// B[q][j] = w_q G_d[q][j] -- so the gradients feed the matrix cores straight
// from the lane's registers. The geometry (W, dW, x, dx/dxi: nine tensor
// sums per Gauss point) is sum-factorised once per point, FP64, and shared
// through LDS; no workgroup barrier is needed.
// Eight waves per SIMD: the kernel fits 64 registers (the accumulators in
// architectural VGPRs) without spilling; left to itself the compiler takes
// 88 (64 + 24 accumulation registers), which caps residency at five waves.
template <int P, int PROB, typename T, bool REG, bool SYM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_element_w(iga_patch_t pt, int nq1, const double* __restrict__ qp,
                const double* __restrict__ qw, const AsmDev a, T* __restrict__ ke_all, int n_e,
                int e_begin, int e_end) {
  constexpr int N1 = P + 1;
  static_assert(N1 <= 4, "at most four functions per direction");
  constexpr int NL = N1 * N1;
  constexpr int VEC = PROB == IGA_ELASTICITY ? 2 : 1;
  constexpr int NLV = NL * VEC;
  constexpr int NCD = PROB == IGA_ELASTICITY ? 4 : 1;
  constexpr int WPB = 4;    // elements (waves) per workgroup
  constexpr int REC = 12;   // doubles per Gauss-point record
  __shared__ double s_tabq[WPB][2][2][4][4];  // [dir][value|derivative][q][a]
  __shared__ double s_taba[WPB][2][2][4][4];  // [dir][value|derivative][a][q]
  __shared__ double s_coef[WPB][3][16];       // w, w x, w y per local function
  __shared__ double s_geo[WPB][16][3][4];     // [q][field][value, d/dxi0, d/dxi1, -]
  __shared__ __align__(16) double s_rec[WPB][16][REC];
  __shared__ int s_fdof[WPB][16 * VEC];
  __shared__ int s_gdof[WPB][16];

  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  // elements [e_begin, e_end) of the n_e of the patch
  const int e = e_begin + blockIdx.x * WPB + w;
  if (e >= e_end) return;
  const int64_t bi = blockIdx.y;
  const int nq = nq1 * nq1;
  const double* coefp = a.coef_points ? a.coef_points + bi * n_e * nq : nullptr;
  const double* srcp = a.src_points ? a.src_points + bi * n_e * nq * VEC : nullptr;
  double* rhsp = a.rhs ? a.rhs + bi * a.n_free : nullptr;
  double* diagp = a.diag ? a.diag + bi * a.n_free : nullptr;
  constexpr int KE = SYM ? sym_ke_values(NL, VEC) : NLV * NLV;  // stored values per element
  T* kout = ke_all ? ke_all + (bi * n_e + e) * (int64_t)KE : nullptr;
  const int ei0 = e % pt.n_elem[0], ei1 = e / pt.n_elem[0];

  // ---- every global read of the element, issued together ------------------
  // Unconditional reads from clamped (valid) addresses, selected afterwards:
  // a read under a condition would be waited for before the next one.
  const double h0 = pt.hi[0][ei0] - pt.lo[0][ei0], h1 = pt.hi[1][ei1] - pt.lo[1][ei1];
  // table lane: direction td, function ta, Gauss point tq (meaningful for l < 32)
  const int td = (l >> 4) & 1, ta = (l >> 2) & 3, tq = l & 3;
  const double xq = qp[tq < nq1 ? tq : nq1 - 1];
  double crow[N1];
  {
    const int fa = ta < N1 ? ta : N1 - 1;
    const double* C = td ? pt.ext[1] + ((size_t)ei1 * N1 + fa) * N1
                         : pt.ext[0] + ((size_t)ei0 * N1 + fa) * N1;
#pragma unroll
    for (int b = 0; b < N1; ++b) crow[b] = C[b];
  }
  // function lane i = l % 16 (i0 = i % 4, i1 = i / 4)
  const int fi0 = l & 3, fi1 = (l >> 2) & 3;
  const bool fvalid = fi0 < N1 && fi1 < N1;
  int sdof;
  {
    const int f0 = REG ? ei0 : pt.first[0][ei0], f1 = REG ? ei1 : pt.first[1][ei1];
    sdof = (f0 + (fi0 < N1 ? fi0 : N1 - 1)) + (f1 + (fi1 < N1 ? fi1 : N1 - 1)) * pt.n_cp[0];
  }
  const double wt_r = pt.wgts[sdof];
  const double px_r = pt.cpts[(size_t)sdof * 2], py_r = pt.cpts[(size_t)sdof * 2 + 1];
  int fd_r[VEC];
#pragma unroll
  for (int c = 0; c < VEC; ++c) fd_r[c] = a.f2f[(size_t)VEC * sdof + c];
  // record lane q = l % 16 (q0 = q % 4, q1 = q / 4)
  const int rq0 = l & 3, rq1 = (l >> 2) & 3;
  const bool qvalid = rq0 < nq1 && rq1 < nq1;
  const int cq0 = rq0 < nq1 ? rq0 : nq1 - 1, cq1 = rq1 < nq1 ? rq1 : nq1 - 1;
  const int qn = cq0 + nq1 * cq1;
  const double qwt = qw[cq0] * qw[cq1];
  // coefficient and source at the point: raw reads now, selected in the
  // record phase (selecting here would place the wait here)
  const double coef_r = *(coefp ? coefp + (size_t)e * nq + qn : qw);
  double src_r[VEC];
  {
    const double* sp = srcp ? srcp + ((size_t)e * nq + qn) * VEC : qw;
#pragma unroll
    for (int c = 0; c < VEC; ++c) src_r[c] = sp[srcp ? c : 0];
  }

  // ---- 1-D extracted tables: lane (dir, a, q) for l < 32 --------------------
  if (l < 32) {
    double v = 0.0, dv = 0.0;
    if (ta < N1 && tq < nq1) {
      double B[N1], dB[N1];
      bernstein<P>(xq, B, dB);
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        v += crow[b] * B[b];
        dv += crow[b] * dB[b];
      }
      dv *= 2.0 / (td ? h1 : h0);
    }
    s_tabq[w][td][0][tq][ta] = v;
    s_tabq[w][td][1][tq][ta] = dv;
    s_taba[w][td][0][ta][tq] = v;
    s_taba[w][td][1][ta][tq] = dv;
  }
  // ---- local functions: weights, weighted control points, dof indices -------
  if (l < 16) {
    const double wt = fvalid ? wt_r : 0.0;
#pragma unroll
    for (int c = 0; c < VEC; ++c) s_fdof[w][l * VEC + c] = fvalid ? fd_r[c] : -1;
    s_gdof[w][l] = fvalid ? sdof : 0;
    s_coef[w][0][l] = wt;
    s_coef[w][1][l] = wt * px_r;
    s_coef[w][2][l] = wt * py_r;
  }
  wave_lds_sync();

  // ---- geometry per Gauss point: lane (field f, point q), sum-factorised -----
  {
    const int q = l & 15, f = l >> 4, q0 = q & 3, q1 = q >> 2;
    if (f < 3 && q0 < nq1 && q1 < nq1) {
      const double* F = s_coef[w][f];
      double tv[4], td[4];
#pragma unroll
      for (int i1 = 0; i1 < N1; ++i1) {
        double sv = 0.0, sd = 0.0;
#pragma unroll
        for (int i0 = 0; i0 < N1; ++i0) {
          const double c = F[i0 + 4 * i1];
          sv += s_tabq[w][0][0][q0][i0] * c;
          sd += s_tabq[w][0][1][q0][i0] * c;
        }
        tv[i1] = sv;
        td[i1] = sd;
      }
      double v = 0.0, d0 = 0.0, d1 = 0.0;
#pragma unroll
      for (int i1 = 0; i1 < N1; ++i1) {
        const double b = s_tabq[w][1][0][q1][i1], db = s_tabq[w][1][1][q1][i1];
        v += b * tv[i1];
        d0 += b * td[i1];
        d1 += db * tv[i1];
      }
      s_geo[w][q][f][0] = v;
      s_geo[w][q][f][1] = d0;
      s_geo[w][q][f][2] = d1;
    }
  }
  wave_lds_sync();
  // ---- Gauss-point records: 1/W, dW, inverse Jacobian, weights ---------------
  if (l < 16) {
    const int q = l;
    double rec[REC];
#pragma unroll
    for (int t = 0; t < REC; ++t) rec[t] = 0.0;
    if (qvalid) {
      const double W = s_geo[w][q][0][0], dW0 = s_geo[w][q][0][1], dW1 = s_geo[w][q][0][2];
      const double iW = 1.0 / W;
      const double x0 = s_geo[w][q][1][0] * iW, x1 = s_geo[w][q][2][0] * iW;
      // Jm[c][d] = d x_c / d xi_d
      const double j00 = (s_geo[w][q][1][1] - x0 * dW0) * iW;
      const double j01 = (s_geo[w][q][1][2] - x0 * dW1) * iW;
      const double j10 = (s_geo[w][q][2][1] - x1 * dW0) * iW;
      const double j11 = (s_geo[w][q][2][2] - x1 * dW1) * iW;
      const double det = j00 * j11 - j01 * j10;
      const double id = 1.0 / det;
      const double wq = 0.25 * h0 * h1 * fabs(det) * qwt;
      rec[0] = iW;
      rec[1] = dW0;
      rec[2] = dW1;
      // inverse Jacobian, d xi_d / d x_c at [2d + c]
      rec[3] = j11 * id;
      rec[4] = -j01 * id;
      rec[5] = -j10 * id;
      rec[6] = j00 * id;
      double coefv = 1.0;
      if constexpr (PROB == IGA_POISSON)
        coefv = a.coef_kind == IGA_COEF_POINTS ? coef_r : a.coef_const;
      rec[7] = wq * coefv;
#pragma unroll
      for (int c = 0; c < VEC; ++c)
        rec[8 + c] = a.src_kind == IGA_SRC_POINTS  ? wq * src_r[c]
                     : a.src_kind == IGA_SRC_CONST ? wq * a.src_const[c]
                                                   : 0.0;
    }
#pragma unroll
    for (int t = 0; t < REC; ++t) s_rec[w][q][t] = rec[t];
  }
  wave_lds_sync();

  // ---- gradient products on the matrix cores --------------------------------
  const int i = l & 15, g = l >> 4;
  const double wi = s_coef[w][0][i];
  // weighted first-direction factors of the lane's function at q0 = g
  const double a0w = s_tabq[w][0][0][g][i & 3] * wi, da0w = s_tabq[w][0][1][g][i & 3] * wi;
  typename Mfma<T>::acc_t acc[NCD];
#pragma unroll
  for (int t = 0; t < NCD; ++t) acc[t] = Mfma<T>::zero();
  double ld[VEC];
#pragma unroll
  for (int c = 0; c < VEC; ++c) ld[c] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double* rc = s_rec[w][4 * k + g];
    const double b = s_taba[w][1][0][i >> 2][k], db = s_taba[w][1][1][i >> 2][k];
    const double iW = rc[0];
    const double R = a0w * b * iW;
    const double r0 = (da0w * b - R * rc[1]) * iW;
    const double r1 = (a0w * db - R * rc[2]) * iW;
    const double gx = r0 * rc[3] + r1 * rc[5];
    const double gy = r0 * rc[4] + r1 * rc[6];
    const double wk = rc[7];
    const T ga0 = (T)gx, ga1 = (T)gy, gw0 = (T)(wk * gx), gw1 = (T)(wk * gy);
    if constexpr (PROB == IGA_ELASTICITY) {
      acc[0] = Mfma<T>::step(ga0, gw0, acc[0]);
      acc[1] = Mfma<T>::step(ga0, gw1, acc[1]);
      acc[2] = Mfma<T>::step(ga1, gw0, acc[2]);
      acc[3] = Mfma<T>::step(ga1, gw1, acc[3]);
    } else {
      acc[0] = Mfma<T>::step(ga0, gw0, acc[0]);
      acc[0] = Mfma<T>::step(ga1, gw1, acc[0]);
    }
#pragma unroll
    for (int c = 0; c < VEC; ++c) ld[c] += R * rc[8 + c];
  }

  // ---- load vector: sum the four Gauss-point groups of each function --------
  if (rhsp && a.src_kind != IGA_SRC_NONE) {
#pragma unroll
    for (int c = 0; c < VEC; ++c) {
      ld[c] += __shfl_xor(ld[c], 16);
      ld[c] += __shfl_xor(ld[c], 32);
    }
    if (l < 16 && (i & 3) < N1 && (i >> 2) < N1) {
#pragma unroll
      for (int c = 0; c < VEC; ++c) {
        const int r = s_fdof[w][i * VEC + c];
        if (r >= 0) unsafeAtomicAdd(rhsp + r, ld[c]);
      }
    }
  }

  // ---- constitutive combination, element matrix out, diagonal, lifting -------
  const bool lift = rhsp != nullptr && a.gfull != nullptr;
  bool pres = false;
  if (l < 16 && (i & 3) < N1 && (i >> 2) < N1)
#pragma unroll
    for (int c = 0; c < VEC; ++c) pres |= s_fdof[w][i * VEC + c] < 0;
  const bool any_pres = __ballot(pres) != 0;
  const int j = i;  // accumulator column
  const bool jok = (j & 3) < N1 && (j >> 2) < N1;
  const int jc = N1 == 4 ? j : (j & 3) + N1 * (j >> 2);  // compact local index
  // constitutive table in the accumulator type, converted once
  T et[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) et[t] = std::is_same_v<T, float> ? (T)a.ecd32[t] : (T)a.ecd[t];
  const bool ortho = a.ortho != 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ir = Mfma<T>::row(g, r);
    const bool ok = jok && (ir & 3) < N1 && (ir >> 2) < N1;
    if (!ok) continue;
    const int ic = N1 == 4 ? ir : (ir & 3) + N1 * (ir >> 2);
    // combined in the accumulator type
    T v[VEC][VEC];
    if constexpr (PROB == IGA_ELASTICITY) {
      // E index ((c*2 + d)*2 + a)*2 + b, acc[c*2 + d] = M_cd
      if (ortho) {
        v[0][0] = et[0] * acc[0][r] + et[12] * acc[3][r];
        v[0][1] = et[5] * acc[1][r] + et[9] * acc[2][r];
        v[1][0] = et[6] * acc[1][r] + et[10] * acc[2][r];
        v[1][1] = et[3] * acc[0][r] + et[15] * acc[3][r];
      } else {
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            T s = (T)0;
#pragma unroll
            for (int cd = 0; cd < 4; ++cd) s += et[(cd * 2 + aa) * 2 + bb] * acc[cd][r];
            v[aa][bb] = s;
          }
      }
    } else {
      v[0][0] = acc[0][r];
    }
    if (kout) {
      if constexpr (SYM) {
        // upper blocks only, packed by rows
        if (ic <= jc) {
          T* dst = kout + (size_t)sym_block(ic, jc, NL) * (VEC * VEC);
          if constexpr (VEC == 2) {
            using T2 = T __attribute__((ext_vector_type(2)));
            *reinterpret_cast<T2*>(dst) = T2{v[0][0], v[0][1]};
            *reinterpret_cast<T2*>(dst + 2) = T2{v[1][0], v[1][1]};
          } else {
            *dst = v[0][0];
          }
        }
      } else {
#pragma unroll
        for (int aa = 0; aa < VEC; ++aa) {
          T* dst = kout + (size_t)(ic * VEC + aa) * NLV + jc * VEC;
          if constexpr (VEC == 2) {
            using T2 = T __attribute__((ext_vector_type(2)));
            *reinterpret_cast<T2*>(dst) = T2{v[aa][0], v[aa][1]};
          } else {
            *dst = v[aa][0];
          }
        }
      }
    }
    if (diagp && ir == j) {
#pragma unroll
      for (int aa = 0; aa < VEC; ++aa) {
        const int rr = s_fdof[w][ir * VEC + aa];
        if (rr >= 0) unsafeAtomicAdd(diagp + rr, (double)v[aa][aa]);
      }
    }
    if (lift && any_pres) {
#pragma unroll
      for (int aa = 0; aa < VEC; ++aa) {
        const int rr = s_fdof[w][ir * VEC + aa];
        if (rr < 0) continue;
#pragma unroll
        for (int bb = 0; bb < VEC; ++bb) {
          if (s_fdof[w][j * VEC + bb] >= 0) continue;
          const double gv = a.gfull[(size_t)VEC * s_gdof[w][j] + bb];
          if (gv != 0.0) unsafeAtomicAdd(rhsp + rr, -(double)v[aa][bb] * gv);
        }
      }
    }
  }
}

// Sign flip of an MFMA operand by a mask (0 or the sign bit of the high
// word), one integer operation instead of an FP64 multiply.
__device__ __forceinline__ double flip_sign(double x, uint32_t m) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)m << 32));
}
__device__ __forceinline__ float flip_sign(float x, uint32_t m) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ m);
}

// Element pass for scalar Poisson on 3-D patches with p <= 2 and p + 1 Gauss
// points per direction: one wavefront per element, four per workgroup, no
// workgroup barrier. The FP64 geometry -- W, w x and their parametric
// derivatives at every Gauss point, sixteen tensor sums -- is sum-factorised
// in two stages: the first direction per (field, a1, a2), then the other two
// per (field, point). With A = J^-1 / W, b = A dW / W and s = sqrt|w_q a(x_q)|,
// the scaled physical gradient of R_i = w_i N_i / W is
//   s grad R_i = (s A) (w_i dN_i) - (w_i N_i) (s b),
// and K_e = sum_q sign(w_q a) (s grad R_i)(s grad R_j): both MFMA operands
// are the same scaled gradient (up to the sign), so one point record (s A,
// s b, the source weight over W, the point's table offsets) and seven
// products per function and point give the operands. Lane l owns the
// functions l % 16 + 16 u (u < TI) at the Gauss point 4 t + l / 16 in step t,
// the operand layout of the 16x16x4 MFMA, so the gradients feed the matrix
// cores from registers. K_e is symmetric: only the tiles I <= J are formed
// (TI (TI + 1) / 2 of TI^2); they are placed (the off-diagonal ones twice)
// in the wave's LDS, over the geometry, and streamed out row-major with
// coalesced stores; the diagonal and the lifting read the staged matrix.
template <int P, typename T, bool REG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
    k_element_w3(iga_patch_t pt, const double* __restrict__ qp, const double* __restrict__ qw,
                 const AsmDev a, T* __restrict__ ke_all, int n_e) {
  constexpr int N1 = P + 1, N2 = N1 * N1, NL = N2 * N1, NQ = NL;
  constexpr int TI = (NL + 15) / 16;       // 16-function tiles
  constexpr int NF = 16 * TI;              // padded function count
  constexpr int NT = (NQ + 3) / 4;         // MFMA steps per gradient component
  constexpr int NTILE = TI * (TI + 1) / 2;  // tiles I <= J
  constexpr int WPB = 4;                   // elements (waves) per workgroup
  // doubles per Gauss-point record: 16 used; the stride of 18 puts the four
  // records read together (one per lane group) on different LDS banks
  constexpr int REC = 18;
  constexpr int T1 = 4 * N2 * N1 * 2 > NF * N1 * 2 ? 4 * N2 * N1 * 2 : NF * N1 * 2;
  // one LDS buffer per wave (offsets in doubles, 16-byte aligned):
  //   tab [dir][q][a][value|derivative]          1-D tables
  //   ctl [4][NL]                                w, w x, w y, w z per function
  //   t1  stage-1 sums [field][a1 + N1 a2][q0][v|d]; afterwards the weighted
  //       first-direction factors w_f (N, dN)(q0) per padded function [f][q0][v|d]
  //   geo [q][4 field + (value, d0, d1, d2)], then in place the point record
  // and, once the products are done, the element matrix [NL][NL] over all of it
  constexpr int O_TAB = 0;
  constexpr int O_CTL = O_TAB + 3 * N1 * N1 * 2;
  constexpr int O_T1 = (O_CTL + 4 * NL + 1) & ~1;
  constexpr int O_GEO = O_T1 + T1;
  constexpr int KE_D = (NL * NL * (int)sizeof(T) + 7) / 8;
  constexpr int SZ = ((O_GEO + NT * 4 * REC > KE_D ? O_GEO + NT * 4 * REC : KE_D) + 1) & ~1;
  __shared__ __align__(16) double s_buf[WPB][SZ];
  __shared__ int s_fdof[WPB][NL];
  __shared__ int s_gdof[WPB][NL];

  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  const int e = blockIdx.x * WPB + w;
  if (e >= n_e) return;
  double* const tab = s_buf[w] + O_TAB;
  double* const ctl = s_buf[w] + O_CTL;
  double* const t1b = s_buf[w] + O_T1;
  double* const geo = s_buf[w] + O_GEO;
  auto TAB = [&](int d, int q, int f) { return tab + ((d * N1 + q) * N1 + f) * 2; };
  const int64_t bi = blockIdx.y;
  const double* coefp = a.coef_points ? a.coef_points + bi * n_e * NQ : nullptr;
  const double* srcp = a.src_points ? a.src_points + bi * n_e * NQ : nullptr;
  double* rhsp = a.rhs ? a.rhs + bi * a.n_free : nullptr;
  double* diagp = a.diag ? a.diag + bi * a.n_free : nullptr;
  T* kout = ke_all ? ke_all + (bi * n_e + e) * (int64_t)(NL * NL) : nullptr;
  const int ei0 = e % pt.n_elem[0], ei1 = (e / pt.n_elem[0]) % pt.n_elem[1],
            ei2 = e / (pt.n_elem[0] * pt.n_elem[1]);

  // ---- every global read of the element, issued together ------------------
  // table lane (direction td, function ta, point tq), meaningful for l < 3 N1^2;
  // the kernel-argument arrays are selected, not indexed at run time
  const int tl = l < 3 * N2 ? l : 0;
  const int td = tl / N2, ta = (tl / N1) % N1, tq = tl % N1;
  const int etd = td == 0 ? ei0 : td == 1 ? ei1 : ei2;
  const double* ext = td == 0 ? pt.ext[0] : td == 1 ? pt.ext[1] : pt.ext[2];
  const double* lo = td == 0 ? pt.lo[0] : td == 1 ? pt.lo[1] : pt.lo[2];
  const double* hi = td == 0 ? pt.hi[0] : td == 1 ? pt.hi[1] : pt.hi[2];
  double crow[N1];
#pragma unroll
  for (int b = 0; b < N1; ++b) crow[b] = ext[((size_t)etd * N1 + ta) * N1 + b];
  const double xq = qp[tq];
  const double hd = hi[etd] - lo[etd];
  // function lane f = l (f0 + N1 f1 + N1^2 f2), meaningful for l < NL;
  // regular patches: the first function of element i is i (no table read)
  const int fl = l < NL ? l : NL - 1;
  const int f0 = REG ? ei0 : pt.first[0][ei0], f1 = REG ? ei1 : pt.first[1][ei1],
            f2 = REG ? ei2 : pt.first[2][ei2];
  const int sdof = (f0 + fl % N1) + pt.n_cp[0] * ((f1 + (fl / N1) % N1) + pt.n_cp[1] * (f2 + fl / N2));
  const double wt_r = pt.wgts[sdof];
  const double px = pt.cpts[(size_t)sdof * 3], py = pt.cpts[(size_t)sdof * 3 + 1],
               pz = pt.cpts[(size_t)sdof * 3 + 2];
  const int fd_r = a.f2f[sdof];
  // point lane q = l, meaningful for l < NQ
  const int ql = l < NQ ? l : NQ - 1;
  const double coef_r = *(coefp ? coefp + (size_t)e * NQ + ql : qw);
  const double src_r = *(srcp ? srcp + (size_t)e * NQ + ql : qw);
  const double jac_ref = 0.125 * (pt.hi[0][ei0] - pt.lo[0][ei0]) * (pt.hi[1][ei1] - pt.lo[1][ei1]) *
                         (pt.hi[2][ei2] - pt.lo[2][ei2]);

  // ---- 1-D extracted tables and the local functions -------------------------
  if (l < 3 * N2) {
    double B[N1], dB[N1], v = 0.0, dv = 0.0;
    bernstein<P>(xq, B, dB);
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      v += crow[b] * B[b];
      dv += crow[b] * dB[b];
    }
    double* tp = TAB(td, tq, ta);
    tp[0] = v;
    tp[1] = dv * (2.0 / hd);
  }
  if (l < NL) {
    ctl[l] = wt_r;
    ctl[NL + l] = wt_r * px;
    ctl[2 * NL + l] = wt_r * py;
    ctl[3 * NL + l] = wt_r * pz;
    s_fdof[w][l] = fd_r;
    s_gdof[w][l] = sdof;
  }
  wave_lds_sync();

  // ---- geometry, stage 1: contract the first direction -----------------------
  if (l < 4 * N2) {
    const int f = l / N2, a12 = l % N2;
    double c[N1];
#pragma unroll
    for (int a0 = 0; a0 < N1; ++a0) c[a0] = ctl[f * NL + a0 + N1 * a12];
#pragma unroll
    for (int q0 = 0; q0 < N1; ++q0) {
      double sv = 0.0, sd = 0.0;
#pragma unroll
      for (int a0 = 0; a0 < N1; ++a0) {
        sv += TAB(0, q0, a0)[0] * c[a0];
        sd += TAB(0, q0, a0)[1] * c[a0];
      }
      t1b[((f * N2 + a12) * N1 + q0) * 2] = sv;
      t1b[((f * N2 + a12) * N1 + q0) * 2 + 1] = sd;
    }
  }
  wave_lds_sync();
  // ---- stage 2: the other two directions, per (field, point) ----------------
#pragma unroll
  for (int it = l; it < 4 * NQ; it += 64) {
    const int f = it / NQ, q = it % NQ;
    const int q0 = q % N1, q1 = (q / N1) % N1, q2 = q / N2;
    double v = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int a2 = 0; a2 < N1; ++a2) {
      const double n2 = TAB(2, q2, a2)[0], dn2 = TAB(2, q2, a2)[1];
#pragma unroll
      for (int a1 = 0; a1 < N1; ++a1) {
        const double n1 = TAB(1, q1, a1)[0], dn1 = TAB(1, q1, a1)[1];
        const double* t1 = t1b + ((f * N2 + a1 + N1 * a2) * N1 + q0) * 2;
        const double tv = t1[0], tdv = t1[1];
        const double b = n1 * n2;
        v += tv * b;
        d0 += tdv * b;
        d1 += tv * (dn1 * n2);
        d2 += tv * (n1 * dn2);
      }
    }
    double* gq = geo + q * REC + 4 * f;
    gq[0] = v;
    gq[1] = d0;
    gq[2] = d1;
    gq[3] = d2;
  }
  wave_lds_sync();
  // ---- point records: s A, s b, the source weight over W, the point's table
  // offsets and the sign of its stiffness weight; and the weighted
  // first-direction factors of every function ----------------------------------
  if (l < NT * 4) {
    double rec[13];
#pragma unroll
    for (int t = 0; t < 13; ++t) rec[t] = 0.0;
    uint32_t sign = 0;
    const int qc = l < NQ ? l : 0;
    double* gq = geo + l * REC;
    if (l < NQ) {
      double G[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) G[t] = gq[t];
      const double W = G[0], iW = 1.0 / W;
      // J[k][c] = d x_c / d xi_k
      double J[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double xc = G[4 * (c + 1)] * iW;
#pragma unroll
        for (int k = 0; k < 3; ++k) J[k][c] = (G[4 * (c + 1) + 1 + k] - xc * G[1 + k]) * iW;
      }
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double wq = jac_ref * fabs(det) * qw[l % N1] * qw[(l / N1) % N1] * qw[l / N2];
      const double wk = wq * (a.coef_kind == IGA_COEF_POINTS ? coef_r : a.coef_const);
      sign = wk < 0.0 ? 0x80000000u : 0u;
      // (s J^-1 / W)[c][k] at 3 c + k
      const double sc = sqrt(fabs(wk)) * iW / det;
      double A[3][3];
      A[0][0] = c00 * sc;
      A[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * sc;
      A[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * sc;
      A[1][0] = c01 * sc;
      A[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * sc;
      A[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * sc;
      A[2][0] = c02 * sc;
      A[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * sc;
      A[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * sc;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 0; k < 3; ++k) rec[3 * c + k] = A[c][k];
        rec[9 + c] = (A[c][0] * G[1] + A[c][1] * G[2] + A[c][2] * G[3]) * iW;
      }
      rec[12] = iW * (a.src_kind == IGA_SRC_POINTS  ? wq * src_r
                      : a.src_kind == IGA_SRC_CONST ? wq * a.src_const[0]
                                                    : 0.0);
    }
#pragma unroll
    for (int t = 0; t < 13; ++t) gq[t] = rec[t];
    // table offsets (in doubles) of the point in each direction, the sign
    int* ofs = reinterpret_cast<int*>(gq + 13);
    ofs[0] = (qc % N1) * 2;
    ofs[1] = ((qc / N1) % N1) * N1 * 2;
    ofs[2] = (qc / N2) * N1 * 2;
    ofs[3] = (int)sign;
  }
#pragma unroll
  for (int it = l; it < NF * N1; it += 64) {
    const int f = it / N1, q0 = it % N1;
    const int fc = f < NL ? f : 0;
    const double wf = f < NL ? ctl[fc] : 0.0;
    t1b[it * 2] = wf * TAB(0, q0, fc % N1)[0];
    t1b[it * 2 + 1] = wf * TAB(0, q0, fc % N1)[1];
  }
  wave_lds_sync();

  // ---- gradient products on the matrix cores --------------------------------
  const int r16 = l & 15, g = l >> 4;
  // the lane's functions: weighted first-direction factors, second- and
  // third-direction tables
  const double* fb[TI];
  const double* tb1[TI];
  const double* tb2[TI];
#pragma unroll
  for (int u = 0; u < TI; ++u) {
    const int f = r16 + 16 * u;
    const int fc = f < NL ? f : 0;
    fb[u] = t1b + f * N1 * 2;
    tb1[u] = TAB(1, 0, (fc / N1) % N1);
    tb2[u] = TAB(2, 0, fc / N2);
  }
  typename Mfma<T>::acc_t acc[NTILE];
#pragma unroll
  for (int n = 0; n < NTILE; ++n) acc[n] = Mfma<T>::zero();
  double ld[TI];
#pragma unroll
  for (int u = 0; u < TI; ++u) ld[u] = 0.0;
  // with no negative stiffness weight in the element (the usual case) both
  // operands are the same registers
  auto products = [&](auto has_neg) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      // padded points (q >= NQ) carry a zero record: the operands vanish there
      const double* rc = geo + (4 * t + g) * REC;
      const int* ofs = reinterpret_cast<const int*>(rc + 13);
      const int o0 = ofs[0], o1 = ofs[1], o2 = ofs[2];
      const uint32_t sg = (uint32_t)ofs[3];
      const double ws = rc[12];
      T ga[TI][3], gw[TI][3];
#pragma unroll
      for (int u = 0; u < TI; ++u) {
        const double a0 = fb[u][o0], e0 = fb[u][o0 + 1];
        const double v1 = tb1[u][o1], e1 = tb1[u][o1 + 1];
        const double v2 = tb2[u][o2], e2 = tb2[u][o2 + 1];
        const double p0 = v1 * v2;
        const double Nw = a0 * p0;  // w N
        const double d0 = e0 * p0, d1 = a0 * (e1 * v2), d2 = a0 * (v1 * e2);  // w dN
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double gc =
              rc[3 * c] * d0 + rc[3 * c + 1] * d1 + rc[3 * c + 2] * d2 - Nw * rc[9 + c];
          ga[u][c] = (T)gc;
          if constexpr (decltype(has_neg)::value) gw[u][c] = flip_sign(ga[u][c], sg);
          else gw[u][c] = ga[u][c];
        }
        ld[u] += Nw * ws;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        int n = 0;
#pragma unroll
        for (int I = 0; I < TI; ++I)
#pragma unroll
          for (int J = I; J < TI; ++J, ++n) acc[n] = Mfma<T>::step(ga[I][c], gw[J][c], acc[n]);
      }
    }
  };
  const bool neg = l < NT * 4 && reinterpret_cast<const int*>(geo + (l < NT * 4 ? l : 0) * REC + 13)[3] != 0;
  if (__ballot(neg) != 0)
    products(std::true_type{});
  else
    products(std::false_type{});

  // ---- load vector: sum the four point groups of each function --------------
  if (rhsp && a.src_kind != IGA_SRC_NONE) {
#pragma unroll
    for (int u = 0; u < TI; ++u) {
      ld[u] += __shfl_xor(ld[u], 16);
      ld[u] += __shfl_xor(ld[u], 32);
    }
    if (g == 0) {
#pragma unroll
      for (int u = 0; u < TI; ++u) {
        const int f = r16 + 16 * u;
        if (f >= NL) continue;
        const int r = s_fdof[w][f];
        if (r >= 0) unsafeAtomicAdd(rhsp + r, ld[u]);
      }
    }
  }

  // ---- element matrix: staged in LDS, streamed out; diagonal, lifting --------
  wave_lds_sync();  // every read of the geometry and tables precedes this
  T* kl = reinterpret_cast<T*>(s_buf[w]);
  {
    int n = 0;
#pragma unroll
    for (int I = 0; I < TI; ++I)
#pragma unroll
      for (int J = I; J < TI; ++J, ++n) {
        const int j = 16 * J + r16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * I + Mfma<T>::row(g, r);
          if (i < NL && j < NL) {
            kl[i * NL + j] = acc[n][r];
            if (I != J) kl[j * NL + i] = acc[n][r];
          }
        }
      }
  }
  wave_lds_sync();
  if (kout)
    for (int t = l; t < NL * NL; t += 64) kout[t] = kl[t];
  if (diagp && l < NL) {
    const int rr = s_fdof[w][l];
    if (rr >= 0) unsafeAtomicAdd(diagp + rr, (double)kl[l * NL + l]);
  }
  const bool lift = rhsp != nullptr && a.gfull != nullptr;
  if (lift && __ballot(l < NL && s_fdof[w][l < NL ? l : 0] < 0) != 0) {
    for (int t = l; t < NL * NL; t += 64) {
      const int i = t / NL, j = t - i * NL;
      const int ri = s_fdof[w][i];
      if (ri < 0 || s_fdof[w][j] >= 0) continue;
      const double gv = a.gfull[s_gdof[w][j]];
      if (gv != 0.0) unsafeAtomicAdd(rhsp + ri, -(double)kl[t] * gv);
    }
  }
}

// Pass 2: vals[k] = sum over the elements shared by row and column of the
// matching element-matrix entries (fixed order -> bitwise reproducible).
template <int DIM, int P, int PROB, typename T>
__global__ void __launch_bounds__(256)
    k_gather(iga_patch_t pt, int vec, int64_t nnz, const int32_t* __restrict__ rows,
             const int32_t* __restrict__ cols, const int32_t* __restrict__ freed,
             const T* __restrict__ ke_all, T* __restrict__ vals_all, int batch, int64_t n_e) {
  using S = Shape<DIM, P, PROB>;
  constexpr int N1 = S::N1, NL = S::NL, NLV = S::NLV;
  for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < nnz * batch;
       kk += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bi = kk / nnz, k = kk - bi * nnz;
    const T* ke = ke_all + (size_t)bi * n_e * NLV * NLV;
    T* vals = vals_all + (size_t)bi * nnz;
    const int gi = freed[rows[k]], gj = freed[cols[k]];
    const int si = gi / vec, ci = gi - si * vec;
    const int sj = gj / vec, cj = gj - sj * vec;
    if (PROB == IGA_POISSON && ci != cj) {
      vals[k] = (T)0;
      continue;
    }
    int id[3] = {0, 0, 0}, jd[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    {
      int ri = si, rj = sj;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        id[d] = ri % pt.n_cp[d];
        ri /= pt.n_cp[d];
        jd[d] = rj % pt.n_cp[d];
        rj /= pt.n_cp[d];
        lo[d] = max(pt.elo[d][id[d]], pt.elo[d][jd[d]]);
        hi[d] = min(pt.ehi[d][id[d]], pt.ehi[d][jd[d]]);
      }
    }
    double sum = 0.0;
    for (int e2 = lo[2]; e2 <= hi[2]; ++e2)
      for (int e1 = lo[1]; e1 <= hi[1]; ++e1)
        for (int e0 = lo[0]; e0 <= hi[0]; ++e0) {
          const int ee[3] = {e0, e1, e2};
          int li = 0, lj = 0, stride = 1, e = 0, estride = 1;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            const int f = pt.first[d][ee[d]];
            li += (id[d] - f) * stride;
            lj += (jd[d] - f) * stride;
            stride *= N1;
            e += ee[d] * estride;
            estride *= pt.n_elem[d];
          }
          const size_t base = (size_t)e * NLV * NLV;
          if constexpr (PROB == IGA_POISSON)
            sum += (double)ke[base + (size_t)li * NL + lj];
          else
            sum += (double)ke[base + (size_t)(li * DIM + ci) * NLV + lj * DIM + cj];
        }
    vals[k] = (T)sum;
  }
}

// Pass 2 (row form): one wavefront per CSR row. For each element supporting
// the row's basis function (fixed order) the wave reads that element's
// matrix row -- NLV contiguous values, a coalesced read -- and accumulates it
// into a dense stencil box of the row held in LDS. Lane groups take
// different elements into separate boxes that are summed in a fixed order:
// bitwise reproducible. The box is then written out in CSR order: columns
// are sorted by dof, and dofs grow with the box order (last direction
// slowest, component fastest), so the slot of a present box entry is the
// number of present entries before it -- a ballot and a popcount. An entry
// is present when its control point lies in the patch, its dof is free and
// the two functions share an element. The row's diagonal (the box centre)
// is added to diag when it is given.
template <int DIM, int P, int PROB, typename T>
__global__ void __launch_bounds__(256)
    k_gather_rows(iga_patch_t pt, int vec, int32_t n_rows, const int32_t* __restrict__ indptr,
                  const int32_t* __restrict__ cols, const int32_t* __restrict__ freed,
                  const int32_t* __restrict__ f2f, const T* __restrict__ ke_all,
                  T* __restrict__ vals_all, double* __restrict__ diag_all, int32_t n_free,
                  int64_t nnz, int64_t n_e) {
  using S = Shape<DIM, P, PROB>;
  constexpr int N1 = S::N1, NLV = S::NLV;
  constexpr int VEC = S::VEC;
  constexpr int BW = 2 * P + 1;                           // stencil box width per direction
  constexpr int BOX = DIM == 2 ? BW * BW : BW * BW * BW;  // control points in the box
  constexpr int BOXV = BOX * VEC;
  constexpr int NLVP = NLV <= 8 ? 8 : NLV <= 16 ? 16 : NLV <= 32 ? 32 : 64;  // lanes per element
  constexpr int GROUPS = kWave / NLVP;                                       // elements per step
  constexpr int WAVES = 4;
  __shared__ double box[WAVES][GROUPS][BOXV];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  const int row = blockIdx.x * WAVES + wave;
  if (row >= n_rows) return;
  const int64_t b = blockIdx.y;
  const T* ke = ke_all + b * n_e * NLV * NLV;
  T* vals = vals_all + b * nnz;
  double* mybox = &box[wave][0][0];
  for (int t = lane; t < GROUPS * BOXV; t += kWave) mybox[t] = 0.0;
  wave_lds_sync();

  const int gi = freed[row];
  const int si = gi / vec, ci = gi - si * vec;
  int id[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, ne[3] = {1, 1, 1};
  {
    int rem = si;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      id[d] = rem % pt.n_cp[d];
      rem /= pt.n_cp[d];
      lo[d] = pt.elo[d][id[d]];
      ne[d] = pt.ehi[d][id[d]] - lo[d] + 1;
    }
  }
  const int grp = lane / NLVP, gl = lane % NLVP;
  // supporting elements as the fixed (P+1)^DIM grid from the first one;
  // every element-matrix read of the row is issued before any accumulation
  constexpr int NSUP = S::NL;
  constexpr int STEPS = (NSUP + GROUPS - 1) / GROUPS;
  constexpr int CPL = (NLV + NLVP - 1) / NLVP;  // columns per lane
  double vbuf[STEPS][CPL];
  int obuf[STEPS][CPL];
#pragma unroll
  for (int st = 0; st < STEPS; ++st) {
    const int sidx = st * GROUPS + grp;
    int es[3] = {0, 0, 0};
    bool ok = sidx < NSUP;
    {
      int rem = sidx;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int sd = rem % N1;
        rem /= N1;
        ok = ok && sd < ne[d];
        es[d] = lo[d] + (ok ? sd : 0);
      }
    }
    int e = 0, estride = 1, li = 0, lstride = 1, fo[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int f = pt.first[d][es[d]];
      e += es[d] * estride;
      estride *= pt.n_elem[d];
      li += (id[d] - f) * lstride;
      lstride *= N1;
      fo[d] = f - id[d] + P;  // box offset of the element's first function
    }
    const int lrow = PROB == IGA_POISSON ? li : li * VEC + ci;
    const T* krow = ke + (size_t)e * NLV * NLV + (size_t)lrow * NLV;
#pragma unroll
    for (int u = 0; u < CPL; ++u) {
      const int c = gl + u * NLVP;
      obuf[st][u] = -1;
      vbuf[st][u] = 0.0;
      if (!ok || c >= NLV) continue;
      const int f = PROB == IGA_POISSON ? c : c / VEC;
      const int cj = PROB == IGA_POISSON ? 0 : c - f * VEC;
      int rem = f, o = 0, ostride = 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        o += (fo[d] + rem % N1) * ostride;
        rem /= N1;
        ostride *= BW;
      }
      obuf[st][u] = grp * BOXV + o * VEC + cj;
      vbuf[st][u] = (double)krow[c];
    }
  }
#pragma unroll
  for (int st = 0; st < STEPS; ++st) {
#pragma unroll
    for (int u = 0; u < CPL; ++u)
      if (obuf[st][u] >= 0) mybox[obuf[st][u]] += vbuf[st][u];
    wave_lds_sync();
  }
  const int k0 = indptr[row], k1 = indptr[row + 1];
  double* diag = diag_all ? diag_all + b * n_free : nullptr;
  if (vec == VEC) {
    constexpr int CENTER = DIM == 2 ? P * BW + P : (P * BW + P) * BW + P;
    int base = 0;
#pragma unroll
    for (int t0 = 0; t0 < BOXV; t0 += kWave) {
      const int t = t0 + lane;
      bool present = false;
      if (t < BOXV) {
        const int node = t / VEC, cj = t - node * VEC;
        int rem = node, full = 0, stride = 1;
        bool ok = true;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const int jd = id[d] + rem % BW - P;
          rem /= BW;
          ok = ok && jd >= 0 && jd < pt.n_cp[d];
          if (ok) ok = max(lo[d], pt.elo[d][jd]) <= min(lo[d] + ne[d] - 1, pt.ehi[d][jd]);
          full += jd * stride;
          stride *= pt.n_cp[d];
        }
        present = ok && f2f[(size_t)full * VEC + cj] >= 0;
      }
      const uint64_t m = __ballot(present);
      const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
      if (present && k0 + slot < k1) {
        double acc = mybox[t];
#pragma unroll
        for (int g = 1; g < GROUPS; ++g) acc += mybox[g * BOXV + t];
        vals[k0 + slot] = (T)acc;
        if (diag && t == CENTER * VEC + ci) diag[row] += acc;
      }
      base += __popcll(m);
    }
  } else {
    // vector Poisson: components decouple, the box holds one per point
    if (GROUPS > 1)
      for (int t = lane; t < BOXV; t += kWave) {
        double acc = mybox[t];
#pragma unroll
        for (int g = 1; g < GROUPS; ++g) acc += mybox[g * BOXV + t];
        mybox[t] = acc;
      }
    wave_lds_sync();
    for (int k = k0 + lane; k < k1; k += kWave) {
      const int gj = freed[cols[k]];
      const int sj = gj / vec, cjj = gj - sj * vec;
      int rem = sj, o = 0, ostride = 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int jd = rem % pt.n_cp[d];
        rem /= pt.n_cp[d];
        o += (jd - id[d] + P) * ostride;
        ostride *= BW;
      }
      const double v = cjj == ci ? mybox[o * VEC] : 0.0;
      vals[k] = (T)v;
      if (diag && gj == gi) diag[row] += v;
    }
  }
}

// Pass 2 (control-point form): one wavefront per control point, covering
// the rows of all its components (VEC rows share the supporting elements,
// the stencil box and the column set). Per supporting element the wave reads
// the VEC adjacent element-matrix rows -- VEC * NLV contiguous values -- and
// adds them into the component boxes in LDS with atomic adds: within one
// instruction every lane hits its own box entry and a wave's LDS operations
// execute in program order, so the summation order is fixed (bitwise
// reproducible) with no read-modify-write round trip per element. Every
// element-matrix read is issued unconditionally and at once, then the
// free-dof reads of the column test (in the patch, free, overlapping
// support); on a regular patch the support tables are closed forms, so no
// dependent read stands between the point index and the element matrices.
// Write-out as in the row form, once per free component row.
template <int DIM, int P, int PROB, typename T, bool REG, bool SYM>
__global__ void __launch_bounds__(256)
    k_gather_points(iga_patch_t pt, int32_t p_begin, int32_t p_end, const int32_t* __restrict__ indptr,
                    const int32_t* __restrict__ f2f, const T* __restrict__ ke_all,
                    T* __restrict__ vals_all, double* __restrict__ diag_all, int32_t n_free,
                    int64_t nnz, int64_t n_e) {
  using S = Shape<DIM, P, PROB>;
  constexpr int N1 = S::N1, NLV = S::NLV;
  constexpr int VEC = S::VEC;
  constexpr int BW = 2 * P + 1;
  constexpr int BOX = DIM == 2 ? BW * BW : BW * BW * BW;
  constexpr int BOXV = BOX * VEC;            // entries of one row's box
  constexpr int VALS = VEC * NLV;            // values per supporting element
  constexpr int NLVP = VALS <= 8 ? 8 : VALS <= 16 ? 16 : VALS <= 32 ? 32 : 64;
  constexpr int GROUPS = kWave / NLVP;
  constexpr int CPL = (VALS + NLVP - 1) / NLVP;
  constexpr int WAVES = 4;
  constexpr int CHUNKS = (BOXV + kWave - 1) / kWave;
  __shared__ double box[WAVES][GROUPS * VEC * BOXV];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  // control points [p_begin, p_end)
  const int pnt = p_begin + blockIdx.x * WAVES + wave;
  if (pnt >= p_end) return;
  const int64_t b = blockIdx.y;
  constexpr int NL = S::NL;
  constexpr int KE = SYM ? sym_ke_values(NL, VEC) : NLV * NLV;  // stored values per element
  const T* ke = ke_all + b * n_e * KE;
  T* vals = vals_all + b * nnz;
  double* mybox = box[wave];
  for (int t = lane; t < GROUPS * VEC * BOXV; t += kWave) mybox[t] = 0.0;
  // support tables, closed forms on a regular patch
  auto elo = [&](int d, int i) {
    if constexpr (REG) return max(0, i - P);
    else return pt.elo[d][i];
  };
  auto ehi = [&](int d, int i) {
    if constexpr (REG) return min(pt.n_elem[d] - 1, i);
    else return pt.ehi[d][i];
  };
  auto first = [&](int d, int e) {
    if constexpr (REG) return e;
    else return pt.first[d][e];
  };

  int id[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, ne[3] = {1, 1, 1}, rowc[VEC];
  {
    int rem = pnt;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      id[d] = rem % pt.n_cp[d];
      rem /= pt.n_cp[d];
      lo[d] = elo(d, id[d]);
      ne[d] = ehi(d, id[d]) - lo[d] + 1;
    }
  }
#pragma unroll
  for (int c = 0; c < VEC; ++c) rowc[c] = f2f[(size_t)pnt * VEC + c];
  // lane part of the box offset of each value it reads
  const int grp = lane / NLVP, gl = lane % NLVP;
  int loff[CPL];
#pragma unroll
  for (int u = 0; u < CPL; ++u) {
    const int v = gl + u * NLVP;
    const int rc = v / NLV, c = v - rc * NLV;
    const int f = c / VEC, cj = c - f * VEC;
    int rem = f, o = 0, ostride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      o += (rem % N1) * ostride;
      rem /= N1;
      ostride *= BW;
    }
    loff[u] = v < VALS ? (grp * VEC + rc) * BOXV + o * VEC + cj : -1;
  }
  constexpr int NSUP = S::NL;
  constexpr int STEPS = (NSUP + GROUPS - 1) / GROUPS;
  // element-matrix reads in flight per batch: at most 16 values per lane
  constexpr int SB = CPL >= 16 ? 1 : (16 / CPL < STEPS ? 16 / CPL : STEPS);
  T vbuf[SB][CPL];
  int sofs[SB];  // element part of the box offset, -1 when no such element
  auto issue = [&](int st, int sb) {
    const int sidx = st * GROUPS + grp;
    int es[3] = {0, 0, 0};
    bool ok = sidx < NSUP;
    {
      int rem = sidx;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int sd = rem % N1;
        rem /= N1;
        ok = ok && sd < ne[d];
        es[d] = lo[d] + (ok ? sd : 0);
      }
    }
    int e = 0, estride = 1, li = 0, lstride = 1, o = 0, ostride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int f = first(d, es[d]);
      e += es[d] * estride;
      estride *= pt.n_elem[d];
      li += (id[d] - f) * lstride;
      lstride *= N1;
      o += (f - id[d] + P) * ostride;  // box offset of the element's first function
      ostride *= BW;
    }
    sofs[sb] = ok ? o * VEC : -1;
    // The reads are unconditional (an absent element reads the first
    // supporting one, skipped when adding): a conditional read would make
    // the compiler wait for each one before issuing the next.
    if constexpr (SYM) {
      // block (li, f) of the row: stored as (li, f) when li <= f, else as
      // the transpose of (f, li)
      const T* kel = ke + (size_t)e * KE;
#pragma unroll
      for (int u = 0; u < CPL; ++u) {
        const int v = min(gl + u * NLVP, VALS - 1);
        const int rc = v / NLV, c = v - rc * NLV;
        const int f = c / VEC, cj = c - f * VEC;
        const bool up = li <= f;
        const int blk = up ? sym_block(li, f, NL) : sym_block(f, li, NL);
        vbuf[sb][u] = kel[blk * (VEC * VEC) + (up ? rc * VEC + cj : cj * VEC + rc)];
      }
    } else {
      // rows li*VEC .. li*VEC + VEC-1 of the element matrix are contiguous
      const T* krow = ke + (size_t)e * NLV * NLV + (size_t)li * VEC * NLV;
#pragma unroll
      for (int u = 0; u < CPL; ++u) {
        const int v = gl + u * NLVP;
        vbuf[sb][u] = krow[v < VALS ? v : VALS - 1];
      }
    }
  };
  auto accumulate = [&](int sb) {
    if (sofs[sb] < 0) return;
#pragma unroll
    for (int u = 0; u < CPL; ++u)
      if (loff[u] >= 0)
        __hip_atomic_fetch_add(mybox + sofs[sb] + loff[u], (double)vbuf[sb][u], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
  };
#pragma unroll
  for (int sb = 0; sb < SB; ++sb) issue(sb, sb);
  // column test, in flight with the element-matrix reads: free index of each
  // box entry, or -1 (unconditional reads: the point's own dof out of the box)
  int colf[CHUNKS];
#pragma unroll
  for (int ch = 0; ch < CHUNKS; ++ch) {
    const int t = ch * kWave + lane;
    colf[ch] = -1;
    if (t < BOXV) {
      const int node = t / VEC, cj = t - node * VEC;
      int rem = node, full = 0, stride = 1;
      bool ok = true;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int jd = id[d] + rem % BW - P;
        rem /= BW;
        ok = ok && jd >= 0 && jd < pt.n_cp[d];
        if (ok) ok = max(lo[d], elo(d, jd)) <= min(lo[d] + ne[d] - 1, ehi(d, jd));
        full += jd * stride;
        stride *= pt.n_cp[d];
      }
      const int fv = f2f[ok ? (size_t)full * VEC + cj : (size_t)pnt * VEC];
      colf[ch] = ok ? fv : -1;
    }
  }
  // the rows' CSR extents, in flight with the element-matrix reads
  // (unconditional: a prescribed row reads row 0's, unused)
  int k0r[VEC], k1r[VEC];
#pragma unroll
  for (int rc = 0; rc < VEC; ++rc) {
    const int row = rowc[rc] >= 0 ? rowc[rc] : 0;
    k0r[rc] = indptr[row];
    k1r[rc] = indptr[row + 1];
  }
  uint64_t mask[CHUNKS];
#pragma unroll
  for (int ch = 0; ch < CHUNKS; ++ch) mask[ch] = __ballot(colf[ch] >= 0);
  wave_lds_sync();
#pragma unroll
  for (int sb = 0; sb < SB; ++sb) accumulate(sb);
  for (int st0 = SB; st0 < STEPS; st0 += SB) {
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
      if (st0 + sb < STEPS) issue(st0 + sb, sb);
#pragma unroll
    for (int sb = 0; sb < SB; ++sb)
      if (st0 + sb < STEPS) accumulate(sb);
  }
  wave_lds_sync();
  constexpr int CENTER = DIM == 2 ? P * BW + P : (P * BW + P) * BW + P;
  double* diag = diag_all ? diag_all + b * n_free : nullptr;
#pragma unroll
  for (int rc = 0; rc < VEC; ++rc) {
    const int row = rowc[rc];
    if (row < 0) continue;
    const int k0 = k0r[rc], k1 = k1r[rc];
    int base = 0;
#pragma unroll
    for (int ch = 0; ch < CHUNKS; ++ch) {
      const int t = ch * kWave + lane;
      const uint64_t m = mask[ch];
      const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
      if (((m >> lane) & 1ull) && k0 + slot < k1) {
        double acc = mybox[rc * BOXV + t];
#pragma unroll
        for (int g = 1; g < GROUPS; ++g) acc += mybox[(g * VEC + rc) * BOXV + t];
        vals[k0 + slot] = (T)acc;
        if (diag && t == CENTER * VEC + rc) diag[row] += acc;
      }
      base += __popcll(m);
    }
  }
}

template <int DIM, int P, int PROB, typename T>
static int launch_assemble(const iga_patch_t& pt, const iga_rule_t& rule, const AsmDev& a,
                           const iga_asm_args_t& args, hipStream_t st) {
  using S = Shape<DIM, P, PROB>;
  constexpr int NL = S::NL, NLP = S::NLP, NLV = S::NLV, N1 = S::N1, NS = S::NS;
  const int nq1 = rule.n_q1;
  const int nq = DIM == 2 ? nq1 * nq1 : nq1 * nq1 * nq1;
  const int nqp = (nq + 3) & ~3;
  constexpr int NLT = (NL + 15) / 16 * 16;
  // 256 threads: lanes map to basis functions, the slots loop over Gauss points
  const int threads = 256;
  static_assert(NLP <= 256, "one slot of basis lanes must fit the workgroup");
  constexpr bool lds_combine = PROB == IGA_ELASTICITY && NLT == 16 && DIM == 2;
  const size_t lds = sizeof(double) * (2 * DIM * nq1 * N1 + nq * NL + nq) +
                     sizeof(T) * (2 * DIM * nqp * NLT) + sizeof(int) * (NL * 3 + NL + 2) +
                     (lds_combine ? sizeof(double) * DIM * DIM * 256 : 0) +
                     sizeof(double) * (NL * (DIM + 1) + nq * (2 + DIM + DIM * DIM)) + 16;
  if (lds > 64 * 1024) return IGA_ERR_UNSUPPORTED;
  int n_e = 1;
  for (int d = 0; d < DIM; ++d) n_e *= pt.n_elem[d];
  if (n_e == 0) return IGA_OK;
  const int batch = args.batch > 1 ? args.batch : 1;
  if (batch > 1 && (a.rhs || a.diag) && args.n_free <= 0) return IGA_ERR_ARG;
  T* ke = reinterpret_cast<T*>(args.ke_work);
  if (args.vals) {
    const int64_t need = (int64_t)batch * n_e * NLV * NLV * sizeof(T);
    if (!ke || args.ke_work_bytes < need) return IGA_ERR_SHAPE;
    if (!args.rows || !args.cols || !args.freed) return IGA_ERR_ARG;
  } else {
    ke = nullptr;
  }
  const int passes = args.passes ? args.passes : 3;
  // with the row-form gather, the diagonal comes from the assembled rows
  const bool row_gather = args.vals && args.nnz > 0 && args.indptr && args.n_free > 0;
  AsmDev ae = a;
  if (row_gather) ae.diag = nullptr;
  int n_pts = 1;
  for (int d = 0; d < DIM; ++d) n_pts *= pt.n_cp[d];
  // planar, at most 4 x 4 functions and Gauss points: one wave per element
  const bool planar_w = DIM == 2 && P <= 3 && nq1 <= 4 && a.vec == S::VEC;
  const bool point_gather = row_gather && args.vec == S::VEC;
  // in FP64 the planar element pass and the point gather share the packed
  // symmetric layout of the element matrices (both passes see the same
  // arguments): half the bytes written and read. In FP32 the gather's reads
  // of the transposed blocks, 16 B each from as many rows, cost more than
  // the halved bytes save (measured), so the rows stay whole.
  const bool sym = std::is_same_v<T, double> && planar_w && point_gather && !args.ke_full;
  auto element_w = [&](int e_begin, int e_end, hipStream_t s) {
    if constexpr (DIM == 2 && P <= 3) {
      const dim3 grid((unsigned)((e_end - e_begin + 3) / 4), batch);
#define IGA_EW(REG, SYM)                                                                          \
  hipLaunchKernelGGL((k_element_w<P, PROB, T, REG, SYM>), grid, dim3(256), 0, s, pt, nq1, rule.pts, \
                     rule.wts, ae, ke, n_e, e_begin, e_end)
      if (pt.regular) {
        if (sym) IGA_EW(true, true);
        else IGA_EW(true, false);
      } else {
        if (sym) IGA_EW(false, true);
        else IGA_EW(false, false);
      }
#undef IGA_EW
    }
  };
  auto gather_points = [&](int p_begin, int p_end, hipStream_t s) {
    const dim3 grid((unsigned)((p_end - p_begin + 3) / 4), batch);
#define IGA_GP(REG, SYM)                                                                           \
  hipLaunchKernelGGL((k_gather_points<DIM, P, PROB, T, REG, SYM>), grid, dim3(256), 0, s, pt, p_begin, \
                     p_end, args.indptr, args.full_to_free, ke, reinterpret_cast<T*>(args.vals),     \
                     args.diag, args.n_free, args.nnz, (int64_t)n_e)
    if constexpr (DIM == 2 && P <= 3) {
      if (sym) {
        if (pt.regular) IGA_GP(true, true);
        else IGA_GP(false, true);
        return;
      }
    }
    if (pt.regular) IGA_GP(true, false);
    else IGA_GP(false, false);
#undef IGA_GP
  };
  if (passes & 1) {
    bool done = false;
    if (planar_w) {
      element_w(0, n_e, st);
      done = true;
    }
    if constexpr (DIM == 3 && P <= 2 && PROB == IGA_POISSON) {
      // scalar Poisson, p + 1 Gauss points per direction: one wave per element
      if (nq1 == P + 1 && a.vec == 1) {
        const dim3 grid((unsigned)((n_e + 3) / 4), batch);
        if (pt.regular)
          hipLaunchKernelGGL((k_element_w3<P, T, true>), grid, dim3(256), 0, st, pt, rule.pts,
                             rule.wts, ae, ke, n_e);
        else
          hipLaunchKernelGGL((k_element_w3<P, T, false>), grid, dim3(256), 0, st, pt, rule.pts,
                             rule.wts, ae, ke, n_e);
        done = true;
      }
    }
    if (!done)
      hipLaunchKernelGGL((k_element<DIM, P, PROB, T>), dim3(n_e, batch), dim3(threads), lds, st,
                         pt, nq1, rule.pts, rule.wts, ae, ke);
    IGA_CHECK_LAUNCH();
  }
  if ((passes & 2) && args.vals && args.nnz > 0) {
    if (point_gather) {
      gather_points(0, n_pts, st);
    } else if (row_gather) {
      const unsigned blocks = (unsigned)((args.n_free + 3) / 4);
      hipLaunchKernelGGL((k_gather_rows<DIM, P, PROB, T>), dim3(blocks, batch), dim3(256), 0, st,
                         pt, args.vec, args.n_free, args.indptr, args.cols, args.freed,
                         args.full_to_free, ke, reinterpret_cast<T*>(args.vals), args.diag,
                         args.n_free, args.nnz, (int64_t)n_e);
    } else {
      const int64_t blocks = std::min<int64_t>((args.nnz * batch + 255) / 256, 256 * 64);
      hipLaunchKernelGGL((k_gather<DIM, P, PROB, T>), dim3((unsigned)blocks), dim3(256), 0, st,
                         pt, args.vec, args.nnz, args.rows, args.cols, args.freed, ke,
                         reinterpret_cast<T*>(args.vals), batch, (int64_t)n_e);
    }
    IGA_CHECK_LAUNCH();
  }
  return IGA_OK;
}

template <int DIM, int P>
static int dispatch_prob(const iga_patch_t& pt, const iga_rule_t& rule, const AsmDev& a,
                         const iga_asm_args_t& args, hipStream_t st) {
  const bool f32 = args.precision == IGA_FP32;
  if (a.problem == IGA_POISSON)
    return f32 ? launch_assemble<DIM, P, IGA_POISSON, float>(pt, rule, a, args, st)
               : launch_assemble<DIM, P, IGA_POISSON, double>(pt, rule, a, args, st);
  return f32 ? launch_assemble<DIM, P, IGA_ELASTICITY, float>(pt, rule, a, args, st)
             : launch_assemble<DIM, P, IGA_ELASTICITY, double>(pt, rule, a, args, st);
}

template <int DIM, int P, int PROB>
static int64_t ke_bytes(int64_t n_e, int precision) {
  constexpr int NLV = Shape<DIM, P, PROB>::NLV;
  return n_e * NLV * NLV * (precision == IGA_FP32 ? 4 : 8);
}

}  // namespace iga

extern "C" int64_t iga_assemble_workspace(const iga_patch_t* patch, int32_t problem,
                                          int32_t precision) {
  using namespace iga;
  if (!patch) return -1;
  int64_t n_e = 1;
  for (int d = 0; d < patch->dim; ++d) n_e *= patch->n_elem[d];
  const int key = patch->dim * 10 + patch->degree;
  const bool el = problem == IGA_ELASTICITY;
#define IGA_W(D, P) \
  return el ? ke_bytes<D, P, IGA_ELASTICITY>(n_e, precision) : ke_bytes<D, P, IGA_POISSON>(n_e, precision)
  switch (key) {
    case 21: IGA_W(2, 1);
    case 22: IGA_W(2, 2);
    case 23: IGA_W(2, 3);
    case 24: IGA_W(2, 4);
    case 31: IGA_W(3, 1);
    case 32: IGA_W(3, 2);
    default: return -1;
  }
#undef IGA_W
}

extern "C" int iga_assemble(const iga_patch_t* patch, const iga_rule_t* rule,
                            const iga_asm_args_t* args, void* stream) {
  using namespace iga;
  if (!patch || !rule || !args) return IGA_ERR_ARG;
  AsmDev a;
  a.problem = args->problem;
  a.vec = args->vec;
  a.precision = args->precision;
  a.coef_kind = args->coef_kind;
  a.coef_const = args->coef_const;
  a.coef_points = args->coef_points;
  for (int i = 0; i < 36; ++i) a.cmat[i] = args->cmat[i];
  // E^cd_ab from the Voigt matrix: the strain s = (p, q) holds d u_p / d x_q
  // (and d u_q / d x_p when p != q)
  for (int t = 0; t < 81; ++t) a.ecd[t] = 0.0;
  if (args->problem == IGA_ELASTICITY) {
    const int D = patch->dim, NSv = D == 2 ? 3 : 6;
    const int vp2[3] = {0, 1, 0}, vq2[3] = {0, 1, 1};
    const int vp3[6] = {0, 1, 2, 1, 0, 0}, vq3[6] = {0, 1, 2, 2, 2, 1};
    const int* vp = D == 2 ? vp2 : vp3;
    const int* vq = D == 2 ? vq2 : vq3;
    auto holds = [&](int s, int u, int x) {
      return (u == vp[s] && x == vq[s]) || (vp[s] != vq[s] && u == vq[s] && x == vp[s]);
    };
    for (int cc = 0; cc < D; ++cc)
      for (int dd = 0; dd < D; ++dd)
        for (int aa = 0; aa < D; ++aa)
          for (int bb = 0; bb < D; ++bb) {
            double v = 0.0;
            for (int s1 = 0; s1 < NSv; ++s1)
              for (int s2 = 0; s2 < NSv; ++s2)
                if (holds(s1, aa, cc) && holds(s2, bb, dd)) v += args->cmat[s1 * NSv + s2];
            a.ecd[((cc * D + dd) * D + aa) * D + bb] = v;
          }
  }
  for (int t = 0; t < 16; ++t) a.ecd32[t] = (float)a.ecd[t];
  // the eight E^cd_ab that vanish when C_02 = C_12 = 0 (planar)
  a.ortho = 0;
  if (args->problem == IGA_ELASTICITY && patch->dim == 2) {
    const int zero[8] = {1, 2, 4, 7, 8, 11, 13, 14};
    a.ortho = 1;
    for (int z : zero) a.ortho &= a.ecd[z] == 0.0;
  }
  a.src_kind = args->src_kind;
  for (int i = 0; i < 3; ++i) a.src_const[i] = args->src_const[i];
  a.src_points = args->src_points;
  a.f2f = args->full_to_free;
  a.gfull = args->gfull;
  a.rhs = args->rhs;
  a.diag = args->diag;
  a.n_free = args->n_free;
  if (!a.f2f) return IGA_ERR_ARG;
  if (a.problem != IGA_POISSON && a.problem != IGA_ELASTICITY) return IGA_ERR_ARG;
  if (a.problem == IGA_ELASTICITY && a.vec != patch->dim) return IGA_ERR_SHAPE;
  if (a.vec < 1 || a.vec > 3) return IGA_ERR_SHAPE;
  if (args->precision != IGA_FP64 && args->precision != IGA_FP32) return IGA_ERR_UNSUPPORTED;
  if (a.coef_kind == IGA_COEF_POINTS && !a.coef_points) return IGA_ERR_ARG;
  if (a.src_kind == IGA_SRC_POINTS && !a.src_points) return IGA_ERR_ARG;
  if (rule->n_q1 < 1 || rule->n_q1 > 8) return IGA_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  const int key = patch->dim * 10 + patch->degree;
  switch (key) {
    case 21: return dispatch_prob<2, 1>(*patch, *rule, a, *args, st);
    case 22: return dispatch_prob<2, 2>(*patch, *rule, a, *args, st);
    case 23: return dispatch_prob<2, 3>(*patch, *rule, a, *args, st);
    case 24: return dispatch_prob<2, 4>(*patch, *rule, a, *args, st);
    case 31: return dispatch_prob<3, 1>(*patch, *rule, a, *args, st);
    case 32: return dispatch_prob<3, 2>(*patch, *rule, a, *args, st);
    default: return IGA_ERR_UNSUPPORTED;
  }
}
