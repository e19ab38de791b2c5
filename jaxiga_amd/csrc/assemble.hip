// Fused Galerkin assembly for a single tensor-product NURBS patch on gfx950.
//
// One workgroup per element. The element's 1-D Bezier-extracted tables
// (N1 = C_d B, dN1 = C_d dB * 2/h for each direction) are staged in LDS; the
// rational basis, the FP64 Jacobian/determinant and the physical gradients
// are formed with one lane per (Gauss point, local basis function) and
// group-of-lanes reductions; the element matrix is contracted from the LDS
// gradient table and scattered straight into the CSR values of the free-free
// block with hardware FP64 atomics (slot found by a search of the row's
// sorted column list). The load (source work) and the Dirichlet lifting
// -K_fp g go to the free right-hand side in the same pass.
#include "common.hpp"

namespace iga {

struct AsmDev {
  int problem, vec;
  int coef_kind;
  double coef_const;
  const double* coef_points;
  double cmat[36];
  int src_kind;
  double src_const[3];
  const double* src_points;
  const int32_t* f2f;
  const double* gfull;
  const int32_t* indptr;
  const int32_t* cols;
  double* vals;
  double* rhs;
  double* diag;
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

// Strain vector of basis function i in component c: B[s] = sum over the
// Voigt pair of the matching gradient entries.
template <int DIM>
__device__ __forceinline__ void strain_col(const double* g, int c, double* Bs) {
  constexpr int NS = DIM == 2 ? 3 : 6;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    int a, b;
    voigt<DIM>(s, a, b);
    double v = 0.0;
    if (c == a) v += g[b];
    if (a != b && c == b) v += g[a];
    Bs[s] = v;
  }
}

template <int DIM, int P>
__global__ void __launch_bounds__(1024)
    k_assemble(iga_patch_t pt, int nq1, const double* __restrict__ qp,
               const double* __restrict__ qw, AsmDev a) {
  constexpr int N1 = P + 1;
  constexpr int NL = ipow(N1, DIM);
  constexpr int NLP = pow2ceil(NL);
  constexpr int NS = DIM == 2 ? 3 : 6;
  const int nq = DIM == 2 ? nq1 * nq1 : nq1 * nq1 * nq1;

  extern __shared__ double smem[];
  double* N1t = smem;                         // [DIM][nq1][N1]
  double* dN1t = N1t + DIM * nq1 * N1;        // [DIM][nq1][N1]
  double* G = dN1t + DIM * nq1 * N1;          // [nq][DIM][NL]
  double* Rs = G + nq * DIM * NL;             // [nq][NL]
  double* Wk = Rs + nq * NL;                  // [nq] weight * coefficient
  double* Wf = Wk + nq;                       // [nq] weight
  int* fdof = reinterpret_cast<int*>(Wf + nq);  // [NL*vec] free index or -1
  int* gdof = fdof + NL * 3;                  // [NL] scalar dofs

  const int e = blockIdx.x;
  int ei[3] = {0, 0, 0};
  {
    int rem = e;
    for (int d = 0; d < DIM; ++d) {
      ei[d] = rem % pt.n_elem[d];
      rem /= pt.n_elem[d];
    }
  }

  // ---- 1-D extracted tables -------------------------------------------
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
  __syncthreads();

  // ---- rational basis and push-forward per Gauss point --------------------
  for (int q = slot; q < nq; q += nslots) {
    int qi[3] = {q % nq1, (q / nq1) % nq1, q / (nq1 * nq1)};
    double N = 1.0, dN[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int o = (d * nq1 + qi[d]) * N1 + ai[d];
      const double v = N1t[o], dv = dN1t[o];
      N *= v;
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] *= (k == d) ? dv : v;
    }
    PointGeo<DIM, NLP> g;
    rational_pushforward<DIM, NLP>(active, N, dN, wgt, Pc, g);
    double wq = jac_ref * fabs(g.detJ);
#pragma unroll
    for (int d = 0; d < DIM; ++d) wq *= qw[qi[d]];
    if (active) {
#pragma unroll
      for (int k = 0; k < DIM; ++k) G[(q * DIM + k) * NL + lane] = g.dR[k];
      Rs[q * NL + lane] = g.R;
    }
    if (lane == 0) {
      const double coef =
          a.coef_kind == IGA_COEF_POINTS ? a.coef_points[(size_t)e * nq + q] : a.coef_const;
      Wk[q] = wq * coef;
      Wf[q] = wq;
    }
  }
  __syncthreads();

  // ---- element matrix, scatter, lifting -----------------------------------
  const bool want_lift = a.rhs != nullptr && a.gfull != nullptr;
  for (int idx = threadIdx.x; idx < NL * NL; idx += blockDim.x) {
    const int i = idx / NL, j = idx % NL;
    if (a.problem == IGA_POISSON) {
      double k = 0.0;
      for (int q = 0; q < nq; ++q) {
        double gg = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) gg += G[(q * DIM + d) * NL + i] * G[(q * DIM + d) * NL + j];
        k += Wk[q] * gg;
      }
      for (int c = 0; c < a.vec; ++c) {
        const int r = fdof[i * a.vec + c];
        if (r < 0) continue;
        const int cc = fdof[j * a.vec + c];
        if (cc >= 0) {
          if (a.vals) {
            const int hi = a.indptr[r + 1];
            const int s = lower_bound(a.cols, a.indptr[r], hi, cc);
            if (s < hi && a.cols[s] == cc) unsafeAtomicAdd(a.vals + s, k);
          }
          if (a.diag && i == j) unsafeAtomicAdd(a.diag + r, k);
        } else if (want_lift) {
          const double gv = a.gfull[(size_t)a.vec * gdof[j] + c];
          if (gv != 0.0) unsafeAtomicAdd(a.rhs + r, -k * gv);
        }
      }
    } else {
      double kb[DIM][DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c)
#pragma unroll
        for (int cp = 0; cp < DIM; ++cp) kb[c][cp] = 0.0;
      for (int q = 0; q < nq; ++q) {
        double gi[DIM], gj[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          gi[d] = G[(q * DIM + d) * NL + i];
          gj[d] = G[(q * DIM + d) * NL + j];
        }
        const double w = Wf[q];
#pragma unroll
        for (int cp = 0; cp < DIM; ++cp) {
          double bj[NS], cb[NS];
          strain_col<DIM>(gj, cp, bj);
#pragma unroll
          for (int s = 0; s < NS; ++s) {
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < NS; ++t) acc += a.cmat[s * NS + t] * bj[t];
            cb[s] = acc;
          }
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            double bi[NS];
            strain_col<DIM>(gi, c, bi);
            double acc = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) acc += bi[s] * cb[s];
            kb[c][cp] += w * acc;
          }
        }
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const int r = fdof[i * DIM + c];
        if (r < 0) continue;
        int row_lo = 0, row_hi = 0;
        if (a.vals) {
          row_lo = a.indptr[r];
          row_hi = a.indptr[r + 1];
        }
#pragma unroll
        for (int cp = 0; cp < DIM; ++cp) {
          const int cc = fdof[j * DIM + cp];
          const double k = kb[c][cp];
          if (cc >= 0) {
            if (a.vals) {
              const int s = lower_bound(a.cols, row_lo, row_hi, cc);
              if (s < row_hi && a.cols[s] == cc) unsafeAtomicAdd(a.vals + s, k);
            }
            if (a.diag && i == j && c == cp) unsafeAtomicAdd(a.diag + r, k);
          } else if (want_lift) {
            const double gv = a.gfull[(size_t)DIM * gdof[j] + cp];
            if (gv != 0.0) unsafeAtomicAdd(a.rhs + r, -k * gv);
          }
        }
      }
    }
  }

  // ---- source work --------------------------------------------------------
  if (a.rhs && a.src_kind != IGA_SRC_NONE) {
    for (int idx = threadIdx.x; idx < NL * a.vec; idx += blockDim.x) {
      const int i = idx / a.vec, c = idx % a.vec;
      const int r = fdof[i * a.vec + c];
      if (r < 0) continue;
      double f = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double fv = a.src_kind == IGA_SRC_CONST
                              ? a.src_const[c]
                              : a.src_points[((size_t)e * nq + q) * a.vec + c];
        f += Wf[q] * fv * Rs[q * NL + i];
      }
      unsafeAtomicAdd(a.rhs + r, f);
    }
  }
}

template <int DIM, int P>
static int launch_assemble(const iga_patch_t& pt, const iga_rule_t& rule, const AsmDev& a,
                           hipStream_t st) {
  constexpr int N1 = P + 1;
  constexpr int NL = ipow(N1, DIM);
  constexpr int NLP = pow2ceil(NL);
  const int nq1 = rule.n_q1;
  const int nq = DIM == 2 ? nq1 * nq1 : nq1 * nq1 * nq1;
  int threads = nq * NLP;
  threads = ((threads + kWave - 1) / kWave) * kWave;
  if (threads > 1024) threads = 1024;
  const size_t lds = sizeof(double) * (2 * DIM * nq1 * N1 + nq * DIM * NL + nq * NL + 2 * nq) +
                     sizeof(int) * (NL * 3 + NL);
  if (lds > 64 * 1024) return IGA_ERR_UNSUPPORTED;
  int n_e = 1;
  for (int d = 0; d < DIM; ++d) n_e *= pt.n_elem[d];
  if (n_e == 0) return IGA_OK;
  hipLaunchKernelGGL((k_assemble<DIM, P>), dim3(n_e), dim3(threads), lds, st, pt, nq1, rule.pts,
                     rule.wts, a);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

}  // namespace iga

extern "C" int iga_assemble(const iga_patch_t* patch, const iga_rule_t* rule,
                            const iga_asm_args_t* args, void* stream) {
  using namespace iga;
  if (!patch || !rule || !args) return IGA_ERR_ARG;
  AsmDev a;
  a.problem = args->problem;
  a.vec = args->vec;
  a.coef_kind = args->coef_kind;
  a.coef_const = args->coef_const;
  a.coef_points = args->coef_points;
  for (int i = 0; i < 36; ++i) a.cmat[i] = args->cmat[i];
  a.src_kind = args->src_kind;
  for (int i = 0; i < 3; ++i) a.src_const[i] = args->src_const[i];
  a.src_points = args->src_points;
  a.f2f = args->full_to_free;
  a.gfull = args->gfull;
  a.indptr = args->indptr;
  a.cols = args->cols;
  a.vals = args->vals;
  a.rhs = args->rhs;
  a.diag = args->diag;
  if (!a.f2f) return IGA_ERR_ARG;
  if (a.vals && (!a.indptr || !a.cols)) return IGA_ERR_ARG;
  if (a.problem == IGA_ELASTICITY && a.vec != patch->dim) return IGA_ERR_SHAPE;
  if (a.vec < 1 || a.vec > 3) return IGA_ERR_SHAPE;
  if (args->precision != IGA_FP64) return IGA_ERR_UNSUPPORTED;
  if (a.coef_kind == IGA_COEF_POINTS && !a.coef_points) return IGA_ERR_ARG;
  if (a.src_kind == IGA_SRC_POINTS && !a.src_points) return IGA_ERR_ARG;
  if (rule->n_q1 < 1 || rule->n_q1 > 8) return IGA_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  const int key = patch->dim * 10 + patch->degree;
  switch (key) {
    case 21: return launch_assemble<2, 1>(*patch, *rule, a, st);
    case 22: return launch_assemble<2, 2>(*patch, *rule, a, st);
    case 23: return launch_assemble<2, 3>(*patch, *rule, a, st);
    case 24: return launch_assemble<2, 4>(*patch, *rule, a, st);
    case 31: return launch_assemble<3, 1>(*patch, *rule, a, st);
    case 32: return launch_assemble<3, 2>(*patch, *rule, a, st);
    default: return IGA_ERR_UNSUPPORTED;
  }
}
