// Basis evaluation at arbitrary reference points of listed elements
// (BasisData: R, physical gradients, points, weights, boundary normals) and
// the point-load scatter used for Neumann work.
//
// One workgroup per listed element, one lane per (point, local basis
// function). Each lane evaluates its own 1-D Bernstein values at its point,
// applies the element's 1-D extraction rows (tensor-factored C = C_w (x) C_v
// (x) C_u), and the group of lanes of a point reduces the NURBS weight
// function, the geometry and the Jacobian in FP64.
#include "common.hpp"

namespace iga {

template <int DIM, int P>
__global__ void __launch_bounds__(1024)
    k_basis(iga_patch_t pt, int n_e, const int32_t* __restrict__ elems, int nq,
            const double* __restrict__ ref, int ref_shared, const double* __restrict__ wref,
            int wref_shared, const int32_t* __restrict__ sides, double* R, double* dR, double* x,
            double* w, double* normal, double* detj) {
  constexpr int N1 = P + 1;
  constexpr int NL = ipow(N1, DIM);
  constexpr int NLP = pow2ceil(NL);
  const int be = blockIdx.x;
  if (be >= n_e) return;
  const int e = elems[be];
  // every loop over directions is unrolled: indexing the by-value patch
  // struct or these small arrays with a run-time index would place them in
  // memory (promoted to LDS, sized for 1024 threads, or scratch)
  int ei[3] = {0, 0, 0};
  {
    int rem = e;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      ei[d] = rem % pt.n_elem[d];
      rem /= pt.n_elem[d];
    }
  }
  const int lane = threadIdx.x % NLP;
  const int slot = threadIdx.x / NLP;
  const int nslots = blockDim.x / NLP;
  const bool active = lane < NL;
  int ai[3] = {0, 0, 0};
  {
    int rem = lane;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      ai[d] = rem % N1;
      rem /= N1;
    }
  }
  double Pc[DIM], wgt = 1.0;
  if (active) {
    int stride = 1, sdof = 0;
#pragma unroll
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
  double h[DIM], Crow[DIM][N1];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    h[d] = pt.hi[d][ei[d]] - pt.lo[d][ei[d]];
    const double* C = pt.ext[d] + ((size_t)ei[d] * N1 + ai[d]) * N1;
#pragma unroll
    for (int b = 0; b < N1; ++b) Crow[d][b] = active ? C[b] : 0.0;
  }
  const int side = sides ? sides[be] : -1;

  for (int q = slot; q < nq; q += nslots) {
    const double* xi = ref + ((ref_shared ? 0 : (size_t)be * nq) + q) * DIM;
    double N = 1.0, dN[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double B[N1], dB[N1];
      bernstein<P>(xi[d], B, dB);
      double v = 0.0, dv = 0.0;
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        v += Crow[d][b] * B[b];
        dv += Crow[d][b] * dB[b];
      }
      dv *= 2.0 / h[d];
      N *= v;
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] *= (k == d) ? dv : v;
    }
    PointGeo<DIM, NLP> g;
    rational_pushforward<DIM, NLP>(active, N, dN, wgt, Pc, g);
    const size_t pq = (size_t)be * nq + q;
    if (active) {
      if (R) R[pq * NL + lane] = g.R;
      if (dR)
#pragma unroll
        for (int k = 0; k < DIM; ++k) dR[(pq * DIM + k) * NL + lane] = g.dR[k];
    }
    if (lane == 0) {
      if (x)
#pragma unroll
        for (int c = 0; c < DIM; ++c) x[pq * DIM + c] = g.x[c];
      if (detj) detj[pq] = g.detJ;
      const double wr = wref ? wref[(wref_shared ? 0 : (size_t)be * nq) + q] : 1.0;
      if (side < 0) {
        double jr = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) jr *= 0.5 * h[d];
        if (w) w[pq] = wr * fabs(g.detJ) * jr;
      } else {
        const int dd = side >> 1;
        const double sgn = (side & 1) ? 1.0 : -1.0;
        // outward co-normal n = J^{-1} e_dd (solves J n = e_dd)
        double n[DIM];
        if constexpr (DIM == 2) {
          const double id = 1.0 / g.detJ;
          const auto& J = g.J;
          n[0] = (dd == 0 ? J[1][1] : -J[0][1]) * id;
          n[1] = (dd == 0 ? -J[1][0] : J[0][0]) * id;
        } else {
          // column dd of the inverse = cross product of the two tangent rows / det
          const int t1 = (dd + 1) % 3, t2 = (dd + 2) % 3;
          auto jr3 = [&](int k, int c) { return k == 0 ? g.J[0][c] : (k == 1 ? g.J[1][c] : g.J[2][c]); };
          n[0] = (jr3(t1, 1) * jr3(t2, 2) - jr3(t1, 2) * jr3(t2, 1)) / g.detJ;
          n[1] = (jr3(t1, 2) * jr3(t2, 0) - jr3(t1, 0) * jr3(t2, 2)) / g.detJ;
          n[2] = (jr3(t1, 0) * jr3(t2, 1) - jr3(t1, 1) * jr3(t2, 0)) / g.detJ;
        }
        double nn = 0.0;
#pragma unroll
        for (int c = 0; c < DIM; ++c) nn += n[c] * n[c];
        nn = sqrt(nn);
        if (normal)
#pragma unroll
          for (int c = 0; c < DIM; ++c) normal[pq * DIM + c] = n[c] / nn * sgn;
        // rows of J and box sizes picked by value: a run-time index into g.J or
        // h would keep them in memory
        auto jr = [&](int k, int c) {
          double v = g.J[0][c];
#pragma unroll
          for (int r = 1; r < DIM; ++r) v = k == r ? g.J[r][c] : v;
          return v;
        };
        auto hs = [&](int k) {
          double v = h[0];
#pragma unroll
          for (int r = 1; r < DIM; ++r) v = k == r ? h[r] : v;
          return v;
        };
        double surf, ts = 1.0;
        if constexpr (DIM == 2) {
          const int t = 1 - dd;
          surf = sqrt(jr(t, 0) * jr(t, 0) + jr(t, 1) * jr(t, 1));
          ts = 0.5 * hs(t);
        } else {
          const int t1 = (dd + 1) % 3, t2 = (dd + 2) % 3;
          const double c0 = jr(t1, 1) * jr(t2, 2) - jr(t1, 2) * jr(t2, 1);
          const double c1 = jr(t1, 2) * jr(t2, 0) - jr(t1, 0) * jr(t2, 2);
          const double c2 = jr(t1, 0) * jr(t2, 1) - jr(t1, 1) * jr(t2, 0);
          surf = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
          ts = 0.25 * hs(t1) * hs(t2);
        }
        if (w) w[pq] = wr * surf * ts;
      }
    }
  }
}

template <int DIM, int P>
static int launch_basis(const iga_patch_t& pt, int n_e, const int32_t* elems, int nq,
                        const double* ref, int ref_shared, const double* wref, int wref_shared,
                        const int32_t* sides, double* R, double* dR, double* x, double* w,
                        double* normal, double* detj, hipStream_t st) {
  constexpr int NL = ipow(P + 1, DIM);
  constexpr int NLP = pow2ceil(NL);
  int threads = nq * NLP;
  threads = ((threads + kWave - 1) / kWave) * kWave;
  if (threads > 1024) threads = 1024;
  if (n_e == 0 || nq == 0) return IGA_OK;
  hipLaunchKernelGGL((k_basis<DIM, P>), dim3(n_e), dim3(threads), 0, st, pt, n_e, elems, nq, ref,
                     ref_shared, wref, wref_shared, sides, R, dR, x, w, normal, detj);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}

// rhs[f(vec*dof + c)] += sum_q w t_c R_i over listed elements; one thread per
// (element, local function, component).
__global__ void k_load_points(int n_e, int nq, int nl, int vec, const double* __restrict__ R,
                              const double* __restrict__ w, const double* __restrict__ t,
                              const int32_t* __restrict__ edofs, const int32_t* __restrict__ f2f,
                              double* rhs) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)n_e * nl * vec;
  if (idx >= total) return;
  const int c = idx % vec;
  const int i = (idx / vec) % nl;
  const int e = idx / ((long long)vec * nl);
  const int r = f2f[(size_t)vec * edofs[(size_t)e * nl + i] + c];
  if (r < 0) return;
  double s = 0.0;
  for (int q = 0; q < nq; ++q) {
    const size_t pq = (size_t)e * nq + q;
    s += w[pq] * t[pq * vec + c] * R[pq * nl + i];
  }
  unsafeAtomicAdd(rhs + r, s);
}

}  // namespace iga

extern "C" int iga_basis_eval(const iga_patch_t* patch, int32_t n_e, const int32_t* elems,
                              int32_t n_q, const double* ref, int32_t ref_shared,
                              const double* wref, int32_t wref_shared, const int32_t* sides,
                              double* R, double* dR, double* x, double* w, double* normal,
                              double* detj, void* stream) {
  using namespace iga;
  if (!patch || !elems || !ref) return IGA_ERR_ARG;
  if (n_e < 0 || n_q < 0) return IGA_ERR_SHAPE;
  hipStream_t st = (hipStream_t)stream;
  const int key = patch->dim * 10 + patch->degree;
#define IGA_B(D, P)                                                                         \
  return launch_basis<D, P>(*patch, n_e, elems, n_q, ref, ref_shared, wref, wref_shared, sides, \
                            R, dR, x, w, normal, detj, st)
  switch (key) {
    case 21: IGA_B(2, 1);
    case 22: IGA_B(2, 2);
    case 23: IGA_B(2, 3);
    case 24: IGA_B(2, 4);
    case 25: IGA_B(2, 5);
    case 31: IGA_B(3, 1);
    case 32: IGA_B(3, 2);
    case 33: IGA_B(3, 3);
    default: return IGA_ERR_UNSUPPORTED;
  }
#undef IGA_B
}

extern "C" int iga_load_points(int32_t n_e, int32_t n_q, int32_t n_l, int32_t vec,
                               const double* R, const double* w, const double* t,
                               const int32_t* edofs, const int32_t* full_to_free, double* rhs,
                               void* stream) {
  using namespace iga;
  if (!R || !w || !t || !edofs || !full_to_free || !rhs) return IGA_ERR_ARG;
  const long long total = (long long)n_e * n_l * vec;
  if (total == 0) return IGA_OK;
  const int threads = 256;
  const long long blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(k_load_points, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream,
                     n_e, n_q, n_l, vec, R, w, t, edofs, full_to_free, rhs);
  IGA_CHECK_LAUNCH();
  return IGA_OK;
}
