// Basis evaluation at arbitrary reference points of listed elements
// (BasisData: R, physical gradients, points, weights, boundary normals) and
// the point-load scatter used for Neumann work.
//
// A block covers several listed elements: first one lane per point forms the
// NURBS weight function, the geometry and the Jacobian in FP64 from the
// element's 1-D extracted rows (tensor-factored C = C_w (x) C_v (x) C_u) and
// the control points, then one lane per (point, function) writes the basis
// and its physical gradients, consecutive lanes to consecutive addresses.
#include "common.hpp"

namespace iga {

// Geometry at one point of one element, shared by the two phases.
template <int DIM>
struct PointFrame {
  double W, dW[DIM], inv[DIM][DIM];  // weight function, its parametric gradient, J^-1
};

// 1-D extracted values and parametric derivatives of the element's functions
// along every direction at the reference point xi.
template <int DIM, int P>
__device__ __forceinline__ void extracted_1d(const iga_patch_t& pt, const int* ei, const double* xi,
                                             double (&v)[DIM][P + 1], double (&dv)[DIM][P + 1],
                                             double (&h)[DIM]) {
  constexpr int N1 = P + 1;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double B[N1], dB[N1];
    bernstein<P>(xi[d], B, dB);
    h[d] = pt.hi[d][ei[d]] - pt.lo[d][ei[d]];
    const double* C = pt.ext[d] + (size_t)ei[d] * N1 * N1;
#pragma unroll
    for (int a = 0; a < N1; ++a) {
      double s = 0.0, ds = 0.0;
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        s += C[a * N1 + b] * B[b];
        ds += C[a * N1 + b] * dB[b];
      }
      v[d][a] = s;
      dv[d][a] = ds * (2.0 / h[d]);
    }
  }
}

// Co-normal J^-1 e_DD (unnormalised) and surface measure of the side normal
// to direction DD: the norm of the tangent row (2-D) or of the cross product
// of the two tangent rows (3-D), times the parametric side size factor.
template <int DIM, int DD>
__device__ __forceinline__ void boundary_frame(const double (&J)[DIM][DIM], const double (&inv)[DIM][DIM],
                                               const double (&h)[DIM], double (&n)[DIM], double& surf,
                                               double& ts) {
#pragma unroll
  for (int c = 0; c < DIM; ++c) n[c] = inv[c][DD];
  if constexpr (DIM == 2) {
    constexpr int T = 1 - DD;
    surf = sqrt(J[T][0] * J[T][0] + J[T][1] * J[T][1]);
    ts = 0.5 * h[T];
  } else {
    constexpr int T1 = (DD + 1) % 3, T2 = (DD + 2) % 3;
    const double c0 = J[T1][1] * J[T2][2] - J[T1][2] * J[T2][1];
    const double c1 = J[T1][2] * J[T2][0] - J[T1][0] * J[T2][2];
    const double c2 = J[T1][0] * J[T2][1] - J[T1][1] * J[T2][0];
    surf = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    ts = 0.25 * h[T1] * h[T2];
  }
}

// Three phases per block of EPB elements, everything a phase shares in LDS:
//  0. one lane per (element, function): the function's weight and weighted
//     control point (w, w x);
//  1. one lane per (element, point): the 1-D extracted values and
//     derivatives of the element's functions at the point (kept for phase
//     2), then the rational weight function, the geometry x, J = dx/dxi,
//     det J and J^-1 accumulated over the functions; x, w, normal and det J
//     are written here;
//  2. one lane per (element, point, function), functions fastest: R and the
//     physical gradients from the kept 1-D values and the point's frame, so
//     consecutive lanes write consecutive doubles.
template <int DIM, int P>
struct BasisLds {
  static constexpr int N1 = P + 1;
  static constexpr int NL = ipow(N1, DIM);
  // bytes for EPB elements of nq points
  __host__ __device__ static size_t ctl(int epb) { return sizeof(double) * (size_t)epb * NL * (DIM + 1); }
  __host__ __device__ static size_t vals(int epb, int nq) {
    return sizeof(double) * (size_t)epb * nq * DIM * N1 * 2;
  }
  __host__ __device__ static size_t frames(int epb, int nq) {
    return sizeof(PointFrame<DIM>) * (size_t)epb * nq;
  }
  __host__ __device__ static size_t total(int epb, int nq, bool basis) {
    return ctl(epb) + (basis ? vals(epb, nq) + frames(epb, nq) : 0);
  }
};

template <int DIM, int P, int EPB>
__global__ void __launch_bounds__(256)
    k_basis(iga_patch_t pt, int n_e, const int32_t* __restrict__ elems, int nq,
            const double* __restrict__ ref, int ref_shared, const double* __restrict__ wref,
            int wref_shared, const int32_t* __restrict__ sides, double* R, double* dR, double* x,
            double* w, double* normal, double* detj) {
  constexpr int N1 = P + 1;
  constexpr int NL = ipow(N1, DIM);
  using L = BasisLds<DIM, P>;
  extern __shared__ double lds[];
  const bool basis = R || dR;
  double* ctl = lds;                                           // [EPB][NL][DIM+1]
  double* vals = lds + L::ctl(EPB) / sizeof(double);           // [EPB*nq][DIM][N1][2]
  PointFrame<DIM>* frames = reinterpret_cast<PointFrame<DIM>*>(
      lds + (L::ctl(EPB) + L::vals(EPB, nq)) / sizeof(double));  // [EPB*nq]
  const int be0 = blockIdx.x * EPB;
  const int n_here = min(EPB, n_e - be0);
  auto elem_index = [&](int e, int (&ei)[3]) {
    int rem = e;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      ei[d] = rem % pt.n_elem[d];
      rem /= pt.n_elem[d];
    }
  };
  // ---- phase 0: the elements' weights and weighted control points ----
  for (int t = threadIdx.x; t < n_here * NL; t += blockDim.x) {
    const int le = t / NL, a = t % NL;
    int ei[3] = {0, 0, 0};
    elem_index(elems[be0 + le], ei);
    const int ai[3] = {a % N1, (a / N1) % N1, a / (N1 * N1)};
    int sdof = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      sdof += (pt.first[d][ei[d]] + ai[d]) * stride;
      stride *= pt.n_cp[d];
    }
    const double wa = pt.wgts[sdof];
    double* c = ctl + t * (DIM + 1);
    c[0] = wa;
#pragma unroll
    for (int k = 0; k < DIM; ++k) c[1 + k] = wa * pt.cpts[(size_t)sdof * DIM + k];
  }
  __syncthreads();
  // ---- phase 1: one lane per point ----
  for (int t = threadIdx.x; t < n_here * nq; t += blockDim.x) {
    const int le = t / nq, q = t % nq, be = be0 + le;
    int ei[3] = {0, 0, 0};
    elem_index(elems[be], ei);
    double v[DIM][N1], dv[DIM][N1], h[DIM];
    extracted_1d<DIM, P>(pt, ei, ref + ((ref_shared ? 0 : (size_t)be * nq) + q) * DIM, v, dv, h);
    if (basis) {
      double* vt = vals + (size_t)t * DIM * N1 * 2;
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int a = 0; a < N1; ++a) {
          vt[(d * N1 + a) * 2] = v[d][a];
          vt[(d * N1 + a) * 2 + 1] = dv[d][a];
        }
    }
    double W = 0.0, dW[DIM] = {}, X[DIM] = {}, Y[DIM][DIM] = {};
    const double* ce = ctl + le * NL * (DIM + 1);
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      const int ai[3] = {a % N1, (a / N1) % N1, a / (N1 * N1)};
      double N = 1.0, dN[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        N *= v[d][ai[d]];
#pragma unroll
        for (int k = 0; k < DIM; ++k) dN[k] *= k == d ? dv[d][ai[d]] : v[d][ai[d]];
      }
      const double* c = ce + a * (DIM + 1);
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
    double xp[DIM], J[DIM][DIM];
    const double iW = 1.0 / W;
#pragma unroll
    for (int c = 0; c < DIM; ++c) xp[c] = X[c] * iW;
#pragma unroll
    for (int k = 0; k < DIM; ++k)
#pragma unroll
      for (int c = 0; c < DIM; ++c) J[k][c] = (Y[k][c] - dW[k] * xp[c]) * iW;
    PointFrame<DIM> F;
    double det;
    F.W = W;
#pragma unroll
    for (int k = 0; k < DIM; ++k) F.dW[k] = dW[k];
    if constexpr (DIM == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double id = 1.0 / det;
      F.inv[0][0] = J[1][1] * id;
      F.inv[0][1] = -J[0][1] * id;
      F.inv[1][0] = -J[1][0] * id;
      F.inv[1][1] = J[0][0] * id;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      F.inv[0][0] = c00 * id;
      F.inv[1][0] = c01 * id;
      F.inv[2][0] = c02 * id;
      F.inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      F.inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      F.inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      F.inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      F.inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      F.inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
    if (basis) frames[t] = F;
    const size_t pq = (size_t)be * nq + q;
    if (x)
#pragma unroll
      for (int c = 0; c < DIM; ++c) x[pq * DIM + c] = xp[c];
    if (detj) detj[pq] = det;
    const double wr = wref ? wref[(wref_shared ? 0 : (size_t)be * nq) + q] : 1.0;
    const int side = sides ? sides[be] : -1;
    if (side < 0) {
      double jr = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) jr *= 0.5 * h[d];
      if (w) w[pq] = wr * fabs(det) * jr;
    } else {
      const double sgn = (side & 1) ? 1.0 : -1.0;
      // outward co-normal n = J^-1 e_dd, normalised; surface measure from the
      // tangent rows of J; the direction is dispatched at compile time so that
      // no array is indexed at run time
      double n[DIM], surf = 0.0, ts = 1.0;
      switch (side >> 1) {
        case 0: boundary_frame<DIM, 0>(J, F.inv, h, n, surf, ts); break;
        case 1: boundary_frame<DIM, 1>(J, F.inv, h, n, surf, ts); break;
        default: boundary_frame<DIM, DIM - 1>(J, F.inv, h, n, surf, ts); break;
      }
      double nn = 0.0;
#pragma unroll
      for (int c = 0; c < DIM; ++c) nn += n[c] * n[c];
      nn = sqrt(nn);
      if (normal)
#pragma unroll
        for (int c = 0; c < DIM; ++c) normal[pq * DIM + c] = n[c] / nn * sgn;
      if (w) w[pq] = wr * surf * ts;
    }
  }
  if (!basis) return;
  __syncthreads();
  // ---- phase 2: one lane per (point, function) ----
  for (int t = threadIdx.x; t < n_here * nq * NL; t += blockDim.x) {
    const int a = t % NL, pt_i = t / NL, le = pt_i / nq;
    const int ai[3] = {a % N1, (a / N1) % N1, a / (N1 * N1)};
    const double* vt = vals + (size_t)pt_i * DIM * N1 * 2;
    double N = 1.0, dN[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) dN[k] = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double vv = vt[(d * N1 + ai[d]) * 2], dd = vt[(d * N1 + ai[d]) * 2 + 1];
      N *= vv;
#pragma unroll
      for (int k = 0; k < DIM; ++k) dN[k] *= k == d ? dd : vv;
    }
    const PointFrame<DIM>& F = frames[pt_i];
    const double wa = ctl[(le * NL + a) * (DIM + 1)];
    const double iW = 1.0 / F.W;
    const double Ra = N * wa * iW;
    const size_t pq = (size_t)be0 * nq + pt_i;
    if (R) R[pq * NL + a] = Ra;
    if (dR) {
      double gp[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) gp[k] = (dN[k] * wa - Ra * F.dW[k]) * iW;
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += F.inv[c][k] * gp[k];
        dR[(pq * DIM + c) * NL + a] = s;
      }
    }
  }
}

template <int DIM, int P>
static int launch_basis(const iga_patch_t& pt, int n_e, const int32_t* elems, int nq,
                        const double* ref, int ref_shared, const double* wref, int wref_shared,
                        const int32_t* sides, double* R, double* dR, double* x, double* w,
                        double* normal, double* detj, hipStream_t st) {
  if (n_e == 0 || nq == 0) return IGA_OK;
  using L = BasisLds<DIM, P>;
  const bool basis = R || dR;
  // elements per block: the most that keep phase 1 within 256 lanes and the
  // block's LDS within 48 KB
  auto go = [&](auto E) -> int {
    constexpr int EPB = decltype(E)::value;
    const size_t lds = L::total(EPB, nq, basis);
    if (lds > 64 * 1024) return (int)IGA_ERR_UNSUPPORTED;
    hipLaunchKernelGGL((k_basis<DIM, P, EPB>), dim3((n_e + EPB - 1) / EPB), dim3(256), lds, st, pt,
                       n_e, elems, nq, ref, ref_shared, wref, wref_shared, sides, R, dR, x, w, normal,
                       detj);
    IGA_CHECK_LAUNCH();
    return (int)IGA_OK;
  };
  // and, for few elements (a boundary), enough blocks to spread over the GPU
  auto fits = [&](int epb) {
    return epb * nq <= 256 && L::total(epb, nq, basis) <= 48 * 1024 && (epb == 1 || n_e / epb >= 512);
  };
  if (fits(32)) return go(std::integral_constant<int, 32>{});
  if (fits(16)) return go(std::integral_constant<int, 16>{});
  if (fits(8)) return go(std::integral_constant<int, 8>{});
  if (fits(4)) return go(std::integral_constant<int, 4>{});
  return go(std::integral_constant<int, 1>{});
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
