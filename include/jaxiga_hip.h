/*
 * jaxiga_hip.h -- C ABI of the MI355X (gfx950) Galerkin hot path.
 *
 * Every entry point takes raw device pointers, sizes and a hipStream_t passed
 * as void*, returns 0 on success or a negative iga_status code, and never
 * throws. Buffers belong to the caller; the library keeps no global state
 * except the per-device autotune table, guarded by a mutex.
 *
 * Reference interfaces replaced (paths relative to the JAXIGA source tree):
 *   iga_basis_eval            space/evaluation.py:305-467  (evaluate -> BasisData)
 *   iga_assemble              methods/galerkin.py:53-148   (element_tangents +
 *                             _assemble_csr_values), forms/library.py:54-65,174-181
 *                             (Poisson / LinearElasticity element_tangent),
 *                             methods/galerkin.py:240-245  (rhs: source work and
 *                             the -K_fp g lifting), methods/galerkin.py:221-237
 *                             (tangent_diagonal)
 *   iga_load_points           methods/_common.py:312-318   (neumann_work -> load)
 *   iga_csr_count/iga_csr_fill methods/_common.py:203-264  (_cached_csr_pattern)
 *   iga_scan_exclusive        methods/_common.py:250-255   (indptr by cumsum)
 *   iga_csr_diag_inv          solvers/linear.py:334-343    (Jacobi inverse diagonal)
 *   iga_spmv                  solvers/linear.py:331-332    (segment-sum matvec)
 *   iga_cg_solve              solvers/linear.py:324-351    (jax cg, Jacobi M)
 *   iga_residual_norms        solvers/diagnostics.py:8-19  (independent ||Au-b||)
 *   iga_mf_apply / iga_mf_diag methods/galerkin.py:207-237 (tangent_matvec /
 *                             tangent_diagonal, sum-factorised)
 *   iga_rff_eval              (no reference counterpart: seeded GRF inputs)
 */
#ifndef JAXIGA_HIP_H
#define JAXIGA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum iga_status {
  IGA_OK = 0,
  IGA_ERR_SHAPE = -1,       /* inconsistent sizes */
  IGA_ERR_UNSUPPORTED = -2, /* degree/dimension/quadrature outside the compiled set */
  IGA_ERR_HIP = -3,         /* a HIP runtime call failed */
  IGA_ERR_ARG = -4          /* null pointer or invalid enum */
};

/* Single tensor-product NURBS patch, all pointers device memory.
 * ext[d]   : (n_elem[d], p+1, p+1) row-major, N_a = sum_b ext[e][a][b] B_b
 * first[d] : (n_elem[d]) first supported basis index of each span
 * lo/hi[d] : (n_elem[d]) span bounds in [0,1]
 * cpts     : (n_cp, dim) unweighted, first direction fastest; wgts: (n_cp)  */
typedef struct iga_patch {
  int32_t dim;
  int32_t degree;
  int32_t n_elem[3];
  int32_t n_cp[3];
  const double* ext[3];
  const int32_t* first[3];
  const double* lo[3];
  const double* hi[3];
  const double* cpts;
  const double* wgts;
} iga_patch_t;

/* Gauss rule: the same n_q1 points per direction, tensor order first fastest. */
typedef struct iga_rule {
  int32_t n_q1;
  const double* pts; /* (n_q1) on [-1,1] */
  const double* wts; /* (n_q1) */
} iga_rule_t;

enum iga_problem { IGA_POISSON = 0, IGA_ELASTICITY = 1 };
enum iga_coef { IGA_COEF_CONST = 0, IGA_COEF_POINTS = 1 };
enum iga_src { IGA_SRC_NONE = 0, IGA_SRC_CONST = 1, IGA_SRC_POINTS = 2 };
enum iga_prec { IGA_FP64 = 0, IGA_FP32 = 1, IGA_BF16 = 2 };

/* Element stiffness + load, fused with the CSR scatter-add.
 * vec        : components per scalar basis (1 Poisson, dim elasticity)
 * coef       : Poisson a0 as constant (coef_const) or per Gauss point (n_e, n_q)
 * cmat       : elasticity constitutive matrix, host, (n_s, n_s) row-major
 * src        : none / constant per component (src_const[vec], host) /
 *              per Gauss point (n_e, n_q, vec) device
 * full_to_free: (n_dofs) free index or -1
 * gfull      : (n_dofs) prescribed values (0 on free dofs), or null
 * indptr/cols: CSR pattern of the free-free block; vals (nnz) += K_e
 * rhs        : (n_free) += source work - K_fp g ; diag (n_free) += diag(K_e)
 * Any of vals/rhs/diag may be null.                                       */
typedef struct iga_asm_args {
  int32_t problem;
  int32_t vec;
  int32_t precision;
  int32_t coef_kind;
  double coef_const;
  const double* coef_points;
  double cmat[36];
  int32_t src_kind;
  double src_const[3];
  const double* src_points;
  const int32_t* full_to_free;
  const double* gfull;
  const int32_t* indptr;
  const int32_t* cols;
  double* vals;
  double* rhs;
  double* diag;
} iga_asm_args_t;

int iga_assemble(const iga_patch_t* patch, const iga_rule_t* rule, const iga_asm_args_t* args,
                 void* stream);

/* Basis at arbitrary reference points of listed elements.
 * elems (n_e) device int32; ref (n_q, dim) if ref_shared else (n_e, n_q, dim);
 * wref (n_q) or (n_e, n_q) or null; sides (n_e) side codes 2*dir+end or null.
 * Outputs (any may be null): R (n_e,n_q,n_l), dR (n_e,n_q,dim,n_l),
 * x (n_e,n_q,dim), w (n_e,n_q), normal (n_e,n_q,dim), detj (n_e,n_q).      */
int iga_basis_eval(const iga_patch_t* patch, int32_t n_e, const int32_t* elems, int32_t n_q,
                   const double* ref, int32_t ref_shared, const double* wref, int32_t wref_shared,
                   const int32_t* sides, double* R, double* dR, double* x, double* w,
                   double* normal, double* detj, void* stream);

/* Load from values at points: rhs[f(g)] += sum_q w t_c R_i for listed elements.
 * R (n_e,n_q,n_l), w (n_e,n_q), t (n_e,n_q,vec), edofs (n_e,n_l) scalar dofs. */
int iga_load_points(int32_t n_e, int32_t n_q, int32_t n_l, int32_t vec, const double* R,
                    const double* w, const double* t, const int32_t* edofs,
                    const int32_t* full_to_free, double* rhs, void* stream);

/* CSR pattern of the free-free block on a single tensor patch.
 * nb_lo/nb_hi[d] : (n_cp[d]) inclusive 1-D neighbour range of each basis
 * free (n_free) free dof list (ascending), full_to_free (n_dofs)            */
int iga_csr_count(int32_t dim, const int32_t* n_cp, const int32_t* const* nb_lo,
                  const int32_t* const* nb_hi, int32_t vec, int32_t n_free, const int32_t* free,
                  const int32_t* full_to_free, int32_t* counts, void* stream);
int iga_csr_fill(int32_t dim, const int32_t* n_cp, const int32_t* const* nb_lo,
                 const int32_t* const* nb_hi, int32_t vec, int32_t n_free, const int32_t* free,
                 const int32_t* full_to_free, const int32_t* indptr, int32_t* cols,
                 int32_t* rows, void* stream);
/* out[0] = 0, out[i+1] = sum_{k<=i} in[k]; workspace of iga_scan_workspace(n) ints */
int64_t iga_scan_workspace(int32_t n);
int iga_scan_exclusive(int32_t n, const int32_t* in, int32_t* out, int32_t* workspace,
                       void* stream);

/* inv[r] = 1/A[r,r], or 1 where the diagonal is zero or absent */
int iga_csr_diag_inv(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                     double* inv, void* stream);
/* y = A x (FP64) */
int iga_spmv(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
             const double* x, double* y, void* stream);

/* Jacobi-PCG from x0 = 0 with the accelerator semantics of jax cg:
 * stop when ||r||_2 <= tol ||b||_2 or after maxiter iterations.
 * work: iga_cg_workspace(n) doubles. out: iterations, final ||r||^2, ||b||^2. */
int64_t iga_cg_workspace(int32_t n);
int iga_cg_solve(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                 const double* dinv, const double* b, double* x, double tol, int32_t maxiter,
                 double* work, int32_t* out_iters, double* out_rr, double* out_bb, void* stream);

/* ||A u - b||_2^2 and ||b||_2^2 in FP64 into out[0], out[1] (device). */
int iga_residual_norms(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                       const double* u, const double* b, double* out, double* work, void* stream);

/* Seeded random-Fourier-feature field g(x) = amp * sum_m cos(omega_m . x + phase_m)
 * at n points (n, dim) (device, FP64 throughout). */
int iga_rff_eval(int32_t n, int32_t dim, const double* pts, int32_t m, const double* omega,
                 const double* phase, double amp, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
