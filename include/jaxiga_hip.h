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
 *   iga_cg1_solve             solvers/linear.py:324-351    (the same Jacobi-PCG in
 *                             single-reduction form)
 *   iga_cg_solve_noalloc      (no reference counterpart: iga_cg_solve without a
 *                             workspace, for callers that cannot allocate)
 *   iga_residual_norms        solvers/diagnostics.py:8-19  (independent ||Au-b||)
 *   iga_mf_setup/iga_mf_apply methods/galerkin.py:207-218 (tangent_matvec,
 *                             sum-factorised instead of a JVP)
 *   iga_mf_setup_elasticity   methods/galerkin.py:207-218 (tangent_matvec of
 *                             LinearElasticity, sum-factorised)
 *   iga_mf_diag               methods/galerkin.py:221-237 (tangent_diagonal)
 *   iga_cg_solve_mf           solvers/linear.py:414-442    (matrix_free_solve, cg)
 *   iga_diff_norms            solvers/diagnostics.py:8-19  (matrix-free diagnostics)
 *   iga_inv_guard             solvers/linear.py:426-430    (Jacobi inverse, zero guard)
 *   iga_residual_vec/iga_axpy solvers/diagnostics.py:8-19 (FP64 outer residual of the
 *                             mixed-precision refinement; no reference counterpart)
 *   iga_cg_solve_mixed        solvers/linear.py:324-351    (the same Jacobi-PCG, on an
 *                             FP32 matrix with FP64 residual replacement; the
 *                             reference runs FP64 only)
 *   iga_elem_pattern / iga_elem_basis / iga_elem_assemble
 *                             space/function_space.py:23-38 (padded elements),
 *                             methods/_common.py:125-264 (chunks, CSR pattern)
 *   iga_neohooke              forms/library.py:206-225 (Hyperelasticity residual and
 *                             tangent for the Newton path, methods/galerkin.py:337-355)
 *   iga_point_forces          methods/galerkin.py:37-148 under jax.grad with
 *                             respect to the control points (shape cotangent)
 *   iga_net_* / iga_nmse / iga_adam / iga_mg_* / iga_masked_spgemm /
 *   iga_stencil_* / iga_dense_inverse / iga_csr_f32 / iga_energy_loss / iga_fcg_mg
 *                             (no reference counterpart: the learned preconditioner,
 *                             a Galerkin multigrid with learned smoothers in flexible
 *                             CG; the reference allows only "none"/"jacobi",
 *                             solvers/linear.py:113-116)
 *   iga_batched_cg            (no reference counterpart: the reference batches by
 *                             vmap over params, tests/test_advanced.py:337-344)
 *   iga_batched_cg_seg / _plan / _workspace
 *                             (the same batched solve with every instance's rows
 *                             spread over the GPU; no reference counterpart)
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
  const int32_t* elo[3]; /* (n_cp[d]) first span supporting each 1-D basis function */
  const int32_t* ehi[3]; /* (n_cp[d]) last span supporting each 1-D basis function */
  int32_t regular;       /* 1 when no interior knot repeats in any direction:
                            first[e] = e, elo[i] = max(0, i - p),
                            ehi[i] = min(n_elem - 1, i) (kernels may use the
                            closed forms instead of the tables) */
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

/* Element stiffness + load, then the CSR values of the free-free block.
 * vec        : components per scalar basis (1 Poisson, dim elasticity)
 * precision  : IGA_FP64 or IGA_FP32 (element contraction on the matrix cores
 *              in that type; vals and the element-matrix workspace in that
 *              type; basis, Jacobian and weights always FP64)
 * coef       : Poisson a0 as constant (coef_const) or per Gauss point (n_e, n_q)
 * cmat       : elasticity constitutive matrix, host, (n_s, n_s) row-major
 * src        : none / constant per component (src_const[vec], host) /
 *              per Gauss point (n_e, n_q, vec) device
 * full_to_free: (n_dofs) free index or -1;  freed: (n_free) free dof list
 * gfull      : (n_dofs) prescribed values (0 on free dofs), or null
 * rows/cols  : (nnz) COO view of the CSR pattern; vals (nnz) is overwritten
 * ke_work    : element-matrix buffer of iga_assemble_workspace() bytes
 * rhs        : (n_free) += source work - K_fp g ; diag (n_free) += diag(K),
 *              added by the row-form gather pass when it runs, else by the
 *              element pass
 * passes     : bit 1 element pass, bit 2 gather pass (0 = both); the
 *              gather reads the element matrices of an earlier element pass
 * batch      : number of coefficient/source instances sharing the patch and
 *              pattern (0 or 1 = one); instance b uses coef_points/src_points
 *              + b*n_e*n_q(*vec), rhs/diag + b*n_free, vals + b*nnz and its
 *              own slice of ke_work (the workspace scales with batch)
 * indptr     : (n_free + 1) CSR row pointers, optional: with it (and n_free)
 *              the gather pass runs one wavefront per row over coalesced
 *              element-matrix rows; without it, one thread per nonzero
 * ke_full    : 0: the element-matrix workspace layout is the library's own
 *              (FP64 on planar p <= 3 patches with indptr: the upper blocks of
 *              each symmetric element matrix only, packed by rows; otherwise
 *              full rows); both passes must then get the same arguments.
 *              1: full rows (n_e, n_l*vec, n_l*vec), rows a*vec + i -- for a
 *              gather pass over element matrices the caller wrote.
 * vals, rhs and diag may each be null.                                    */
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
  const int32_t* freed;
  const double* gfull;
  int64_t nnz;
  const int32_t* rows;
  const int32_t* cols;
  void* vals;
  void* ke_work;
  int64_t ke_work_bytes;
  double* rhs;
  double* diag;
  int32_t passes;
  int32_t batch;
  int32_t n_free;
  const int32_t* indptr;
  int32_t ke_full;
} iga_asm_args_t;

/* bytes of element-matrix workspace iga_assemble needs for vals, or -1 */
int64_t iga_assemble_workspace(const iga_patch_t* patch, int32_t problem, int32_t precision);
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
 * stop when ||r||_2 <= tol ||b||_2 or after maxiter iterations. The CSR
 * values are FP64 or FP32 (`precision`); vectors, dot products and all
 * recurrences are FP64. work: iga_cg_workspace(n) doubles.
 * out: iterations, final recursive ||r||^2, ||b||^2.
 * iga_cg1_solve: the same Jacobi-PCG in its single-reduction form
 * (Chronopoulos-Gear; no reference counterpart): the same iterates in exact
 * arithmetic, two kernels per iteration (the SpMV with the w.u partials
 * fused in, one vector kernel) instead of three.                             */
int64_t iga_cg_workspace(int32_t n);
int iga_cg_solve(int32_t n, const int32_t* indptr, const int32_t* cols, const void* vals,
                 int32_t precision, const double* dinv, const double* b, double* x, double tol,
                 int32_t maxiter, double* work, int32_t* out_iters, double* out_rr, double* out_bb,
                 void* stream);
int iga_cg1_solve(int32_t n, const int32_t* indptr, const int32_t* cols, const void* vals,
                  int32_t precision, const double* dinv, const double* b, double* x, double tol,
                  int32_t maxiter, double* work, int32_t* out_iters, double* out_rr,
                  double* out_bb, void* stream);

/* Jacobi-PCG exactly as iga_cg_solve (same kernels, same results), with no
 * workspace argument and no allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its
 * workspace is a file-scope device array of 16 MiB per GPU, held until the
 * process ends; n above IGA_CG_NOALLOC_MAX_N returns IGA_ERR_SHAPE. Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * host threads, other streams) race and must be serialised by the caller.  */
#define IGA_CG_NOALLOC_MAX_N 400000
int iga_cg_solve_noalloc(int32_t n, const int32_t* indptr, const int32_t* cols, const void* vals,
                         int32_t precision, const double* dinv, const double* b, double* x,
                         double tol, int32_t maxiter, int32_t* out_iters, double* out_rr,
                         double* out_bb, void* stream);

/* Jacobi-PCG on the FP32 matrix vals32 with FP64 vectors, dot products and
 * recurrences, and reliable updates: every replace_every iterations (>= 1),
 * and whenever the recursion reports convergence, r is recomputed as
 * b - A x with the FP64 matrix vals64 (same pattern) while the search
 * direction is kept; the solve stops only on the true FP64 residual
 * ||b - A64 x||_2 <= tol ||b||_2 (or at maxiter). work: iga_cg_workspace(n).
 * out: iterations, final true ||r||^2, ||b||^2.                               */
int iga_cg_solve_mixed(int32_t n, const int32_t* indptr, const int32_t* cols,
                       const float* vals32, const double* vals64, const double* dinv,
                       const double* b, double* x, double tol, int32_t maxiter,
                       int32_t replace_every, double* work, int32_t* out_iters, double* out_rr,
                       double* out_bb, void* stream);

/* r = b - A u (FP64) and ||r||^2, ||b||^2 into out[0], out[1]; work 2048 doubles */
int iga_residual_vec(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                     const double* u, const double* b, double* r, double* out, double* work,
                     void* stream);
/* y += a x */
int iga_axpy(int32_t n, double a, const double* x, double* y, void* stream);

/* ||A u - b||_2^2 and ||b||_2^2 in FP64 into out[0], out[1] (device). */
int iga_residual_norms(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                       const double* u, const double* b, double* out, double* work, void* stream);

/* Matrix-free sum-factorised Poisson operator (scalar, vec = 1).
 * flags: IGA_MF_RATIONAL  -- NURBS quotient (weights not all equal; with
 *                            equal weights, any value, R_i = N_i and the
 *                            polynomial apply is exact);
 *        IGA_MF_SEPARABLE -- polynomial patch whose map is a tensor product
 *                            of 1-D maps (control-point coordinate c depends
 *                            on the c-th index only, e.g. a refined box): the
 *                            geometry is recomputed from 1-D tables and only
 *                            the coefficient is stored per Gauss point.
 *        (not both)
 * precision: IGA_FP64 or IGA_FP32 arithmetic (geometry always built in FP64);
 *            IGA_BF16 (separable only) stores the coefficient in BF16 and
 *            runs the apply in FP32.
 * iga_mf_workspace : bytes of the operator workspace, or -1
 * iga_mf_setup     : fill the workspace once; coefficient constant or per
 *                    Gauss point (n_e, n_q). The workspace starts with a
 *                    header; the library also records the layout under the
 *                    workspace address, so an apply on a workspace that was
 *                    not set up, or for another patch, returns an error.
 * iga_mf_apply     : y = K x over the free dofs, x and y FP64 (n_free).
 * iga_mf_diag      : diag(K) over the free dofs (FP64), sum-factorised from
 *                    squared 1-D tables (separable layout only; replaces the
 *                    scattered element diagonals of tangent_diagonal,
 *                    methods/galerkin.py:221-237).                          */
#define IGA_MF_RATIONAL 1
#define IGA_MF_SEPARABLE 2
#define IGA_MF_ELASTICITY 4
int64_t iga_mf_workspace(const iga_patch_t* patch, int32_t precision, int32_t flags);
int iga_mf_setup(const iga_patch_t* patch, const iga_rule_t* rule, int32_t precision,
                 int32_t flags, int32_t coef_kind, double coef_const, const double* coef_points,
                 void* work, int64_t work_bytes, void* stream);
/* Linear elasticity (vec = dim) on the general layout: set-up stores
 * sqrt(w |det J|) J^-1 per Gauss point (and W, grad W when rational); cmat is
 * the Voigt matrix (3x3 in 2-D, 6x6 in 3-D, engineering shear, row-major,
 * host memory). The workspace size is iga_mf_workspace(patch, precision,
 * flags | IGA_MF_ELASTICITY); iga_mf_apply and iga_cg_solve_mf then act on
 * the vector free dofs (full dof = vec * control point + component). Replaces
 * tangent_matvec for LinearElasticity (methods/galerkin.py:207-218). */
int iga_mf_setup_elasticity(const iga_patch_t* patch, const iga_rule_t* rule, int32_t precision,
                            int32_t flags, const double* cmat, void* work, int64_t work_bytes,
                            void* stream);
int iga_mf_apply(const iga_patch_t* patch, const void* work, int64_t work_bytes, int32_t n_free,
                 const int32_t* freed, const int32_t* full_to_free, const void* x, void* y,
                 void* stream);
int iga_mf_diag(const iga_patch_t* patch, const void* work, int64_t work_bytes, int32_t n_free,
                const int32_t* freed, const int32_t* full_to_free, double* diag, void* stream);
/* Jacobi-PCG (semantics of iga_cg_solve) on the matrix-free operator. */
int iga_cg_solve_mf(const iga_patch_t* patch, const void* mf_work, int64_t mf_work_bytes,
                    int32_t n_free, const int32_t* freed, const int32_t* full_to_free,
                    const double* dinv, const double* b, double* x, double tol, int32_t maxiter,
                    double* work, int32_t* out_iters, double* out_rr, double* out_bb, void* stream);

/* ||y - b||^2 and ||b||^2 in FP64 into out[0], out[1] (device); work: 2048 doubles */
int iga_diff_norms(int32_t n, const double* y, const double* b, double* out, double* work,
                   void* stream);
/* inv[i] = 1/d[i], or 1 where d[i] == 0 (Jacobi zero guard) */
int iga_inv_guard(int32_t n, const double* d, double* inv, void* stream);

/* Implicit differentiation of the CSR solve (replaces the reverse rule of
 * _csr_sparse_solve, src/jaxiga/solvers/linear.py:376-386): after one
 * adjoint solve K lam = u_bar on the same symmetric CSR, the cotangent of
 * the CSR values is data_bar[k] = alpha * a[rows[k]] * b[cols[k]] with
 * alpha = -1, a = lam, b = u.                                              */
int iga_csr_outer(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* a,
                  const double* b, double alpha, double* out, void* stream);

/* Fields at points from basis data (replaces Solution.at / Solution.grad,
 * src/jaxiga/post/solution.py:34-50, and gives the point sensitivities of
 * the element kernels in the reverse pass of a Galerkin solve,
 * src/jaxiga/methods/galerkin.py:37-148 under jax.grad): for the one or two
 * coefficient vectors u0, u1 (n_cp * vec doubles, component fastest; u1 may
 * be null),
 *   val  (n_f, n_e, n_q, vec)      = sum_a R[e,q,a] u[dofs[e,a]]
 *   grad (n_f, n_e, n_q, vec, dim) = sum_a dR[e,q,d,a] u[dofs[e,a]]
 * R (n_e,n_q,n_l) and dR (n_e,n_q,dim,n_l) as iga_basis_eval writes them,
 * dofs (n_e,n_l) control-point indices. val or grad may be null (not both).
 * dim 2 with vec 1 or 2, dim 3 with vec 1 or 3.                            */
int iga_field_points(int32_t n_e, int32_t n_q, int32_t n_l, int32_t dim, int32_t vec,
                     const double* R, const double* dR, const int32_t* dofs, const double* u0,
                     const double* u1, double* val, double* grad, void* stream);

/* Shape cotangent of the solve (replaces the reverse pass of JAX autodiff
 * through the element kernels to the control points, exercised by the
 * reference's tests/test_galerkin.py:233-262): point quantities pulled back
 * to the control points,
 *   G[dofs[e][k]][c] += sum_q ( sum_j P[e][q][c][j] dR[e][q][j][k]
 *                               + R[e][q][k] xbar[e][q][c] )
 * with R (n_e,n_q,n_l), dR (n_e,n_q,dim,n_l) as iga_basis_eval writes them,
 * P (n_e,n_q,dim,dim), xbar (n_e,n_q,dim) or null (R may then be null),
 * dofs (n_e,n_l) control-point indices, G (n_cp,dim) accumulated (FP64
 * atomics: the summation order is not fixed).                              */
int iga_point_forces(int32_t n_e, int32_t n_q, int32_t n_l, int32_t dim, const double* R,
                     const double* dR, const double* P, const double* xbar, const int32_t* dofs,
                     double* G, void* stream);

/* The element-list path: the padded-element contract of the reference's
 * FunctionSpace (space/function_space.py:23-38: elem_dofs (n_e, n_lmax)
 * padded with the sentinel n_scalar, extraction (n_e, n_lmax, (p+1)^dim)
 * with zero rows under the padding) and its chunked evaluation
 * (space/evaluation.py:256-302, methods/_common.py:125-160).
 * iga_elem_pattern  (methods/_common.py:203-264, _build_csr_pattern): sorted
 *   unique keys row*n_free + col over every (element, local row, local column)
 *   with free dofs; first call with cols == NULL writes nnz to *out_nnz (the
 *   stream is synchronised), a second call with the same arguments and
 *   workspace fills indptr (n_free+1), cols/rows (nnz) and scatter (n_e,
 *   n_lmax*vec, n_lmax*vec): the slot of each entry or nnz for padding and
 *   prescribed dofs. Workspace: iga_elem_pattern_workspace bytes (-1 when
 *   n_e (n_lmax vec)^2 exceeds 2^31 - 1).
 * iga_elem_basis (space/evaluation.py:305-467 on a range of elements): R
 *   (n_c,n_q,n_lmax) or null, dR (n_c,n_q,dim,n_lmax), x (n_c,n_q,dim) or
 *   null, w (n_c,n_q) of elements lo .. lo+n_c-1 at the tensor Gauss points;
 *   boxes (n_e, 2 dim): the parametric lower corners then upper corners.
 * iga_elem_assemble (methods/galerkin.py:122-148 and 240-245 on a range of
 *   elements): Poisson or elasticity tangents from that basis data,
 *   scatter-added into args->vals through `scatter` (FP64 atomics), and the
 *   source work and lifting -K_fp g into args->rhs; coefficient and source
 *   per point are (n_c, n_q[, vec]) for the range. args: problem, vec,
 *   coef_*, cmat, src_*, full_to_free, gfull, nnz, vals, rhs.              */
int64_t iga_elem_pattern_workspace(int32_t n_e, int32_t n_lmax, int32_t vec);
int iga_elem_pattern(int32_t n_e, int32_t n_lmax, int32_t vec, const int32_t* elem_dofs,
                     int32_t n_scalar, const int32_t* full_to_free, int32_t n_free, void* work,
                     int64_t work_bytes, int32_t* indptr, int32_t* cols, int32_t* rows,
                     int32_t* scatter, int64_t* out_nnz, void* stream);
int iga_elem_basis(int32_t dim, int32_t degree, int32_t n_lmax, int64_t lo, int64_t n_c,
                   const iga_rule_t* rule, const double* ext, const int32_t* elem_dofs,
                   int32_t n_scalar, const double* cpts, const double* wgts, const double* boxes,
                   double* R, double* dR, double* x, double* w, void* stream);
int iga_elem_assemble(int32_t dim, int32_t n_q, int32_t n_lmax, int64_t lo, int64_t n_c,
                      const double* R, const double* dR, const double* w, const int32_t* elem_dofs,
                      int32_t n_scalar, const iga_asm_args_t* args, const int32_t* scatter,
                      void* stream);

/* Compressible neo-Hookean element residuals and tangents at a displacement
 * (replaces the residual and the Hessian of Hyperelasticity.energy,
 * forms/library.py:206-225, that the reference's Newton method re-assembles
 * every step, methods/galerkin.py:337-355, solvers/newton.py:27-145):
 * psi = mu/2 (tr F^T F - d) - mu ln J + lam/2 (ln J)^2, F = I + grad u.
 * R (n_e,n_q,n_l), dR (n_e,n_q,dim,n_l) and w (n_e,n_q) as iga_basis_eval
 * writes them at the interior Gauss points; edofs (n_e,n_l) control points;
 * u (n_cp*dim, component fastest) the full displacement; src (n_e,n_q,dim)
 * body force at the points or null. rfree (n_free) += sum_q w (P:grad R_a
 * - f.R_a) over the free dofs (atomics), and/or ke (n_e, n_l*dim, n_l*dim)
 * the element tangents, row a*dim+i, the layout iga_assemble's gather pass
 * (passes = 2) reads. dim 2 (p <= 4) or 3 (p <= 2).                        */
int iga_neohooke(int32_t n_e, int32_t n_q, int32_t n_l, int32_t dim, const double* R,
                 const double* dR, const double* w, const int32_t* edofs, const double* u,
                 const int32_t* full_to_free, double mu, double lam, const double* src,
                 double* rfree, double* ke, void* stream);

/* Learned networks on the control-point grid: 3x3 convolutions, float32 NCHW.
 * kind[l]   : 0 forward convolution (weights (cout, cin, 3, 3)),
 *             1 transposed convolution + skip add (weights (cin, cout, 3, 3))
 * src[l]    : index of the input layer (-1 = network input, c_in channels)
 * skip[l]   : layer added to a transposed convolution's output (or -1)
 * woff[l]   : offset (floats) of the layer's weights in the parameter buffer
 * boff[l]   : offset of the layer's bias (cout floats), or -1 for none
 * relu[l]   : 1 applies max(0, .) after the bias (and after the skip add)
 * H/W[0]    : input size; H/W[l+1]: output size of layer l
 * precision : 0 FP32 throughout; 1 FP16 operands (FP32 accumulation) for the
 *             stride-1 convolutions of the forward pass (inference); 2 as 1
 *             with those layers' FP16 weights already in act               */
typedef struct iga_net {
  int32_t n_layers;
  int32_t kind[16], cin[16], cout[16], stride[16], src[16], skip[16];
  int64_t woff[16];
  int32_t H[17], W[17];
  int32_t c_in;
  int32_t relu[16];
  int64_t boff[16];
  int32_t precision;
} iga_net_t;

/* floats of activation buffer for `batch` samples (input at offset 0, the
 * output is the last layer's slot), of weight-gradient scratch */
int64_t iga_net_act_floats(const iga_net_t* net, int32_t batch);
int64_t iga_net_act_offset(const iga_net_t* net, int32_t batch, int32_t layer);
int64_t iga_net_scratch_floats(const iga_net_t* net, int32_t batch);
/* forward: reads the input slot of `act`, writes every layer's slot */
int iga_net_forward(const iga_net_t* net, int32_t batch, const float* params, float* act,
                    void* stream);
/* inference: as iga_net_forward, reading the input from `input` and writing
 * the last layer to `output` (the other layers use their slots of act) */
int iga_net_infer(const iga_net_t* net, int32_t batch, const float* params, float* act,
                  const float* input, float* output, void* stream);
/* packs the FP16 weights of the FP16 layers (precision 1 or 2) into act: a
 * precision-1 call does it itself each time; after this call, inference at
 * precision 2 with the same act skips it until params change */
int iga_net_pack16(const iga_net_t* net, int32_t batch, const float* params, float* act,
                   void* stream);
/* backward: `gact` (same layout as act) holds d loss/d output in the output
 * slot on entry; on exit `grads` (parameter layout) and the input slot hold
 * d loss / d parameters (weights and biases) and d loss / d input */
int iga_net_backward(const iga_net_t* net, int32_t batch, const float* params,
                     const float* act, float* gact, float* grads, float* scratch, void* stream);
/* per-sample normalised MSE mean_n ||p_n - t_n||^2 / ||t_n||^2, its gradient,
 * sums (2*batch doubles) and loss (1 double) on the device */
int iga_nmse(int32_t batch, int64_t per_sample, const float* pred, const float* target,
             float* grad, double* sums, double* loss, void* stream);
/* one Adam step (bias-corrected, step >= 1) over n parameters */
int iga_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
             float eps, int32_t step, void* stream);

/* Geometric multigrid with learned smoothers (no reference counterpart).
 * Level l has n free dofs on an h x w grid (row-major, direction 0 fastest),
 * the FP32 operator in CSR, dsq = D^-1/2, and a smoother: per-dof stencil
 * weights (ks*ks, n) acting as S v = D^-1/2 T_W(D^-1/2 v),
 * (T_W u)_i = sum_o W[o][i] u[i + off(o)], or, when stencil is null, damped
 * Jacobi omega D^-1. P (n x n_next, CSR) maps the next coarser level's
 * coefficients to this level's; pt is its transpose. The coarsest level is
 * solved with the dense FP32 inverse coarse_inv. One V-cycle: pre-smooth S,
 * residual, restrict, recurse, prolong, residual, post-smooth S^T (FP32).
 * work: iga_mg_work_floats() floats.                                       */
typedef struct iga_mg_level {
  int32_t n, h, w, nnz;
  const int32_t* indptr;
  const int32_t* cols;
  const float* vals;
  const float* dsq;
  const float* stencil;
  float omega;
  int32_t p_nnz, pt_nnz;
  const int32_t* p_ptr;
  const int32_t* p_cols;
  const float* p_vals;
  const int32_t* pt_ptr;
  const int32_t* pt_cols;
  const float* pt_vals;
  /* optional (level 0, grid cycle): an FP16 copy of `stencil` (binary16,
   * same layout) that the grid cycle's level-0 smoothers read in its place */
  const void* stencil16;
} iga_mg_level_t;

/* Grid form of a hierarchy (optional). The free dofs of every level form an
 * h x w grid and every level operator couples grid points at most `radius`
 * apart, so the operators are kept in stencil layout, op[o][i] = K[i, i +
 * off(o)] (o over the (2 radius + 1)^2 offsets, row-major, zero off the
 * grid), and each transfer P_l = P_l^(1) (x) P_l^(0) as two 1-D bands:
 * c_lo[f], pw[f][k] = P[f][c_lo[f] + k] (k < 4) for the fine rows and
 * f_lo[c], rw[c][k] = P[f_lo[c] + k][c] (k < 6) for the coarse columns.
 * Direction 0 runs along a grid row (fastest index), direction 1 across.
 * With it, iga_mg_vcycle from level 0 and the flexible-CG V-cycle run as
 * tile kernels that fuse every transfer with the smoother beside it (radius
 * 3 and ks = 7 or no level-0 stencil; damped Jacobi on levels 1 .. tail-1;
 * a dense tail at tail_level >= 2 whose finer level has <= 8192 dofs). */
typedef struct iga_mg_band {
  const int32_t* c_lo;
  const float* pw;
  const int32_t* f_lo;
  const float* rw;
} iga_mg_band_t;

typedef struct iga_mg_grid {
  int32_t n_levels; /* levels 0 .. n_levels-1 carry op (filled by iga_mg_grid_ops) */
  int32_t radius;
  float* op[16];
  iga_mg_band_t band[16][2]; /* [l][d]: level l -> l + 1, direction d */
  double* op64; /* optional FP64 stencil layout of level 0 (the flexible-CG operator) */
  float* feats; /* optional, with op64: level 0's op64[o][i] / op64[centre][i] in FP32 */
} iga_mg_grid_t;

typedef struct iga_mg {
  int32_t n_levels; /* 1..16 */
  int32_t ks;       /* stencil width: 3, 5, 7 or 9 */
  iga_mg_level_t level[16];
  const float* coarse_inv;
  float* work;
  /* optional: the V-cycle below tail_level (1 <= tail_level < n_levels - 1,
   * damped-Jacobi levels only, n <= 2048) as one dense n x n FP32 matrix
   * (row-major, rows padded with zeros to iga_mg_dense_ld(n) floats, from
   * iga_mg_tail); null = none */
  int32_t tail_level;
  const float* tail;
  const iga_mg_grid_t* grid; /* optional grid form (tile kernels); null = CSR kernels */
} iga_mg_t;

/* Numeric set-up of a hierarchy for a new level-0 operator: Galerkin levels
 * K_{l+1} = P_l^T K_l P_l (FP64: tensor passes with setup->grid, else one
 * wavefront per coarse row, where the fine stencil of every coarse row must
 * fit a 16 x 16 box of fine grid points),
 * FP32 copies, inverse diagonals (1 where zero) and D^-1/2, and the FP32
 * dense inverse of the coarsest level (n <= 4096). pt: the CSR of P_l^T
 * (coarse x fine) in FP64; radius: the operators' stencil radius. */
typedef struct iga_mg_setup_level {
  int32_t n, h, w, nnz;
  const int32_t* indptr;
  const int32_t* cols;
  double* vals64;
  float* vals32;
  double* dinv;
  float* dsq;
  const int32_t* pt_ptr;
  const int32_t* pt_cols;
  const double* pt_vals;
} iga_mg_setup_level_t;

/* Optional tensor form of the set-up (every level's free dofs an h x w grid,
 * operator radius 3, each transfer P_l = P_l^(1) (x) P_l^(0)): the coarse
 * column bands of the 1-D transfers in FP64 (f_lo[c], rw[c][k] = P[f_lo[c] +
 * k][c], k < 6; d = 0 along a grid row) and the operators in stencil layout
 * (as iga_mg_grid_t). Each Galerkin level is then two 1-D passes over FP64
 * stencils; the second writes the coarse stencil (FP64, and FP32 where op32
 * is set), its CSR values and scalings. Level 0's stencils (and feats) are
 * formed from its CSR first. work: 49 max_l(h_l w_{l+1}) doubles. */
typedef struct iga_mg_setup_grid {
  const int32_t* f_lo[16][2];
  const double* rw[16][2];
  double* op64[16];
  float* op32[16];
  float* feats;
  double* work;
  /* optional (tail_level >= 1): the dense tail of the V-cycle from tail_level
   * (damped Jacobi with weight omega below it, exact coarsest solve) formed
   * in FP64 from the stencils, V_l = 2M - MAM + G V_{l+1} G^T, G = (I - MA)P,
   * into tail (n x iga_mg_dense_ld(n) FP32), with the coarse inverse; the
   * coarsest level has <= 48 dofs and level tail_level+1 <= 128.
   * tail_work: iga_mg_tail_dense_doubles(setup, tail_level) doubles. */
  int32_t tail_level;
  double omega;
  float* tail;
  double* tail_work;
  /* regular[l] != 0: every CSR row of level l holds exactly the in-grid
   * points of its 7 x 7 window (checked by the caller; values are then
   * written at computed slots) */
  int32_t regular[16];
  /* which part of the set-up a call runs: bit 0 the level-0 stencil layouts,
   * features, copies and scalings; bit 1 the Galerkin levels, the dense tail
   * and the coarse inverse (which read what bit 0 wrote); 0 = both. Bit 1's
   * work may also run as two calls: bit 2 its first launch, bit 3 the rest;
   * and bit 0's as bit 4 (without level 0's CSR copies) and bit 5 (the copies
   * alone, which only the flexible CG's FP64 residual reads). A caller may run
   * bit 0, start work that reads the level-0 features on another stream, and
   * run bit 1 (or bits 2 and 3) beside it. */
  int32_t stages;
  /* optional (gal_multi != 0): coarse levels formed up to three at a time.
   * Level m (1 <= m < n_levels) is formed from level gal_src[m] (m-3 <= src
   * < m) through the product of the transfers in between, whose coarse-column
   * band is gal_f_lo[m][d], gal_rw[m][d] (rows of 6, 14 or 30 FP64 values for
   * one, two or three levels); the levels formed from one source are the
   * consecutive levels below it and take one pass-0 and one pass-1 launch in
   * all. gal_work: per launch the sum over its levels of 49 h_src w_m doubles,
   * then 32 max_l w_l zeros; a level that is the source of levels two or
   * three down carries 32 zeros past the end of its op64 stencil. The caller
   * checks that the wide bands' neighbour offsets stay inside the LDS rows
   * the passes stage them in (16 and 28 zeros around the 14 and 30 values). */
  int32_t gal_multi;
  int32_t gal_src[16];
  const int32_t* gal_f_lo[16][2];
  const double* gal_rw[16][2];
  double* gal_work;
  /* optional: device memory holding the address of the level-0 values; the
   * level-0 stage then reads them through it (vals0 of iga_mg_load only has
   * to be non-null), so that a captured graph of the set-up takes new
   * values by one 8-byte write instead of a copy of the values. */
  const double* const* vals0_slot;
} iga_mg_setup_grid_t;

typedef struct iga_mg_setup {
  int32_t n_levels;
  int32_t radius;
  iga_mg_setup_level_t level[16];
  float* coarse_inv;
  double* inv_work; /* 2 n_c^2 doubles */
  const iga_mg_setup_grid_t* grid; /* optional tensor form; null = row-wise Galerkin kernel */
} iga_mg_setup_t;

int iga_mg_load(const iga_mg_setup_t* setup, const double* vals0, void* stream);
int64_t iga_mg_tail_dense_doubles(const iga_mg_setup_t* setup, int32_t level);

int64_t iga_mg_work_floats(const iga_mg_t* mg);
/* Stencil layout of the FP32 level operators 0 .. mg->grid->n_levels-1 into
 * mg->grid->op (from each level's CSR: indptr, cols, vals), and of the FP64
 * level-0 values vals64 (CSR of level 0) into mg->grid->op64 when both are set. */
int iga_mg_grid_ops(const iga_mg_t* mg, const double* vals64, void* stream);
/* x = V(b) on level `level` (FP32 vectors of that level's size) */
/* The V-cycle started at `level` (damped-Jacobi levels down to the dense
 * coarsest solve; mg->tail ignored) is linear: out (n x ld, row-major,
 * n = mg->level[level].n <= 2048, ld = iga_mg_dense_ld(n), zero beyond column
 * n) = its matrix, formed by running the same FP32 kernels on the n unit
 * vectors at once. Setting mg->tail = out and mg->tail_level = level then
 * replaces the levels below `level` by one dense product per cycle.
 * work: iga_mg_tail_floats(mg, level) floats. */
int iga_mg_dense_ld(int32_t n);
int64_t iga_mg_tail_floats(const iga_mg_t* mg, int32_t level);
int iga_mg_tail(const iga_mg_t* mg, int32_t level, float* work, float* out, void* stream);
int iga_mg_vcycle(const iga_mg_t* mg, int32_t level, const float* b, float* x, void* stream);
/* C = A B on C's pattern (COO c_rows/c_cols, nnz_c entries): A in CSR, B as the
 * CSR of B^T, sorted columns; FP64. Two calls form a Galerkin product P^T K P. */
int iga_masked_spgemm(int64_t nnz_c, const int32_t* c_rows, const int32_t* c_cols,
                      const int32_t* a_ptr, const int32_t* a_cols, const double* a_vals,
                      const int32_t* bt_ptr, const int32_t* bt_cols, const double* bt_vals,
                      double* c_vals, void* stream);
/* feats[o][i] = K[i, i+off(o)] / K[i,i], o over (2 radius + 1)^2 grid offsets */
int iga_stencil_features(int32_t n, int32_t w, int32_t radius, const int32_t* indptr,
                         const int32_t* cols, const double* vals, float* feats, void* stream);
/* FP32 inverse of a small SPD CSR matrix (n <= 4096); work: 2 n^2 doubles */
int iga_dense_inverse(int32_t n, const int32_t* indptr, const int32_t* cols, const double* vals,
                      double* work, float* inv, void* stream);
/* y = A x (mode 0), b - A x (1), y + A x (2) for `batch` vectors (strides in floats) */
int iga_csr_f32(int32_t n, int32_t nnz, const int32_t* indptr, const int32_t* cols,
                const float* vals, const float* x, const float* b, float* y, int32_t mode,
                int32_t batch, int64_t x_stride, int64_t y_stride, void* stream);
/* out (=|+=) D^-1/2 T (D^-1/2 v) for `batch` vectors, T = T_W or T_W^T */
int iga_stencil_apply(int32_t h, int32_t w, int32_t ks, const float* stencil, const float* dsq,
                      const float* v, float* out, int32_t transpose, int32_t accumulate,
                      int32_t batch, void* stream);
/* dW = d(sum_b g_b . y_b)/dW for y_b = D^-1/2 T_W^T (D^-1/2 v_b) */
int iga_stencil_wgrad_t(int32_t h, int32_t w, int32_t ks, int32_t batch, const float* dsq,
                        const float* v, const float* g, float* dW, void* stream);
/* loss = mean_b z_b . kz_b (FP64), per-sample values, grad = -2 kz / batch */
int iga_energy_loss(int32_t batch, int64_t n, const float* z, const float* kz, float* grad,
                    double* per_sample, double* loss, void* stream);
/* Flexible CG (one stored direction) on the FP64 operator with one FP32
 * V-cycle as the preconditioner; stops at ||r|| <= tol ||b|| or maxiter.
 * graph_iters > 0 captures one graph for every iteration count up to
 * graph_iters (at most 16) and replays them. work: iga_fcg_workspace(n)
 * doubles; work32: 2n floats. */
int64_t iga_fcg_workspace(int32_t n);
/* A reusable plan: the iterations are captured once over fixed buffers (b, x,
 * vals, work, and everything the iga_mg_t points to); each run re-initialises
 * the solver state on the device and replays the graphs, so one plan serves
 * every right-hand side written into b and every operator whose values are
 * rewritten in place. A run replays as many iterations as the plan's recent
 * runs (the last eight) most often needed before it first reads the
 * convergence flag, then one per read, and never more than maxiter in all.
 * The plan owns only the graphs; destroy it before the buffers are freed.
 * b is written by a run given b_in (iga_fcg_mg_run). */
typedef struct iga_fcg_plan iga_fcg_plan_t;
int iga_fcg_mg_plan(const iga_mg_t* mg, int32_t n, int32_t nnz, const int32_t* indptr,
                    const int32_t* cols, const double* vals, const double* b, double* x,
                    double* work, float* work32, int32_t graph_iters, void* stream,
                    iga_fcg_plan_t** out);
/* out_res (optional): ||b - A x|| recomputed in FP64 from the returned x
 * (independent of the recursion), read back with the solver state;
 * x_out (optional): n doubles receiving x (the plan's x is overwritten by
 * its next run): written by the last kernel of every replay on the grid
 * cycle, else copied on the stream behind the iterations;
 * b_in (optional): n doubles of the right-hand side, read by the run's first
 * kernel and kept in the plan's b (which the plan then writes: b must not
 * be read-only memory); null: the plan's b as it stands. A run may use any
 * stream, also one other than the plan's capture stream. */
int iga_fcg_mg_run(iga_fcg_plan_t* plan, double tol, int32_t maxiter, int32_t* out_iters,
                   double* out_rr, double* out_bb, double* out_res, double* x_out,
                   const double* b_in, void* stream);
void iga_fcg_mg_destroy(iga_fcg_plan_t* plan);
int iga_fcg_mg(const iga_mg_t* mg, int32_t n, const int32_t* indptr, const int32_t* cols,
               const double* vals, const double* b, double* x, double tol, int32_t maxiter,
               int32_t graph_iters, double* work, float* work32, int32_t* out_iters,
               double* out_rr, double* out_bb, void* stream);

/* Batched Jacobi-PCG: one workgroup per instance runs the whole solve in one
 * persistent launch. insts: device array of n_inst records
 *   { int32 n, pad; const int32* indptr, *cols; const double* vals, *dinv,
 *     *b; double* x; double* work (4n) }
 * order: device permutation (largest first); lanes_per_row in {8,16,32,64};
 * threads a multiple of 64 up to 1024. When the recursion meets tol, the true
 * residual b - A x is recomputed in FP64 and the instance restarts from it
 * while it is above tol (at most three restarts, within maxiter). Outputs per
 * instance: iterations and the true relative residual ||b - A x|| / ||b||.
 * No host synchronisation (graph-capturable). */
int iga_batched_cg(int32_t n_inst, const void* insts, const int32_t* order,
                   int32_t lanes_per_row, int32_t threads, double tol, int32_t maxiter,
                   int32_t* out_iters, double* out_rel, void* stream);

/* Segmented batched Jacobi-PCG, the same iteration and outputs as
 * iga_batched_cg (FP64 recurrences, true-residual restarts, at most three),
 * with the rows of every instance cut into blocks of rows_per_block rows,
 * one 256-thread workgroup each: every kernel covers the blocks of all
 * instances, so a large instance spreads over the whole GPU; three launches
 * per iteration for the batch; a converged instance's blocks return at once.
 * The host enqueues 16 iterations between reads of the active-instance count
 * (not graph-capturable; concurrent calls are serialised).
 * iga_batched_cg_seg_plan (host only) fills the block map from the host row
 * counts n_rows[n_inst]: blk_inst[b] (instance of block b), blk_row0[b]
 * (its first row) and inst_blk0[i] (first block of instance i, n_inst + 1
 * entries); it returns the number of blocks (any output may be null to count
 * only). The three arrays are passed to the solve in device memory.
 * iga_batched_cg_seg_workspace: bytes of `work` (64-byte aligned). */
int32_t iga_batched_cg_seg_plan(int32_t n_inst, const int32_t* n_rows, int32_t rows_per_block,
                                int32_t* blk_inst, int32_t* blk_row0, int32_t* inst_blk0);
int64_t iga_batched_cg_seg_workspace(int32_t n_inst, int32_t n_blocks);
int iga_batched_cg_seg(int32_t n_inst, const void* insts, int32_t n_blocks,
                       const int32_t* blk_inst, const int32_t* blk_row0,
                       const int32_t* inst_blk0, int32_t rows_per_block,
                       int32_t lanes_per_row, double tol, int32_t maxiter, void* work,
                       int64_t work_bytes, int32_t* out_iters, double* out_rel, void* stream);

/* Seeded random-Fourier-feature field g(x) = amp * sum_m cos(omega_m . x + phase_m)
 * at n points (n, dim) (device, FP64 throughout). */
int iga_rff_eval(int32_t n, int32_t dim, const double* pts, int32_t m, const double* omega,
                 const double* phase, double amp, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
