"""FP64 CPU oracle for the Galerkin hot path -- TEST INFRASTRUCTURE ONLY.

This package is the parity checker. It restates, in NumPy/SciPy FP64, the
algorithms of the JAXIGA Galerkin linear path (reference files cited per
function as ``path:line`` relative to the reference repository root,
``src/jaxiga/...``). It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
and there only as the checker or the timed CPU baseline -- never as part of
the shipped product path. The product package (``jaxiga_amd``) never imports
it and fails loudly when its HIP library is missing.

Modules
-------
geometry   Patch, primitives, knot insertion and degree elevation
space      Bezier extraction, Bernstein bases, Gauss rules, IEN/DOF maps,
           basis evaluation (``evaluate``)
forms      boundary-condition resolution, element kernels, materials
assembly   CSR pattern, scatter-add assembly, load vector
solver     SciPy-faithful Jacobi-CG, textbook PCG, diagnostics, error norms
pipeline   end-to-end ``solve`` for the configurations used by the tests
grf        seeded random-Fourier-feature Gaussian random fields (inputs)
learned    networks of the learned preconditioner (forward, exact gradients)
multigrid  knot coarsening, Galerkin levels, stencil smoothers, V-cycle, FCG
adjoint    reverse pass of the solve: CSR reverse rule, element and point cotangents,
           control-point (shape) cotangent
nonlinear  neo-Hookean residual and tangent, Newton with backtracking

Pinning status of every function is listed in ``PINNING`` below; each entry
names the test module that pins it against closed forms, SciPy, or brute
force. A function absent from ``PINNING`` would be "parity unpinned".
"""

PINNING = {
    "geometry.insert_knots": "tests/test_oracle_geometry.py (surface invariance, refinement operator)",
    "geometry.elevate": "tests/test_oracle_geometry.py (surface invariance, closed-form Bezier elevation)",
    "space.bezier_extraction": "tests/test_oracle_space.py (SciPy BSpline.design_matrix per span)",
    "space.bernstein": "tests/test_oracle_space.py (SciPy BSpline on a Bezier knot vector, FD)",
    "space.gauss_rule": "tests/test_oracle_space.py (exact polynomial integration)",
    "space.FunctionSpace": "tests/test_oracle_space.py (closed-form IEN and boundary dofs)",
    "space.evaluate": "tests/test_oracle_space.py (partition of unity, exact areas, divergence theorem, affine reproduction)",
    "forms.element_tangents": "tests/test_oracle_forms.py (finite-difference Hessian of the element energy)",
    "assembly.csr_pattern": "tests/test_oracle_assembly.py (brute-force dense pattern, sorted unique keys)",
    "assembly.assemble_csr": "tests/test_oracle_assembly.py (dense np.add.at assembly)",
    "assembly.load_vector": "tests/test_oracle_assembly.py (sum of loads = integral of f, lifting)",
    "solver.scipy_csr_cg": "tests/test_oracle_solver.py (dense solve agreement, restart budget)",
    "solver.pcg_textbook": "tests/test_oracle_solver.py (dense solve agreement)",
    "solver.errornorm": "tests/test_oracle_solver.py (||sin sin||_L2 = 1/2, convergence rates)",
    "pipeline.solve": "tests/test_oracle_pipeline.py (patch test, convergence rate, Lame annulus)",
    "grf.rff_field": "tests/test_oracle_grf.py (sample covariance vs squared-exponential kernel)",
    "learned.forward/backward": "tests/test_oracle_learned.py (torch-CPU autograd, finite differences)",
    "multigrid.refinement_matrix": "tests/test_oracle_multigrid.py (coarse B-splines reproduced exactly)",
    "multigrid.hierarchy": "tests/test_oracle_multigrid.py (Galerkin = coarse stiffness assembled directly)",
    "multigrid.stencil_apply/wgrad_t": "tests/test_oracle_multigrid.py (adjointness, finite differences, dense stencil)",
    "multigrid.stencil_features": "tests/test_oracle_multigrid.py (dense-matrix restatement)",
    "multigrid.vcycle/fcg": "tests/test_oracle_multigrid.py (symmetry, exact two-level solve, dense solve)",
    "adjoint.csr_solve_vjp": "tests/test_oracle_adjoint.py (central finite differences of the direct solve in data and b)",
    "adjoint.element_cotangent/point_cotangents/boundary_cotangent": "tests/test_oracle_adjoint.py (finite differences in a0 at points, source amplitude, E, nu, traction; non-zero Dirichlet data)",
    "nonlinear.element_residual_tangent": "tests/test_oracle_nonlinear.py (central finite differences of the neo-Hookean energy and of the residual, with a body force)",
    "nonlinear.newton": "tests/test_oracle_nonlinear.py (quadratic convergence and the prescribed top displacement on the reference's nonlinear case)",
    "adjoint.shape_cotangent": "tests/test_oracle_adjoint.py (central finite differences of direct solves on moved control points, Poisson with callable a0(x), f(x) and constant non-zero Dirichlet data, rational elasticity with a body force of x; planted mistakes caught)",
}
