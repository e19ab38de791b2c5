"""LinearOptions validation (CPU): the reference's checks
(tests/test_solver_robustness.py:13-22, tests/test_advanced.py:234-238,286-288)
and the multigrid preconditioner this build adds."""

import numpy as np
import pytest

from jaxiga_amd.linear import LinearOptions


@pytest.mark.parametrize("tol", [0, -1e-6, np.inf, np.nan])
def test_linear_options_reject_invalid_tolerance(tol):
    with pytest.raises(ValueError, match="tol"):
        LinearOptions(tol=tol)


@pytest.mark.parametrize("maxiter", [0, -1, 1.5, True])
def test_linear_options_reject_invalid_iteration_limit(maxiter):
    with pytest.raises(ValueError, match="maxiter"):
        LinearOptions(maxiter=maxiter)


def test_superlu_ordering_and_matrix_free_method_are_validated():
    with pytest.raises(ValueError, match="SuperLU ordering"):
        LinearOptions(superlu_ordering="not-an-ordering")
    with pytest.raises(ValueError, match="iterative"):
        LinearOptions(method="scipy", matrix_free=True)
    with pytest.raises(ValueError, match="unknown linear method"):
        LinearOptions(method="lu")


def test_multigrid_preconditioner_needs_assembled_cg():
    LinearOptions(method="cg", preconditioner="multigrid", tol=1e-10)
    for kwargs in ({"method": "dense"}, {"method": "gmres"}, {"method": "cg", "matrix_free": True}):
        with pytest.raises(ValueError, match="multigrid"):
            LinearOptions(preconditioner="multigrid", **kwargs)
    with pytest.raises(ValueError, match="unknown preconditioner"):
        LinearOptions(method="cg", preconditioner="ilu")
