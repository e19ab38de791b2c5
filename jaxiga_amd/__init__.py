"""jaxiga_amd -- the JAXIGA Galerkin hot path, MI355X-native.

A drop-in for the part of the JAXIGA API that the Galerkin linear path uses::

    import jaxiga_amd as jx

    patch = jx.primitives.quadrilateral([[0, 0], [1, 0], [1, 1], [0, 1]]).elevate(3).refine(5)
    V = jx.FunctionSpace(patch)
    problem = jx.Poisson(V, dirichlet=[jx.DirichletBC(0.0, where=lambda x: x[:, 0] > -1)],
                         source=lambda x, p: 8 * torch.pi**2 * torch.sin(2 * torch.pi * x[0])
                                             * torch.sin(2 * torch.pi * x[1]))
    sol = jx.solve(problem, params={"a0": 1.0}, linear=jx.LinearOptions(method="cg"))

Setup (geometry refinement, extraction operators, DOF maps, boundary
resolution) runs on the host as in the reference; basis evaluation, element
assembly, the CSR scatter, the operator applies and the Krylov solves run in
hand-written HIP kernels for gfx950 behind the C ABI of ``include/jaxiga_hip.h``.
"""

__version__ = "0.1.0"

from jaxiga_amd.evaluation import BasisData, evaluate
from jaxiga_amd.forms import (
    DirichletBC,
    Hyperelasticity,
    LinearElasticity,
    Neumann,
    PointField,
    Poisson,
    Problem,
    materials,
)
from jaxiga_amd.geometry import Patch, primitives
from jaxiga_amd.linear import LinearOptions
from jaxiga_amd.methods import solve
from jaxiga_amd.pointset import PointSet, boundary_gauss, gauss
from jaxiga_amd.post import Solution, errornorm, integrate, norm
from jaxiga_amd.space import DofMap, FunctionSpace, expand_dofs

__all__ = [
    "Patch", "primitives", "FunctionSpace", "DofMap", "expand_dofs", "PointSet", "gauss",
    "boundary_gauss", "BasisData", "evaluate", "Problem", "Poisson", "LinearElasticity",
    "Hyperelasticity",
    "DirichletBC", "Neumann", "PointField", "materials", "LinearOptions", "solve", "Solution",
    "errornorm", "integrate", "norm",
]
