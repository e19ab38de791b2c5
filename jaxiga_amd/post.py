"""Solutions, error norms and integrals, evaluated on the GPU.

Mirrors ``post/solution.py`` (``Solution`` :22-50: ``at``, ``grad``) and
``post/norms.py`` (``integrate`` :35-39, ``norm`` :42-56, ``errornorm``
:59-111: relative by default, on the ``p+1`` Gauss rule unless ``quad`` is
given). User functions follow the point-callable convention: ``x`` is a
``(dim, n)`` float64 GPU tensor and the result a tensor of shape ``(n,)`` or
``(vec, n)``; ``exact_grad`` returns ``(vec, dim, n)`` (or ``(dim, n)`` for a
scalar field).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from jaxiga_amd.adjoint import field_points
from jaxiga_amd.evaluation import evaluate
from jaxiga_amd.pointset import gauss
from jaxiga_amd.space import expand_dofs


@dataclasses.dataclass(frozen=True)
class Solution:
    space: object
    u: object  # (n_dofs,) float64 tensor on the GPU
    params: object = None
    problem: object = None
    stats: dict = dataclasses.field(default_factory=dict)

    def _local(self, basis):
        import torch

        vec = self.space.vec
        idx = torch.as_tensor(expand_dofs(np.asarray(basis.dofs), vec), device=self.u.device)
        return self.u[idx].reshape(basis.n_elems, basis.n_local, vec)

    def _traced(self):
        import torch

        return torch.is_grad_enabled() and self.u.requires_grad

    def at(self, ps=None, basis=None):
        """Values ``(n_e, n_q, vec)`` at a point set (the field kernel; the
        differentiable contraction when ``u`` carries a gradient)."""
        import torch

        basis = basis if basis is not None else evaluate(self.space, ps)
        if self._traced():
            return torch.einsum("eqn,enc->eqc", basis.R, self._local(basis))
        return field_points(basis, self.u, vec=self.space.vec, gradients=False)[0][0]

    __call__ = at

    def grad(self, ps=None, basis=None):
        """Gradients ``(n_e, n_q, vec, dim)`` at a point set."""
        import torch

        basis = basis if basis is not None else evaluate(self.space, ps)
        if self._traced():
            return torch.einsum("eqdn,enc->eqcd", basis.dR, self._local(basis))
        return field_points(basis, self.u, vec=self.space.vec, values=False)[1][0]


def _user(fn, x, shape):
    import torch

    flat = x.reshape(-1, x.shape[-1]).T
    out = torch.as_tensor(fn(flat), dtype=torch.float64, device=x.device)
    n = flat.shape[1]
    # put the point axis first, then restore the requested trailing shape
    out = torch.movedim(out, -1, 0) if out.shape[-1] == n else out.reshape(n, -1)
    return out.reshape(shape)


def integrate(fn, space, quad=None) -> float:
    basis = evaluate(space, gauss(space, quad), outputs=("x", "w"))
    vals = _user(fn, basis.x, basis.w.shape)
    return float((basis.w * vals).sum())


def norm(sol, norm: str = "L2", quad=None) -> float:
    basis = evaluate(sol.space, gauss(sol.space, quad))
    u = sol.at(basis=basis)
    if norm == "L2":
        return float((basis.w * (u**2).sum(-1)).sum().sqrt())
    if norm == "H1":
        g = sol.grad(basis=basis)
        return float((basis.w * ((u**2).sum(-1) + (g**2).sum((-1, -2)))).sum().sqrt())
    raise ValueError(f"unknown norm {norm!r}")


def errornorm(sol, exact, norm: str = "L2", exact_grad=None, quad=None, relative=True) -> float:
    space = sol.space
    basis = evaluate(space, gauss(space, quad))
    uh = sol.at(basis=basis)
    w = basis.w
    if norm == "L2":
        ue = _user(exact, basis.x, uh.shape)
        num = (w * ((uh - ue) ** 2).sum(-1)).sum()
        den = (w * (ue**2).sum(-1)).sum()
    elif norm == "H1":
        if exact_grad is None:
            raise ValueError("the 'H1' norm needs exact_grad")
        gh = sol.grad(basis=basis)
        ge = _user(exact_grad, basis.x, gh.shape)
        num = (w * ((gh - ge) ** 2).sum((-1, -2))).sum()
        den = (w * (ge**2).sum((-1, -2))).sum()
    elif norm == "energy":
        if exact_grad is None:
            raise ValueError("the 'energy' norm needs exact_grad")
        num, den = _energy_norm(sol, basis, exact_grad)
    else:
        raise ValueError(f"unknown norm {norm!r}")
    if not relative:
        return float(num.sqrt())
    return float((num / den).sqrt())


def _energy_norm(sol, basis, exact_stress):
    """Strain energy of the stress error through the constitutive law
    (``post/norms.py:114-140``); ``exact_stress`` returns Voigt ``(n_s, n)``."""
    import torch

    problem = sol.problem
    if not hasattr(problem, "C"):
        raise ValueError(f"the energy norm needs a problem with a constitutive matrix; "
                         f"{type(problem).__name__} has none")
    C = torch.as_tensor(problem.C(sol.params), dtype=torch.float64, device=sol.u.device)
    Cinv = torch.linalg.inv(C)
    g = sol.grad(basis=basis)  # (e, q, c, d)
    if g.shape[-1] == 2:
        strain = torch.stack([g[..., 0, 0], g[..., 1, 1], g[..., 0, 1] + g[..., 1, 0]], -1)
    else:
        strain = torch.stack([g[..., 0, 0], g[..., 1, 1], g[..., 2, 2], g[..., 1, 2] + g[..., 2, 1],
                              g[..., 0, 2] + g[..., 2, 0], g[..., 0, 1] + g[..., 1, 0]], -1)
    sh = torch.einsum("ij,eqj->eqi", C, strain)
    se = _user(exact_stress, basis.x, sh.shape)
    d = sh - se
    num = (basis.w * torch.einsum("eqi,ij,eqj->eq", d, Cinv, d)).sum()
    den = (basis.w * torch.einsum("eqi,ij,eqj->eq", se, Cinv, se)).sum()
    return num, den
