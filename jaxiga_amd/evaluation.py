"""Basis evaluation on the GPU (``iga_basis_eval``).

Mirrors ``space/evaluation.py`` of the reference: ``BasisData`` :37-92 and
``evaluate`` :305-467 (Bezier extraction, chain-rule scaling to the patch
parameters, rational quotient, push-forward through the FP64 Jacobian,
physical weights ``w_ref |det J| prod(h/2)``, and on boundary sets the outward
unit normal ``J^-1 e_d / |J^-1 e_d|`` with weights ``w_ref |tangent| h_t/2``).

Every array of the returned :class:`BasisData` is a float64 tensor on the GPU
with the reference's shapes (element axis first).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from jaxiga_amd import _lib
from jaxiga_amd.runtime import device, lib, stream


@dataclasses.dataclass
class BasisData:
    R: object  # (n_e, n_q, n_l)
    dR: object  # (n_e, n_q, dim, n_l)
    x: object  # (n_e, n_q, dim)
    dofs: np.ndarray  # (n_e, n_l)
    w: object | None = None  # (n_e, n_q)
    normal: object | None = None  # (n_e, n_q, dim)
    detJ: object | None = None  # (n_e, n_q)
    tag: str = "interior"

    @property
    def n_elems(self) -> int:
        return self.R.shape[0]

    @property
    def n_q(self) -> int:
        return self.R.shape[1]

    @property
    def n_local(self) -> int:
        return self.R.shape[-1]

    def flat(self):
        n = self.n_elems * self.n_q
        return self.R.reshape(n, -1), self.dR.reshape(n, self.dR.shape[2], -1), self.x.reshape(n, -1)


def _point_inputs(ps, dev):
    """Device copies of a point set's element list, reference coordinates,
    weights and sides, uploaded once per point set (a point set is immutable)
    so repeated evaluations issue no host-to-device copies."""
    import torch

    cached = getattr(ps, "_device_inputs", None)
    if cached is not None and cached[0] == dev:
        return cached[1]
    f64 = dict(dtype=torch.float64, device=dev)
    elems = torch.as_tensor(np.asarray(ps.elems), dtype=torch.int32, device=dev)
    ref = torch.as_tensor(np.ascontiguousarray(ps.ref), **f64)
    wref = None if ps.weights is None else torch.as_tensor(np.ascontiguousarray(ps.weights), **f64)
    sides = None if ps.sides is None else torch.as_tensor(np.asarray(ps.sides), dtype=torch.int32, device=dev)
    out = (elems, ref, wref, sides)
    object.__setattr__(ps, "_device_inputs", (dev, out))
    return out


def evaluate(space, ps, order: int = 1, outputs=("R", "dR", "x", "w", "normal", "detJ")) -> BasisData:
    """Evaluate ``space``'s basis at the points of ``ps`` on the GPU."""
    import torch

    if order not in (0, 1):
        raise NotImplementedError("second derivatives are outside the supported hot path")
    dev = device()
    ds = space.device()
    dim = space.dim
    n_e = ps.n_elems
    n_q = ps.n_q
    n_l = space.n_local
    f64 = dict(dtype=torch.float64, device=dev)
    elems, ref, wref, sides = _point_inputs(ps, dev)
    want = set(outputs)
    R = torch.empty((n_e, n_q, n_l), **f64) if "R" in want else None
    dR = torch.empty((n_e, n_q, dim, n_l), **f64) if "dR" in want else None
    x = torch.empty((n_e, n_q, dim), **f64) if "x" in want else None
    w = torch.empty((n_e, n_q), **f64) if ("w" in want and wref is not None) else None
    normal = torch.empty((n_e, n_q, dim), **f64) if ("normal" in want and sides is not None) else None
    detJ = torch.empty((n_e, n_q), **f64) if "detJ" in want else None
    rc = lib().iga_basis_eval(
        ds.struct, n_e, _lib.ptr(elems), n_q, _lib.ptr(ref), int(ps.shared_ref), _lib.ptr(wref),
        int(wref is not None and wref.dim() == 1), _lib.ptr(sides), _lib.ptr(R), _lib.ptr(dR),
        _lib.ptr(x), _lib.ptr(w), _lib.ptr(normal), _lib.ptr(detJ), stream())
    _lib.check(rc, "iga_basis_eval")
    return BasisData(R=R, dR=dR, x=x, dofs=_element_dofs(space, ps), w=w, normal=normal, detJ=detJ,
                     tag=ps.tag)


def _element_dofs(space, ps):
    """``space.elem_dofs`` of the point set's elements, formed once per point
    set and space (the full interior set is the table itself, not a copy)."""
    cached = getattr(ps, "_dofs_cache", None)
    if cached is not None and cached[0] is space.elem_dofs:
        return cached[1]
    elems = np.asarray(ps.elems)
    n = space.elem_dofs.shape[0]
    if len(elems) == n and elems[0] == 0 and np.array_equal(elems, np.arange(n)):
        dofs = space.elem_dofs
    else:
        dofs = space.elem_dofs[elems]
    object.__setattr__(ps, "_dofs_cache", (space.elem_dofs, dofs))
    return dofs
