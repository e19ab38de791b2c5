"""Point sets: tensor Gauss rules on every element and on boundary faces.

Mirrors ``space/pointset.py`` of the reference: ``PointSet`` :27-75,
``gauss_1d``/``_tensor_rule``/``gauss`` :98-136 (``p+1`` points per direction
by default, first direction fastest) and ``boundary_gauss`` :148-200 (one face
rule per side code ``2*direction + end``, gathered per element).
"""

from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass(frozen=True)
class PointSet:
    elems: np.ndarray
    ref: np.ndarray  # (n_q, dim) shared or (n_e, n_q, dim)
    weights: np.ndarray | None = None  # (n_q,) or (n_e, n_q)
    sides: np.ndarray | None = None
    tag: str = "interior"

    @property
    def n_elems(self) -> int:
        return len(self.elems)

    @property
    def n_q(self) -> int:
        return self.ref.shape[-2]

    @property
    def dim(self) -> int:
        return self.ref.shape[-1]

    @property
    def shared_ref(self) -> bool:
        return self.ref.ndim == 2


def gauss_1d(n: int):
    return np.polynomial.legendre.leggauss(int(n))


def tensor_rule(n_per_dir):
    """Tensor Gauss-Legendre points/weights on [-1,1]^dim, first direction fastest."""
    rules = [gauss_1d(n) for n in n_per_dir]
    grids = np.meshgrid(*[r[0] for r in rules], indexing="ij")
    wgrids = np.meshgrid(*[r[1] for r in rules], indexing="ij")
    pts = np.stack([g.ravel(order="F") for g in grids], axis=1)
    wts = np.prod([w.ravel(order="F") for w in wgrids], axis=0)
    return pts, wts


def n_points(space, n=None):
    return [(p + 1) if n is None else int(n) for p in space.degree]


def gauss(space, n: int | None = None) -> PointSet:
    pts, wts = tensor_rule(n_points(space, n))
    return PointSet(elems=np.arange(space.n_elems), ref=pts, weights=wts, tag="interior")


def boundary_gauss(space, target, n: int | None = None) -> PointSet:
    bset = space.boundary(target) if isinstance(target, str) else target
    label = target if isinstance(target, str) else "set"
    dim = space.dim
    n_per = n_points(space, n)
    rules = {}
    for code in np.unique(bset.sides):
        d, end = divmod(int(code), 2)
        tang = [k for k in range(dim) if k != d]
        if tang:
            fp, fw = tensor_rule([n_per[k] for k in tang])
        else:
            fp, fw = np.zeros((1, 0)), np.ones(1)
        pts = np.zeros((len(fp), dim))
        for j, k in enumerate(tang):
            pts[:, k] = fp[:, j]
        pts[:, d] = 1.0 if end else -1.0
        rules[int(code)] = (pts, fw)
    if not rules:
        return PointSet(np.zeros(0, np.int64), np.zeros((0, 1, dim)), np.zeros((0, 1)),
                        np.zeros(0, np.int64), f"boundary:{label}")
    ref = np.stack([rules[int(c)][0] for c in bset.sides])
    wts = np.stack([rules[int(c)][1] for c in bset.sides])
    return PointSet(elems=np.asarray(bset.elems), ref=ref, weights=wts,
                    sides=np.asarray(bset.sides), tag=f"boundary:{label}")
