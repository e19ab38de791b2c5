"""NURBS patches and the named primitives (setup stage, host).

Mirrors ``geometry/nurbs.py`` (``Patch.create`` :298-322, ``elevate``
:384-420, ``insert_knots`` :422-444, ``refine`` :446-455) and
``geometry/primitives.py`` (``quadrilateral`` :41-69, ``rectangle`` :72-74,
``quarter_annulus`` :107-133, ``cuboid`` :285-309) of the reference.

Conventions are the reference's: knots normalised to [0, 1], control points
stored unweighted and flattened with the first parametric direction fastest,
sides named ``u0``/``u1``/``v0``/``v1``/``w0``/``w1`` with user labels.

Both refinements (knot insertion and degree elevation) are exact changes of
basis into a richer spline space that contains the old one. They are computed
the same way: the new homogeneous coefficients are the unique interpolant of
the old spline at the Greville abscissae of the new space (Schoenberg-Whitney
holds there), with B-splines evaluated by a vectorised Cox-de Boor recursion.
"""

from __future__ import annotations

import dataclasses

import numpy as np

DIR_NAMES = ("u", "v", "w")
LEGACY_SIDES = {"left": "u0", "right": "u1", "down": "v0", "up": "v1"}


def side_name(direction: int, end: int) -> str:
    return f"{DIR_NAMES[direction]}{end}"


def side_code(name: str) -> int:
    return 2 * DIR_NAMES.index(name[0]) + int(name[1:])


def canonical_side(name: str) -> str:
    return LEGACY_SIDES.get(name, name)


def bspline_values(knots, p, x):
    """All degree-``p`` B-splines at ``x``: ``(len(x), n)``; right end closed."""
    t = np.asarray(knots, float)
    x = np.atleast_1d(np.asarray(x, float))
    nspan = len(t) - 1
    spans = np.nonzero(t[1:] > t[:-1])[0]
    # span of each point: last non-empty span whose left knot <= x
    k = spans[np.clip(np.searchsorted(t[spans], x, side="right") - 1, 0, len(spans) - 1)]
    N = np.zeros((len(x), nspan))
    N[np.arange(len(x)), k] = 1.0
    for q in range(1, p + 1):
        left = t[: nspan - q]
        right = t[q + 1 : nspan + 1]
        d1 = t[q : nspan] - left
        d2 = right - t[1 : nspan - q + 1]
        a = np.where(d1 > 0, (x[:, None] - left) / np.where(d1 > 0, d1, 1.0), 0.0)
        b = np.where(d2 > 0, (right - x[:, None]) / np.where(d2 > 0, d2, 1.0), 0.0)
        N = a * N[:, : nspan - q] + b * N[:, 1 : nspan - q + 1]
    return N[:, : len(t) - p - 1]


def greville(knots, p):
    t = np.asarray(knots, float)
    n = len(t) - p - 1
    if p == 0:
        return 0.5 * (t[:n] + t[1 : n + 1])
    c = np.cumsum(np.concatenate([[0.0], t]))
    return (c[p + 1 : n + p + 1] - c[1 : n + 1]) / p


def change_of_basis(old_knots, p_old, new_knots, p_new):
    """Matrix ``T`` with ``new_coefs = T @ old_coefs`` (exact re-representation)."""
    g = greville(new_knots, p_new)
    A_new = bspline_values(new_knots, p_new, g)
    A_old = bspline_values(old_knots, p_old, g)
    return np.linalg.solve(A_new, A_old)


def _normalize(kv):
    kv = np.asarray(kv, float)
    lo, hi = kv[0], kv[-1]
    if hi - lo <= 0:
        raise ValueError(f"degenerate knot vector with span [{lo}, {hi}]")
    return (kv - lo) / (hi - lo)


@dataclasses.dataclass(frozen=True, eq=False)
class Patch:
    """A single NURBS patch (host arrays)."""

    ctrl_pts: np.ndarray
    weights: np.ndarray
    knots: tuple
    degree: tuple
    labels: tuple = ()

    @classmethod
    def create(cls, knots, degree, ctrl_pts, weights=None, labels=None) -> "Patch":
        degree = tuple(int(p) for p in degree)
        knots = tuple(tuple(_normalize(kv).tolist()) for kv in knots)
        if len(knots) != len(degree):
            raise ValueError(f"{len(knots)} knot vectors but {len(degree)} degrees")
        ctrl_pts = np.asarray(ctrl_pts, float)
        if ctrl_pts.ndim != 2:
            raise ValueError(f"ctrl_pts must be (n_cp, dim_phys), got shape {ctrl_pts.shape}")
        expected = [len(kv) - p - 1 for kv, p in zip(knots, degree)]
        if int(np.prod(expected)) != ctrl_pts.shape[0]:
            raise ValueError(
                f"knots/degree imply {expected} = {int(np.prod(expected))} control points "
                f"but {ctrl_pts.shape[0]} were given")
        w = np.ones(ctrl_pts.shape[0]) if weights is None else np.asarray(weights, float)
        if w.shape != (ctrl_pts.shape[0],):
            raise ValueError(f"weights must be ({ctrl_pts.shape[0]},), got {w.shape}")
        if isinstance(labels, dict):
            labels = labels.items()
        labels = tuple(sorted((canonical_side(k), v) for k, v in (labels or ())))
        return cls(ctrl_pts, w, knots, degree, labels)

    @property
    def dim(self) -> int:
        return len(self.degree)

    @property
    def dim_phys(self) -> int:
        return self.ctrl_pts.shape[1]

    @property
    def n_cp_per_dir(self) -> tuple:
        return tuple(len(kv) - p - 1 for kv, p in zip(self.knots, self.degree))

    @property
    def n_cp(self) -> int:
        return int(np.prod(self.n_cp_per_dir))

    @property
    def label_map(self) -> dict:
        return dict(self.labels)

    def knot_arrays(self):
        return [np.asarray(kv, float) for kv in self.knots]

    def homogeneous(self):
        return np.concatenate([self.ctrl_pts * self.weights[:, None], self.weights[:, None]], 1)

    def side_names(self):
        return tuple(side_name(d, e) for d in range(self.dim) for e in (0, 1))

    def with_labels(self, **side_labels) -> "Patch":
        merged = self.label_map
        for side, label in side_labels.items():
            side = canonical_side(side)
            if side not in self.side_names():
                raise ValueError(f"unknown side {side!r}; expected one of {self.side_names()}")
            merged[side] = label
        return Patch.create(self.knots, self.degree, self.ctrl_pts, self.weights, merged)

    def _apply(self, ops):
        """Apply per-direction matrices ``ops[d] (n_new_d, n_old_d)`` (or None)."""
        n = list(self.n_cp_per_dir)
        coefs = self.homogeneous()
        m = coefs.shape[1]
        grid = coefs.reshape(*reversed(n), m)  # axes (n_{D-1}, ..., n_0, m)
        for d, T in enumerate(ops):
            if T is None:
                continue
            ax = self.dim - 1 - d
            grid = np.moveaxis(np.tensordot(T, grid, axes=([1], [ax])), 0, ax)
        return grid.reshape(-1, m)

    def _rebuild(self, coefs, knots, degree) -> "Patch":
        w = coefs[:, -1]
        return Patch.create(knots, degree, coefs[:, :-1] / w[:, None], w, self.labels)

    def elevate(self, target_degree) -> "Patch":
        """Raise the degree **to** ``target_degree`` (a target, not an increment)."""
        if np.isscalar(target_degree):
            target = (int(target_degree),) * self.dim
        else:
            target = tuple(int(p) for p in target_degree)
        if len(target) != self.dim:
            raise ValueError(f"expected {self.dim} target degrees, got {len(target)}")
        if any(t < p for t, p in zip(target, self.degree)):
            raise ValueError(f"cannot lower degree: patch has degree {self.degree}, target {target}")
        if target == self.degree:
            return self
        ops, knots = [], []
        for d, kv in enumerate(self.knot_arrays()):
            t = target[d] - self.degree[d]
            if t == 0:
                ops.append(None)
                knots.append(kv)
                continue
            vals, counts = np.unique(kv, return_counts=True)
            new = np.repeat(vals, counts + t)
            ops.append(change_of_basis(kv, self.degree[d], new, target[d]))
            knots.append(new)
        return self._rebuild(self._apply(ops), knots, target)

    def insert_knots(self, *knots_per_dir) -> "Patch":
        if len(knots_per_dir) != self.dim:
            raise ValueError(f"expected {self.dim} knot sequences, got {len(knots_per_dir)}")
        ops, knots = [], []
        for d, kv in enumerate(self.knot_arrays()):
            new = np.asarray(knots_per_dir[d], float)
            if new.size == 0:
                ops.append(None)
                knots.append(kv)
                continue
            if new.min() <= 0.0 or new.max() >= 1.0:
                raise ValueError(f"interior knots must lie in (0, 1); got {new}")
            merged = np.sort(np.concatenate([kv, new]))
            ops.append(change_of_basis(kv, self.degree[d], merged, self.degree[d]))
            knots.append(merged)
        return self._rebuild(self._apply(ops), knots, self.degree)

    def refine(self, n: int = 1) -> "Patch":
        """``n`` rounds of uniform bisection of every non-empty knot span."""
        if int(n) <= 0:
            return self
        # repeated bisection gives the same knot values as refining n times;
        # all new knots are then inserted in one change of basis per direction
        new = []
        for kv in self.knot_arrays():
            u = np.unique(kv)
            for _ in range(int(n)):
                u = np.unique(np.concatenate([u, 0.5 * (u[:-1] + u[1:])]))
            new.append(np.setdiff1d(u, np.unique(kv)))
        return self.insert_knots(*new)

    def evaluate(self, params) -> np.ndarray:
        """Geometry map at parametric points ``(n, dim)`` in [0,1]^dim."""
        params = np.atleast_2d(np.asarray(params, float))
        cw = self.homogeneous()
        n = self.n_cp_per_dir
        mats = [bspline_values(self.knots[d], self.degree[d], params[:, d]) for d in range(self.dim)]
        grid = cw.reshape(*reversed(n), cw.shape[1])
        out = np.empty((len(params), cw.shape[1]))
        for k in range(len(params)):
            v = grid
            for d in range(self.dim - 1, -1, -1):
                v = np.tensordot(mats[d][k], v, axes=([0], [0]))
            out[k] = v
        return out[:, :-1] / out[:, -1:]

    def summary(self) -> str:
        return (f"Patch(dim={self.dim}, degree={self.degree}, n_cp={self.n_cp} "
                f"{self.n_cp_per_dir}, labels={self.label_map})")


class primitives:  # namespace mirroring ``jaxiga.primitives``
    """Named patch constructors (``geometry/primitives.py``)."""

    @staticmethod
    def quadrilateral(corners) -> Patch:
        corners = np.asarray(corners, float)
        if corners.shape != (4, 2):
            raise ValueError(f"expected 4 corners of dimension 2, got shape {corners.shape}")
        a, b, c, d = corners
        return Patch.create([[0.0, 0.0, 1.0, 1.0]] * 2, (1, 1), np.array([a, b, d, c]), np.ones(4),
                            {"u0": "left", "u1": "right", "v0": "bottom", "v1": "top"})

    @staticmethod
    def rectangle(x0, y0, x1, y1) -> Patch:
        return primitives.quadrilateral([[x0, y0], [x1, y0], [x1, y1], [x0, y1]])

    @staticmethod
    def quarter_annulus(radius_int, radius_ext) -> Patch:
        ri, ro = float(radius_int), float(radius_ext)
        s = 1.0 / np.sqrt(2.0)
        cpts = np.array([[ri, 0.0], [ro, 0.0], [ri, ri], [ro, ro], [0.0, ri], [0.0, ro]])
        w = np.array([1.0, 1.0, s, s, 1.0, 1.0])
        return Patch.create([[0, 0, 1, 1], [0, 0, 0, 1, 1, 1]], (1, 2), cpts, w,
                            {"u0": "inner", "u1": "outer", "v0": "sym_y", "v1": "sym_x"})

    @staticmethod
    def cuboid(lower, upper) -> Patch:
        lo, hi = np.asarray(lower, float), np.asarray(upper, float)
        if lo.shape != (3,) or hi.shape != (3,):
            raise ValueError("lower and upper must each have 3 components")
        pts = np.array([[(lo, hi)[i][0], (lo, hi)[j][1], (lo, hi)[k][2]]
                        for k in (0, 1) for j in (0, 1) for i in (0, 1)])
        return Patch.create([[0, 0, 1, 1]] * 3, (1, 1, 1), pts, np.ones(8),
                            {"u0": "left", "u1": "right", "v0": "bottom", "v1": "top",
                             "w0": "back", "w1": "front"})
