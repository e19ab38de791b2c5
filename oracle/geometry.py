"""Oracle geometry: NURBS patches, knot insertion, degree elevation.

Restates ``src/jaxiga/geometry/nurbs.py`` (Patch.create :298-322, elevate
:384-420, insert_knots :422-444, refine :446-455, _apply_along_dir :535-556,
_from_homogeneous :559-564) and the three primitives of
``src/jaxiga/geometry/primitives.py`` used by the hot path (quadrilateral
:41-69, quarter_annulus :107-133, cuboid :285-309).

Conventions kept from the reference: knots normalised to [0, 1]; control points
stored unweighted, flattened with the first parametric direction fastest;
refinement acts on homogeneous coordinates.

The univariate operators are written from the textbook, not from the
reference's Octave-derived routines: knot insertion is Boehm's single-knot
rule applied once per knot, degree elevation is collocation of the elevated
space at its Greville abscissae. A B-spline's coefficients in a given space
are unique, so both give the reference's control points up to rounding; the
tests pin them by geometry invariance and by the closed-form Bezier elevation.
"""

from __future__ import annotations

import dataclasses

import numpy as np

TOL = 1e-10
DIR_NAMES = ("u", "v", "w")


def side_name(direction: int, end: int) -> str:
    return f"{DIR_NAMES[direction]}{end}"


def side_code(name: str) -> int:
    return 2 * DIR_NAMES.index(name[0]) + int(name[1:])


# ---------------------------------------------------------------------------
# univariate B-spline tools
# ---------------------------------------------------------------------------


def basis_matrix(knots, p, x):
    """Cox-de Boor B-spline values, ``(len(x), n)``; the last span is closed."""
    t = np.asarray(knots, float)
    x = np.atleast_1d(np.asarray(x, float))
    n = len(t) - p - 1
    # degree-0 indicator on half-open spans, the final non-empty span closed
    N = np.zeros((len(x), len(t) - 1))
    for i in range(len(t) - 1):
        if t[i + 1] > t[i]:
            N[:, i] = (x >= t[i]) & (x < t[i + 1])
    last = max(i for i in range(len(t) - 1) if t[i + 1] > t[i])
    N[x == t[-1], :] = 0.0
    N[x == t[-1], last] = 1.0
    for k in range(1, p + 1):
        M = np.zeros((len(x), len(t) - 1 - k))
        for i in range(len(t) - 1 - k):
            d1 = t[i + k] - t[i]
            d2 = t[i + k + 1] - t[i + 1]
            if d1 > 0:
                M[:, i] += (x - t[i]) / d1 * N[:, i]
            if d2 > 0:
                M[:, i] += (t[i + k + 1] - x) / d2 * N[:, i + 1]
        N = M
    return N[:, :n]


def boehm_insert(p, ctrl, knots, u):
    """Insert one knot ``u`` (Boehm). ``ctrl`` is ``(n, m)``."""
    t = np.asarray(knots, float)
    n = ctrl.shape[0]
    k = int(np.searchsorted(t, u, side="right") - 1)
    k = min(max(k, p), n - 1)
    out = np.zeros((n + 1, ctrl.shape[1]))
    out[: k - p + 1] = ctrl[: k - p + 1]
    for i in range(k - p + 1, k + 1):
        alpha = (u - t[i]) / (t[i + p] - t[i])
        out[i] = (1.0 - alpha) * ctrl[i - 1] + alpha * ctrl[i]
    out[k + 1 :] = ctrl[k:]
    return out, np.insert(t, k + 1, u)


def insert_knots_1d(p, ctrl, knots, new):
    for u in np.sort(np.asarray(new, float)):
        ctrl, knots = boehm_insert(p, ctrl, knots, u)
    return ctrl, knots


def elevate_1d(p, ctrl, knots, times):
    """Raise the degree by ``times``: every distinct knot gains ``times``
    multiplicity; coefficients by collocation at the new Greville points."""
    if times == 0:
        return ctrl, np.asarray(knots, float)
    t = np.asarray(knots, float)
    values, counts = np.unique(t, return_counts=True)
    new_t = np.repeat(values, counts + times)
    q = p + times
    n_new = len(new_t) - q - 1
    greville = np.array([new_t[i + 1 : i + q + 1].mean() for i in range(n_new)])
    old = basis_matrix(t, p, greville) @ ctrl
    A = basis_matrix(new_t, q, greville)
    return np.linalg.solve(A, old), new_t


# ---------------------------------------------------------------------------
# Patch
# ---------------------------------------------------------------------------


def _normalize(kv):
    kv = np.asarray(kv, float)
    return (kv - kv[0]) / (kv[-1] - kv[0])


@dataclasses.dataclass
class Patch:
    """Single NURBS patch (``nurbs.py:279-322``)."""

    ctrl_pts: np.ndarray  # (n_cp, dim_phys) unweighted, first direction fastest
    weights: np.ndarray  # (n_cp,)
    knots: list
    degree: tuple
    labels: dict

    @classmethod
    def create(cls, knots, degree, ctrl_pts, weights=None, labels=None):
        knots = [_normalize(k) for k in knots]
        degree = tuple(int(p) for p in degree)
        ctrl_pts = np.asarray(ctrl_pts, float)
        n = [len(k) - p - 1 for k, p in zip(knots, degree)]
        if int(np.prod(n)) != ctrl_pts.shape[0]:
            raise ValueError("knots/degree do not match the control-point count")
        w = np.ones(ctrl_pts.shape[0]) if weights is None else np.asarray(weights, float)
        return cls(ctrl_pts, w, knots, degree, dict(labels or {}))

    @property
    def dim(self):
        return len(self.degree)

    @property
    def n_cp_per_dir(self):
        return tuple(len(k) - p - 1 for k, p in zip(self.knots, self.degree))

    def homogeneous(self):
        return np.concatenate([self.ctrl_pts * self.weights[:, None], self.weights[:, None]], 1)

    def _along(self, coefs, n_per_dir, d, fn):
        """Apply a univariate operator along direction ``d`` (``nurbs.py:535-556``)."""
        dim = len(n_per_dir)
        m = coefs.shape[1]
        grid = coefs.reshape(*reversed(n_per_dir), m)
        grid = np.moveaxis(grid, dim - 1 - d, 0)  # (n_d, ..., m)
        rest = grid.shape[1:]
        out, kv = fn(grid.reshape(grid.shape[0], -1))
        out = out.reshape(out.shape[0], *rest)
        out = np.moveaxis(out, 0, dim - 1 - d)
        return out.reshape(-1, m), kv

    def _rebuild(self, coefs, knots, degree):
        w = coefs[:, -1]
        return Patch.create(knots, degree, coefs[:, :-1] / w[:, None], w, self.labels)

    def elevate(self, target):
        target = (int(target),) * self.dim if np.isscalar(target) else tuple(target)
        coefs = self.homogeneous()
        knots = [k.copy() for k in self.knots]
        n = list(self.n_cp_per_dir)
        deg = list(self.degree)
        for d in range(self.dim):
            t = target[d] - deg[d]
            if t < 0:
                raise ValueError("cannot lower the degree")
            if t == 0:
                continue
            coefs, knots[d] = self._along(coefs, n, d, lambda c: elevate_1d(deg[d], c, knots[d], t))
            deg[d] = target[d]
            n[d] = len(knots[d]) - deg[d] - 1
        return self._rebuild(coefs, knots, deg)

    def insert_knots(self, *per_dir):
        coefs = self.homogeneous()
        knots = [k.copy() for k in self.knots]
        n = list(self.n_cp_per_dir)
        for d in range(self.dim):
            new = np.asarray(per_dir[d], float)
            if new.size == 0:
                continue
            coefs, knots[d] = self._along(
                coefs, n, d, lambda c: insert_knots_1d(self.degree[d], c, knots[d], new)
            )
            n[d] = len(knots[d]) - self.degree[d] - 1
        return self._rebuild(coefs, knots, self.degree)

    def refine(self, n=1):
        patch = self
        for _ in range(int(n)):
            mids = []
            for kv in patch.knots:
                u = np.unique(kv)
                mids.append(0.5 * (u[:-1] + u[1:]))
            patch = patch.insert_knots(*mids)
        return patch

    def evaluate(self, params):
        """Geometry map at parametric points ``(n, dim)`` (tensor Cox-de Boor)."""
        params = np.atleast_2d(np.asarray(params, float))
        cw = self.homogeneous()
        n = self.n_cp_per_dir
        out = np.zeros((len(params), cw.shape[1]))
        mats = [basis_matrix(self.knots[d], self.degree[d], params[:, d]) for d in range(self.dim)]
        grid = cw.reshape(*reversed(n), cw.shape[1])
        for k in range(len(params)):
            # contract directions one at a time: axes are (n_{D-1}, ..., n_0, m)
            val = grid
            for d in range(self.dim - 1, -1, -1):
                val = np.tensordot(mats[d][k], val, axes=([0], [0]))
            out[k] = val
        return out[:, :-1] / out[:, -1:]


# ---------------------------------------------------------------------------
# primitives (primitives.py)
# ---------------------------------------------------------------------------


def quadrilateral(corners):
    """Bilinear quad from counter-clockwise corners (``primitives.py:41-69``)."""
    a, b, c, d = np.asarray(corners, float)
    return Patch.create(
        [[0, 0, 1, 1], [0, 0, 1, 1]], (1, 1), np.array([a, b, d, c]), np.ones(4),
        {"u0": "left", "u1": "right", "v0": "bottom", "v1": "top"},
    )


def rectangle(x0, y0, x1, y1):
    return quadrilateral([[x0, y0], [x1, y0], [x1, y1], [x0, y1]])


def quarter_annulus(r_in, r_out):
    """Quarter annulus, u radial (degree 1), v circumferential (rational
    quadratic) (``primitives.py:107-133``)."""
    s = 1.0 / np.sqrt(2.0)
    # u fastest: (inner, outer) for each of the three angular control points
    pts = np.array([[r_in, 0], [r_out, 0], [r_in, r_in], [r_out, r_out], [0, r_in], [0, r_out]], float)
    w = np.array([1.0, 1.0, s, s, 1.0, 1.0])
    return Patch.create(
        [[0, 0, 1, 1], [0, 0, 0, 1, 1, 1]], (1, 2), pts, w,
        {"u0": "inner", "u1": "outer", "v0": "sym_y", "v1": "sym_x"},
    )


def cuboid(lower, upper):
    """Trilinear box (``primitives.py:285-309``)."""
    lo, hi = np.asarray(lower, float), np.asarray(upper, float)
    pts = np.array(
        [[(lo, hi)[i][0], (lo, hi)[j][1], (lo, hi)[k][2]] for k in (0, 1) for j in (0, 1) for i in (0, 1)]
    )
    return Patch.create(
        [[0, 0, 1, 1]] * 3, (1, 1, 1), pts, np.ones(8),
        {"u0": "left", "u1": "right", "v0": "bottom", "v1": "top", "w0": "back", "w1": "front"},
    )
