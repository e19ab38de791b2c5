"""Oracle space: Bezier extraction, Bernstein bases, Gauss rules, DOF maps and
basis evaluation.

Restates
* ``src/jaxiga/space/bernstein.py`` -- bezier_extraction :16-76 (Borden et
  al. 2011, Algorithm 1), Bernstein values/derivatives :79-156, tensor
  numbering :159-232, element_spans :261-284;
* ``src/jaxiga/space/pointset.py`` -- gauss_1d/_tensor_rule/gauss :98-136,
  boundary_gauss :148-200;
* ``src/jaxiga/space/function_space.py`` -- expand_dofs :300-312,
  _tensor_extraction :315-323, _patch_connectivity :326-369,
  _classify_boundaries :490-546, make_dofmap/DofMap :554-603;
* ``src/jaxiga/geometry/multipatch.py`` -- side_indices :99-131;
* ``src/jaxiga/space/evaluation.py`` -- _rational_and_pushforward :95-172,
  _surface_measure :175-193, evaluate :305-467.

Everything is FP64 NumPy. ``evaluate`` follows the reference's own data path:
the full per-element Kronecker extraction operator times the tensor Bernstein
table, then the rational quotient and a per-point linear solve.
"""

from __future__ import annotations

import dataclasses
from math import comb

import numpy as np

from oracle.geometry import side_code, side_name


# ---------------------------------------------------------------------------
# Bezier extraction (Borden et al. 2011, Algorithm 1)
# ---------------------------------------------------------------------------


def bezier_extraction(knots, p):
    """Per-element ``(p+1, p+1)`` operators with ``N_e = C_e @ B``."""
    U = np.asarray(knots, float)
    m = len(U)
    end = m - p - 1  # index of the first copy of the last knot value
    a, b = p, p + 1
    cur = np.eye(p + 1)
    ops = []
    while b < end:
        nxt = np.eye(p + 1)
        first = b
        while b + 1 < end and U[b + 1] == U[b]:
            b += 1
        mult = b - first + 1
        if mult < p:
            span = U[b] - U[a]
            alphas = np.array([span / (U[a + j] - U[a]) for j in range(mult + 1, p + 1)])
            r = p - mult
            for j in range(1, r + 1):
                save = r - j
                s = mult + j
                for k in range(p, s - 1, -1):
                    al = alphas[k - s]
                    cur[:, k] = al * cur[:, k] + (1.0 - al) * cur[:, k - 1]
                nxt[save : save + j + 1, save] = cur[p - j : p + 1, p]
        ops.append(cur)
        cur = nxt
        a, b = b, b + 1
    ops.append(cur)
    return ops


def element_spans(knots, p):
    """Non-empty spans ``lo, hi`` and the first supported basis index."""
    U = np.asarray(knots, float)
    idx = np.nonzero(U[1:] > U[:-1])[0]
    return U[idx], U[idx + 1], idx - p


# ---------------------------------------------------------------------------
# Bernstein polynomials on [-1, 1]
# ---------------------------------------------------------------------------


def _bern(u, p):
    u = np.atleast_1d(np.asarray(u, float))
    if p < 0:
        return np.zeros((len(u), 0))
    s, t = 0.5 * (1.0 - u), 0.5 * (1.0 + u)
    return np.stack([comb(p, k) * s ** (p - k) * t**k for k in range(p + 1)], axis=1)


def bernstein(u, p, order=0):
    """Degree-``p`` Bernstein values (order 0), first or second derivative."""
    u = np.atleast_1d(np.asarray(u, float))
    if order == 0:
        return _bern(u, p)
    if order == 1:
        low = np.concatenate([np.zeros((len(u), 1)), _bern(u, p - 1), np.zeros((len(u), 1))], 1)
        return 0.5 * p * (low[:, :-1] - low[:, 1:])
    if p < 2:
        return np.zeros((len(u), p + 1))
    low = np.concatenate([np.zeros((len(u), 2)), _bern(u, p - 2), np.zeros((len(u), 2))], 1)
    return 0.25 * p * (p - 1) * (low[:, :-2] - 2 * low[:, 1:-1] + low[:, 2:])[:, : p + 1]


def _tensor(factors):
    """Per-point tensor product, first direction fastest."""
    out = factors[0]
    for f in factors[1:]:
        out = (f[:, :, None] * out[:, None, :]).reshape(out.shape[0], -1)
    return out


def bernstein_tensor(pts, degree):
    """``B (n, n_l)`` and ``dB (n, dim, n_l)`` at reference points ``(n, dim)``."""
    pts = np.atleast_2d(pts)
    dim = len(degree)
    vals = [bernstein(pts[:, d], degree[d]) for d in range(dim)]
    ders = [bernstein(pts[:, d], degree[d], 1) for d in range(dim)]
    B = _tensor(vals)
    dB = np.stack([_tensor([ders[k] if k == d else vals[k] for k in range(dim)]) for d in range(dim)], 1)
    return B, dB


# ---------------------------------------------------------------------------
# Gauss rules
# ---------------------------------------------------------------------------


def gauss_rule(n_per_dir):
    """Tensor Gauss-Legendre rule on [-1,1]^dim, first direction fastest."""
    rules = [np.polynomial.legendre.leggauss(int(n)) for n in n_per_dir]
    pts = rules[0][0][:, None]
    wts = rules[0][1]
    for x, w in rules[1:]:
        pts = np.concatenate(
            [np.tile(pts, (len(x), 1)), np.repeat(x, len(pts))[:, None]], axis=1
        )
        wts = np.tile(wts, len(x)) * np.repeat(w, len(wts))
    return pts, wts


def boundary_rule(degree, code, n=None):
    """Face Gauss rule for side ``code`` embedded in the reference cube."""
    dim = len(degree)
    n_per = [(p + 1) if n is None else int(n) for p in degree]
    d, end = divmod(int(code), 2)
    tang = [k for k in range(dim) if k != d]
    if tang:
        fp, fw = gauss_rule([n_per[k] for k in tang])
    else:
        fp, fw = np.zeros((1, 0)), np.ones(1)
    pts = np.zeros((len(fp), dim))
    for j, k in enumerate(tang):
        pts[:, k] = fp[:, j]
    pts[:, d] = 1.0 if end else -1.0
    return pts, fw


# ---------------------------------------------------------------------------
# Function space
# ---------------------------------------------------------------------------


def expand_dofs(scalar_dofs, vec):
    scalar_dofs = np.asarray(scalar_dofs)
    if vec == 1:
        return scalar_dofs
    return (vec * scalar_dofs[..., :, None] + np.arange(vec)).reshape(*scalar_dofs.shape[:-1], -1)


def _tensor_index(ranges, strides):
    """Flat indices of a tensor block, first listed direction fastest."""
    idx = np.zeros(1, dtype=np.int64)
    for r, s in zip(ranges, strides):
        idx = (np.asarray(r, dtype=np.int64)[:, None] * s + idx[None, :]).reshape(-1)
    return idx


def side_indices(n_per_dir, side):
    d = "uvw".index(side[0])
    end = int(side[1:])
    strides = np.concatenate([[1], np.cumprod(n_per_dir[:-1])]).astype(np.int64)
    ranges = [np.arange(n) for n in n_per_dir]
    ranges[d] = np.array([0 if end == 0 else n_per_dir[d] - 1])
    return _tensor_index(ranges, strides)


@dataclasses.dataclass
class BoundarySet:
    elems: np.ndarray
    sides: np.ndarray
    dofs: np.ndarray


@dataclasses.dataclass
class FunctionSpace:
    patch: object
    vec: int
    dim: int
    degree: tuple
    elem_dofs: np.ndarray  # (n_e, n_l)
    elem_vertex: np.ndarray  # (n_e, 2*dim) [mins..., maxs...]
    extraction_1d: list  # per direction: list of (p+1, p+1)
    elem_index: np.ndarray  # (n_e, dim) per-direction element index
    cpts: np.ndarray
    wgts: np.ndarray
    boundaries: dict
    n_elem_per_dir: tuple

    @property
    def n_elems(self):
        return len(self.elem_dofs)

    @property
    def n_scalar_basis(self):
        return len(self.wgts)

    @property
    def n_dofs(self):
        return self.n_scalar_basis * self.vec

    @property
    def n_local(self):
        return self.elem_dofs.shape[1]

    def extraction(self, elems=None):
        """Full Kronecker operators ``(n, n_l, n_l)``, later directions outermost."""
        elems = np.arange(self.n_elems) if elems is None else np.asarray(elems)
        out = []
        for e in elems:
            C = self.extraction_1d[-1][self.elem_index[e, -1]]
            for d in range(self.dim - 2, -1, -1):
                C = np.kron(C, self.extraction_1d[d][self.elem_index[e, d]])
            out.append(C)
        return np.array(out)

    def boundary(self, label):
        return self.boundaries[label]


def function_space(patch, vec=1):
    dim = patch.dim
    deg = patch.degree
    n_cp = patch.n_cp_per_dir
    strides = np.concatenate([[1], np.cumprod(n_cp[:-1])]).astype(np.int64)
    spans = [element_spans(patch.knots[d], deg[d]) for d in range(dim)]
    ext = [bezier_extraction(patch.knots[d], deg[d]) for d in range(dim)]
    ne = tuple(len(s[0]) for s in spans)
    n_e = int(np.prod(ne))

    # element numbering, first direction fastest
    eidx = np.zeros((n_e, dim), dtype=np.int64)
    rem = np.arange(n_e)
    for d in range(dim):
        eidx[:, d] = rem % ne[d]
        rem //= ne[d]
    local = _tensor_index([np.arange(p + 1) for p in deg], strides)  # (n_l,)
    base = sum(spans[d][2][eidx[:, d]] * strides[d] for d in range(dim))
    elem_dofs = base[:, None] + local[None, :]
    boxes = np.concatenate(
        [np.stack([spans[d][0][eidx[:, d]] for d in range(dim)], 1),
         np.stack([spans[d][1][eidx[:, d]] for d in range(dim)], 1)], 1)

    estr = np.concatenate([[1], np.cumprod(ne[:-1])]).astype(np.int64)
    collected = {}
    for d in range(dim):
        for end in (0, 1):
            side = side_name(d, end)
            label = patch.labels.get(side, f"patch0/{side}")
            rng = [np.arange(n) for n in ne]
            rng[d] = np.array([0 if end == 0 else ne[d] - 1])
            elems = _tensor_index(rng, estr)
            dofs = side_indices(n_cp, side)
            entry = collected.setdefault(label, ([], [], []))
            entry[0].append(elems)
            entry[1].append(np.full(len(elems), side_code(side)))
            entry[2].append(dofs)
    boundaries = {
        k: BoundarySet(np.concatenate(v[0]), np.concatenate(v[1]), np.unique(np.concatenate(v[2])))
        for k, v in collected.items()
    }
    return FunctionSpace(
        patch=patch, vec=int(vec), dim=dim, degree=deg, elem_dofs=elem_dofs, elem_vertex=boxes,
        extraction_1d=ext, elem_index=eidx, cpts=np.asarray(patch.ctrl_pts, float).copy(),
        wgts=np.asarray(patch.weights, float).copy(), boundaries=boundaries, n_elem_per_dir=ne,
    )


@dataclasses.dataclass
class DofMap:
    free: np.ndarray
    prescribed: np.ndarray
    values: np.ndarray
    n_dofs: int

    def lift(self, u_free):
        u = np.zeros(self.n_dofs)
        u[self.free] = u_free
        u[self.prescribed] = self.values
        return u


def make_dofmap(n_dofs, prescribed, values):
    prescribed = np.asarray(prescribed, dtype=np.int64)
    order = np.argsort(prescribed, kind="stable")
    prescribed = prescribed[order]
    values = np.asarray(values, float)[order] if len(prescribed) else np.zeros(0)
    free = np.setdiff1d(np.arange(n_dofs), prescribed)
    return DofMap(free, prescribed, values, int(n_dofs))


# ---------------------------------------------------------------------------
# evaluation
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class BasisData:
    R: np.ndarray  # (n_e, n_q, n_l)
    dR: np.ndarray  # (n_e, n_q, dim, n_l)
    x: np.ndarray  # (n_e, n_q, dim_phys)
    dofs: np.ndarray  # (n_e, n_l)
    w: np.ndarray | None = None
    normal: np.ndarray | None = None
    detJ: np.ndarray | None = None


def evaluate(space, elems, ref, weights=None, sides=None):
    """Basis, physical gradients, points and weights (``evaluation.py:305-467``).

    ``ref`` is ``(n_q, dim)`` shared by all elements or ``(n_e, n_q, dim)``.
    """
    elems = np.asarray(elems, dtype=np.int64)
    dim = space.dim
    n_e = len(elems)
    ref = np.asarray(ref, float)
    if ref.ndim == 2:
        B, dB = bernstein_tensor(ref, space.degree)
        B = np.broadcast_to(B, (n_e,) + B.shape)
        dB = np.broadcast_to(dB, (n_e,) + dB.shape)
    else:
        B, dB = bernstein_tensor(ref.reshape(-1, dim), space.degree)
        B = B.reshape(n_e, ref.shape[1], -1)
        dB = dB.reshape(n_e, ref.shape[1], dim, -1)
    boxes = space.elem_vertex[elems]
    scale = 2.0 / (boxes[:, dim:] - boxes[:, :dim])
    C = space.extraction(elems)
    dofs = space.elem_dofs[elems]
    cp = space.cpts[dofs]  # (n_e, n_l, dim_phys)
    wg = space.wgts[dofs]  # (n_e, n_l)

    N = np.einsum("eij,eqj->eqi", C, B)
    dN = np.einsum("eij,eqdj->eqdi", C, dB) * scale[:, None, :, None]

    M = N * wg[:, None, :]
    dM = dN * wg[:, None, None, :]
    W = M.sum(-1)
    dW = dM.sum(-1)
    R = M / W[..., None]
    dRp = dM / W[..., None, None] - M[:, :, None, :] * dW[..., None] / W[..., None, None] ** 2
    x = np.einsum("eqi,eic->eqc", R, cp)
    J = np.einsum("eqai,eic->eqac", dRp, cp)  # J[a, c] = d x_c / d xi_a
    detJ = np.linalg.det(J)
    dR = np.linalg.solve(J, dRp)

    w = None
    normal = None
    if sides is not None:
        sides = np.asarray(sides)
        d_e, end_e = sides // 2, sides % 2
        sign = np.where(end_e == 1, 1.0, -1.0)
        normal = np.zeros_like(x)
        surf = np.zeros(x.shape[:2])
        tscale = np.ones(n_e)
        for i in range(n_e):
            d = int(d_e[i])
            tang = [a for a in range(dim) if a != d]
            Ji = J[i]
            if len(tang) == 1:
                surf[i] = np.linalg.norm(Ji[:, tang[0], :], axis=-1)
            elif len(tang) == 2:
                surf[i] = np.linalg.norm(np.cross(Ji[:, tang[0], :], Ji[:, tang[1], :]), axis=-1)
            else:
                surf[i] = 1.0
            e = np.zeros(dim)
            e[d] = 1.0
            nn = np.linalg.solve(Ji, np.broadcast_to(e, (Ji.shape[0], dim))[..., None])[..., 0]
            normal[i] = nn / np.linalg.norm(nn, axis=-1, keepdims=True) * sign[i]
            tscale[i] = np.prod([1.0 / scale[i, a] for a in tang]) if tang else 1.0
        if weights is not None:
            w = np.asarray(weights) * surf * tscale[:, None]
    elif weights is not None:
        jac = np.prod(2.0 / scale, axis=1) / 2.0**dim
        w = np.asarray(weights)[None, :] * np.abs(detJ) * jac[:, None]
    return BasisData(R=R, dR=dR, x=x, dofs=dofs, w=w, normal=normal, detJ=detJ)


def gauss_basis(space, n=None):
    pts, wts = gauss_rule([(p + 1) if n is None else int(n) for p in space.degree])
    return evaluate(space, np.arange(space.n_elems), pts, wts)


def boundary_basis(space, bset, n=None):
    codes = np.asarray(bset.sides)
    rules = {int(c): boundary_rule(space.degree, c, n) for c in np.unique(codes)}
    ref = np.stack([rules[int(c)][0] for c in codes])
    wts = np.stack([rules[int(c)][1] for c in codes])
    return evaluate(space, bset.elems, ref, wts, codes)
