"""Oracle for the learned multigrid preconditioner (test infrastructure only).

The reference has no multigrid (``src/jaxiga/solvers/linear.py:113-116``);
this module states the project's own definition in NumPy/SciPy FP64:
dyadic knot coarsening, the knot-insertion transfer (built here by Boehm
insertion, ``oracle.geometry.boehm_insert``, independently of the product's
collocation), the Galerkin levels ``P^T K P``, the operator-stencil features,
the stencil smoother and its transpose, the V-cycle, flexible CG with one
stored direction (Notay 2000) and the smoother-training loss. Pinned by
``tests/test_oracle_multigrid.py``: the transfer reproduces every coarse
B-spline exactly, the Galerkin operator equals the coarse-space stiffness
assembled directly, stencil transpose/gradients against adjointness and
finite differences, the V-cycle is symmetric and converges.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from oracle import geometry as G


def coarsen_knots(knots, p):
    t = np.asarray(knots, float)
    u, mult = np.unique(t, return_counts=True)
    m = len(u) - 1
    if m % 2 or m < 4:
        return None
    keep = np.ones(len(u), bool)
    keep[1:-1:2] = False
    return np.repeat(u[keep], mult[keep])


def refinement_matrix(coarse_knots, fine_knots, p):
    """``T`` with fine coefficients ``= T @`` coarse coefficients, by Boehm
    insertion of the missing knots into the identity coefficient matrix."""
    tc = np.asarray(coarse_knots, float)
    nc = len(tc) - p - 1
    new = list(np.asarray(fine_knots, float))
    for u in tc:  # the multiset difference fine - coarse
        new.remove(u)
    ctrl, knots = G.insert_knots_1d(p, np.eye(nc), tc, new)
    assert np.allclose(knots, fine_knots)
    return ctrl


def free_rectangle(n_cp, prescribed):
    n0, n1 = n_cp
    pres = np.zeros(n0 * n1, bool)
    pres[np.asarray(prescribed, int)] = True
    grid = pres.reshape(n1, n0)
    fixed = {(0, 0): grid[:, 0].all(), (0, 1): grid[:, -1].all(),
             (1, 0): grid[0, :].all(), (1, 1): grid[-1, :].all()}
    lo = [1 if fixed[(d, 0)] else 0 for d in range(2)]
    hi = [n_cp[d] - 2 if fixed[(d, 1)] else n_cp[d] - 1 for d in range(2)]
    return lo, hi, fixed


def hierarchy(K, knots, p, prescribed, max_coarse=64, max_levels=16):
    """Levels ``[{K, P, h, w}]`` (P maps the next coarser level to this one)."""
    n_cp = [len(t) - p - 1 for t in knots]
    lo, hi, fixed = free_rectangle(n_cp, prescribed)
    levels = [{"K": sp.csr_matrix(K), "h": hi[1] - lo[1] + 1, "w": hi[0] - lo[0] + 1,
               "knots": [np.asarray(t, float) for t in knots]}]
    while len(levels) < max_levels and levels[-1]["K"].shape[0] > max_coarse:
        fine = levels[-1]
        ck = [coarsen_knots(t, p) for t in fine["knots"]]
        if any(c is None for c in ck):
            break
        P1 = []
        for d in range(2):
            T = refinement_matrix(ck[d], fine["knots"][d], p)
            nf, nc = T.shape
            r0, r1 = (1 if fixed[(d, 0)] else 0), (nf - 2 if fixed[(d, 1)] else nf - 1)
            c0, c1 = (1 if fixed[(d, 0)] else 0), (nc - 2 if fixed[(d, 1)] else nc - 1)
            P1.append(T[r0 : r1 + 1, c0 : c1 + 1])
        P = sp.csr_matrix(np.kron(P1[1], P1[0]))
        fine["P"] = P
        levels.append({"K": sp.csr_matrix(P.T @ fine["K"] @ P), "h": P1[1].shape[1],
                       "w": P1[0].shape[1], "knots": ck})
    return levels


def stencil_features(K, h, w, R):
    """``feats[o, i] = K[i, i + off(o)] / K[i, i]`` on the ``h x w`` grid."""
    K = sp.coo_matrix(K)
    d = sp.csr_matrix(K).diagonal()
    F = np.zeros(((2 * R + 1) ** 2, h * w))
    dr = K.col // w - K.row // w
    dc = K.col % w - K.row % w
    ok = (np.abs(dr) <= R) & (np.abs(dc) <= R)
    o = (dr + R) * (2 * R + 1) + dc + R
    F[o[ok], K.row[ok]] = K.data[ok] / d[K.row[ok]]
    return F


def _offsets(ks):
    r = ks // 2
    return [(dy, dx) for dy in range(-r, r + 1) for dx in range(-r, r + 1)]


def stencil_apply(W, dsq, v, h, w, ks, transpose=False):
    """``D^-1/2 T (D^-1/2 v)`` with ``(T_W u)_i = sum_o W[o, i] u[i + off(o)]``
    (or its transpose); ``v`` is ``(n,)`` or ``(B, n)``."""
    u = (dsq * v).reshape(-1, h, w)
    out = np.zeros_like(u)
    Wg = W.reshape(ks * ks, h, w)
    for o, (dy, dx) in enumerate(_offsets(ks)):
        # target rows r, source rows r + dy (both in range)
        r0, r1 = max(0, -dy), min(h, h - dy)
        c0, c1 = max(0, -dx), min(w, w - dx)
        if r0 >= r1 or c0 >= c1:
            continue
        tgt = (slice(None), slice(r0, r1), slice(c0, c1))
        src = (slice(None), slice(r0 + dy, r1 + dy), slice(c0 + dx, c1 + dx))
        if not transpose:
            out[tgt] += Wg[o][tgt[1:]] * u[src]
        else:
            out[src] += Wg[o][tgt[1:]] * u[tgt]
    return (dsq * out.reshape(u.shape[0], -1)).reshape(np.shape(v))


def stencil_wgrad_t(dsq, v, g, h, w, ks):
    """d(sum_b g_b . y_b)/dW for ``y_b = D^-1/2 T_W^T (D^-1/2 v_b)``."""
    u = (dsq * np.atleast_2d(v)).reshape(-1, h, w)
    gs = (dsq * np.atleast_2d(g)).reshape(-1, h, w)
    # y_i = dsq_i sum_o W[o, j] u_j over i = j + off(o): dW[o, j] = sum_b gs_b[j + off(o)] u_b[j]
    out = np.zeros((ks * ks, h, w))
    for o, (dy, dx) in enumerate(_offsets(ks)):
        r0, r1 = max(0, -dy), min(h, h - dy)
        c0, c1 = max(0, -dx), min(w, w - dx)
        if r0 >= r1 or c0 >= c1:
            continue
        out[o, r0:r1, c0:c1] = (gs[:, r0 + dy : r1 + dy, c0 + dx : c1 + dx] * u[:, r0:r1, c0:c1]).sum(0)
    return out.reshape(ks * ks, h * w)


def vcycle(levels, stencils, b, omega=0.5, ks=7, level=0):
    """One V-cycle (pre-smoother S, post-smoother S^T, dense coarsest solve);
    ``stencils[l]`` is ``(ks*ks, n_l)`` or None for damped Jacobi."""
    L = len(levels) - 1
    lv = levels[level]
    K = lv["K"]
    if level == L:
        return np.linalg.solve(K.toarray(), b)
    dsq = 1.0 / np.sqrt(K.diagonal())
    W = stencils[level] if stencils is not None else None

    def S(v, transpose):
        if W is None:
            return omega * dsq * dsq * v
        return stencil_apply(W, dsq, v, lv["h"], lv["w"], ks, transpose)

    x = S(b, False)
    r = b - K @ x
    x = x + lv["P"] @ vcycle(levels, stencils, lv["P"].T @ r, omega, ks, level + 1)
    r = b - K @ x
    return x + S(r, True)


def fcg(K, b, M, tol=1e-10, maxiter=1000):
    """Flexible CG with one stored direction (Notay 2000), x0 = 0; stops at
    ||r|| <= tol ||b||. Returns ``(x, iterations)``."""
    x = np.zeros_like(b)
    r = b.copy()
    bb = b @ b
    p_prev = q_prev = None
    for it in range(maxiter + 1):
        if not (r @ r > tol * tol * bb) or it == maxiter:
            return x, it
        z = M(r)
        p = z.copy()
        if p_prev is not None:
            p -= (z @ q_prev) / (p_prev @ q_prev) * p_prev
        q = K @ p
        alpha = (p @ r) / (p @ q)
        x += alpha * p
        r -= alpha * q
        p_prev, q_prev = p, q
    return x, maxiter


def energy_loss(K, W, dsq, E, R, h, w, ks):
    """``mean_b ||E_b - S^T R_b||_K^2`` and its gradient with respect to W."""
    Y = stencil_apply(W, dsq, R, h, w, ks, transpose=True)
    Z = E - Y
    KZ = (K @ Z.T).T
    B = E.shape[0]
    loss = float(np.einsum("bi,bi->b", Z, KZ).mean())
    gy = -2.0 * KZ / B
    return loss, stencil_wgrad_t(dsq, R, gy, h, w, ks)
