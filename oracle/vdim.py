"""Volume density interpolation (vdim_correction, Inti.jl src/vdim.jl:30-129)
restated in numpy for 2D and 3D, on top of the C oracle's direct sums and
boundary correction.

TEST INFRASTRUCTURE ONLY (a CPU reference; the product never imports it).

Independent of the product's closed-form particular solutions: polynomials
are numpy coefficient arrays c[i, j(, l)] (x^i y^j z^l) and P is obtained by
solving the linear equations on its coefficients:
  Laplace   Delta P = p with P = |x|^2 q, q homogeneous of degree |I|;
  Helmholtz (Delta + k^2) P = p with deg P = |I|.
The per-element least squares use numpy's SVD-based lstsq, as the reference
solves with svd(L) (vdim.jl:104-107).

Parity of P: for Helmholtz (k != 0) the polynomial solution is unique
((Delta + k^2) is invertible on polynomials of degree <= d), so P and every
correction built from it are fixed by the mathematics. For Laplace P is not
unique (any harmonic polynomial may be added): the reference takes it from
ElementaryPDESolutions.solve_laplace (compat 0.2 - 0.3 in the reference's
Project.toml, not vendored), whose particular choice no fixture fixes, so the
Laplace correction values are "parity unpinned" against the reference; the
choice here is the unique solution of the form |x|^2 q. The corrected volume
potential itself is pinned for both: its convergence against a manufactured
solution's Green representation.
"""
import math

import numpy as np
from numpy.polynomial import polynomial as npoly

from . import bdim, csr_matvec, matvec, nearest_lists

LAP, HEL = 0, 1


def multiindices(order, dim=2):
    """I with |I| <= order in Iterators.product order (first index fastest)."""
    import itertools
    out = []
    for t in itertools.product(range(order + 1), repeat=dim):
        I = tuple(reversed(t))  # product's last index is fastest: reverse
        if sum(I) <= order:
            out.append(I)
    return out


def _lap(c):
    out = np.zeros_like(c)
    for ax in range(c.ndim):
        if c.shape[ax] > 2:
            d = npoly.polyder(c, 2, axis=ax)
            out[tuple(slice(0, n) for n in d.shape)] += d
    return out


def _r2(q):
    """|x|^2 q for a coefficient array q (the top entries must be zero)."""
    P = np.zeros_like(q)
    for ax in range(q.ndim):
        src = [slice(None)] * q.ndim
        dst = [slice(None)] * q.ndim
        src[ax] = slice(0, -2)
        dst[ax] = slice(2, None)
        P[tuple(dst)] += q[tuple(src)]
    return P


def _monomials(deg, dim, exact):
    import itertools
    return [e for e in itertools.product(range(deg + 1), repeat=dim) if (sum(e) == deg if exact else sum(e) <= deg)]


def particular_solution(I, pde, k, size):
    """Coefficient array (size^dim) of P for p = x^I / I!."""
    dim = len(I)
    d = sum(I)
    p = np.zeros((size,) * dim)
    p[I] = 1.0 / np.prod([math.factorial(i) for i in I])
    if pde == LAP:
        # unknowns: coefficients of q (homogeneous of degree d); P = |x|^2 q
        cols = []
        for e in _monomials(d, dim, True):
            q = np.zeros((size,) * dim)
            q[e] = 1.0
            cols.append(_r2(q))
        M = np.stack([_lap(P).ravel() for P in cols], axis=1)
        sol = np.linalg.lstsq(M, p.ravel(), rcond=None)[0]
        return sum(s * P for s, P in zip(sol, cols))
    # Helmholtz: unknown coefficients of all monomials of degree <= d
    basis = _monomials(d, dim, False)
    cols = []
    for e in basis:
        P = np.zeros((size,) * dim)
        P[e] = 1.0
        cols.append(P)
    M = np.stack([(_lap(P) + k * k * P).ravel() for P in cols], axis=1)
    flat = [np.ravel_multi_index(e, (size,) * dim) for e in basis]
    sol = np.linalg.solve(M[flat], p.ravel()[flat])
    return sum(s * P for s, P in zip(sol, cols))


def _eval(c, x):
    if c.ndim == 2:
        return npoly.polyval2d(x[0], x[1], c)
    return npoly.polyval3d(x[0], x[1], x[2], c)


def _grad_n(c, x, n):
    out = 0.0
    for ax in range(c.ndim):
        out = out + _eval(npoly.polyder(c, 1, axis=ax), x) * n[ax]
    return out


def boundary_operators(pde, k, target, boundary, mu, maxdist=np.inf, self_pattern=False,
                       variant=None, near=None):
    """Corrected S and D (boundary -> target) as callables on (n, r) arrays;
    variant="fma" takes the boundary correction from the FMA-contracted build;
    near: element -> target lists (default: the oracle's nearest-node lists)."""
    dim = boundary.coords.shape[0]
    if near is None and not self_pattern:
        near = nearest_lists(target, boundary, maxdist)
    rp, ci, vS, vD, _ = bdim(pde, dim, k, target, boundary, mu, near=near, variant=variant)
    m, n = target.coords.shape[1], boundary.coords.shape[1]

    def S(X):
        return matvec(pde, dim, 0, k, target, boundary, X) + csr_matvec(m, n, rp, ci, vS, X)

    def D(X):
        return matvec(pde, dim, 1, k, target, boundary, X) + csr_matvec(m, n, rp, ci, vD, X)
    return S, D


def vdim(pde, k, target, source, boundary, mu, order, S, D, near=None):
    """CSR (rowptr, colind, values) of dV."""
    dim = source.coords.shape[0]
    size = order + 3
    idx = multiindices(order, dim)
    ps, Ps = [], []
    for I in idx:
        p = np.zeros((size,) * dim)
        p[I] = 1.0 / np.prod([math.factorial(i) for i in I])
        ps.append(p)
        Ps.append(particular_solution(I, pde, k, size))
    b = np.stack([_eval(p, source.coords) for p in ps], axis=1)
    g0 = np.stack([_eval(P, boundary.coords) for P in Ps], axis=1)
    g1 = np.stack([_grad_n(P, boundary.coords, boundary.normals) for P in Ps], axis=1)
    Pt = np.stack([_eval(P, target.coords) for P in Ps], axis=1)
    V = matvec(pde, dim, 0, k, target, source, b)
    Theta = S(g1) - D(g0) - V + mu[:, None] * Pt
    if pde == LAP:
        Theta = Theta.real
    m = target.coords.shape[1]
    rows, cols, vals = [], [], []
    types = getattr(source, "etype_qtags", None) or [source.qtags]
    lists = [near] if (near is None or len(types) == 1) else list(near)
    if near is None:
        lists = [None] * len(types)
    for qtags, nr in zip(types, lists):  # vdim.jl:76-121, type by type
        nq, ne = qtags.shape
        for e in range(ne):
            jg = qtags[:, e]
            L = b[jg].T  # nb x nq
            tg = jg if nr is None else nr[1][nr[0][e]:nr[0][e + 1]]
            for i in tg:
                w = np.linalg.lstsq(L, Theta[i], rcond=None)[0]
                rows += [i] * nq
                cols += list(jg)
                vals += list(w)
    order_ = np.lexsort((np.array(cols), np.array(rows)))
    rows = np.array(rows)[order_]
    cols = np.array(cols)[order_]
    vals = np.array(vals)[order_]
    rowptr = np.zeros(m + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    return np.cumsum(rowptr), cols, vals
