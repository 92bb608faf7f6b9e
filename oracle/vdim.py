"""Volume density interpolation (vdim_correction, Inti.jl src/vdim.jl:30-129)
restated in numpy for 2D, on top of the C oracle's direct sums and boundary
correction.

TEST INFRASTRUCTURE ONLY (a CPU reference; the product never imports it).

Independent of the product's closed-form particular solutions: polynomials
are numpy coefficient arrays c[i, j] (x^i y^j) and P is obtained by solving the
linear equations on its coefficients:
  Laplace   Delta P = p with P = (x^2 + y^2) q, q homogeneous of degree |I|;
  Helmholtz (Delta + k^2) P = p with deg P = |I|.
The per-element least squares use numpy's SVD-based lstsq, as the reference
solves with svd(L) (vdim.jl:104-107).

Parity unpinned: which particular solution P the reference picks. It comes
from ElementaryPDESolutions (compat 0.2 - 0.3 in the reference's Project.toml,
not vendored), and P is not unique, so no fixture fixes that choice. The
corrected volume potential itself is pinned: its convergence rate against a
manufactured solution's Green representation.
"""
import math

import numpy as np
from numpy.polynomial import polynomial as npoly

from . import bdim, csr_matvec, matvec, nearest_lists

LAP, HEL = 0, 1


def multiindices(order):
    """(i, j) with i + j <= order, i fastest (Iterators.product order)."""
    return [(i, j) for j in range(order + 1) for i in range(order + 1) if i + j <= order]


def _lap(c):
    out = np.zeros_like(c)
    if c.shape[0] > 2:
        d = npoly.polyder(c, 2, axis=0)
        out[:d.shape[0], :d.shape[1]] += d
    if c.shape[1] > 2:
        d = npoly.polyder(c, 2, axis=1)
        out[:d.shape[0], :d.shape[1]] += d
    return out


def particular_solution(I, pde, k, size):
    """Coefficient array (size x size) of P for p = x^I / I!."""
    d = I[0] + I[1]
    p = np.zeros((size, size))
    p[I] = 1.0 / (math.factorial(I[0]) * math.factorial(I[1]))
    if pde == LAP:
        # unknowns: coefficients of q = sum_a q_a x^a y^(d-a); P = r^2 q
        cols = []
        for a in range(d + 1):
            q = np.zeros((size, size))
            q[a, d - a] = 1.0
            P = np.zeros((size, size))
            P[2:, :] += q[:-2, :]
            P[:, 2:] += q[:, :-2]
            cols.append(P)
        M = np.stack([_lap(P).ravel() for P in cols], axis=1)
        sol = np.linalg.lstsq(M, p.ravel(), rcond=None)[0]
        return sum(s * P for s, P in zip(sol, cols))
    # Helmholtz: unknown coefficients of all monomials of degree <= d
    basis = [(a, b) for a in range(d + 1) for b in range(d + 1 - a)]
    cols = []
    for (a, b) in basis:
        P = np.zeros((size, size))
        P[a, b] = 1.0
        cols.append(P)
    M = np.stack([(_lap(P) + k * k * P).ravel() for P in cols], axis=1)
    sol = np.linalg.solve(M[[a * size + b for (a, b) in basis]], p.ravel()[[a * size + b for (a, b) in basis]])
    return sum(s * P for s, P in zip(sol, cols))


def _eval(c, x):
    return npoly.polyval2d(x[0], x[1], c)


def _grad_n(c, x, n):
    cx = npoly.polyder(c, 1, axis=0)
    cy = npoly.polyder(c, 1, axis=1)
    return npoly.polyval2d(x[0], x[1], cx) * n[0] + npoly.polyval2d(x[0], x[1], cy) * n[1]


def boundary_operators(pde, k, target, boundary, mu, maxdist=np.inf, self_pattern=False,
                       variant=None):
    """Corrected S and D (boundary -> target) as callables on (n, r) arrays;
    variant="fma" takes the boundary correction from the FMA-contracted build."""
    near = None if self_pattern else nearest_lists(target, boundary, maxdist)
    rp, ci, vS, vD, _ = bdim(pde, 2, k, target, boundary, mu, near=near, variant=variant)
    m, n = target.coords.shape[1], boundary.coords.shape[1]

    def S(X):
        return matvec(pde, 2, 0, k, target, boundary, X) + csr_matvec(m, n, rp, ci, vS, X)

    def D(X):
        return matvec(pde, 2, 1, k, target, boundary, X) + csr_matvec(m, n, rp, ci, vD, X)
    return S, D


def vdim(pde, k, target, source, boundary, mu, order, S, D, near=None):
    """CSR (rowptr, colind, values) of dV."""
    size = order + 3
    idx = multiindices(order)
    ps, Ps = [], []
    for I in idx:
        p = np.zeros((size, size))
        p[I] = 1.0 / (math.factorial(I[0]) * math.factorial(I[1]))
        ps.append(p)
        Ps.append(particular_solution(I, pde, k, size))
    b = np.stack([_eval(p, source.coords) for p in ps], axis=1)
    g0 = np.stack([_eval(P, boundary.coords) for P in Ps], axis=1)
    g1 = np.stack([_grad_n(P, boundary.coords, boundary.normals) for P in Ps], axis=1)
    Pt = np.stack([_eval(P, target.coords) for P in Ps], axis=1)
    V = matvec(pde, 2, 0, k, target, source, b)
    Theta = S(g1) - D(g0) - V + mu[:, None] * Pt
    if pde == LAP:
        Theta = Theta.real
    nq, ne = source.qtags.shape
    m = target.coords.shape[1]
    rows, cols, vals = [], [], []
    for e in range(ne):
        jg = source.qtags[:, e]
        L = b[jg].T  # nb x nq
        tg = jg if near is None else near[1][near[0][e]:near[0][e + 1]]
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
