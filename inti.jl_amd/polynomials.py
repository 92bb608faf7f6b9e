"""Monomial bases and polynomial particular solutions for the volume density
interpolation (Inti.jl src/vdim.jl:322-379).

For every multi-index I with |I| <= order, in the order of
Iterators.product(0:order, ..., 0:order) (first index fastest, vdim.jl:339-341),
the basis monomial is p_I(x) = x^I / I! and P_I is a polynomial with
    Laplace   (-Delta u):        Delta P = p,
    Helmholtz (-Delta u - k^2 u): (Delta + k^2) P = p,
so that mu(x) P(x) + S[dP/dn](x) - D[P](x) - V[p](x) = 0 for the Green
multiplier mu of x (Green's third identity, the sign convention of
_vdim_auxiliary_quantities, vdim.jl:260-288).

Particular solutions (homogeneous p of degree d in N dimensions):
    Laplace:   P = sum_j (-1)^j c_j |x|^(2j+2) Delta^j p,
               c_0 = 1/a_0, c_j = c_(j-1)/a_j, a_j = (2j+2)(N + 2d - 2j);
    Helmholtz: P = sum_j (-1)^j Delta^j p / k^(2j+2).
The Laplace one is the solution of the form |x|^2 q (q of degree d), which is
unique; the Helmholtz one is the only polynomial solution.
"""
import itertools
import math

import numpy as np

# a polynomial is a dict {exponent tuple: coefficient}


def _add(P, Q, s=1.0):
    R = dict(P)
    for e, c in Q.items():
        R[e] = R.get(e, 0.0) + s * c
    return {e: c for e, c in R.items() if c != 0.0}


def _scale(P, s):
    return {e: c * s for e, c in P.items()}


def laplacian(P):
    R = {}
    for e, c in P.items():
        for d in range(len(e)):
            if e[d] >= 2:
                f = list(e)
                f[d] -= 2
                f = tuple(f)
                R[f] = R.get(f, 0.0) + c * e[d] * (e[d] - 1)
    return {e: c for e, c in R.items() if c != 0.0}


def times_r2(P, power=1):
    """P * |x|^(2 power)."""
    N = len(next(iter(P))) if P else 0
    R = dict(P)
    for _ in range(power):
        S = {}
        for e, c in R.items():
            for d in range(N):
                f = list(e)
                f[d] += 2
                f = tuple(f)
                S[f] = S.get(f, 0.0) + c
        R = S
    return R


def gradient(P, N):
    G = []
    for d in range(N):
        R = {}
        for e, c in P.items():
            if e[d] >= 1:
                f = list(e)
                f[d] -= 1
                f = tuple(f)
                R[f] = R.get(f, 0.0) + c * e[d]
        G.append(R)
    return G


def evaluate(P, x):
    """Values of P at the points x (N, n)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros(x.shape[1])
    for e, c in P.items():
        t = np.full(x.shape[1], float(c))
        for d, k in enumerate(e):
            if k:
                t = t * x[d] ** k
        out = out + t
    return out


def monomial_multiindices(order, N):
    """Multi-indices of total degree <= order, first index fastest."""
    out = []
    for I in itertools.product(range(order + 1), repeat=N):
        I = tuple(reversed(I))  # itertools varies the last index fastest
        if sum(I) <= order:
            out.append(I)
    return out


def monomial(I):
    return {tuple(I): 1.0 / math.prod(math.factorial(i) for i in I)}


def laplace_solution(p):
    """Delta P = p for a homogeneous polynomial p."""
    if not p:
        return {}
    N = len(next(iter(p)))
    d = sum(next(iter(p)))
    P, h, c, j = {}, dict(p), 1.0, 0
    while h:
        a = (2 * j + 2) * (N + 2 * d - 2 * j)
        c = c / a
        P = _add(P, times_r2(h, j + 1), (-1) ** j * c)
        h = laplacian(h)
        j += 1
    return P


def helmholtz_solution(p, k):
    """(Delta + k^2) P = p."""
    P, h, j = {}, dict(p), 0
    while h:
        P = _add(P, h, (-1) ** j / k ** (2 * j + 2))
        h = laplacian(h)
        j += 1
    return P


def polynomial_solutions_vdim(op, order, center=None):
    """(monomials p, solutions P, multi-indices) of polynomial_solutions_vdim
    (vdim.jl:322-361) for a Laplace or Helmholtz operator."""
    N = op.dim
    idx = monomial_multiindices(order, N)
    ps, Ps = [], []
    for I in idx:
        p = monomial(I)
        P = laplace_solution(p) if op.k == 0 else helmholtz_solution(p, op.k)
        ps.append(p)
        Ps.append(P)
    return ps, Ps, idx


def evaluate_basis(polys, x, center=None):
    """Matrix [p(x_j)]: (n, nbasis) at points x (N, n), shifted by center."""
    x = np.asarray(x, dtype=np.float64)
    if center is not None:
        x = x - np.asarray(center, dtype=np.float64)[:, None]
    return np.stack([evaluate(p, x) for p in polys], axis=1)


def normal_derivative(polys, x, nrm, center=None):
    """[grad P(x_j) . n_j]: (n, nbasis)."""
    x = np.asarray(x, dtype=np.float64)
    if center is not None:
        x = x - np.asarray(center, dtype=np.float64)[:, None]
    N = x.shape[0]
    cols = []
    for P in polys:
        G = gradient(P, N)
        cols.append(sum(evaluate(G[d], x) * nrm[d] for d in range(N)))
    return np.stack(cols, axis=1)
