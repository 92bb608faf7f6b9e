"""Sixteen right-hand sides: the direct sum whose products run on the FP64
matrix cores. Bars against the oracle's direct sum: FP64 <= 1e-12 and MIXED
<= 1e-6 (relative 2-norm per column), for every (PDE, dim, kind) and the
far-field kinds, with target sets other than the sources, node counts that
are not multiples of a tile, several source chunks, and against the
one-column apply of the same operator (1e-13); repeated applies are
bit-identical."""
import itertools
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAP, HEL = 0, 1
KNAMES = ["SL", "DL", "ADL", "HS", "COMB_SD", "COMB_AH"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cloud(inti, dim, n, seed, shift=0.0):
    """n points on a wavy closed curve / surface with unit normals and
    positive weights; qtags one node per element."""
    rng = np.random.default_rng(seed)
    if dim == 2:
        t = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = 1 + 0.2 * np.cos(3 * t) + shift
        X = np.stack([r * np.cos(t), r * np.sin(t)])
    else:
        X = rng.normal(size=(3, n))
        X /= np.linalg.norm(X, axis=0)
        X *= 1 + 0.1 * X[0] * X[1] + shift
    N = rng.normal(size=(dim, n))
    N /= np.linalg.norm(N, axis=0)
    w = rng.uniform(0.5, 1.5, n) / n
    return inti.Quadrature(X, N, w, np.arange(n)[None, :])


def _kernel(inti, pde, dim, kind, k):
    op = inti.Laplace(dim=dim) if pde == LAP else inti.Helmholtz(k=k, dim=dim)
    if kind in (4, 5):
        c0, c1 = (0.7, -1.3) if pde == LAP else (0.4 - 1.1j, 0.9 + 0.3j)
        return inti.CombinedFieldKernel(op, c0, c1, "SD" if kind == 4 else "AH"), (c0, c1)
    cls = [inti.SingleLayerKernel, inti.DoubleLayerKernel, inti.AdjointDoubleLayerKernel,
           inti.HyperSingularKernel][kind]
    return cls(op), None


def _dens(pde, n, r, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, r))
    return X if pde == LAP else X + 1j * rng.normal(size=(n, r))


def _colerr(Y, ref):
    return (np.linalg.norm(Y - ref, axis=0) / np.linalg.norm(ref, axis=0)).max()


@pytest.mark.parametrize("pde,dim,kind", list(itertools.product([LAP, HEL], [2, 3], range(6))))
def test_sixteen_rhs_matches_oracle(inti, oracle, pde, dim, kind):
    m, n = 1000, 1300  # neither a multiple of 64 nor of 256
    say(f"[gpu] 16 rhs pde={pde} dim={dim} kind={KNAMES[kind]} m={m} n={n}")
    src = cloud(inti, dim, n, 3 + dim)
    trg = cloud(inti, dim, m, 5 + dim, shift=0.05)
    k = 0.0 if pde == LAP else (4.0 if dim == 2 else 3.0)
    K, coefs = _kernel(inti, pde, dim, kind, k)
    X = _dens(pde, n, 16, kind + 10 * dim)
    ref = oracle.matvec(pde, dim, kind, k, trg, src, X, coefs=coefs)
    A = inti.IntegralOperator(K, trg, src)
    Y = A.apply(X)
    assert _colerr(Y, ref) <= 1e-12
    assert np.array_equal(Y, A.apply(X))
    cols = np.stack([A.apply(X[:, r]) for r in (0, 7, 15)], axis=1)
    assert _colerr(Y[:, [0, 7, 15]], cols) <= 1e-13
    Am = inti.IntegralOperator(K, trg, src, precision="mixed")
    assert _colerr(Am.apply(X), ref) <= 1e-6


@pytest.mark.parametrize("dim,kind", [(2, 6), (2, 7), (3, 6), (3, 7), (3, 8)])
def test_sixteen_rhs_far_field(inti, oracle, dim, kind):
    """Far-field patterns of 16 densities at 200 directions."""
    say(f"[gpu] 16 rhs far field dim={dim} kind={kind}")
    src = cloud(inti, dim, 900, 11)
    rng = np.random.default_rng(12)
    d = rng.normal(size=(dim, 200))
    d /= np.linalg.norm(d, axis=0)
    op = inti.Helmholtz(k=5.0, dim=dim)
    coefs = None
    if kind == 6:
        K = inti.FarFieldSingleLayerKernel(op)
    elif kind == 7:
        K = inti.FarFieldDoubleLayerKernel(op)
    else:
        coefs = (0.3 - 0.9j, 1.1 + 0.2j)
        K = inti.CombinedFieldKernel(op, coefs[0], coefs[1], "FF")
    X = _dens(HEL, 900, 16, 13)
    ref = oracle.matvec(HEL, dim, kind, 5.0, inti.points(d), src, X, coefs=coefs)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        Y = inti.IntegralOperator(K, inti.points(d), src, precision=prec).apply(X)
        assert _colerr(Y, ref) <= tol, prec


@pytest.mark.parametrize("pde,dim,kind", [(HEL, 3, 4), (HEL, 2, 1), (LAP, 3, 0), (LAP, 2, 3)])
def test_sixteen_rhs_source_chunks(inti, oracle, pde, dim, kind):
    """The source range cut into 1, 3 and 8 chunks (partial sums added in
    the epilogue) with a ragged last tile; 17 and 33 columns (blocks of 16
    plus one)."""
    n = 2305
    say(f"[gpu] 16 rhs chunks pde={pde} dim={dim} kind={KNAMES[kind]} n={n}")
    Q = cloud(inti, dim, n, 21)
    k = 0.0 if pde == LAP else 2.5
    K, coefs = _kernel(inti, pde, dim, kind, k)
    X = _dens(pde, n, 33, 22)
    ref = oracle.matvec(pde, dim, kind, k, Q, Q, X, coefs=coefs)
    A = inti.IntegralOperator(K, Q, Q)
    for split in (1, 3, 8):
        A.ctx.check(A.ctx.lib.inti_op_set_launch(A.handle, 256, 1, split))
        assert _colerr(A.apply(X[:, :17]), ref[:, :17]) <= 1e-12, split
        assert _colerr(A.apply(X), ref) <= 1e-12, split


@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (17, 64), (64, 5), (300, 4)])
def test_sixteen_rhs_tiny_sizes(inti, oracle, m, n):
    """Fewer targets than a fragment and fewer sources than a step of 4:
    clamped target rows, padded sources, 16 and 19 columns; and the
    sigma-only pullback (transposed one-direction sum) at the same sizes."""
    say(f"[gpu] 16 rhs tiny m={m} n={n}")
    src = cloud(inti, 3, n, 31 + n)
    trg = cloud(inti, 3, m, 41 + m, shift=0.07)
    K, coefs = _kernel(inti, HEL, 3, 4, 2.0)
    X = _dens(HEL, n, 19, 5)
    ref = oracle.matvec(HEL, 3, 4, 2.0, trg, src, X, coefs=coefs)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        A = inti.IntegralOperator(K, trg, src, precision=prec)
        assert _colerr(A.apply(X[:, :16]), ref[:, :16]) <= tol, prec
        assert _colerr(A.apply(X), ref) <= tol, prec
    u = np.resize(_dens(HEL, max(m, n), 1, 6)[:, 0], m)
    g = inti.IntegralOperator(K, trg, src).pullback(u, X[:, 0])["sigma"]
    o = oracle.adjoint(HEL, 3, 4, 2.0, trg, src, u, X[:, 0], coefs=coefs, dcoefs=(0.0, 0.0), want_xbar=False)
    assert np.linalg.norm(g - o["sigma_bar"]) <= 1e-12 * np.linalg.norm(o["sigma_bar"])
