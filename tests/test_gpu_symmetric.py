"""Symmetric evaluation of square operators (target is source): each
unordered pair evaluated once and applied in both directions. Bars: against
the oracle's direct sum, FP64 <= 1e-12 and MIXED <= 1e-6 (relative 2-norm);
against the plain (one direction per pair) GPU sum <= 1e-13; bit-identical
results on repeated applies; every work-group size (1, 2, 4 tiles of 256)
and ragged node counts (one node past a tile, whole tiles, fewer tiles than
a group)."""
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


def cloud(inti, dim, n, seed):
    """n points on a wavy closed curve / surface with unit normals and
    positive weights; qtags one node per element."""
    rng = np.random.default_rng(seed)
    if dim == 2:
        t = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = 1 + 0.2 * np.cos(3 * t)
        X = np.stack([r * np.cos(t), r * np.sin(t)])
    else:
        X = rng.normal(size=(3, n))
        X /= np.linalg.norm(X, axis=0)
        X *= 1 + 0.1 * X[0] * X[1]
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


def _density(pde, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    return x if pde == LAP else x + 1j * rng.normal(size=n)


@pytest.mark.parametrize("pde,dim,kind", list(itertools.product([LAP, HEL], [2, 3], range(6))))
def test_symmetric_matches_oracle(inti, oracle, pde, dim, kind):
    n = 1100 if dim == 2 else 1300
    say(f"[gpu] symmetric pde={pde} dim={dim} kind={KNAMES[kind]} n={n}")
    Q = cloud(inti, dim, n, 7 + dim)
    k = 0.0 if pde == LAP else (4.0 if dim == 2 else 3.0)
    K, coefs = _kernel(inti, pde, dim, kind, k)
    x = _density(pde, n, kind + 10 * dim)
    ref = oracle.matvec(pde, dim, kind, k, Q, Q, x, coefs=coefs)
    nref = np.linalg.norm(ref)
    plain = inti.IntegralOperator(K, Q, Q)
    plain.set_symmetric(0)
    y0 = plain @ x
    assert np.linalg.norm(y0 - ref) <= 1e-12 * nref
    for group in (1, 2, 4):
        A = inti.IntegralOperator(K, Q, Q)
        A.set_symmetric(2, group)
        y = A @ x
        assert np.linalg.norm(y - ref) <= 1e-12 * nref, group
        assert np.linalg.norm(y - y0) <= 1e-13 * nref, group
        assert np.array_equal(y, A @ x)
        Am = inti.IntegralOperator(K, Q, Q, precision="mixed")
        Am.set_symmetric(2, group)
        assert np.linalg.norm(Am @ x - ref) <= 1e-6 * nref, group


@pytest.mark.parametrize("n", [257, 511, 512, 513, 1024, 1281, 2305])
def test_symmetric_ragged_sizes(inti, oracle, n):
    say(f"[gpu] symmetric ragged n={n}")
    Q = cloud(inti, 3, n, n)
    op = inti.Helmholtz(k=2.5, dim=3)
    c0, c1 = 0.3 + 0.8j, -1.1 + 0.2j
    K = inti.CombinedFieldKernel(op, c0, c1, "SD")
    x = _density(HEL, n, n + 1)
    ref = oracle.matvec(HEL, 3, 4, 2.5, Q, Q, x, coefs=(c0, c1))
    for group in (1, 2, 4):
        A = inti.IntegralOperator(K, Q, Q)
        A.set_symmetric(2, group)
        assert np.linalg.norm(A @ x - ref) <= 1e-12 * np.linalg.norm(ref), (n, group)


def test_symmetric_with_correction_identity_alpha_beta(inti):
    """The corrected CFIE I/2 + D - i k S with its DIM correction: symmetric
    vs plain evaluation through the same fused epilogue, with alpha/beta."""
    say("[gpu] symmetric corrected CFIE")
    Q = inti.sphere_quadrature(6)
    op = inti.Helmholtz(k=2 * np.pi, dim=3)
    L, _ = inti.combined_field(op=op, quad=Q, c_single=-1j * op.k, c_double=1.0, c_identity=0.5)
    rng = np.random.default_rng(3)
    n = len(Q)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    y0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    L.iop.set_symmetric(0)
    ref = L.apply(x, y0, alpha=0.3 - 0.7j, beta=1.1 + 0.4j)
    L.iop.set_symmetric(2)
    got = L.apply(x, y0, alpha=0.3 - 0.7j, beta=1.1 + 0.4j)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)
    # multi-RHS applies keep the plain kernels and agree column by column
    X = np.stack([x, y0], axis=1)
    Y = L.apply(X)
    assert np.linalg.norm(Y[:, 0] - L.apply(x)) <= 1e-13 * np.linalg.norm(Y[:, 0])


@pytest.mark.parametrize("pde,dim,kind", [(LAP, 3, 0), (HEL, 3, 4), (HEL, 2, 1), (LAP, 2, 1)])
def test_symmetric_with_duplicate_nodes(inti, oracle, pde, dim, kind):
    """Nodes repeated in other tiles (coincident pairs across tiles): the
    quadrature is not 'distinct', the kernels keep the coincidence mask, and
    the symmetric sum still equals the oracle's (which zeroes those pairs)."""
    say(f"[gpu] symmetric with duplicate nodes pde={pde} dim={dim} kind={KNAMES[kind]}")
    base = cloud(inti, dim, 900, 21)
    dup = np.arange(0, 900, 7)
    X = np.concatenate([base.coords, base.coords[:, dup]], axis=1)
    N = np.concatenate([base.normals, base.normals[:, dup]], axis=1)
    w = np.concatenate([base.weights, base.weights[dup]])
    n = X.shape[1]
    Q = inti.Quadrature(X, N, w, np.arange(n)[None, :])
    k = 0.0 if pde == LAP else 2.0
    K, coefs = _kernel(inti, pde, dim, kind, k)
    x = _density(pde, n, 5)
    ref = oracle.matvec(pde, dim, kind, k, Q, Q, x, coefs=coefs)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        A = inti.IntegralOperator(K, Q, Q, precision=prec)
        A.set_symmetric(2)
        y = A @ x
        assert np.all(np.isfinite(y))
        assert np.linalg.norm(y - ref) <= tol * np.linalg.norm(ref), prec


@pytest.mark.parametrize("n", [700, 1300, 2305])
def test_symmetric_mixed_cfie_minus_ik(inti, oracle, n):
    """The combined field D - i k S (c0 = -i k c1, the examples' CFIE) in
    mixed precision: the symmetric kernel's staged ch q then shares the
    k q slot; against the oracle (1e-6) and the one-direction mixed sum."""
    say(f"[gpu] symmetric mixed CFIE c0 = -ik n={n}")
    Q = cloud(inti, 3, n, 50 + n)
    k = 3.5
    op = inti.Helmholtz(k=k, dim=3)
    for c1 in (1.0, 0.7 - 0.2j):
        c0 = -1j * k * c1
        K = inti.CombinedFieldKernel(op, c0, c1, "SD")
        x = _density(HEL, n, n + 3)
        ref = oracle.matvec(HEL, 3, 4, k, Q, Q, x, coefs=(c0, c1))
        A = inti.IntegralOperator(K, Q, Q, precision="mixed")
        A.set_symmetric(2)
        y = A @ x
        assert np.linalg.norm(y - ref) <= 1e-6 * np.linalg.norm(ref), c1
        assert np.array_equal(y, A @ x)
        B = inti.IntegralOperator(K, Q, Q, precision="mixed")
        B.set_symmetric(0)
        assert np.linalg.norm(y - B @ x) <= 1e-6 * np.linalg.norm(ref), c1
