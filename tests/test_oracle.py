"""Pins of the CPU oracle (no GPU): mpmath goldens, closed forms, derivative
identities between the kernels, an independent least-squares solve, and the
Green-identity / DIM thresholds of the reference's own tests."""
import numpy as np
import pytest

LAP, HEL = 0, 1
SL, DL, ADL, HS, COMB_SD, COMB_AH = range(6)


def test_hankel_matches_mpmath(oracle, golden):
    x = golden["hankel_x"]
    for n in (0, 1, 2):
        ref = golden[f"hankel_h{n}"]
        got = np.array([oracle.hankel1(n, v) for v in x])
        rel = np.abs(got - ref) / np.abs(ref)
        assert rel.max() < 1e-13, (n, x[rel.argmax()], rel.max())


def _pt(rng, dim):
    v = rng.normal(size=dim)
    return v


def _unit(rng, dim):
    v = rng.normal(size=dim)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("dim", [2, 3])
def test_single_layer_closed_forms(oracle, golden, dim):
    rng = np.random.default_rng(1)
    for _ in range(5):
        x, y = _pt(rng, dim), _pt(rng, dim)
        d = np.linalg.norm(x - y)
        lap = oracle.kernel(LAP, dim, SL, 0.0, x, None, y, None)
        ref = -np.log(d) / (2 * np.pi) if dim == 2 else 1 / (4 * np.pi * d)
        assert abs(lap - ref) <= 1e-15 * max(1, abs(ref))
        k = 2.3
        hel = oracle.kernel(HEL, dim, SL, k, x, None, y, None)
        if dim == 3:
            ref = np.exp(1j * k * d) / (4 * np.pi * d)
        else:
            import mpmath as mp
            ref = 0.25j * complex(mp.besselj(0, k * d) + 1j * mp.bessely(0, k * d))
        assert abs(hel - ref) <= 1e-14 * abs(ref)


def _fd(f, h=1e-5):
    return (f(h) - f(-h)) / (2 * h)


@pytest.mark.parametrize("pde,dim,k", [(LAP, 2, 0.0), (LAP, 3, 0.0), (HEL, 2, 1.7), (HEL, 3, 1.7),
                                       (HEL, 2, 12.0), (HEL, 3, 9.0)])
def test_kernel_derivative_identities(oracle, pde, dim, k):
    """DL = dG/dn_y, ADL = dG/dn_x, HS = d2G/dn_x dn_y (kernels.jl:59-112)."""
    rng = np.random.default_rng(7)
    for _ in range(6):
        x, y = _pt(rng, dim), _pt(rng, dim)
        nx, ny = _unit(rng, dim), _unit(rng, dim)
        G = lambda a, b: oracle.kernel(pde, dim, SL, k, a, None, b, None)  # noqa: E731
        dl = oracle.kernel(pde, dim, DL, k, x, nx, y, ny)
        adl = oracle.kernel(pde, dim, ADL, k, x, nx, y, ny)
        hs = oracle.kernel(pde, dim, HS, k, x, nx, y, ny)
        dl_fd = _fd(lambda h: G(x, y + h * ny))
        adl_fd = _fd(lambda h: G(x + h * nx, y))
        hs_fd = _fd(lambda h: oracle.kernel(pde, dim, DL, k, x + h * nx, nx, y, ny))
        scale = max(abs(dl), abs(adl), abs(hs), 1e-3)
        assert abs(dl - dl_fd) < 1e-7 * scale
        assert abs(adl - adl_fd) < 1e-7 * scale
        assert abs(hs - hs_fd) < 1e-7 * scale


def test_combined_kinds_and_same_point(oracle):
    rng = np.random.default_rng(3)
    c0, c1 = 0.3 - 1.1j, -0.7 + 0.2j
    for dim in (2, 3):
        x, y = _pt(rng, dim), _pt(rng, dim)
        nx, ny = _unit(rng, dim), _unit(rng, dim)
        for pde, k in ((HEL, 2.0), (LAP, 0.0)):
            cc = (c0, c1) if pde == HEL else (c0.real, c1.real)
            sd = oracle.kernel(pde, dim, COMB_SD, k, x, nx, y, ny, coefs=cc)
            ah = oracle.kernel(pde, dim, COMB_AH, k, x, nx, y, ny, coefs=cc)
            ref_sd = cc[0] * oracle.kernel(pde, dim, SL, k, x, nx, y, ny) + cc[1] * oracle.kernel(pde, dim, DL, k, x, nx, y, ny)
            ref_ah = cc[0] * oracle.kernel(pde, dim, ADL, k, x, nx, y, ny) + cc[1] * oracle.kernel(pde, dim, HS, k, x, nx, y, ny)
            assert abs(sd - ref_sd) <= 1e-15 * abs(ref_sd)
            assert abs(ah - ref_ah) <= 1e-15 * abs(ref_ah)
            for kind in range(6):
                assert oracle.kernel(pde, dim, kind, k, x, nx, x, ny, coefs=cc) == 0


def test_assemble_and_matvec_consistent(oracle, inti):
    Q = inti.circle_quadrature(12, qorder=5)
    A = oracle.assemble(HEL, 2, DL, 1.3, Q, Q)
    i, j = 4, 17
    ref = oracle.kernel(HEL, 2, DL, 1.3, Q.coords[:, i], Q.normals[:, i], Q.coords[:, j],
                        Q.normals[:, j]) * Q.weights[j]
    assert A[i, j] == ref
    rng = np.random.default_rng(0)
    s = rng.normal(size=len(Q)) + 1j * rng.normal(size=len(Q))
    y = oracle.matvec(HEL, 2, DL, 1.3, Q, Q, s)
    assert np.linalg.norm(y - A @ s) <= 1e-14 * np.linalg.norm(y)


def test_laplace_dl_on_circle_closed_form(oracle):
    """Points on a circle of radius a: (r.n_y)/(2 pi d^2) = -1/(4 pi a)."""
    a = 1.7
    th = np.linspace(0, 2 * np.pi, 40, endpoint=False)
    P = np.stack([a * np.cos(th), a * np.sin(th)])

    class Q:
        coords = P
        normals = P / a
        weights = np.ones(40)
    A = oracle.assemble(LAP, 2, DL, 0.0, Q, Q).real
    off = A[~np.eye(40, dtype=bool)]
    assert np.allclose(off, -1 / (4 * np.pi * a), rtol=1e-13, atol=0)
    assert np.all(np.diag(A) == 0)


def test_monopoles(oracle):
    pts = np.array([[0.0, 2.0, 1.0], [-1.0, 1.0, 3.0]])
    z = oracle.monopoles(2, pts, 5)
    assert z.shape == (2, 15)
    c = np.array([1.0, 1.0])
    R = 1.5 * np.hypot(2, 4) / 2
    assert np.allclose(np.linalg.norm(z - c[:, None], axis=0), R, rtol=1e-15)
    assert np.allclose(z[:, 0], c + [R, 0])
    p3 = np.array([[0.0, 2.0], [0.0, 2.0], [0.0, 2.0]])
    z3 = oracle.monopoles(3, p3, 6)
    assert z3.shape == (3, 18)
    assert np.allclose(np.linalg.norm(z3 - 1.0, axis=0), 1.5 * np.sqrt(12) / 2, rtol=1e-14)
    assert np.isclose(z3[1, 0], 1 + 1.5 * np.sqrt(12) / 2) and np.isclose(z3[1, -1], 1 - 1.5 * np.sqrt(12) / 2)


def _bdim_pieces(oracle, pde, dim, k, Q, mu, derivative=False):
    """A = M^T per element and Theta rows, rebuilt from kernel values
    (bdim.jl:94-118)."""
    nq = Q.qtags.shape[0]
    zs = oracle.monopoles(dim, Q.coords, nq)
    ns = zs.shape[1]
    n = len(Q.weights)
    B0 = np.array([[oracle.kernel(pde, dim, SL, k, Q.coords[:, j], Q.normals[:, j], zs[:, q], None)
                    for q in range(ns)] for j in range(n)])
    B1 = np.array([[oracle.kernel(pde, dim, ADL, k, Q.coords[:, j], Q.normals[:, j], zs[:, q], None)
                    for q in range(ns)] for j in range(n)])
    kG, kS, kD = (ADL, ADL, HS) if derivative else (SL, SL, DL)
    G = np.array([[oracle.kernel(pde, dim, kG, k, Q.coords[:, i], Q.normals[:, i], zs[:, q], None)
                   for q in range(ns)] for i in range(n)])
    Th = mu[:, None] * G + oracle.matvec(pde, dim, kS, k, Q, Q, B1) - oracle.matvec(pde, dim, kD, k, Q, Q, B0)
    return B0, B1, Th


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
def test_bdim_weights_solve_the_least_squares_problem(oracle, inti, pde, k):
    """W from the oracle's Householder QR equals numpy's lstsq of the same
    (M^T, Theta_i) (bdim.jl:159-182), and dD/dS take W[:nq] and -W[nq:]."""
    Q = inti.circle_quadrature(10, qorder=5)
    mu = np.full(len(Q), -0.5)
    rowptr, colind, vS, vD, diag = oracle.bdim(pde, 2, k, Q, Q, mu)
    B0, B1, Th = _bdim_pieces(oracle, pde, 2, k, Q, mu)
    nq = Q.qtags.shape[0]
    for e in (0, 3, 9):
        jg = Q.qtags[:, e]
        A = np.concatenate([B0[jg, :], B1[jg, :]], axis=0).T
        for i in jg:
            W = np.linalg.lstsq(A, Th[i], rcond=None)[0]
            cols = colind[rowptr[i]:rowptr[i + 1]]
            assert np.array_equal(cols, np.sort(jg))
            d = vD[rowptr[i]:rowptr[i + 1]]
            s = vS[rowptr[i]:rowptr[i + 1]]
            scale = np.abs(W).max()
            assert np.abs(d - W[:nq]).max() < 1e-7 * scale
            assert np.abs(s + W[nq:]).max() < 1e-7 * scale
    assert diag[0] > 1


def test_bdim_self_pattern(oracle, inti):
    Q = inti.sphere_quadrature(1, base="octa")
    mu = np.full(len(Q), -0.5)
    rowptr, colind, vS, vD, _ = oracle.bdim(LAP, 3, 0.0, Q, Q, mu)
    nq, ne = Q.qtags.shape
    assert len(colind) == nq * nq * ne
    for e in range(ne):
        for i in Q.qtags[:, e]:
            assert np.array_equal(colind[rowptr[i]:rowptr[i + 1]], np.sort(Q.qtags[:, e]))


def _green_errors(oracle, pde, dim, k, Q, xs, sigma, derivative=False):
    n = len(Q.weights)
    src = np.asarray(xs, dtype=float)
    u = np.array([oracle.kernel(pde, dim, SL, k, Q.coords[:, j], Q.normals[:, j], src, None) for j in range(n)])
    dudn = np.array([oracle.kernel(pde, dim, ADL, k, Q.coords[:, j], Q.normals[:, j], src, None) for j in range(n)])
    kS, kD = (ADL, HS) if derivative else (SL, DL)
    S = oracle.assemble(pde, dim, kS, k, Q, Q)
    D = oracle.assemble(pde, dim, kD, k, Q, Q)
    mu = np.full(n, -0.5)
    rowptr, colind, vS, vD, _ = oracle.bdim(pde, dim, k, Q, Q, mu, derivative=derivative)
    import scipy.sparse as sp
    dS = sp.csr_matrix((vS, colind, rowptr), shape=(n, n)).toarray()
    dD = sp.csr_matrix((vD, colind, rowptr), shape=(n, n)).toarray()
    ref = dudn if derivative else u
    nrm = np.abs(ref).max()
    e0 = np.abs(S @ dudn - D @ u - sigma * ref).max() / nrm
    e1 = np.abs((S + dS) @ dudn - (D + dD) @ u - sigma * ref).max() / nrm
    return e0, e1


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 1.2)])
@pytest.mark.parametrize("side", ["interior", "exterior"])
def test_green_identity_2d(oracle, inti, pde, k, side):
    """green_identities_test.jl:93-126: e1 < 1e-2 (5e-2 for ADL/HS) and e1 < e0."""
    Q = inti.circle_quadrature(13, qorder=5)
    sigma = 0.5 if side == "interior" else -0.5
    xs = (3.0, 3.0) if side == "interior" else (0.1, 0.1)
    e0, e1 = _green_errors(oracle, pde, 2, k, Q, xs, sigma)
    assert e1 < 1e-2 and e1 < e0
    e0, e1 = _green_errors(oracle, pde, 2, k, Q, xs, sigma, derivative=True)
    assert e1 < 5e-2 and e1 < e0


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 1.2)])
def test_green_identity_3d(oracle, inti, pde, k):
    Q = inti.sphere_quadrature(2, base="ico")
    e0, e1 = _green_errors(oracle, pde, 3, k, Q, (3.0, 3.0, 3.0), 0.5)
    assert e1 < 1e-2 and e1 < e0


def test_dim_offsurface_targets(oracle, inti):
    """dim_test.jl:42-90: with targets off the circle, D+ 1 ~ 0 outside,
    D- 1 ~ -1 inside (tolerance 1e-3)."""
    Q = inti.circle_quadrature(32, qorder=3)
    g = np.linspace(-1.5, 1.5, 13)
    P = np.array([[a, b] for a in g for b in g]).T
    r = np.linalg.norm(P, axis=0)
    for inside in (True, False):
        idx = np.where((r < 0.97) if inside else (r > 1.03))[0]

        class T:
            coords = np.ascontiguousarray(P[:, idx])
            normals = None
            weights = np.ones(len(idx))
        near = oracle.nearest_lists(T, Q)
        mu = np.full(len(idx), -1.0 if inside else 0.0)
        rowptr, colind, vS, vD, _ = oracle.bdim(LAP, 2, 0.0, T, Q, mu, near=near)
        ones = np.ones(len(Q.weights))
        v = oracle.matvec(LAP, 2, DL, 0.0, T, Q, ones) + oracle.csr_matvec(len(idx), len(Q.weights), rowptr, colind, vD, ones)
        target = -1.0 if inside else 0.0
        assert np.abs(v - target).max() < 1e-3


def _mixed_curve(inti, split=6, q=(3, 8), nel=16):
    """An ellipse whose first `split` P2 elements carry one Gauss rule and the
    rest another: two element types with different nq."""
    E = inti.ellipse_quadrature(1.0, 0.6, nel).elements
    return inti.join_quadratures(inti.curve_quadrature(E[:split], q[0]), inti.curve_quadrature(E[split:], q[1]))


def _mixed_sphere(inti, split=30, q=(2, 4), freq=2):
    E = inti.sphere_quadrature(freq, base="ico").elements
    return inti.join_quadratures(inti.surface_quadrature(E[:split], q[0], center=(0, 0, 0)),
                                 inti.surface_quadrature(E[split:], q[1], center=(0, 0, 0)))


@pytest.mark.parametrize("dim", [2, 3])
def test_bdim_types_split_reproduces_one_type(oracle, inti, dim):
    """bdim.jl:79-80 counts monopoles over the sum of nq of the types. Cutting
    a one-type quadrature's elements into two types of the same nq, with the
    oversample factor halved, gives the same monopoles and the same
    per-element problems, so the correction must be bit-identical; a wrong
    per-type offset into the tags, near lists or triplets breaks it."""
    if dim == 2:
        E = inti.ellipse_quadrature(1.0, 0.6, 20).elements
        one = inti.curve_quadrature(E, 5)
        two = inti.join_quadratures(inti.curve_quadrature(E[:7], 5), inti.curve_quadrature(E[7:], 5))
        pde, k = HEL, 2.0
    else:
        E = inti.sphere_quadrature(1, base="ico").elements
        one = inti.surface_quadrature(E, 2, center=(0, 0, 0))
        two = inti.join_quadratures(inti.surface_quadrature(E[:9], 2, center=(0, 0, 0)),
                                    inti.surface_quadrature(E[9:], 2, center=(0, 0, 0)))
        pde, k = LAP, 0.0
    assert np.array_equal(one.coords, two.coords) and len(two.etype_qtags) == 2
    mu = np.full(len(one), -0.5)
    a = oracle.bdim(pde, dim, k, one, one, mu, oversample=3.0)
    b = oracle.bdim(pde, dim, k, two, two, mu, oversample=1.5)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    # off-surface targets with one near list per type
    T = inti.points(one.coords[:, ::3] * 1.02)
    mu = np.zeros(T.coords.shape[1])
    near1, near2 = oracle.nearest_lists(T, one), oracle.nearest_lists(T, two)
    assert len(near2) == 2
    a = oracle.bdim(pde, dim, k, T, one, mu, near=near1, oversample=3.0)
    b = oracle.bdim(pde, dim, k, T, two, mu, near=near2, oversample=1.5)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("dim", [2, 3])
def test_bdim_types_green_identity(oracle, inti, dim):
    """Green's identity (green_identities_test.jl:93-126 bar: e1 < 1e-2 and
    e1 < e0) on a boundary with two element types of different nq."""
    Q = _mixed_curve(inti) if dim == 2 else _mixed_sphere(inti)
    assert len({q.shape[0] for q in Q.etype_qtags}) == 2
    xs = (3.0, 3.0) if dim == 2 else (3.0, 3.0, 3.0)
    for pde, k in ((LAP, 0.0), (HEL, 1.2)):
        e0, e1 = _green_errors(oracle, pde, dim, k, Q, xs, 0.5)
        assert e1 < 1e-2 and e1 < e0 / 10


def test_near_lists_per_type_match_product(oracle, inti):
    """The product's etype_to_nearest_points and the oracle's brute force give
    the same per-type lists (targets placed away from nearest-node ties)."""
    Q = _mixed_curve(inti)
    rng = np.random.default_rng(4)
    th = rng.uniform(0, 2 * np.pi, 60)
    T = inti.points(np.stack([1.1 * np.cos(th), 0.7 * np.sin(th)]))
    a = inti.etype_to_nearest_points(T, Q)
    b = oracle.nearest_lists(T, Q)
    assert len(a) == len(b) == 2
    for (p, i), (q, j) in zip(a, b):
        assert np.array_equal(p, q) and np.array_equal(i, j)
    assert sum(len(i) for _, i in a) == 60
