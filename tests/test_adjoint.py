"""Reverse mode of the direct sum (y = K sigma) and the Helmholtz far-field
kernels. No reference counterpart (the reference differentiates with
ForwardDiff, which is not available here): the oracle's closed-form
derivatives are pinned by central finite differences of the oracle's own
forward sums, the far-field kernels by the large-R limit of the potentials,
and the GPU pullback is compared with the oracle (1e-12) and with finite
differences of the GPU forward."""
import numpy as np
import pytest

LAP, HEL = 0, 1
SL, DL, COMB, FF_SL, FF_DL, FF_COMB = 0, 1, 4, 6, 7, 8


class Pts:
    def __init__(self, coords, normals=None, weights=None):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)


def _setup(dim, m, n, seed, farfield=False):
    rng = np.random.default_rng(seed)
    y = rng.normal(size=(dim, n))
    nv = rng.normal(size=(dim, n))
    nv /= np.linalg.norm(nv, axis=0)
    w = rng.uniform(0.5, 1.5, size=n)
    if farfield:
        x = rng.normal(size=(dim, m))
        x /= np.linalg.norm(x, axis=0)
    else:
        x = rng.normal(size=(dim, m)) + 3.0  # away from the sources
    u = rng.normal(size=m) + 1j * rng.normal(size=m)
    s = rng.normal(size=n) + 1j * rng.normal(size=n)
    return x, y, nv, w, u, s


CASES = [(LAP, 2, SL), (LAP, 2, DL), (LAP, 3, SL), (LAP, 3, DL), (LAP, 3, COMB),
         (HEL, 2, SL), (HEL, 2, DL), (HEL, 2, COMB), (HEL, 3, SL), (HEL, 3, DL), (HEL, 3, COMB),
         (HEL, 2, FF_SL), (HEL, 2, FF_DL), (HEL, 3, FF_SL), (HEL, 3, FF_DL), (HEL, 3, FF_COMB)]


def _coefs(pde, kind, k):
    """((c0, c1), (dc0/dk, dc1/dk)): c0 = -i k (CFIE) for Helmholtz, 0.7 for
    Laplace; c1 = 1."""
    if kind not in (COMB, FF_COMB):
        return None, None
    if pde == LAP:
        return (0.7, 1.0), (0.0, 0.0)
    return (-1j * k, 1.0), (-1j, 0.0)


@pytest.mark.parametrize("pde,dim,kind", CASES)
def test_oracle_adjoint_matches_finite_differences(oracle, pde, dim, kind):
    """d/dt Re<ubar, y(theta + t v)> at t = 0 by fourth-order central
    differences (h = 1e-3 relative) against <grad, v> from the oracle's
    closed-form pullback, for theta = sigma, source nodes / normals /
    weights, target nodes and k."""
    ff = kind >= FF_SL
    k = 2.3 if pde == HEL else 0.0
    x, y, nv, w, u, s = _setup(dim, 7, 11, 100 * pde + 10 * dim + kind, farfield=ff)
    if pde == LAP:
        u, s = u.real.astype(complex), s.real.astype(complex)
    co, dco = _coefs(pde, kind, k)

    def J(x_, y_, n_, w_, k_, s_):
        c = co
        if co is not None and pde == HEL:
            c = (-1j * k_, 1.0)
        yv = oracle.matvec(pde, dim, kind, k_, Pts(x_), Pts(y_, n_, w_), s_, coefs=c)
        return np.real(np.vdot(u, yv))

    g = oracle.adjoint(pde, dim, kind, k, Pts(x), Pts(y, nv, w), u, s, coefs=co, dcoefs=dco,
                       want_xbar=not ff)
    rng = np.random.default_rng(7)
    base = (x, y, nv, w, k, s)
    checks = [(1, "ybar"), (2, "nbar"), (3, "wbar")] + ([] if ff else [(0, "xbar")])
    if pde == HEL:
        checks.append((4, "kbar"))
    checks.append((5, "sigma_bar"))
    for idx, name in checks:
        v = rng.normal(size=np.shape(base[idx])) if idx != 4 else 1.0
        if idx == 5:
            v = v + (1j * rng.normal(size=v.shape) if pde == HEL else 0)
            pred = np.real(np.vdot(g["sigma_bar"], v))
        elif idx == 4:
            pred = g["kbar"]
        else:
            pred = np.sum(g[name] * v)
        h = 1e-3
        vals = []
        for t in (-2 * h, -h, h, 2 * h):
            args = list(base)
            args[idx] = base[idx] + t * v
            vals.append(J(*args))
        fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
        scale = max(abs(fd), abs(pred), 1e-3 * abs(J(*base)), 1e-12)
        assert abs(fd - pred) <= 1e-7 * scale, (name, fd, pred)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_farfield_is_the_limit_of_the_potential(oracle, dim):
    """u(R xh) R^{(dim-1)/2} e^{-ikR} -> u_inf(xh) as R grows (error O(1/R)),
    for the single- and double-layer potentials of a random source set."""
    k = 3.0
    x, y, nv, w, _, s = _setup(dim, 5, 20, 42 + dim, farfield=True)
    y = 0.3 * y
    for kind, ffkind in ((SL, FF_SL), (DL, FF_DL)):
        uinf = oracle.matvec(HEL, dim, ffkind, k, Pts(x), Pts(y, nv, w), s)
        errs = []
        for R in (1e3, 1e4):
            u = oracle.matvec(HEL, dim, kind, k, Pts(R * x), Pts(y, nv, w), s)
            errs.append(np.abs(u * R ** ((dim - 1) / 2) * np.exp(-1j * k * R) - uinf).max())
        assert errs[1] < 2e-3 * np.abs(uinf).max()
        assert 5 < errs[0] / errs[1] < 20  # first order in 1/R


# ---------------------------------------------------------------- GPU ----

def _kernel(inti, pde, dim, kind, k):
    op = inti.Laplace(dim=dim) if pde == LAP else inti.Helmholtz(k=k, dim=dim)
    co, dco = _coefs(pde, kind, k)
    if kind == SL:
        return op, inti.SingleLayerKernel(op), None
    if kind == DL:
        return op, inti.DoubleLayerKernel(op), None
    if kind == FF_SL:
        return op, inti.FarFieldSingleLayerKernel(op), None
    if kind == FF_DL:
        return op, inti.FarFieldDoubleLayerKernel(op), None
    pair = "SD" if kind == COMB else "FF"
    return op, inti.CombinedFieldKernel(op, co[0], co[1], pair), dco


def _quad(inti, y, nv, w):
    return inti.Quadrature(y, nv, w, np.arange(y.shape[1], dtype=np.int64).reshape(1, -1))


@pytest.mark.gpu
@pytest.mark.parametrize("pde,dim,kind", CASES)
def test_gpu_pullback_matches_oracle(inti, oracle, pde, dim, kind):
    """Every gradient of the GPU pullback against the oracle's, 1e-12
    relative (norm-wise per output; the scalar k gradient 1e-11)."""
    print(f"pullback pde={pde} dim={dim} kind={kind}", flush=True)
    ff = kind >= FF_SL
    k = 2.3 if pde == HEL else 0.0
    x, y, nv, w, u, s = _setup(dim, 300, 700, 7 + kind, farfield=ff)
    if pde == LAP:
        u, s = u.real, s.real
    op, K, dco = _kernel(inti, pde, dim, kind, k)
    A = inti.IntegralOperator(K, inti.points(x), _quad(inti, y, nv, w))
    wrt = ["sigma", "source_nodes", "source_normals", "source_weights"] + ([] if ff else ["target_nodes"])
    if pde == HEL:
        wrt.append("k")
    g = A.pullback(u, s, wrt=tuple(wrt), dcoefs=dco)
    co, _ = _coefs(pde, kind, k)
    o = oracle.adjoint(pde, dim, kind, k, Pts(x), Pts(y, nv, w), u, s, coefs=co, dcoefs=dco,
                       want_xbar=not ff)
    pairs = {"sigma": "sigma_bar", "source_nodes": "ybar", "source_normals": "nbar",
             "source_weights": "wbar", "target_nodes": "xbar"}
    for a, b in pairs.items():
        if a not in g:
            continue
        ref = o[b] if a != "sigma" or pde == HEL else o[b].real
        assert np.linalg.norm(g[a] - ref) <= 1e-12 * np.linalg.norm(ref), a
    if pde == HEL:
        assert abs(g["k"] - o["kbar"]) <= 1e-11 * abs(o["kbar"])


@pytest.mark.gpu
@pytest.mark.parametrize("pde,dim,kind", CASES)
@pytest.mark.parametrize("m,n", [(300, 700), (1300, 257)])
def test_gpu_sigma_only_pullback_one_direction(inti, oracle, pde, dim, kind, m, n):
    """sigma_bar alone of an operator whose targets are not its sources (and
    of the far fields): the transposed one-direction sum (kernel values only,
    one thread per source) against the oracle, 1e-12; bit-identical on
    repeat; ragged tiles on both sides."""
    print(f"sigma-only pullback one direction pde={pde} dim={dim} kind={kind} m={m} n={n}", flush=True)
    ff = kind >= FF_SL
    k = 2.3 if pde == HEL else 0.0
    x, y, nv, w, u, s = _setup(dim, m, n, 17 + kind + m, farfield=ff)
    if pde == LAP:
        u, s = u.real, s.real
    op, K, dco = _kernel(inti, pde, dim, kind, k)
    A = inti.IntegralOperator(K, inti.points(x), _quad(inti, y, nv, w))
    g = A.pullback(u, s)["sigma"]
    assert np.array_equal(g, A.pullback(u, s)["sigma"])
    co, _ = _coefs(pde, kind, k)
    o = oracle.adjoint(pde, dim, kind, k, Pts(x), Pts(y, nv, w), u, s, coefs=co, dcoefs=dco, want_xbar=False)
    ref = o["sigma_bar"] if pde == HEL else o["sigma_bar"].real
    assert np.linalg.norm(g - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_gpu_farfield_objective_gradient_by_finite_differences(inti):
    """The shape-gradient objective at a small size: sound-soft-like density
    sigma on a sphere perturbed by 0.1 sin(3 theta) cos(2 phi), k = 10,
    J = sum over 64 Fibonacci directions of |u_inf|^2 for the far field of
    D[sigma] - i k S[sigma]. Its gradient in the density, node coordinates,
    normals, weights and k (ubar = 2 u_inf) against central differences
    (h = 1e-6) of the GPU forward along 20 seeded random directions."""
    print("far-field objective FD", flush=True)
    k = 10.0
    Q = inti.sphere_quadrature(8, radial=lambda th, ph: 1.0 + 0.1 * np.sin(3 * th) * np.cos(2 * ph))
    n = len(Q)
    nd = 64
    i = np.arange(nd) + 0.5
    phi = np.arccos(1 - 2 * i / nd)
    theta = np.pi * (1 + 5 ** 0.5) * i
    X = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)])
    rng = np.random.default_rng(20261202)
    sig = rng.normal(size=n) + 1j * rng.normal(size=n)
    op = inti.Helmholtz(k=k, dim=3)

    def J(y, nv, w, kk, s):
        o = inti.Helmholtz(k=kk, dim=3)
        A = inti.IntegralOperator(inti.CombinedFieldKernel(o, -1j * kk, 1.0, "FF"), inti.points(X),
                                  _quad(inti, y, nv, w))
        u = A.apply(s)
        return float(np.sum(np.abs(u) ** 2)), u

    base = (Q.coords, Q.normals, Q.weights, k, sig)
    J0, u0 = J(*base)
    A = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0, "FF"), inti.points(X),
                              _quad(inti, *base[:3]))
    g = A.pullback(2 * u0, sig, wrt=("sigma", "source_nodes", "source_normals", "source_weights", "k"),
                   dcoefs=(-1j, 0.0))
    worst = 0.0
    for t in range(20):
        which = t % 5
        v = rng.normal(size=np.shape(base[which])) if which != 3 else 1.0
        if which == 4:
            v = v + 1j * rng.normal(size=n)
            pred = np.real(np.vdot(g["sigma"], v))
        elif which == 3:
            pred = g["k"]
        else:
            pred = np.sum(g[("source_nodes", "source_normals", "source_weights")[which]] * v)
        h = 1e-6
        ap, am = list(base), list(base)
        ap[which] = base[which] + h * v
        am[which] = base[which] - h * v
        fd = (J(*ap)[0] - J(*am)[0]) / (2 * h)
        rel = abs(fd - pred) / max(abs(pred), 1e-8 * J0)
        worst = max(worst, rel)
        assert rel < 1e-5, (t, which, fd, pred)
    print(f"far-field FD: nodes {n}, J {J0:.6e}, worst rel {worst:.2e}", flush=True)


@pytest.mark.gpu
def test_gpu_pullback_of_the_corrected_operator(inti):
    """sigma_bar of I/2 + D - i k S with its DIM correction equals the
    conjugate transpose of the assembled corrected matrix applied to ubar;
    geometry gradients are refused while a correction is attached."""
    print("pullback corrected", flush=True)
    k = 3.0
    Q = inti.sphere_quadrature(3)
    op = inti.Helmholtz(k=k, dim=3)
    L, (dS, dD) = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, c_identity=0.5)
    rng = np.random.default_rng(5)
    u = rng.normal(size=len(Q)) + 1j * rng.normal(size=len(Q))
    s = rng.normal(size=len(Q)) + 1j * rng.normal(size=len(Q))
    g = L.iop.pullback(u, s, wrt=("sigma",))
    Sm = inti.assemble_matrix(inti.IntegralOperator(inti.SingleLayerKernel(op), Q))
    Dm = inti.assemble_matrix(inti.IntegralOperator(inti.DoubleLayerKernel(op), Q))
    M = 0.5 * np.eye(len(Q)) + Dm + dD.to_scipy().toarray() - 1j * k * (Sm + dS.to_scipy().toarray())
    ref = M.conj().T @ u
    assert np.linalg.norm(g["sigma"] - ref) <= 1e-12 * np.linalg.norm(ref)
    # the correction's adjoint is a gather (one writer per entry, fixed
    # order): the same bits on every call, in both summation modes
    for mode in (0, 2):
        L.iop.set_symmetric(mode)
        a = L.iop.pullback(u, s, wrt=("sigma",))["sigma"]
        for _ in range(3):
            assert np.array_equal(L.iop.pullback(u, s, wrt=("sigma",))["sigma"], a)
        assert np.linalg.norm(a - ref) <= 1e-12 * np.linalg.norm(ref)
    with pytest.raises(inti.IntiError):
        L.iop.pullback(u, s, wrt=("source_nodes",))


@pytest.mark.gpu
def test_gpu_pullback_error_paths(inti):
    op = inti.Helmholtz(k=1.0, dim=3)
    Q = inti.sphere_quadrature(2)
    u = np.ones(len(Q), dtype=complex)
    A = inti.IntegralOperator(inti.AdjointDoubleLayerKernel(op), Q)
    with pytest.raises(inti.IntiError):
        A.pullback(u, u)
    F = inti.IntegralOperator(inti.FarFieldSingleLayerKernel(op), inti.points(Q.coords[:, :5]), Q)
    with pytest.raises(inti.IntiError):
        F.pullback(np.ones(5, dtype=complex), u, wrt=("target_nodes",))
    M = inti.IntegralOperator(inti.SingleLayerKernel(op), Q, precision="mixed")
    with pytest.raises(inti.IntiError):
        M.pullback(u, u)
    # the combined node gradient needs target is source, and excludes the parts
    T = inti.points(Q.coords[:, :5] * 2.0)
    with pytest.raises(ValueError):
        inti.IntegralOperator(inti.SingleLayerKernel(op), T, Q).pullback(np.ones(5, dtype=complex), u,
                                                                           wrt=("nodes",))
    with pytest.raises(ValueError):
        inti.IntegralOperator(inti.SingleLayerKernel(op), Q).pullback(u, u, wrt=("nodes", "source_nodes"))


SYM_CASES = [(LAP, 2, SL), (LAP, 2, DL), (LAP, 3, SL), (LAP, 3, DL), (LAP, 3, COMB),
             (HEL, 2, SL), (HEL, 2, DL), (HEL, 2, COMB), (HEL, 3, SL), (HEL, 3, DL), (HEL, 3, COMB)]


@pytest.mark.gpu
@pytest.mark.parametrize("pde,dim,kind", SYM_CASES)
@pytest.mark.parametrize("n", [700, 1300])
def test_gpu_symmetric_pullback(inti, oracle, pde, dim, kind, n):
    """An operator on one quadrature (target is source): the symmetric
    pullback (one pass over unordered pairs) against the oracle (1e-12 per
    output, k gradient 1e-11) and against the two-pass pullback (1e-13),
    bit-identical on repeat; node counts not a multiple of the tile."""
    print(f"symmetric pullback pde={pde} dim={dim} kind={kind} n={n}", flush=True)
    k = 2.3 if pde == HEL else 0.0
    _, y, nv, w, _, s = _setup(dim, 1, n, 11 + kind + n)
    rng = np.random.default_rng(n + kind)
    u = rng.normal(size=n) + 1j * rng.normal(size=n)
    if pde == LAP:
        u, s = u.real, s.real
    op, K, dco = _kernel(inti, pde, dim, kind, k)
    Q = _quad(inti, y, nv, w)
    wrt = ("sigma", "source_nodes", "source_normals", "source_weights", "target_nodes") + \
        (("k",) if pde == HEL else ())
    A = inti.IntegralOperator(K, Q)
    A.set_symmetric(2)
    g = A.pullback(u, s, wrt=wrt, dcoefs=dco)
    g2 = A.pullback(u, s, wrt=wrt, dcoefs=dco)
    B = inti.IntegralOperator(K, Q)
    B.set_symmetric(0)
    h = B.pullback(u, s, wrt=wrt, dcoefs=dco)
    co, _ = _coefs(pde, kind, k)
    o = oracle.adjoint(pde, dim, kind, k, Pts(y), Pts(y, nv, w), u, s, coefs=co, dcoefs=dco, want_xbar=True)
    pairs = {"sigma": "sigma_bar", "source_nodes": "ybar", "source_normals": "nbar",
             "source_weights": "wbar", "target_nodes": "xbar"}
    for a, b in pairs.items():
        ref = o[b] if a != "sigma" or pde == HEL else o[b].real
        nr = np.linalg.norm(ref)
        assert np.linalg.norm(g[a] - ref) <= 1e-12 * nr, a
        assert np.linalg.norm(g[a] - h[a]) <= 1e-13 * nr, a
        assert np.array_equal(g[a], g2[a]), a
    if pde == HEL:
        assert abs(g["k"] - o["kbar"]) <= 1e-11 * abs(o["kbar"])
        assert abs(g["k"] - h["k"]) <= 1e-12 * abs(o["kbar"])
    # the node gradient in one output (source and target parts summed in the
    # pass): symmetric and two-pass, against the sum of the oracle's parts
    wrt_c = ("sigma", "nodes", "source_normals", "source_weights") + (("k",) if pde == HEL else ())
    gc = A.pullback(u, s, wrt=wrt_c, dcoefs=dco)
    hc = B.pullback(u, s, wrt=wrt_c, dcoefs=dco)
    ref = o["ybar"] + o["xbar"]
    nr = np.linalg.norm(ref)
    assert np.linalg.norm(gc["nodes"] - ref) <= 1e-12 * nr
    assert np.linalg.norm(hc["nodes"] - ref) <= 1e-12 * nr
    for a in ("sigma", "source_normals", "source_weights"):
        assert np.linalg.norm(gc[a] - g[a]) <= 1e-13 * np.linalg.norm(g[a]), a
    assert np.array_equal(A.pullback(u, s, wrt=wrt_c, dcoefs=dco)["nodes"], gc["nodes"])


@pytest.mark.gpu
@pytest.mark.parametrize("pde,dim,kind", SYM_CASES)
@pytest.mark.parametrize("n", [700, 1300])
def test_gpu_sigma_only_pullback_transposed_sum(inti, oracle, pde, dim, kind, n):
    """sigma_bar alone (A^H ubar) of an operator on one quadrature: the
    transposed symmetric sum (kernel values only) against the oracle (1e-12)
    and the general pullback (1e-13), bit-identical on repeat."""
    print(f"sigma-only pullback pde={pde} dim={dim} kind={kind} n={n}", flush=True)
    k = 2.3 if pde == HEL else 0.0
    _, y, nv, w, _, s = _setup(dim, 1, n, 31 + kind + n)
    rng = np.random.default_rng(n + 7 * kind)
    u = rng.normal(size=n) + 1j * rng.normal(size=n)
    if pde == LAP:
        u, s = u.real, s.real
    op, K, dco = _kernel(inti, pde, dim, kind, k)
    A = inti.IntegralOperator(K, _quad(inti, y, nv, w))
    A.set_symmetric(2)
    g = A.pullback(u, s)["sigma"]
    assert np.array_equal(g, A.pullback(u, s)["sigma"])
    B = inti.IntegralOperator(K, _quad(inti, y, nv, w))
    B.set_symmetric(0)
    h = B.pullback(u, s)["sigma"]
    co, _ = _coefs(pde, kind, k)
    o = oracle.adjoint(pde, dim, kind, k, Pts(y), Pts(y, nv, w), u, s, coefs=co, dcoefs=dco, want_xbar=False)
    ref = o["sigma_bar"] if pde == HEL else o["sigma_bar"].real
    assert np.linalg.norm(g - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.linalg.norm(g - h) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_gpu_sigma_only_pullback_of_the_corrected_operator_transposed(inti):
    """The corrected CFIE I/2 + D - i k S (DIM correction, identity) with the
    transposed symmetric sum: A^H ubar equals the general pullback's and the
    assembled matrix's conjugate transpose applied to ubar."""
    print("sigma-only pullback, corrected operator, transposed sum", flush=True)
    k = 3.0
    Q = inti.sphere_quadrature(4)
    op = inti.Helmholtz(k=k, dim=3)
    L, (dS, dD) = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, c_identity=0.5)
    rng = np.random.default_rng(6)
    u = rng.normal(size=len(Q)) + 1j * rng.normal(size=len(Q))
    s = rng.normal(size=len(Q)) + 1j * rng.normal(size=len(Q))
    L.iop.set_symmetric(2)
    g = L.iop.pullback(u, s, wrt=("sigma",))["sigma"]
    L.iop.set_symmetric(0)
    h = L.iop.pullback(u, s, wrt=("sigma",))["sigma"]
    Sm = inti.assemble_matrix(inti.IntegralOperator(inti.SingleLayerKernel(op), Q))
    Dm = inti.assemble_matrix(inti.IntegralOperator(inti.DoubleLayerKernel(op), Q))
    M = 0.5 * np.eye(len(Q)) + Dm + dD.to_scipy().toarray() - 1j * k * (Sm + dS.to_scipy().toarray())
    ref = M.conj().T @ u
    assert np.linalg.norm(g - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.linalg.norm(g - h) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_gpu_symmetric_pullback_automatic_on_a_large_operator(inti):
    """A 3D Helmholtz combined field on a 20,280-node sphere: the automatic
    mode takes the symmetric pullback (enough tile pairs to fill the GPU);
    every output equals the two-pass pullback's to 1e-12."""
    print("symmetric pullback, automatic mode, 20,280 nodes", flush=True)
    Q = inti.sphere_quadrature(13)
    n = len(Q)
    k = 3.0
    op = inti.Helmholtz(k=k, dim=3)
    K = inti.CombinedFieldKernel(op, -1j * k, 1.0)
    rng = np.random.default_rng(13)
    u = rng.normal(size=n) + 1j * rng.normal(size=n)
    s = rng.normal(size=n) + 1j * rng.normal(size=n)
    wrt = ("sigma", "source_nodes", "source_normals", "source_weights", "target_nodes", "k")
    A = inti.IntegralOperator(K, Q)
    g = A.pullback(u, s, wrt=wrt, dcoefs=(-1j, 0.0))
    B = inti.IntegralOperator(K, Q)
    B.set_symmetric(0)
    h = B.pullback(u, s, wrt=wrt, dcoefs=(-1j, 0.0))
    for a in wrt[:-1]:
        assert np.linalg.norm(g[a] - h[a]) <= 1e-12 * np.linalg.norm(h[a]), a
    assert abs(g["k"] - h["k"]) <= 1e-12 * abs(h["k"])


@pytest.mark.gpu
def test_gpu_shape_gradient_through_the_corrected_solve(inti):
    """The adjoint-state gradient of J = sum |u_inf|^2 through the corrected
    CFIE solve (bench.shape_gradient: sound-soft scattering of e^{ikz}, k = 4,
    perturbed sphere at icosphere frequency 8, 7,680 nodes, 16 directions)
    against central differences (h = 1e-6) of the whole solve along 8 seeded
    directions in nodes, normals, weights and k: the full gradient (the DIM
    corrections differentiated too) against FD of the solve with the
    corrections rebuilt on every perturbed geometry, and the gradient with the
    corrections held fixed against FD with them frozen. Errors are measured
    against the largest directional derivative of the same kind (a random
    node direction can give a derivative far below the gradient's size; the
    FD error, h^2 times the third derivative plus the solver's noise over h,
    does not shrink with it): 1e-5 frozen, 1e-4 full (the rebuilt
    corrections' least squares are ill-conditioned)."""
    import bench
    print("shape gradient through the solve, 7,680 nodes", flush=True)
    Q = inti.sphere_quadrature(8, radial=lambda th, ph: 1.0 + 0.1 * np.sin(3 * th) * np.cos(2 * ph))
    k = 4.0
    dirs = bench.fibonacci_directions(16)
    J, grad, info = bench.shape_gradient(inti, Q.coords, Q.normals, Q.weights, Q.qtags, k, dirs)
    print(f"  J {J:.6e}, solve {info['solve']}, adjoint {info['adjoint_solve']}", flush=True)
    assert info["adjoint_solve"]["relres"] <= 1e-11
    checks = bench.shape_fd_check(inti, Q.coords, Q.normals, Q.weights, Q.qtags, k, dirs, grad,
                                  info["corrections"], 8, 7, sigma0=info["sigma"], grad_frozen=info["grad_frozen"])
    sfull, sfroz = {}, {}
    for w, _, fz, fl, _ in checks:
        sfull[w] = max(sfull.get(w, 0.0), abs(fl))
        sfroz[w] = max(sfroz.get(w, 0.0), abs(fz))
    worst_full = worst_frozen = 0.0
    for w, a, fz, fl, af in checks:
        el, ez = abs(a - fl) / sfull[w], abs(af - fz) / sfroz[w]
        print(f"  {w:8s} gradient {a:+.9e} fd(full) {fl:+.9e} {el:.1e}; frozen {af:+.9e} fd {fz:+.9e} {ez:.1e}",
              flush=True)
        worst_full, worst_frozen = max(worst_full, el), max(worst_frozen, ez)
    assert worst_frozen <= 1e-5
    assert worst_full <= 1e-4
