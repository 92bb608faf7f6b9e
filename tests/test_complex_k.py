"""Complex wavenumbers (Helmholtz with k = kr + i ki, src/kernels.jl:322-413,
3D): the numpy oracle pinned on the CPU (against the C oracle at ki = 0, by
central differences and by the Helmholtz equation for complex k), the host
errors of what takes real wavenumbers only, and on the GPU the assembled
matrices, the one-direction sums (1, 4 and 16 right-hand sides) and the DIM
correction against that oracle."""
import sys

import numpy as np
import pytest

LAP, HEL = 0, 1
SL, DL, ADL, HS, COMB_SD, COMB_AH = 0, 1, 2, 3, 4, 5
K = 2.3 + 0.7j


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ck():
    from oracle import helmholtz_ck
    return helmholtz_ck


def _pts(rng, n, scale=1.0):
    x = scale * rng.normal(size=(3, n))
    nv = rng.normal(size=(3, n))
    return x, nv / np.linalg.norm(nv, axis=0)


@pytest.mark.parametrize("kind", [SL, DL, ADL, HS, COMB_SD, COMB_AH])
def test_ck_oracle_matches_c_oracle_at_real_k(ck, oracle, kind):
    """ki = 0: every kind equals the C oracle's value (the reference's sign
    and normal conventions) to 1e-13 relative."""
    rng = np.random.default_rng(1)
    X, NX = _pts(rng, 5)
    Y, NY = _pts(rng, 6, 2.0)
    coefs = (0.4 - 1.1j, 0.8 + 0.2j)
    got = ck.kernel_matrix(kind, 1.7, X, NX, Y, NY, coefs if kind >= COMB_SD else None)
    for i in range(5):
        for j in range(6):
            ref = oracle.kernel(HEL, 3, kind, 1.7, X[:, i], NX[:, i], Y[:, j], NY[:, j],
                                coefs=coefs if kind >= COMB_SD else None)
            assert abs(got[i, j] - ref) <= 1e-13 * abs(ref)


def test_ck_oracle_derivatives_and_pde(ck):
    """For complex k: DL = dSL/dn_y, ADL = dSL/dn_x, HS = dDL/dn_x by central
    differences (h = 1e-5, 1e-7 relative), and (Delta_x + k^2) G = 0 by a
    second-order five-point Laplacian (1e-5 relative to k^2 G)."""
    rng = np.random.default_rng(2)
    X, NX = _pts(rng, 4)
    Y, NY = _pts(rng, 4, 2.5)
    h = 1e-5
    f = lambda kind, x, y: ck.kernel_matrix(kind, K, x, NX, y, NY)  # noqa: E731
    fd_dl = (f(SL, X, Y + h * NY) - f(SL, X, Y - h * NY)) / (2 * h)
    np.testing.assert_allclose(np.diag(fd_dl), np.diag(f(DL, X, Y)), rtol=1e-7)
    fd_adl = (f(SL, X + h * NX, Y) - f(SL, X - h * NX, Y)) / (2 * h)
    np.testing.assert_allclose(np.diag(fd_adl), np.diag(f(ADL, X, Y)), rtol=1e-7)
    fd_hs = (f(DL, X + h * NX, Y) - f(DL, X - h * NX, Y)) / (2 * h)
    np.testing.assert_allclose(np.diag(fd_hs), np.diag(f(HS, X, Y)), rtol=1e-7)
    hh = 1e-3
    G = f(SL, X, Y)
    lap = -6 * G
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = hh
        lap = lap + f(SL, X + e, Y) + f(SL, X - e, Y)
    lap /= hh * hh
    np.testing.assert_allclose(np.diag(lap + K * K * G), 0, atol=1e-5 * np.abs(K * K * np.diag(G)).max())
    # the damping: |G| = e^{-ki d} / (4 pi d)
    d = np.linalg.norm(X[:, :, None] - Y[:, None, :], axis=0)
    np.testing.assert_allclose(np.abs(G), np.exp(-K.imag * d) / (4 * np.pi * d), rtol=1e-14)


@pytest.mark.parametrize("derivative", [False, True])
def test_ck_oracle_bdim_matches_c_oracle_at_real_k(ck, oracle, inti, derivative):
    """The numpy restatement of the DIM correction at ki = 0 against the C
    oracle's build on a sphere of 120 nodes: values on the pattern within 10x
    the spread of the two least-squares solvers (QR, SVD), zero off it."""
    Q = inti.sphere_quadrature(1, base="ico")
    n = len(Q)
    mu = np.full(n, -0.5)
    rp, ci, vS, vD, _ = oracle.bdim(HEL, 3, 2.3, Q, Q, mu, derivative=derivative)
    z = oracle.monopoles(3, Q.coords, Q.qtags.shape[0])
    oS, oD = ck.bdim(2.3, Q, mu, z, derivative=derivative, solver="qr")
    fS, fD = ck.bdim(2.3, Q, mu, z, derivative=derivative, solver="svd")
    rows = np.repeat(np.arange(n), np.diff(rp))
    for v, o, f in ((vS, oS, fS), (vD, oD, fD)):
        spread = np.abs(f - o).max()
        assert np.abs(v - o[rows, ci]).max() <= 10 * spread + 1e-13 * np.abs(o).max()
        off = o.copy()
        off[rows, ci] = 0
        assert np.abs(off).max() == 0


def test_ck_host_errors(inti):
    """What takes real wavenumbers only says so: 2D, purely imaginary k,
    batches and the DIM pullback."""
    with pytest.raises(NotImplementedError):
        inti.Helmholtz(k=1.0 + 0.5j, dim=2)
    with pytest.raises(ValueError):
        inti.Helmholtz(k=0.5j, dim=3)
    op = inti.Helmholtz(k=K, dim=3)
    assert op.k == K and op.kr == K.real and op.ki == K.imag
    assert inti.Helmholtz(k=2.0 + 0j, dim=3).k == 2.0
    Q = inti.sphere_quadrature(1, base="ico")
    with pytest.raises(NotImplementedError):
        inti.BatchOperator([inti.DoubleLayerKernel(op)], [Q])
    import importlib
    bp = importlib.import_module(inti.__name__ + ".bdim_pullback")
    with pytest.raises(NotImplementedError):
        bp.bdim_pullback(op, Q, np.zeros(1), np.zeros(1), np.zeros(2, dtype=np.int64), np.zeros(1, dtype=np.int64),
                         green_multiplier=np.full(len(Q), -0.5))


@pytest.mark.gpu
def test_ck_errors_on_the_device(inti):
    """The library refuses a complex wavenumber where it has no kernel: mixed
    precision, far fields, the adjoint."""
    say("[gpu] complex k: refusals")
    op = inti.Helmholtz(k=K, dim=3)
    Q = inti.sphere_quadrature(1, base="ico")
    x = np.ones(len(Q), dtype=np.complex128)
    with pytest.raises(inti.IntiError):
        inti.IntegralOperator(inti.DoubleLayerKernel(op), Q, Q, precision="mixed") @ x
    A = inti.IntegralOperator(inti.SingleLayerKernel(op), Q, Q)
    with pytest.raises(inti.IntiError):
        A.pullback(x, x, wrt=("sigma",))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [SL, DL, ADL, HS, COMB_SD, COMB_AH])
def test_ck_assemble_and_apply_match_oracle(inti, ck, kind):
    """Dense assembly and the direct sums with 1, 4 and 16 right-hand sides
    (the last on the FP64 matrix cores) against the numpy oracle on a sphere
    of 480 nodes, 1e-12 relative; k = 2.3 + 0.7i."""
    say(f"[gpu] complex k kind={kind}")
    Q = inti.sphere_quadrature(2, base="ico")
    n = len(Q)
    op = inti.Helmholtz(k=K, dim=3)
    coefs = (-1j * K, 1.0)
    if kind == COMB_SD:
        Kr = inti.CombinedFieldKernel(op, coefs[0], coefs[1], "SD")
    elif kind == COMB_AH:
        Kr = inti.CombinedFieldKernel(op, coefs[0], coefs[1], "AH")
    else:
        Kr = [inti.SingleLayerKernel, inti.DoubleLayerKernel, inti.AdjointDoubleLayerKernel,
              inti.HyperSingularKernel][kind](op)
    A = inti.IntegralOperator(Kr, Q, Q)
    ref = ck.assemble(kind, K, Q, Q, coefs if kind >= COMB_SD else None)
    M = inti.assemble_matrix(A)
    assert np.linalg.norm(M - ref) <= 1e-12 * np.linalg.norm(ref)
    rng = np.random.default_rng(3)
    for nr in (1, 4, 16):
        X = rng.normal(size=(n, nr)) + 1j * rng.normal(size=(n, nr))
        Y = A @ (X[:, 0] if nr == 1 else X)
        Yr = ref @ X
        Y = Y.reshape(n, -1)
        assert np.linalg.norm(Y - Yr) <= 1e-12 * np.linalg.norm(Yr), nr


@pytest.mark.gpu
@pytest.mark.parametrize("derivative", [False, True])
def test_ck_bdim_matches_oracle(inti, ck, oracle, derivative):
    """The DIM correction with k = 2.3 + 0.7i on a sphere of 480 nodes:
    pattern bit-exact against the C oracle's (it does not depend on k);
    weights within 1e-12 of the largest weight, or 10x the oracle's own
    spread where that is larger (two oracle builds: QR least squares with
    Theta summed forwards, SVD with Theta summed backwards); the corrected
    operator on a smooth density within 1e-12 (or 10x that spread)."""
    say(f"[gpu] complex k bdim derivative={derivative}")
    Q = inti.sphere_quadrature(2, base="ico")
    n = len(Q)
    op = inti.Helmholtz(k=K, dim=3)
    if derivative:
        Sop = inti.IntegralOperator(inti.AdjointDoubleLayerKernel(op), Q, Q)
        Dop = inti.IntegralOperator(inti.HyperSingularKernel(op), Q, Q)
    else:
        Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), Q, Q)
        Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q, Q)
    mu = np.full(n, -0.5)
    dS, dD = inti.bdim_correction(op, Q, Q, Sop, Dop, green_multiplier=mu, derivative=derivative)
    rp, ci, vS = dS.export()
    _, _, vD = dD.export()
    orp, oci, _, _, _ = oracle.bdim(HEL, 3, 2.3, Q, Q, mu, derivative=derivative)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    z = oracle.monopoles(3, Q.coords, Q.qtags.shape[0])
    oS, oD = ck.bdim(K, Q, mu, z, derivative=derivative, solver="qr")
    fS, fD = ck.bdim(K, Q, mu, z, derivative=derivative, solver="svd", order="backward")
    rows = np.repeat(np.arange(n), np.diff(rp))
    for v, o, f in ((vS, oS, fS), (vD, oD, fD)):
        spread = np.abs(f[rows, ci] - o[rows, ci]).max()
        assert np.abs(v - o[rows, ci]).max() <= max(10 * spread, 1e-12 * np.abs(o).max())
    x = np.exp(1j * (1.3 * Q.coords[0] - 0.8 * Q.coords[1])) + Q.coords[2] ** 2
    kS, kD = (ADL, HS) if derivative else (SL, DL)
    for A, dA, kk, o, f in ((Dop, dD, kD, oD, fD), (Sop, dS, kS, oS, fS)):
        A.attach(dA)
        y = A @ x
        base = ck.assemble(kk, K, Q, Q) @ x
        ref, alt = base + o @ x, base + f @ x
        spread = np.linalg.norm(alt - ref) / np.linalg.norm(ref)
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= max(1e-12, 10 * spread)


@pytest.mark.gpu
def test_ck_green_identity(inti, ck):
    """Green's identity with a complex wavenumber through
    single_double_layer (green_identities_test.jl:93-104 bar: e1 < 1e-2 and
    below the uncorrected e0): u = G(., x_s) of an exterior source, an
    interior solution for any k; the same check at the reference's
    thresholds as the real-k tests."""
    say("[gpu] complex k Green identity")
    Q = inti.sphere_quadrature(2, base="ico")
    op = inti.Helmholtz(k=K, dim=3)
    xs = np.full((3, 1), 3.0)
    u = ck.kernel_matrix(SL, K, Q.coords, Q.normals, xs, None)[:, 0]
    dudn = ck.kernel_matrix(ADL, K, Q.coords, Q.normals, xs, None)[:, 0]
    Sm = ck.assemble(SL, K, Q, Q)
    Dm = ck.assemble(DL, K, Q, Q)
    e0 = np.abs(Sm @ dudn - Dm @ u - 0.5 * u).max() / np.abs(u).max()
    S, D = inti.single_double_layer(op=op, target=Q, source=Q, compression={"method": "gpu"},
                                    correction={"method": "dim"})
    e1 = np.abs(S @ dudn - D @ u - 0.5 * u).max() / np.abs(u).max()
    say(f"  e0 {e0:.2e} e1 {e1:.2e}")
    assert e1 < 1e-2 and e1 < e0
