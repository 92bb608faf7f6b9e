"""The reference's documented examples on the GPU path, checked against the
analytic solutions they use (docs/src/examples/helmholtz_scattering.jl):
sound-soft scattering by the unit circle (k = 4 pi, dense corrected
operators, direct solve; er < 1e-3 on the circle of radius 2) and by the
unit sphere (k = 4 pi, matrix-free corrected CFIE, GMRES abstol 1e-6;
er < 1e-3 at 100 random points on the sphere of radius 2), plus the
off-surface potentials (nystrom.jl:20-38, api.jl:220-226) against the
oracle's direct sum."""
import numpy as np
import pytest
from scipy import special

LAP, HEL = 0, 1


def circle_soundsoft(x, k, radius=1.0, thin=0.0):
    """Separation of variables in polar coordinates (plane wave e^{ik d.x})."""
    r = np.hypot(x[0], x[1])
    th = np.arctan2(x[1], x[0])
    u = np.zeros(r.shape, dtype=np.complex128)
    for n in range(-80, 81):
        c = -np.exp(1j * n * (np.pi / 2 - thin)) * special.jv(n, k * radius) / special.hankel1(n, k * radius)
        if abs(c) < 1e-16:
            continue
        u += c * special.hankel1(n, k * r) * np.exp(1j * n * th)
    return u


def sphere_soundsoft(x, k, d, radius=1.0, lmax=70):
    """Sound-soft sphere, incident e^{ik d.x}: the Legendre expansion
    -sum_l i^l (2l+1) j_l(ka)/h_l(ka) h_l(kr) P_l(cos gamma)."""
    r = np.linalg.norm(x, axis=0)
    cg = (d @ x) / r
    u = np.zeros(r.shape, dtype=np.complex128)
    for l in range(lmax + 1):
        hl_a = special.spherical_jn(l, k * radius) + 1j * special.spherical_yn(l, k * radius)
        hl_r = special.spherical_jn(l, k * r) + 1j * special.spherical_yn(l, k * r)
        u -= (1j ** l) * (2 * l + 1) * special.spherical_jn(l, k * radius) / hl_a * hl_r \
            * special.eval_legendre(l, cg)
    return u


@pytest.mark.gpu
def test_potentials_match_oracle(inti, oracle):
    """IntegralPotential evaluation = the oracle's direct sum at off-surface
    points, 2D and 3D, Laplace and Helmholtz, SL and DL (<= 1e-12)."""
    print("potentials vs oracle", flush=True)
    rng = np.random.default_rng(11)
    for dim, Q in ((2, inti.circle_quadrature(40, qorder=5)), (3, inti.sphere_quadrature(3))):
        X = rng.normal(size=(dim, 50))
        X = X / np.linalg.norm(X, axis=0) * rng.uniform(1.2, 3.0, size=50)
        sig = rng.normal(size=len(Q)) + 1j * rng.normal(size=len(Q))
        P = inti.points(X)
        for pde, k in ((LAP, 0.0), (HEL, 3.0)):
            op = inti.Laplace(dim=dim) if pde == LAP else inti.Helmholtz(k=k, dim=dim)
            s = sig.real if pde == LAP else sig
            Sp, Dp = inti.single_double_layer_potential(op=op, source=Q)
            for kind, pot in ((0, Sp), (1, Dp)):
                got = pot[s](X)
                ref = oracle.matvec(pde, dim, kind, k, P, Q, s)
                assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
                assert abs(pot[s](X[:, 3]) - ref[3]) <= 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
def test_helmholtz_scattering_circle(inti):
    """2D part of helmholtz_scattering.jl: meshsize lambda/5 on the unit
    circle, P2 elements, quadrature order 4, L = I/2 + D - i k S assembled
    dense with the DIM correction and solved directly."""
    print("scattering circle", flush=True)
    k = 4 * np.pi
    lam = 2 * np.pi / k
    nel = int(np.ceil(2 * np.pi / (lam / 5)))
    Q = inti.circle_quadrature(nel, qorder=4)
    op = inti.Helmholtz(k=k, dim=2)
    S, D = inti.single_double_layer(op=op, target=Q, source=Q, compression={"method": "none"},
                                    correction={"method": "dim"})
    L = 0.5 * np.eye(len(Q)) + D - 1j * k * S
    rhs = -np.exp(1j * k * Q.coords[0])
    sigma = np.linalg.solve(L, rhs)
    Sp, Dp = inti.single_double_layer_potential(op=op, source=Q)
    th = np.arange(0, 2 * np.pi, 0.01)
    X = 2 * np.stack([np.cos(th), np.sin(th)])
    us = Dp[sigma](X) - 1j * k * Sp[sigma](X)
    er = np.abs(us - circle_soundsoft(X, k)).max()
    print(f"circle: nodes {len(Q)}, er {er:.3e}", flush=True)
    assert er < 1e-3
    # the matrix-free fused operator and GPU GMRES give the same density
    Lop, _ = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, c_identity=0.5)
    s2, info = inti.gmres(Lop, rhs, restart=200, maxiter=200, rtol=1e-12)
    assert np.linalg.norm(s2 - sigma) <= 1e-9 * np.linalg.norm(sigma)


@pytest.mark.gpu
def test_helmholtz_scattering_sphere(inti):
    """3D part of helmholtz_scattering.jl: unit sphere, k = 4 pi, P2
    triangles of size about lambda/5, VR4N6, L = I/2 + D - i k S (matrix-free
    corrected) by GMRES (abstol 1e-6, restart 100); the scattered field
    D[sigma] - i k S[sigma] against the series at 100 random points of the
    sphere of radius 2, incident e^{i k x1}."""
    print("scattering sphere", flush=True)
    k = 4 * np.pi
    Q = inti.sphere_quadrature(12)
    op = inti.Helmholtz(k=k, dim=3)
    L, _ = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, c_identity=0.5)
    rhs = -np.exp(1j * k * Q.coords[0])
    sigma, info = inti.gmres(L, rhs, restart=100, maxiter=100, rtol=0.0, atol=1e-6)
    assert info["iterations"] < 100
    rng = np.random.default_rng(2026)
    X = rng.normal(size=(3, 100))
    X = 2 * X / np.linalg.norm(X, axis=0)
    Sp, Dp = inti.single_double_layer_potential(op=op, source=Q)
    us = Dp[sigma](X) - 1j * k * Sp[sigma](X)
    er = np.abs(us - sphere_soundsoft(X, k, np.array([1.0, 0.0, 0.0]))).max()
    print(f"sphere: nodes {len(Q)}, gmres its {info['iterations']}, er {er:.3e}", flush=True)
    assert er < 1e-3
