"""Quadrature construction (host side of the path), no GPU."""
import math

import numpy as np
import pytest


def test_gauss_legendre_matches_mpmath(inti, golden):
    for n in range(1, 13):
        x, w = inti.gauss_legendre(n)
        assert np.abs(x - golden[f"gl_{n}_x"]).max() <= 2.3e-16
        assert np.abs(w - golden[f"gl_{n}_w"]).max() <= 4.5e-16
    # GaussLegendre(; order) -> ceil((order+1)/2) nodes
    assert len(inti.gauss_legendre_order(8)[0]) == 5
    assert len(inti.gauss_legendre_order(9)[0]) == 5
    assert len(inti.gauss_legendre_order(10)[0]) == 6
    # jldoctest of reference_integration.jl:127-137
    x, w = inti.gauss_legendre_order(10)
    assert abs(np.dot(np.cos(x), w) - math.sin(1)) < 1e-15


def test_vr4n6_exactness(inti):
    """Exact for total degree <= 4, not for all of degree 5
    (reference_integration_test.jl:76-101 pattern)."""
    uv, w = inti.vioreanu_rokhlin_triangle(4)
    assert abs(w.sum() - 0.5) < 1e-15
    worst5 = 0
    for a in range(6):
        for b in range(6 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = np.dot(uv[:, 0] ** a * uv[:, 1] ** b, w)
            if a + b <= 4:
                assert abs(got - exact) < 1e-15
            else:
                worst5 = max(worst5, abs(got - exact))
    assert worst5 > 1e-6
    with pytest.raises(ValueError):
        inti.vioreanu_rokhlin_triangle(3)  # no 3rd-order rule in the table


@pytest.mark.parametrize("order", [2, 5])
def test_other_vr_rules_exactness(inti, order):
    uv, w = inti.vioreanu_rokhlin_triangle(order)
    assert abs(w.sum() - 0.5) < 1e-15
    for a in range(order + 1):
        for b in range(order + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.dot(uv[:, 0] ** a * uv[:, 1] ** b, w) - exact) < 1e-15


def test_disk_quadrature(inti):
    from inti_jl_amd import quadrature as q
    errs = []
    for R in (4, 8):
        V, B = q.disk_quadratures(R, 2.0, 1.0, qorder=5)
        assert V.ne == 6 * R * R and V.nq == 10 and B.ne == 6 * R
        assert np.all(V.weights > 0)
        errs.append(abs(V.weights.sum() - 2 * np.pi))
        # volume and boundary discretise the same curve: divergence theorem
        flux = np.dot(np.sum(B.coords * B.normals, axis=0), B.weights)
        assert abs(flux - 2 * V.weights.sum()) < 1e-9
    assert errs[1] < errs[0] / 12


def test_circle_geometry(inti):
    errs = []
    for nel in (16, 32, 64):
        Q = inti.circle_quadrature(nel, qorder=8)
        assert Q.qtags.shape == (5, nel)
        assert np.array_equal(Q.qtags[:, 3], np.arange(15, 20))
        assert np.allclose(np.linalg.norm(Q.normals, axis=0), 1, atol=1e-15)
        assert np.all(np.sum(Q.coords * Q.normals, axis=0) > 0.99)  # outward
        errs.append(abs(Q.weights.sum() - 2 * np.pi))
    assert errs[1] < errs[0] / 10 and errs[2] < errs[1] / 10
    # element nodes lie on the circle exactly
    assert np.allclose(np.linalg.norm(Q.elements, axis=2), 1, atol=1e-15)


def test_divergence_theorem(inti):
    """Integral of x.n over the boundary = dim * volume
    (normal_orientation_test.jl:48-65); the P2 geometry error falls as h^4."""
    C = inti.ellipse_quadrature(2.0, 0.5, 64, qorder=8)
    assert abs(np.dot(np.sum(C.coords * C.normals, axis=0), C.weights) - 2 * np.pi * 2.0 * 0.5) < 1e-5
    errs = []
    for f in (4, 8):
        S = inti.sphere_quadrature(f, base="ico")
        errs.append(abs(np.dot(np.sum(S.coords * S.normals, axis=0), S.weights) - 4 * np.pi))
    assert errs[1] < 2e-4 and errs[1] < errs[0] / 12


def test_sphere_counts_and_orientation(inti):
    for base, f, ne in (("ico", 1, 20), ("ico", 3, 180), ("octa", 2, 32)):
        S = inti.sphere_quadrature(f, base=base)
        assert S.ne == ne and len(S) == 6 * ne
        assert np.all(np.sum(S.coords * S.normals, axis=0) > 0)
    bumpy = inti.sphere_quadrature(4, radial=lambda th, ph: 1 + 0.15 * np.sin(5 * th) * np.cos(4 * ph))
    assert np.all(np.sum(bumpy.coords * bumpy.normals, axis=0) > 0)
    val = np.dot(np.sum(bumpy.coords * bumpy.normals, axis=0), bumpy.weights)
    assert val > 0


def test_ellipse_disk_mesh(inti):
    """Quasi-uniform Delaunay P2 mesh of the elliptical disk: area and
    perimeter of the 2 x 1 ellipse (curved P2 edges: h^4-accurate), every
    triangle counter-clockwise with a positive Jacobian at the nodes, the
    boundary rebuilt from the volume mesh equals the one returned."""
    from scipy import special
    a, b = 2.0, 1.0
    perim = 4 * a * special.ellipe(1 - (b / a) ** 2)
    V, B = inti.ellipse_disk_quadratures(0.1, a=a, b=b, qorder=5, boundary_qorder=10)
    assert abs(V.weights.sum() - np.pi * a * b) < 1e-5
    assert abs(B.weights.sum() - perim) < 1e-6
    from inti_jl_amd.quadrature import p2_triangle_eval, vioreanu_rokhlin_triangle
    uv, _ = vioreanu_rokhlin_triangle(5)
    _, xu, xv = p2_triangle_eval(V.elements, uv)
    assert (xu[..., 0] * xv[..., 1] - xu[..., 1] * xv[..., 0]).min() > 0
    B2 = inti.boundary_quadrature(V, 10)
    np.testing.assert_array_equal(B2.coords, B.coords)
    # divergence theorem on the boundary: int x . n = 2 |Omega|
    assert abs(np.sum(B.weights * np.sum(B.coords * B.normals, axis=0)) - 2 * np.pi * a * b) < 1e-5
