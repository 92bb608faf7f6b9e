"""Volume density interpolation (Inti.jl src/vdim.jl, src/api.jl:282-378):
polynomial bases and particular solutions, the oracle's correction against
the mathematics (Green's representation of a manufactured solution), and the
GPU correction against the oracle. The Poisson test follows
docs/src/examples/poisson.jl (Laplace, elliptical disk 2 x 1, interpolation
order 3 on Vioreanu-Rokhlin order-5 triangles, norm(er) < 1e-5)."""
import math

import numpy as np
import pytest

LAP, HEL = 0, 1


@pytest.fixture(scope="module")
def ov():
    import oracle.vdim as OV
    return OV


def _op(inti, pde, k, dim=2):
    return inti.Laplace(dim=dim) if pde == LAP else inti.Helmholtz(k=k, dim=dim)


# ---------------------------------------------------------------- CPU ----

def test_multiindices_first_index_fastest(inti, ov):
    """Iterators.product(0:n, 0:n[, 0:n]) order filtered to |I| <= n (vdim.jl:339-341)."""
    P = inti.polynomials
    for order in range(5):
        assert P.monomial_multiindices(order, 2) == ov.multiindices(order)
        want = [(i, j, l) for l in range(order + 1) for j in range(order + 1)
                for i in range(order + 1) if i + j + l <= order]
        assert P.monomial_multiindices(order, 3) == want
    assert len(P.monomial_multiindices(3, 2)) == 10
    assert len(P.monomial_multiindices(3, 3)) == 20


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 1.7)])
def test_particular_solutions_match_oracle_2d(inti, ov, pde, k):
    """Closed-form P against the oracle's linear solve on coefficients."""
    P = inti.polynomials
    x = np.random.default_rng(3).normal(size=(2, 9))
    ps, Ps, idx = P.polynomial_solutions_vdim(_op(inti, pde, k), 5)
    for I, p, Pp in zip(idx, ps, Ps):
        c = ov.particular_solution(I, pde, k, 8)
        np.testing.assert_allclose(P.evaluate(Pp, x), ov._eval(c, x), rtol=1e-13, atol=1e-13)
        assert P.evaluate(p, x) == pytest.approx(np.prod(x ** np.array(I)[:, None], axis=0)
                                                 / math.prod(math.factorial(i) for i in I))


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 1.3)])
def test_particular_solutions_solve_the_pde_3d(inti, pde, k):
    """Delta P (+ k^2 P) = p by fourth-order central differences."""
    P = inti.polynomials
    x = np.random.default_rng(5).uniform(-0.7, 0.7, size=(3, 6))
    h = 1e-2
    ps, Ps, _ = P.polynomial_solutions_vdim(_op(inti, pde, k, 3), 4)
    for p, Pp in zip(ps, Ps):
        lap = -30 * P.evaluate(Pp, x) * 3
        for d in range(3):
            e = np.zeros((3, 1))
            e[d] = h
            lap = lap + 16 * (P.evaluate(Pp, x + e) + P.evaluate(Pp, x - e)) \
                - (P.evaluate(Pp, x + 2 * e) + P.evaluate(Pp, x - 2 * e))
        lap = lap / (12 * h * h)
        lhs = lap + k * k * P.evaluate(Pp, x)
        np.testing.assert_allclose(lhs, P.evaluate(p, x), atol=1e-6)


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 1.3)])
def test_particular_solutions_match_oracle_3d(inti, ov, pde, k):
    """3D: the oracle's P (linear solve on coefficient arrays) solves the PDE
    exactly on the coefficients, and equals the closed form: for Helmholtz
    the polynomial solution is unique; for Laplace both are the solution of
    the form |x|^2 q."""
    P = inti.polynomials
    x = np.random.default_rng(4).normal(size=(3, 9))
    ps, Ps, idx = P.polynomial_solutions_vdim(_op(inti, pde, k, 3), 3)
    assert [tuple(I) for I in idx] == ov.multiindices(3, 3)
    for I, Pp in zip(idx, Ps):
        c = ov.particular_solution(tuple(I), pde, k, 6)
        p = np.zeros_like(c)
        p[tuple(I)] = 1.0 / math.prod(math.factorial(i) for i in I)
        np.testing.assert_allclose(ov._lap(c) + k * k * c, p, atol=1e-13)
        np.testing.assert_allclose(P.evaluate(Pp, x), ov._eval(c, x), rtol=1e-12, atol=1e-13)


def _manufactured_3d(inti, ov, ncell, pde, k, qorder, order):
    """The cube [-1/2, 1/2]^3 (box_quadratures): (corrected V f, Green
    reference, uncorrected V f, nodes with |x|_inf < 0.3) for u = cos(k0 th.x),
    f = (k^2 - k0^2) u, all from the oracle."""
    from oracle import csr_matvec, matvec
    Vq, Bq = inti.box_quadratures(ncell, qorder=qorder, boundary_qorder=4)
    m = len(Vq)
    mu = np.full(m, -1.0)
    k0, th = np.pi, np.array([1.0, 2.0, 2.0]) / 3
    u = lambda x: np.cos(k0 * (th @ x))
    f = (k * k - k0 * k0) * u(Vq.coords)
    S, D = ov.boundary_operators(pde, k, Vq, Bq, mu)
    rp, ci, val = ov.vdim(pde, k, Vq, Vq, Bq, mu, order, S, D)
    V0 = matvec(pde, 3, 0, k, Vq, Vq, f)
    Vf = V0 + csr_matvec(m, m, rp, ci, val, f)
    du = -k0 * (th @ Bq.normals) * np.sin(k0 * (th @ Bq.coords))
    vref = -u(Vq.coords) - D(u(Bq.coords)) + S(du)
    inner = np.abs(Vq.coords).max(axis=0) < 0.3
    return Vf, vref, V0, inner


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
@pytest.mark.parametrize("qorder,order", [(3, 2), (5, 3)])
def test_oracle_vdim_3d_greens_representation(inti, ov, pde, k, qorder, order):
    """3D VDIM on the tetrahedral cube (27 cells x 6 tetrahedra, ten or twenty
    Vioreanu-Rokhlin nodes each, interpolation order 2 or 3): at the nodes of
    the inner region |x|_inf < 0.3 the corrected volume potential meets
    Green's representation of the manufactured solution 20x better than the
    plain quadrature (the cube's edges and corners limit the boundary
    integrals elsewhere)."""
    Vf, vref, V0, inner = _manufactured_3d(inti, ov, 3, pde, k, qorder, order)
    ec, ep = np.abs(Vf - vref)[inner].max(), np.abs(V0 - vref)[inner].max()
    assert ec < 5e-4 and ep > 20 * ec, (ec, ep)


def _mixed_disk(inti, nrings, qorders=(4, 5), split=None):
    """The disk of disk_quadratures(nrings) with its first `split` triangles
    (default: half) on one Vioreanu-Rokhlin rule and the rest on another: a
    volume quadrature with two element types."""
    Vq, Bq = inti.disk_quadratures(nrings, qorder=5, boundary_qorder=10)
    E = Vq.elements
    s = len(E) // 2 if split is None else split
    V2 = inti.join_quadratures(inti.area_quadrature(E[:s], qorders[0]), inti.area_quadrature(E[s:], qorders[1]))
    return V2, Bq


def _manufactured(inti, ov, nrings, pde, k, order=3, mixed=False):
    """(corrected V f, Green-representation reference, uncorrected V f) at
    the volume nodes for u = cos(k0 theta.x), f = (k^2 - k0^2) u."""
    from oracle import csr_matvec, matvec
    if mixed:
        Vq, Bq = _mixed_disk(inti, nrings)
    else:
        Vq, Bq = inti.disk_quadratures(nrings, qorder=5, boundary_qorder=10)
    m = len(Vq)
    mu = np.full(m, -1.0)
    k0, th = np.pi, np.array([np.cos(np.pi / 3), np.sin(np.pi / 3)])
    u = lambda x: np.cos(k0 * (th @ x))
    f = (k * k - k0 * k0) * u(Vq.coords)
    S, D = ov.boundary_operators(pde, k, Vq, Bq, mu)
    rp, ci, val = ov.vdim(pde, k, Vq, Vq, Bq, mu, order, S, D)
    V0 = matvec(pde, 2, 0, k, Vq, Vq, f)
    Vf = V0 + csr_matvec(m, m, rp, ci, val, f)
    du = -k0 * (th @ Bq.normals) * np.sin(k0 * (th @ Bq.coords))
    vref = -u(Vq.coords) - D(u(Bq.coords)) + S(du)
    return Vf, vref, V0


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
def test_oracle_vdim_converges_to_greens_representation(inti, ov, pde, k):
    """V f + dV f = -u - D u + S du/dn inside (the check of
    test/convergence/vdim_potential_script.jl): the corrected error falls at
    high order (h^5 or better) while the plain quadrature stays O(h)."""
    e = []
    for nr in (2, 4):
        Vf, vref, V0 = _manufactured(inti, ov, nr, pde, k)
        e.append(np.abs(Vf - vref).max())
        assert np.abs(V0 - vref).max() > 1e3 * e[-1]
    assert e[1] < 5e-6
    assert e[0] / e[1] > 2 ** 5


@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
def test_oracle_vdim_element_types(inti, ov, pde, k):
    """A disk meshed with two element types (6- and 10-node triangle rules,
    vdim.jl:76-121 loops over the types): the corrected volume potential
    meets Green's representation of the manufactured solution far better than
    the plain quadrature, and converges under refinement."""
    e = []
    for nr in (2, 4):
        Vf, vref, V0 = _manufactured(inti, ov, nr, pde, k, mixed=True)
        e.append(np.abs(Vf - vref).max())
        assert np.abs(V0 - vref).max() > 1e2 * e[-1]
    assert e[1] < 1e-4 and e[0] / e[1] > 2 ** 4, e


# ---------------------------------------------------------------- GPU ----

def _csr(corr):
    M = corr.to_scipy().tocsr()
    M.sort_indices()
    return M


def _oracle_csr(m, n, rp, ci, val):
    import scipy.sparse as sp
    M = sp.csr_matrix((val, ci, rp), shape=(m, n))
    M.sort_indices()
    return M


@pytest.mark.gpu
@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
@pytest.mark.parametrize("case,order", [("self", 3), ("on", 3), ("self", 2), ("self", 4)])
def test_vdim_correction_matches_oracle(inti, ov, pde, k, case, order):
    """CSR pattern bit-exact; weights within 10x the oracle's own rounding
    spread (default vs FMA-contracted boundary correction) plus 1e-12 of
    their size. order 2 gives a wide system (6 basis functions, 10 nodes:
    minimum-norm weights), order 4 a tall one (15 x 10: least squares)."""
    print(f"vdim parity pde={pde} case={case} order={order}", flush=True)
    Vq, Bq = inti.disk_quadratures(3, qorder=5, boundary_qorder=10)
    op = _op(inti, pde, k)
    trg = Vq if case == "self" else Bq
    mu = np.full(len(trg), -1.0 if case == "self" else -0.5)
    maxdist = np.inf if case == "self" else 0.4
    # near lists from the oracle on both sides: nodes on the disk's symmetry
    # axes are equidistant from two boundary nodes, and the tie-break of a
    # nearest-neighbour search is implementation-defined
    nearB = None if trg is Bq else ov.nearest_lists(trg, Bq, maxdist)
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), trg, Bq)
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), trg, Bq)
    dS, dD = inti.bdim_correction(op, trg, Bq, Sop, Dop, green_multiplier=mu, near=nearB)
    Sop.attach(dS)
    Dop.attach(dD)
    Vop = inti.IntegralOperator(inti.SingleLayerKernel(op), trg, Vq)
    near = None if trg is Vq else ov.nearest_lists(trg, Vq, maxdist)
    dV = inti.vdim_correction(op, trg, Vq, Bq, Sop, Dop, Vop, green_multiplier=mu,
                              interpolation_order=order, maxdist=maxdist, near=near)
    G = _csr(dV)
    res = []
    for variant in (None, "fma"):
        oS, oD = ov.boundary_operators(pde, k, trg, Bq, mu, maxdist=maxdist,
                                       self_pattern=trg is Bq, variant=variant)
        rp, ci, val = ov.vdim(pde, k, trg, Vq, Bq, mu, order, oS, oD, near=near)
        res.append(_oracle_csr(len(trg), len(Vq), rp, ci, val))
    O, F = res
    assert np.array_equal(G.indptr, O.indptr) and np.array_equal(G.indices, O.indices)
    g = G.data.astype(np.complex128)
    spread = np.abs(F.data - O.data).max()
    assert np.abs(g - O.data).max() <= 10 * spread + 1e-12 * np.abs(O.data).max()


@pytest.mark.gpu
@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
@pytest.mark.parametrize("qorder,order", [(3, 2), (5, 3)])
def test_vdim_correction_3d_matches_oracle(inti, ov, pde, k, qorder, order):
    """3D VDIM on the tetrahedral cube (volume targets, interpolation order 2
    on ten-node or 3 on twenty-node tetrahedra: square Vandermonde systems):
    CSR pattern bit-exact; weights within 10x the oracle's own rounding spread
    (default vs FMA-contracted boundary correction) plus 1e-12 of their size;
    the corrected volume potential on the manufactured solution as accurate
    as the oracle's in the inner region. Helmholtz values are fixed by the
    mathematics (the polynomial particular solution is unique); Laplace ones
    follow the |x|^2 q particular solution (the reference's choice of P is
    not pinned)."""
    from oracle import csr_matvec, matvec
    print(f"vdim 3D parity pde={pde} qorder={qorder} order={order}", flush=True)
    Vq, Bq = inti.box_quadratures(3, qorder=qorder, boundary_qorder=4)
    op = _op(inti, pde, k, 3)
    m = len(Vq)
    mu = np.full(m, -1.0)
    nearB = ov.nearest_lists(Vq, Bq)  # the oracle's tie-break on both sides
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), Vq, Bq)
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), Vq, Bq)
    dS, dD = inti.bdim_correction(op, Vq, Bq, Sop, Dop, green_multiplier=mu, near=nearB)
    Sop.attach(dS)
    Dop.attach(dD)
    Vop = inti.IntegralOperator(inti.SingleLayerKernel(op), Vq, Vq)
    dV = inti.vdim_correction(op, Vq, Vq, Bq, Sop, Dop, Vop, green_multiplier=mu,
                              interpolation_order=order)
    G = _csr(dV)
    res = []
    for variant in (None, "fma"):
        oS, oD = ov.boundary_operators(pde, k, Vq, Bq, mu, variant=variant, near=nearB)
        rp, ci, val = ov.vdim(pde, k, Vq, Vq, Bq, mu, order, oS, oD)
        res.append(_oracle_csr(m, m, rp, ci, val))
    O, F = res
    assert np.array_equal(G.indptr, O.indptr) and np.array_equal(G.indices, O.indices)
    g = G.data.astype(np.complex128)
    spread = np.abs(F.data - O.data).max()
    err = np.abs(g - O.data).max()
    print(f"  weights: GPU vs oracle {err:.2e}, oracle spread {spread:.2e}, size {np.abs(O.data).max():.2e}",
          flush=True)
    assert err <= 10 * spread + 1e-12 * np.abs(O.data).max()
    # the corrected volume potential of the manufactured solution
    Vf_o, vref, V0, inner = _manufactured_3d(inti, ov, 3, pde, k, qorder, order)
    k0, th = np.pi, np.array([1.0, 2.0, 2.0]) / 3
    f = (k * k - k0 * k0) * np.cos(k0 * (th @ Vq.coords))
    Vop.attach(dV)
    Vf = Vop @ f
    e_gpu, e_or = np.abs(Vf - vref)[inner].max(), np.abs(Vf_o - vref)[inner].max()
    print(f"  inner error: GPU {e_gpu:.2e}, oracle {e_or:.2e}", flush=True)
    assert e_gpu <= 2 * e_or + 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
def test_volume_potential_greens_representation(inti, ov, pde, k):
    """volume_potential (matrix-free, corrected) on the manufactured solution:
    as accurate as the oracle's corrected V f (the two may pick different
    elements for targets tied between two nearest nodes, so they agree to the
    discretisation error, not to rounding); dense and matrix-free agree."""
    print(f"volume_potential pde={pde}", flush=True)
    Vq, Bq = inti.disk_quadratures(4, qorder=5, boundary_qorder=10)
    op = _op(inti, pde, k)
    V = inti.volume_potential(op=op, target=Vq, source=Vq,
                              correction={"method": "dim", "boundary": Bq, "interpolation_order": 3})
    k0, th = np.pi, np.array([np.cos(np.pi / 3), np.sin(np.pi / 3)])
    f = (k * k - k0 * k0) * np.cos(k0 * (th @ Vq.coords))
    Vf = V @ f
    Of, vref, _ = _manufactured(inti, ov, 4, pde, k)
    assert np.abs(Vf - vref).max() < max(5e-6, 2 * np.abs(Of - vref).max())
    Vd = inti.volume_potential(op=op, target=Vq, source=Vq, compression={"method": "none"},
                               correction={"method": "dim", "boundary": Bq, "interpolation_order": 3})
    np.testing.assert_allclose(Vd @ f, Vf, rtol=0, atol=1e-12 * np.abs(Vf).max())


@pytest.mark.gpu
def test_poisson_example(inti):
    """docs/src/examples/poisson.jl on the GPU: Delta u = f in the elliptical
    disk (2 x 1, quasi-uniform P2 mesh of size 0.1 with curved boundary
    edges), u = cos(2x) sin(2y); sigma from (-I/2 + D) sigma = V f + g by
    GMRES (abstol 1e-10), then u_h = -V f + D sigma; the example asserts
    norm(er) < 1e-5. The volume potentials build their boundary from the
    volume mesh at twice the volume order (qorder 10, six Gauss nodes per
    edge), as the reference's volume_potential does when no boundary is
    given; the boundary operators here use that same rule. With the example's
    three-node boundary rule (qorder 5) for the boundary operators, the
    near-boundary evaluation of D_b2d on these meshes leaves norm(er) about
    1e-4 (the oracle gives the same figure: the error is the discretisation's,
    concentrated at volume nodes about 0.006 from the boundary), so that
    variant is not asserted."""
    print("poisson example", flush=True)
    meshsize = 0.1
    Vq, Bq = inti.ellipse_disk_quadratures(meshsize, a=2.0, b=1.0, qorder=5, boundary_qorder=10)
    op = inti.Laplace(dim=2)
    S_b2b, D_b2b = inti.single_double_layer(op=op, target=Bq, source=Bq,
                                            compression={"method": "gpu"},
                                            correction={"method": "dim"})
    S_b2d, D_b2d = inti.single_double_layer(op=op, target=Vq, source=Bq,
                                            compression={"method": "gpu"},
                                            correction={"method": "dim", "maxdist": 5 * meshsize,
                                                        "target_location": "inside"})
    V_d2d = inti.volume_potential(op=op, target=Vq, source=Vq, compression={"method": "gpu"},
                                  correction={"method": "dim", "interpolation_order": 3})
    V_d2b = inti.volume_potential(op=op, target=Bq, source=Vq, compression={"method": "gpu"},
                                  correction={"method": "dim", "maxdist": 5 * meshsize,
                                              "interpolation_order": 3, "target_location": "on"})
    ue = lambda x: np.cos(2 * x[0]) * np.sin(2 * x[1])
    f = -8 * ue(Vq.coords)
    g = ue(Bq.coords)
    rhs = V_d2b @ f + g
    L = D_b2b.iop
    L.set_identity(-0.5)
    sigma, info = inti.gmres(L, rhs, restart=100, maxiter=100, rtol=0.0, atol=1e-10)
    assert info["iterations"] < 100
    uh = -(V_d2d @ f) + D_b2d @ sigma
    er = np.abs(uh - ue(Vq.coords))
    print(f"poisson: nodes {len(Vq)}, gmres its {info['iterations']}, norm(er) {np.linalg.norm(er):.3e}, "
          f"max {er.max():.3e}", flush=True)
    assert np.linalg.norm(er) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("pde,k", [(LAP, 0.0), (HEL, 2.0)])
@pytest.mark.parametrize("order", [2, 3])
def test_vdim_element_types_match_oracle(inti, ov, pde, k, order):
    """Two volume element types of 6 and 10 nodes (order 2: a square and a
    wide system; order 3: a tall and a square one), one least-squares launch
    per type: CSR pattern bit-exact, weights within 10x the oracle's own
    rounding spread plus 1e-12 of their size."""
    print(f"vdim element types pde={pde} order={order}", flush=True)
    Vq, Bq = _mixed_disk(inti, 3)
    assert [q.shape[0] for q in Vq.etype_qtags] == [6, 10]
    op = _op(inti, pde, k)
    m = len(Vq)
    mu = np.full(m, -1.0)
    nearB = ov.nearest_lists(Vq, Bq)
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), Vq, Bq)
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), Vq, Bq)
    dS, dD = inti.bdim_correction(op, Vq, Bq, Sop, Dop, green_multiplier=mu, near=nearB)
    Sop.attach(dS)
    Dop.attach(dD)
    Vop = inti.IntegralOperator(inti.SingleLayerKernel(op), Vq, Vq)
    dV = inti.vdim_correction(op, Vq, Vq, Bq, Sop, Dop, Vop, green_multiplier=mu, interpolation_order=order)
    G = _csr(dV)
    res = []
    for variant in (None, "fma"):
        oS, oD = ov.boundary_operators(pde, k, Vq, Bq, mu, variant=variant, near=nearB)
        rp, ci, val = ov.vdim(pde, k, Vq, Vq, Bq, mu, order, oS, oD)
        res.append(_oracle_csr(m, m, rp, ci, val))
    O, F = res
    assert np.array_equal(G.indptr, O.indptr) and np.array_equal(G.indices, O.indices)
    g = G.data.astype(np.complex128)
    spread = np.abs(F.data - O.data).max()
    assert np.abs(g - O.data).max() <= 10 * spread + 1e-12 * np.abs(O.data).max()
    # targets off the volume nodes: the boundary nodes, one near list per type
    near = ov.nearest_lists(Bq, Vq, 0.4)
    assert len(near) == 2
    muB = np.full(len(Bq), -0.5)
    S2 = inti.IntegralOperator(inti.SingleLayerKernel(op), Bq, Bq)
    D2 = inti.IntegralOperator(inti.DoubleLayerKernel(op), Bq, Bq)
    eS, eD = inti.bdim_correction(op, Bq, Bq, S2, D2, green_multiplier=muB)
    S2.attach(eS)
    D2.attach(eD)
    V2 = inti.IntegralOperator(inti.SingleLayerKernel(op), Bq, Vq)
    dV2 = inti.vdim_correction(op, Bq, Vq, Bq, S2, D2, V2, green_multiplier=muB, interpolation_order=order,
                               near=near)
    G2 = _csr(dV2)
    res = []
    for variant in (None, "fma"):
        oS, oD = ov.boundary_operators(pde, k, Bq, Bq, muB, self_pattern=True, variant=variant)
        rp, ci, val = ov.vdim(pde, k, Bq, Vq, Bq, muB, order, oS, oD, near=near)
        res.append(_oracle_csr(len(Bq), m, rp, ci, val))
    O2, F2 = res
    assert np.array_equal(G2.indptr, O2.indptr) and np.array_equal(G2.indices, O2.indices)
    spread = np.abs(F2.data - O2.data).max()
    assert np.abs(G2.data - O2.data).max() <= 10 * spread + 1e-12 * np.abs(O2.data).max()
