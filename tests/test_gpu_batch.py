"""Batched problems (one persistent launch, batched hipGraph GMRES) against
the single-operator path and dense solves of each problem."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ellipses(inti, sizes, seed=3):
    rng = np.random.default_rng(seed)
    qs, ks = [], []
    for ne in sizes:
        a, b = rng.uniform(1, 3), rng.uniform(0.3, 1)
        qs.append(inti.ellipse_quadrature(a, b, ne, qorder=11))  # 6 Gauss nodes per element
        ks.append(rng.uniform(1, 12))
    return qs, ks


def _rand(n, r, cplx, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, r))
    return x + 1j * rng.normal(size=(n, r)) if cplx else x


@pytest.mark.parametrize("case", ["helm2_dl", "helm2_comb", "lap3_sl", "helm3_dl"])
@pytest.mark.parametrize("nrhs", [16, 5])
def test_batch_apply_matches_single_operators(inti, case, nrhs):
    """Every problem's rows of the batched apply equal its own operator's
    apply (FP64, 1e-13), with per-problem identity and a correction."""
    print(f"batch apply {case} nrhs={nrhs}", flush=True)
    if case.startswith("helm2"):
        qs, ks = _ellipses(inti, [9, 60, 23, 130, 41])
        ops = [inti.Helmholtz(k=k, dim=2) for k in ks]
    elif case == "lap3_sl":
        qs = [inti.sphere_quadrature(f) for f in (2, 5, 3)]
        ops = [inti.Laplace(dim=3) for _ in qs]
    else:
        qs = [inti.sphere_quadrature(f) for f in (2, 5, 3)]
        ops = [inti.Helmholtz(k=k, dim=3) for k in (1.0, 3.0, 2.0)]
    if case == "helm2_comb":
        Ks = [inti.CombinedFieldKernel(o, -1j * o.k, 1.0) for o in ops]
    elif case == "lap3_sl":
        Ks = [inti.SingleLayerKernel(o) for o in ops]
    else:
        Ks = [inti.DoubleLayerKernel(o) for o in ops]
    Bt = inti.BatchOperator(Ks, qs)
    cplx = Bt.dtype == np.complex128
    X = _rand(Bt.ntot, nrhs, cplx, 1)
    cid = np.arange(1, len(qs) + 1) * (0.5 - 0.25j if cplx else 0.5)
    Bt.set_identity(cid)
    singles = []
    for p, (K, Q) in enumerate(zip(Ks, qs)):
        A = inti.IntegralOperator(K, Q)
        A.set_identity(cid[p])
        singles.append(A)
    # a correction on problem 1 (the DIM one of its own SL/DL pair)
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(ops[1]), qs[1])
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(ops[1]), qs[1])
    mu = np.full(len(qs[1]), -0.5)
    dS, dD = inti.bdim_correction(ops[1], qs[1], qs[1], Sop, Dop, green_multiplier=mu)
    Bt.attach(1, dD, 0.75)
    singles[1].attach(dD, 0.75)
    Y = Bt.apply(X)
    for p, A in enumerate(singles):
        ref = A.apply(Bt.split(X)[p])
        got = Bt.split(Y)[p]
        assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref), p
    # alpha / beta
    Y0 = _rand(Bt.ntot, nrhs, cplx, 2)
    Y2 = Bt.apply(X, Y0, alpha=2.0, beta=-1.0)
    assert np.linalg.norm(Y2 - (2 * Y - Y0)) <= 1e-13 * np.linalg.norm(Y2)


@pytest.mark.parametrize("kind", ["DL", "COMB"])
def test_batch_apply_matches_oracle(inti, oracle, kind):
    """Every problem of a 6-ellipse batch with k up to 40 (kd up to ~200)
    against the oracle's direct sum on all rows, 16 right-hand sides:
    FP64 <= 1e-12, MIXED <= 1e-6 (src/kernels.jl:360-371)."""
    print(f"batch vs oracle {kind}", flush=True)
    rng = np.random.default_rng(11)
    qs, ks = [], []
    for ne, k in ((50, 40.0), (300, 33.0), (17, 1.0), (140, 12.5), (333, 25.0), (90, 38.0)):
        qs.append(inti.ellipse_quadrature(rng.uniform(1, 3), rng.uniform(0.3, 1), ne, qorder=11))
        ks.append(k)
    ops = [inti.Helmholtz(k=k, dim=2) for k in ks]
    if kind == "DL":
        Ks = [inti.DoubleLayerKernel(o) for o in ops]
        coefs = [None] * len(ops)
    else:
        Ks = [inti.CombinedFieldKernel(o, -1j * o.k, 1.0) for o in ops]
        coefs = [(-1j * o.k, 1.0) for o in ops]
    B64 = inti.BatchOperator(Ks, qs)
    B32 = inti.BatchOperator(Ks, qs, precision="mixed")
    X = _rand(B64.ntot, 16, True, 12)
    Y64, Y32 = B64.split(B64.apply(X)), B32.split(B32.apply(X))
    for p, (Q, k) in enumerate(zip(qs, ks)):
        ref = oracle.matvec(1, 2, 1 if kind == "DL" else 4, k, Q, Q, B64.split(X)[p], coefs=coefs[p])
        e64 = np.linalg.norm(Y64[p] - ref) / np.linalg.norm(ref)
        e32 = np.linalg.norm(Y32[p] - ref) / np.linalg.norm(ref)
        print(f"  problem {p}: N={len(Q)} k={k}: fp64 {e64:.2e}, mixed {e32:.2e}", flush=True)
        assert e64 <= 1e-12 and e32 <= 1e-6, (p, e64, e32)


def test_batch_mixed_precision(inti):
    print("batch mixed", flush=True)
    qs, ks = _ellipses(inti, [50, 300, 120])
    Ks = [inti.DoubleLayerKernel(inti.Helmholtz(k=k, dim=2)) for k in ks]
    B64 = inti.BatchOperator(Ks, qs)
    B32 = inti.BatchOperator(Ks, qs, precision="mixed")
    X = _rand(B64.ntot, 16, True, 4)
    Y64, Y32 = B64.apply(X), B32.apply(X)
    assert np.linalg.norm(Y32 - Y64) <= 1e-6 * np.linalg.norm(Y64)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_gmres_matches_dense_solves(inti, use_graph):
    """(I/2 + D + dD) x = b for 5 ellipses x 16 right-hand sides in one
    batched solve; each system against numpy's dense solve of its problem
    (1e-10 relative), and every relative residual below the tolerance."""
    print(f"batch gmres graph={use_graph}", flush=True)
    qs, ks = _ellipses(inti, [20, 70, 35, 90, 50], seed=8)
    ops = [inti.Helmholtz(k=k, dim=2) for k in ks]
    Ks = [inti.DoubleLayerKernel(o) for o in ops]
    Bt = inti.BatchOperator(Ks, qs)
    Bt.set_identity(0.5)
    dense = []
    for p, (o, Q) in enumerate(zip(ops, qs)):
        Sop = inti.IntegralOperator(inti.SingleLayerKernel(o), Q)
        Dop = inti.IntegralOperator(inti.DoubleLayerKernel(o), Q)
        dS, dD = inti.bdim_correction(o, Q, Q, Sop, Dop, green_multiplier=np.full(len(Q), -0.5))
        Bt.attach(p, dD)
        dense.append(0.5 * np.eye(len(Q)) + inti.assemble_matrix(Dop) + dD.to_scipy().toarray())
    b = _rand(Bt.ntot, 16, True, 9)
    X, info = Bt.gmres(b, restart=60, maxiter=300, rtol=1e-12, use_graph=use_graph)
    assert info["relres"].max() <= 1e-12
    for p, M in enumerate(dense):
        ref = np.linalg.solve(M, Bt.split(b)[p])
        assert np.linalg.norm(Bt.split(X)[p] - ref) <= 1e-10 * np.linalg.norm(ref), p
    print(f"batch gmres: {Bt.nprob} problems x 16 rhs, {info['iterations']} steps, "
          f"max relres {info['relres'].max():.2e}", flush=True)


def test_batch_refined_solve_matches_dense_solves(inti):
    """Iterative refinement of a batch (FP64 residuals, GMRES corrections
    with the mixed-precision batch of the same problems): every system
    against numpy's dense solve of its problem (1e-10 relative) with every
    FP64 residual below 1e-12."""
    print("batch refined solve", flush=True)
    qs, ks = _ellipses(inti, [20, 70, 35, 90, 50], seed=8)
    ops = [inti.Helmholtz(k=k, dim=2) for k in ks]
    Ks = [inti.DoubleLayerKernel(o) for o in ops]
    B64 = inti.BatchOperator(Ks, qs)
    B32 = inti.BatchOperator(Ks, qs, precision="mixed")
    dense = []
    for Bt in (B64, B32):
        Bt.set_identity(0.5)
    for p, (o, Q) in enumerate(zip(ops, qs)):
        Sop = inti.IntegralOperator(inti.SingleLayerKernel(o), Q)
        Dop = inti.IntegralOperator(inti.DoubleLayerKernel(o), Q)
        dS, dD = inti.bdim_correction(o, Q, Q, Sop, Dop, green_multiplier=np.full(len(Q), -0.5))
        B64.attach(p, dD)
        B32.attach(p, dD)
        dense.append(0.5 * np.eye(len(Q)) + inti.assemble_matrix(Dop) + dD.to_scipy().toarray())
    b = _rand(B64.ntot, 16, True, 9)
    X, info = B64.refine_solve(B32, b, restart=60, maxiter_inner=200, inner_rtol=1e-6, rtol=1e-12)
    print(f"batch refined: {info['outer']} outer, {info['inner_iterations']} inner, "
          f"max relres {info['relres'].max():.2e}", flush=True)
    assert info["relres"].max() <= 1e-12
    assert info["outer"] >= 2  # the mixed corrections alone stop near 1e-7
    for p, M in enumerate(dense):
        ref = np.linalg.solve(M, B64.split(b)[p])
        assert np.linalg.norm(B64.split(X)[p] - ref) <= 1e-10 * np.linalg.norm(ref), p


def test_batch_gmres_graph_is_deterministic(inti):
    print("batch gmres determinism", flush=True)
    qs, ks = _ellipses(inti, [30, 64, 45], seed=11)
    Ks = [inti.CombinedFieldKernel(inti.Helmholtz(k=k, dim=2), -1j * k, 1.0) for k in ks]
    Bt = inti.BatchOperator(Ks, qs)
    Bt.set_identity(0.5)
    b = _rand(Bt.ntot, 4, True, 12)
    X1, i1 = Bt.gmres(b, restart=20, maxiter=60, rtol=1e-9)
    X2, i2 = Bt.gmres(b, restart=20, maxiter=60, rtol=1e-9)
    X3, i3 = Bt.gmres(b, restart=20, maxiter=60, rtol=1e-9, use_graph=False)
    assert np.array_equal(X1, X2) and i1["iterations"] == i2["iterations"]
    assert np.array_equal(X1, X3)


def test_batch_autotune_keeps_results(inti):
    print("batch autotune", flush=True)
    qs, ks = _ellipses(inti, [200, 500, 90, 350], seed=5)
    Ks = [inti.DoubleLayerKernel(inti.Helmholtz(k=k, dim=2)) for k in ks]
    Bt = inti.BatchOperator(Ks, qs)
    X = _rand(Bt.ntot, 16, True, 6)
    Y0 = Bt.apply(X)
    t = Bt.autotune(nrhs=16, reps=2)
    assert t["chunk_len"] in (1024, 2048, 4096, 8192) and t["ms"] > 0
    Y1 = Bt.apply(X)
    assert np.linalg.norm(Y1 - Y0) <= 1e-13 * np.linalg.norm(Y0)


def test_batch_error_paths(inti):
    Q = inti.circle_quadrature(10, qorder=5)
    op = inti.Helmholtz(k=1.0, dim=2)
    with pytest.raises(inti.IntiError):
        inti.BatchOperator([inti.HyperSingularKernel(op)], [Q])
    Bt = inti.BatchOperator([inti.DoubleLayerKernel(op)], [Q])
    with pytest.raises(inti.IntiError):
        Bt.gmres(np.ones((len(Q), 3), dtype=complex))
    with pytest.raises(ValueError):
        inti.BatchOperator([inti.DoubleLayerKernel(op), inti.SingleLayerKernel(op)], [Q, Q])


@pytest.mark.parametrize("case", ["helm2", "lap2", "helm3", "lap3"])
def test_batch_bdim_matches_per_problem_builds(inti, oracle, case):
    """The DIM corrections of all problems at once (traces in one launch,
    Theta in one launch over (problem, target tile) work items, the least
    squares of all elements in one launch) against the oracle's build of each problem: pattern bit-exact;
    the corrected operator on a smooth density within 1e-12 of the oracle's
    (or 10x the oracle's own rounding spread, with and without FMA), and the
    weights within the spread the single-problem build is held to."""
    print(f"batch bdim {case}", flush=True)
    LAP, HEL = 0, 1
    if case == "lap2":
        qs, _ = _ellipses(inti, [9, 40, 23], seed=5)
        ops = [inti.Laplace(dim=2) for _ in qs]
    elif case == "helm2":
        qs, ks = _ellipses(inti, [9, 60, 23, 41], seed=7)
        ops = [inti.Helmholtz(k=k, dim=2) for k in ks]
    elif case == "lap3":
        qs = [inti.sphere_quadrature(f) for f in (3, 1, 2)]
        ops = [inti.Laplace(dim=3) for _ in qs]
    else:
        qs = [inti.sphere_quadrature(f) for f in (2, 3, 1)]
        ops = [inti.Helmholtz(k=k, dim=3) for k in (1.5, 2.5, 0.7)]
    dim = ops[0].dim
    pde = LAP if isinstance(ops[0], inti.Laplace) else HEL
    Bt = inti.BatchOperator([inti.DoubleLayerKernel(o) for o in ops], qs)
    got = Bt.bdim_correction()
    for p, (op, Q) in enumerate(zip(ops, qs)):
        n = len(Q)
        mu = np.full(n, -0.5)
        k = op.k
        orp, oci, ovS, ovD, _ = oracle.bdim(pde, dim, k, Q, Q, mu)
        _, _, fvS, fvD, _ = oracle.bdim(pde, dim, k, Q, Q, mu, variant="fma")
        if pde == LAP:
            ovS, ovD, fvS, fvD = ovS.real, ovD.real, fvS.real, fvD.real
        rp, ci, vS = got[p][0].export()
        _, _, vD = got[p][1].export()
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci), p
        for v, o, f in ((vS, ovS, fvS), (vD, ovD, fvD)):
            spread = np.abs(f - o).max()
            assert np.abs(v - o).max() <= 10 * spread + 1e-14 * np.abs(o).max(), p
        x = np.exp(np.sin(Q.coords[0]) + 0.5 * Q.coords[1])
        if pde == HEL:
            x = x * np.exp(0.3j * Q.coords[0])
        for kind, v, o, f in ((1, vD, ovD, fvD), (0, vS, ovS, fvS)):
            y0 = oracle.matvec(pde, dim, kind, k, Q, Q, x)
            y = y0 + oracle.csr_matvec(n, n, rp, ci, v, x)
            ref = y0 + oracle.csr_matvec(n, n, orp, oci, o, x)
            alt = y0 + oracle.csr_matvec(n, n, orp, oci, f, x)
            spread = np.linalg.norm(alt - ref) / np.linalg.norm(ref)
            assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= max(1e-12, 10 * spread), (p, kind)
