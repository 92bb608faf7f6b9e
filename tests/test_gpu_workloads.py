"""Parity at the full sizes of the five workloads, through the C-ABI, against
the oracle on the same seeded inputs (the builders are bench.py's).

  C1  2D Laplace DL, unit circle, 128 P2 x 5 Gauss = 640 nodes,
      sigma = cos(3 theta) + 0.1 N(0,1): bdim + assemble + matvec.
  C2  3D Helmholtz combined field D - i k S, k = 2 pi, unit sphere,
      icosphere 23: 63,480 nodes, complex Gaussian density.
  C3  256 2D Helmholtz DL problems on ellipses (k up to 40, kd up to ~240),
      I/2 + D + dD, 16 complex right-hand sides, one batched apply.
  C4  3D Laplace SL, bumpy sphere, 491,520 nodes (the automatic tile-group
      choice of the symmetric sum: its partial slots exceed 4 GiB at G = 1).
  C5  far-field objective of the sound-soft k = 10 scattering problem,
      100,920 nodes: pullbacks against the oracle at full size and central
      differences (h = 1e-6) along 20 seeded directions.

Bars (each stated where it is asserted):
  * the direct sum (no correction) on the sampled rows: <= 1e-12 relative;
  * the correction's sparsity pattern: bit-exact;
  * DIM weights: backward error eta <= 1e-13 against the oracle's own
    least-squares data (tests/_dimcheck.py explains why the forward error of
    these ill-conditioned solves is not a usable bar);
  * corrected rows: <= max(1e-12, 10 x the oracle's own rounding spread
    (built with FMA contraction vs without), the spread of the rows under
    perturbations of Theta at the size of FP64 summation errors of its
    n-term sums, sqrt(n) eps tau (tests/_dimcheck.py: theta_spread));
  * MIXED: <= 1e-6 against FP64."""
import sys

import numpy as np
import pytest

from tests import _dimcheck as dc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bench():
    import bench as b
    return b


def _sampled_dim_check(oracle, pde, dim, k, Q, mu, elems, corr, label):
    """Pattern, backward error and oracle CSR of the sampled elements."""
    near = dc.near_lists(Q, elems)
    rows = dc.rows_of(Q, elems)
    rp, ci, vS, vD, _ = oracle.bdim(pde, dim, k, Q, Q, mu, near=near)
    rpf, cif, vSf, vDf, _ = oracle.bdim(pde, dim, k, Q, Q, mu, near=near, variant="fma")
    grpS, gciS, gvS = corr[0].export()
    grpD, gciD, gvD = corr[1].export()
    assert dc.same_pattern(grpS, gciS, rp, ci, rows) and dc.same_pattern(grpD, gciD, rp, ci, rows), label
    lrows, A, th, tau = dc.ls_data(oracle, pde, dim, k, Q, mu, elems)
    assert np.array_equal(lrows, rows)
    eta_gpu = dc.weights_backward_error(Q, elems, A, th, tau, rows, grpS, gciS, gvS, grpD, gciD, gvD)
    eta_or = dc.weights_backward_error(Q, elems, A, th, tau, rows, rp, ci, vS, rp, ci, vD)
    say(f"  {label}: weights backward error GPU max {eta_gpu.max():.2e} (oracle {eta_or.max():.2e}), "
        f"{len(rows)} rows of {len(elems)} elements")
    assert eta_gpu.max() <= dc.ETA_BAR, eta_gpu.max()
    return rows, (rp, ci, vS, vD), (rpf, cif, vSf, vDf), (grpS, gciS, gvS), (grpD, gciD, gvD), (A, tau)


# ------------------------------------------------------------------- C1
def test_c1_circle_pipeline(inti, oracle, bench):
    """The whole 640-node pipeline (bdim_correction + assemble_matrix +
    matvec, dense, as single_double_layer with compression none). The
    seeded density's corrected matvec is compared with the extended-
    precision result of the same algorithm (tests/_extended.py): the GPU
    must be as close to it as the oracle is (within 2x the oracle's own
    error, itself ~6e-5 here because the DIM least squares have R-diagonal
    ratios ~5e8). Smooth density and uncorrected sum: 1e-12 vs the oracle."""
    from tests import _extended
    say("[gpu] C1 circle pipeline 640 nodes")
    Q, sig = bench.circle_workload(inti)
    n = len(Q)
    mu = np.full(n, -0.5)
    op = inti.Laplace(dim=2)
    S, D = inti.single_double_layer(op=op, target=Q, source=Q, compression={"method": "none"},
                                    correction={"method": "dim"})
    Dop = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)
    Sop = inti.IntegralOperator(inti.SingleLayerKernel(op), Q)
    corr = inti.bdim_correction(op, Q, Q, Sop, Dop, green_multiplier=mu)
    # pattern (every row) and weights' backward error (every element)
    elems = np.arange(Q.ne)
    rows, orc, orf, gS, gD, (A_, tau_) = _sampled_dim_check(oracle, 0, 2, 0.0, Q, mu, elems, corr, "C1")
    rp, ci, vS, vD = oracle.bdim(0, 2, 0.0, Q, Q, mu)[:4]
    assert np.array_equal(gD[0], rp) and np.array_equal(gD[1], ci)
    Am = oracle.assemble(0, 2, 1, 0.0, Q, Q).real
    Dplain = inti.assemble_matrix(Dop)
    assert np.abs(Dplain - Am).max() <= 1e-12 * np.abs(Am).max()
    th = np.arctan2(Q.coords[1], Q.coords[0])
    smooth = np.exp(np.cos(th))
    ys_or = Am @ smooth + oracle.csr_matvec(n, n, rp, ci, vD, smooth).real
    e_smooth = dc.rel(D @ smooth, ys_or)
    # seeded density
    y_gpu = D @ sig
    Dop.attach(corr[1])
    y_mv = Dop @ sig
    y_or = Am @ sig + oracle.csr_matvec(n, n, rp, ci, vD, sig).real
    zs = oracle.monopoles(2, Q.coords, Q.nq)
    y_ext = _extended.corrected_dl_laplace2d(Q, zs, mu, sig)
    e_or, e_gpu, e_mv = dc.rel(y_or, y_ext), dc.rel(y_gpu, y_ext), dc.rel(y_mv, y_ext)
    say(f"  C1: smooth density GPU vs oracle {e_smooth:.2e}; seeded density vs extended precision: "
        f"oracle {e_or:.2e}, GPU dense {e_gpu:.2e}, GPU matrix-free {e_mv:.2e}; GPU vs oracle "
        f"{dc.rel(y_gpu, y_or):.2e}")
    assert e_smooth <= 1e-12
    assert e_gpu <= 2 * e_or and e_mv <= 2 * e_or


# ------------------------------------------------------------------- C2
def test_c2_cfie_63480(inti, oracle, bench):
    say("[gpu] C2 CFIE 63,480 nodes")
    k = 2 * np.pi
    Q, x = bench.make_workload(inti, 23)
    n = len(Q)
    mu = np.full(n, -0.5)
    op = inti.Helmholtz(k=k, dim=3)
    L64, corr = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0)
    Lmx, _ = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, precision="mixed",
                                 corrections=corr)
    y64 = L64 @ x
    ymx = Lmx @ x
    elems = dc.sample_elements(Q.ne, 48, 2, must=(0, Q.ne - 1))
    rows, orc, orf, gS, gD, (A_, tau_) = _sampled_dim_check(oracle, 1, 3, k, Q, mu, elems, corr, "C2")
    T = dc.subset(Q, rows)
    direct = oracle.matvec(1, 3, 4, k, T, Q, x, coefs=(-1j * k, 1.0))
    c_gpu = dc.csr_rows_apply(gD[0], gD[1], gD[2] - 1j * k * gS[2], rows, x)
    e_direct = dc.rel(y64[rows] - c_gpu, direct)
    y_or = direct + dc.csr_rows_apply(orc[0], orc[1], orc[3] - 1j * k * orc[2], rows, x)
    y_orf = direct + dc.csr_rows_apply(orf[0], orf[1], orf[3] - 1j * k * orf[2], rows, x)
    spread = dc.rel(y_orf, y_or)
    tspread = dc.theta_spread(Q, elems, A_, tau_, rows, x, -1j * k, 1.0, y_or)
    e_corr = dc.rel(y64[rows], y_or)
    e_mx = dc.rel(ymx[rows], y_or)
    e_mx_full = dc.rel(ymx, y64)
    say(f"  C2: {len(rows)} rows: direct sum {e_direct:.2e}, corrected {e_corr:.2e} (oracle spread "
        f"{spread:.2e}, Theta spread {tspread:.2e}), mixed vs oracle {e_mx:.2e}, mixed vs FP64 (all rows) "
        f"{e_mx_full:.2e}")
    assert len(rows) >= 256
    assert e_direct <= 1e-12
    assert e_corr <= max(1e-12, 10 * spread, tspread)
    assert e_mx <= 1e-6 and e_mx_full <= 1e-6


# ------------------------------------------------------------------- C3
def test_c3_batch_256(inti, oracle, bench):
    say("[gpu] C3 batch of 256 problems")
    qs, a, b, ks, rng = bench.batch_workload(inti, 256)
    P = len(qs)
    ops = [inti.Helmholtz(k=ks[p], dim=2) for p in range(P)]
    corr = []
    for p in range(P):
        Sop = inti.IntegralOperator(inti.SingleLayerKernel(ops[p]), qs[p])
        Dop = inti.IntegralOperator(inti.DoubleLayerKernel(ops[p]), qs[p])
        corr.append(inti.bdim_correction(ops[p], qs[p], qs[p], Sop, Dop,
                                         green_multiplier=np.full(len(qs[p]), -0.5)))
    say("  C3: 256 DIM corrections built")
    Ks = [inti.DoubleLayerKernel(o) for o in ops]
    B64 = inti.BatchOperator(Ks, qs)
    B32 = inti.BatchOperator(Ks, qs, precision="mixed")
    for Bt in (B64, B32):
        Bt.set_identity(0.5)
        for p in range(P):
            Bt.attach(p, corr[p][1])
    R = 16
    X = (rng.normal(size=(B64.ntot, R)) + 1j * rng.normal(size=(B64.ntot, R))) * np.sqrt(0.5)
    Y64 = B64.apply(X)
    Y32 = B32.apply(X)
    e_mx_full = dc.rel(Y32, Y64)
    say(f"  C3: ntot {B64.ntot}, mixed vs FP64 (all problems, 16 RHS) {e_mx_full:.2e}")
    assert e_mx_full <= 1e-6
    sizes = np.diff(B64.offsets)
    kd = ks * 2 * np.maximum(a, b)
    pick = {int(np.argmax(kd)), int(np.argmax(sizes)), int(np.argmin(sizes))}
    pick.update(int(v) for v in rng.choice(P, 8, replace=False))
    pick = sorted(pick)
    Xs, Ys, Ym = B64.split(X), B64.split(Y64), B64.split(Y32)
    for p in pick:
        Q, k = qs[p], ks[p]
        mu = np.full(len(Q), -0.5)
        elems = dc.sample_elements(Q.ne, 12, p, must=(0,))
        rows, orc, orf, gS, gD, (A_, tau_) = _sampled_dim_check(oracle, 1, 2, k, Q, mu, elems, corr[p], f"C3 p={p}")
        T = dc.subset(Q, rows)
        direct = oracle.matvec(1, 2, 1, k, T, Q, Xs[p])
        c_gpu = dc.csr_rows_apply(gD[0], gD[1], gD[2], rows, Xs[p])
        e_direct = dc.rel(Ys[p][rows] - 0.5 * Xs[p][rows] - c_gpu, direct)
        e_mx_direct = dc.rel(Ym[p][rows] - 0.5 * Xs[p][rows] - c_gpu, direct)
        y_or = 0.5 * Xs[p][rows] + direct + dc.csr_rows_apply(orc[0], orc[1], orc[3], rows, Xs[p])
        y_orf = 0.5 * Xs[p][rows] + direct + dc.csr_rows_apply(orf[0], orf[1], orf[3], rows, Xs[p])
        spread = dc.rel(y_orf, y_or)
        tspread = dc.theta_spread(Q, elems, A_, tau_, rows, Xs[p], 0.0, 1.0, y_or)
        e_corr = dc.rel(Ys[p][rows], y_or)
        say(f"  C3 p={p}: N={len(Q)}, k={k:.2f}, kd={kd[p]:.1f}: direct sum {e_direct:.2e} (mixed "
            f"{e_mx_direct:.2e}), corrected {e_corr:.2e} (oracle spread {spread:.2e}, Theta spread "
            f"{tspread:.2e})")
        assert e_direct <= 1e-12 and e_mx_direct <= 1e-6
        assert e_corr <= max(1e-12, 10 * spread, tspread)
    assert len(pick) >= 8


# ------------------------------------------------------------------- C4
def test_c4_bumpy_491520(inti, oracle, bench):
    say("[gpu] C4 bumpy sphere 491,520 nodes")
    Q = bench.bumpy_quadrature(inti)
    n = len(Q)
    assert n == 491520
    mu = np.full(n, -0.5)
    op = inti.Laplace(dim=3)
    S64 = inti.IntegralOperator(inti.SingleLayerKernel(op), Q)
    D64 = inti.IntegralOperator(inti.DoubleLayerKernel(op), Q)
    corr = inti.bdim_correction(op, Q, Q, S64, D64, green_multiplier=mu)
    S64.attach(corr[0])
    S32 = inti.IntegralOperator(inti.SingleLayerKernel(op), Q, precision="mixed")
    S32.attach(corr[0])
    th = np.arccos(np.clip(Q.coords[2] / np.linalg.norm(Q.coords, axis=0), -1, 1))
    ph = np.arctan2(Q.coords[1], Q.coords[0])
    rhs = bench._real_harmonics(th, ph, 8, np.random.default_rng(bench.SEED))
    y64 = S64 @ rhs  # automatic symmetric sum, automatic tile group
    y32 = S32 @ rhs
    say("  C4: applies done")
    elems = dc.sample_elements(Q.ne, 48, 4, must=(0, Q.ne - 1))
    rows, orc, orf, gS, gD, (A_, tau_) = _sampled_dim_check(oracle, 0, 3, 0.0, Q, mu, elems, corr, "C4")
    T = dc.subset(Q, rows)
    direct = oracle.matvec(0, 3, 0, 0.0, T, Q, rhs).real
    c_gpu = dc.csr_rows_apply(gS[0], gS[1], gS[2], rows, rhs)
    e_direct = dc.rel(y64[rows] - c_gpu, direct)
    y_or = direct + dc.csr_rows_apply(orc[0], orc[1], orc[2], rows, rhs).real
    y_orf = direct + dc.csr_rows_apply(orf[0], orf[1], orf[2], rows, rhs).real
    spread = dc.rel(y_orf, y_or)
    tspread = dc.theta_spread(Q, elems, A_, tau_, rows, rhs, 1.0, 0.0, y_or)
    e_corr = dc.rel(y64[rows], y_or)
    e_mx, e_mx_full = dc.rel(y32[rows], y_or), dc.rel(y32, y64)
    say(f"  C4: {len(rows)} rows: direct sum {e_direct:.2e}, corrected {e_corr:.2e} (oracle spread "
        f"{spread:.2e}, Theta spread {tspread:.2e}), mixed vs oracle {e_mx:.2e}, mixed vs FP64 (all rows) "
        f"{e_mx_full:.2e}")
    assert e_direct <= 1e-12
    assert e_corr <= max(1e-12, 10 * spread, tspread)
    assert e_mx <= 1e-6 and e_mx_full <= 1e-6


# ------------------------------------------------------------------- C5
def test_c5_far_field_gradient_100920(inti, oracle, bench):
    """sigma from the corrected CFIE solve; J = sum_64 |u_inf|^2. The
    pullback of the far-field operator (sigma, nodes, normals, weights, k)
    against the oracle's at full size (<= 1e-12; k <= 1e-11), and against
    central differences of the GPU objective at h = 1e-6 along 20 seeded
    directions (<= 1e-5 relative)."""
    say("[gpu] C5 far-field objective 100,920 nodes")
    k = 10.0
    Q = bench.shape_quadrature(inti)
    n = len(Q)
    assert n == 100920
    op = inti.Helmholtz(k=k, dim=3)
    L, _ = inti.combined_field(op=op, quad=Q, c_single=-1j * k, c_double=1.0, c_identity=0.5)
    b = -np.exp(1j * k * Q.coords[2])
    sig, info = inti.gmres(L, b, restart=100, maxiter=300, rtol=1e-8)
    say(f"  C5: gmres {info}")
    assert info["relres"] <= 1e-8
    P = inti.points(bench.fibonacci_directions(64))

    def quad(y, nv, w):
        return inti.Quadrature(y, nv, w, np.arange(n, dtype=np.int64).reshape(1, -1))

    def J(y, nv, w, kk, s):
        o = inti.Helmholtz(k=kk, dim=3)
        F = inti.IntegralOperator(inti.CombinedFieldKernel(o, -1j * kk, 1.0, "FF"), P, quad(y, nv, w))
        u = F.apply(s)
        return float(np.sum(np.abs(u) ** 2)), u

    base = (Q.coords, Q.normals, Q.weights, k, sig)
    J0, u0 = J(*base)
    F = inti.IntegralOperator(inti.CombinedFieldKernel(op, -1j * k, 1.0, "FF"), P, quad(*base[:3]))
    wrt = ("sigma", "source_nodes", "source_normals", "source_weights", "k")
    g = F.pullback(2 * u0, sig, wrt=wrt, dcoefs=(-1j, 0.0))
    Qs = quad(*base[:3])
    ref = oracle.adjoint(1, 3, 8, k, P, Qs, 2 * u0, sig, coefs=(-1j * k, 1.0), dcoefs=(-1j, 0.0),
                         want_xbar=False)
    errs = {"sigma": dc.rel(g["sigma"], ref["sigma_bar"]), "nodes": dc.rel(g["source_nodes"], ref["ybar"]),
            "normals": dc.rel(g["source_normals"], ref["nbar"]),
            "weights": dc.rel(g["source_weights"], ref["wbar"]),
            "k": abs(g["k"] - ref["kbar"]) / abs(ref["kbar"])}
    say(f"  C5: pullback vs oracle {errs}")
    assert max(v for kk, v in errs.items() if kk != "k") <= 1e-12 and errs["k"] <= 1e-11
    rng = np.random.default_rng(bench.SEED)
    worst = 0.0
    for t in range(20):
        which = t % 5
        v = rng.normal(size=np.shape(base[which])) if which != 3 else 1.0
        if which == 4:
            v = v + 1j * rng.normal(size=n)
            pred = float(np.real(np.vdot(g["sigma"], v)))
        elif which == 3:
            pred = float(g["k"])
        else:
            pred = float(np.sum(g[("source_nodes", "source_normals", "source_weights")[which]] * v))
        h = 1e-6
        ap, am = list(base), list(base)
        ap[which] = base[which] + h * v
        am[which] = base[which] - h * v
        fd = (J(*ap)[0] - J(*am)[0]) / (2 * h)
        worst = max(worst, abs(fd - pred) / max(abs(pred), 1e-12))
    say(f"  C5: J = {J0:.6e}, worst FD relative error over 20 directions {worst:.2e}")
    assert worst <= 1e-5
