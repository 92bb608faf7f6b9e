"""3D Helmholtz kernels with a complex wavenumber k = kr + i ki (Inti.jl
src/kernels.jl:322-413: the Helmholtz kernels take any k::Number) and the
density-interpolation correction built from them (src/bdim.jl:50-196),
restated in numpy.

TEST INFRASTRUCTURE ONLY (a CPU reference; the product never imports it).

Kernels, with r = x - y, d = |r|, G = e^{ikd} / (4 pi d) (kernels.jl:347-352)
and the radial derivatives G' = G (ik - 1/d), G'' = G ((ik - 1/d)^2 + 1/d^2):
  SL  G                        DL  dG/dn_y = -G' (r.n_y) / d
  ADL dG/dn_x = G' (r.n_x)/d   HS  d^2G/dn_x dn_y
      = -[G'' (r.n_x)(r.n_y)/d^2 + G' (n_x.n_y/d - (r.n_x)(r.n_y)/d^3)]
(kernels.jl:354-413), zero for coincident points. Pinned in
tests/test_complex_k.py: at ki = 0 against the C oracle (the reference's
sign conventions, kernels and DIM weights), and for complex k by central
differences (DL, ADL, HS as normal derivatives of SL and DL), the Helmholtz
equation and the modulus e^{-ki d} / (4 pi d).

The correction follows bdim.jl:50-196 for one element type: Fibonacci
monopoles from the C oracle, traces B0 = G(y_j, z_k), B1 = dG/dn_j (the ADL
kernel with target normal n_j), Theta = mu G(x_i, z_k) + S B1 - D B0 (mu ADL
and the (ADL, HS) pair with derivative=True), and per element the least
squares M^T W = Theta_i^T with M = [B0[jglob]; B1[jglob]]: dD[i, jglob] =
W[:nq], dS[i, jglob] = -W[nq:]. `solver` picks numpy's lstsq (SVD) or a
Householder QR, and `order` sums Theta's direct sums forwards or backwards
over the sources: two builds that differ in both bound the rounding the
problem allows.
"""
import numpy as np

SL, DL, ADL, HS, COMB_SD, COMB_AH = 0, 1, 2, 3, 4, 5
_INV4PI = 1.0 / (4.0 * np.pi)


def kernel_matrix(kind, k, X, NX, Y, NY, coefs=None):
    """K(x_i, y_j) for target points X (3, m) with normals NX and source
    points Y (3, n) with normals NY; coefs (c0, c1) for the combined kinds
    (c0 SL + c1 DL, c0 ADL + c1 HS)."""
    if kind == COMB_SD:
        c0, c1 = coefs
        return c0 * kernel_matrix(SL, k, X, NX, Y, NY) + c1 * kernel_matrix(DL, k, X, NX, Y, NY)
    if kind == COMB_AH:
        c0, c1 = coefs
        return c0 * kernel_matrix(ADL, k, X, NX, Y, NY) + c1 * kernel_matrix(HS, k, X, NX, Y, NY)
    k = complex(k)
    r = X[:, :, None] - Y[:, None, :]
    d = np.sqrt(np.sum(r * r, axis=0))
    same = d <= 1e-14
    d = np.where(same, 1.0, d)
    G = np.exp(1j * k * d) * _INV4PI / d
    t = 1j * k - 1.0 / d
    Gp = G * t
    if kind == SL:
        K = G
    elif kind == DL:
        K = -Gp * np.einsum("dij,dj->ij", r, NY) / d
    elif kind == ADL:
        K = Gp * np.einsum("dij,di->ij", r, NX) / d
    elif kind == HS:
        Gpp = G * (t * t + 1.0 / (d * d))
        rnx = np.einsum("dij,di->ij", r, NX)
        rny = np.einsum("dij,dj->ij", r, NY)
        nxny = (NX.T @ NY)
        K = -(Gpp * rnx * rny / (d * d) + Gp * (nxny / d - rnx * rny / d ** 3))
    else:
        raise ValueError(kind)
    return np.where(same, 0.0, K)


def assemble(kind, k, trg, src, coefs=None):
    """Nystrom matrix K(x_i, y_j) w_j (nystrom.jl:93-96)."""
    K = kernel_matrix(kind, k, trg.coords, trg.normals, src.coords, src.normals, coefs)
    return K * src.weights[None, :]


def _lsq(M, rhs, solver):
    """Least-squares W of M^T W = rhs for each column of rhs (M: (2nq, ns))."""
    A = M.T
    if solver == "svd":
        return np.linalg.lstsq(A, rhs, rcond=None)[0]
    Q, R = np.linalg.qr(A)
    return np.linalg.solve(R, Q.conj().T @ rhs)


def bdim(k, quad, mu, monopoles, derivative=False, solver="qr", order="forward"):
    """Density-interpolation corrections (dS, dD) as dense (n, n) arrays for
    a quadrature with itself, one element type, the self pattern;
    monopoles (3, ns) are the oracle's Fibonacci points."""
    x, nrm, w = quad.coords, quad.normals, quad.weights
    nq, ne = quad.qtags.shape
    n = x.shape[1]
    B0 = kernel_matrix(SL, k, x, nrm, monopoles, None)
    B1 = kernel_matrix(ADL, k, x, nrm, monopoles, None)
    if derivative:
        Sm, Dm = assemble(ADL, k, quad, quad), assemble(HS, k, quad, quad)
        Gt = B1
    else:
        Sm, Dm = assemble(SL, k, quad, quad), assemble(DL, k, quad, quad)
        Gt = B0
    if order == "forward":
        Theta = mu[:, None] * Gt + Sm @ B1 - Dm @ B0
    else:
        Theta = (Sm[:, ::-1] @ B1[::-1] - Dm[:, ::-1] @ B0[::-1]) + mu[:, None] * Gt
    dS = np.zeros((n, n), dtype=np.complex128)
    dD = np.zeros((n, n), dtype=np.complex128)
    for e in range(ne):
        jg = quad.qtags[:, e]
        M = np.concatenate([B0[jg], B1[jg]], axis=0)
        W = _lsq(M, Theta[jg].T, solver)  # (2nq, nq): one column per target of the element
        for t, i in enumerate(jg):
            dD[i, jg] += W[:nq, t]
            dS[i, jg] += -W[nq:, t]
    return dS, dD
