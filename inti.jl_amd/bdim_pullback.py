"""Reverse mode of bdim_correction (Inti.jl src/bdim.jl:50-196) for a 3D
quadrature with itself (self pattern, one element type, the SL/DL pair):
the gradient of a real function of the correction values with respect to the
nodes, normals, weights and the wavenumber.

The forward build, restated as the chain the gradient runs back through:
  z_k      monopoles on a sphere of radius R around the centre xc of the
           nodes' bounding box (bdim.jl:79-92, utils.jl:158-171);
  B0[j,k]  = G(y_j - z_k), B1[j,k] = dG/dn_j (y_j - z_k)   (bdim.jl:94-103);
  Theta    = mu G(x_i, z_k) + S B1 - D B0                   (bdim.jl:106-118);
  per element e with nodes jg and A_e = [B0[jg,:]; B1[jg,:]]^T (ns x 2nq):
           W = argmin |A_e W - Theta[i,:]^T| for each node i of e, and
           dD[i, jg] = W[:nq], dS[i, jg] = -W[nq:]          (bdim.jl:137-185).
Backwards, for a least-squares solution W = A^+ b of full column rank with
residual r = b - A W and cotangent g (dJ = Re sum conj(g) dW):
  z = (A^H A)^{-1} g,  b_bar = A z,  A_bar = r z^H - b_bar W^H.
Theta's cotangent goes through S B1 and D B0 by the GPU pullback of those two
direct sums, one right-hand side (monopole) at a time; the traces, the
mu G term and the monopoles (through the bounding box's extreme nodes) are
differentiated here in closed form. Cotangent convention throughout: for a
complex quantity v, v_bar is such that dJ = Re sum conj(v_bar) dv.
"""
import numpy as np

from .bdim import DimParameters
from .kernels import DoubleLayerKernel, Helmholtz, SingleLayerKernel
from .nystrom import IntegralOperator


def _radial(k, d):
    """G, G', G'' of G(d) = e^{ikd} / (4 pi d) and dG/dk, dG'/dk."""
    ik = 1j * k
    G = np.exp(ik * d) / (4 * np.pi * d)
    a = ik - 1.0 / d
    Gp = G * a
    Gpp = G * (a * a + 1.0 / (d * d))
    return G, Gp, Gpp, 1j * d * G, -k * d * G


def monopoles_3d(coords, ns, radius_mult):
    """Fibonacci sphere around the bounding box of coords (utils.jl:158-171),
    with the pieces its derivative needs: (z (3, ns), unit directions u,
    lo, hi, argmin, argmax, R)."""
    lo, hi = coords.min(axis=1), coords.max(axis=1)
    imin, imax = coords.argmin(axis=1), coords.argmax(axis=1)
    xc = (lo + hi) / 2
    R = radius_mult * np.linalg.norm(hi - lo) / 2
    i = np.arange(1, ns + 1)
    phi = np.pi * (3 - np.sqrt(5.0))
    y = 1 - ((i - 1) / (ns - 1)) * 2
    rad = np.sqrt(1 - y * y)
    th = phi * i
    u = np.stack([np.cos(th) * rad, y, np.sin(th) * rad])
    return xc[:, None] + R * u, u, lo, hi, imin, imax, R


def traces_3d(k, y, n, z):
    """B0 = G(y_j - z_k), B1 = dG/dn_j (n x ns) and the pieces of their
    derivatives."""
    r = y[:, :, None] - z[:, None, :]           # (3, n, ns)
    d = np.sqrt(np.sum(r * r, axis=0))
    G, Gp, Gpp, Gk, Gpk = _radial(k, d)
    rn = np.einsum("dj,djk->jk", n, r)
    B0 = G
    B1 = Gp * rn / d
    return B0, B1, (r, d, G, Gp, Gpp, Gk, Gpk, rn)


def lsq_solve(A, b):
    """W = A^+ b by QR (batched: A (..., m, p), b (..., m)); returns (W, Q, R)."""
    Q, R = np.linalg.qr(A)
    W = np.linalg.solve(R, np.einsum("...ki,...k->...i", Q.conj(), b)[..., None])[..., 0]
    return W, Q, R


def lsq_backward(A, Q, R, W, b, g):
    """Cotangents (b_bar, A_bar) of W = A^+ b for dJ = Re sum conj(g) dW,
    with A = Q R: b_bar = A z = Q R^{-H} g (one triangular solve: the
    matrices are ill-conditioned, and (A^H A)^{-1} squares that),
    z = R^{-1} R^{-H} g, A_bar = r z^H - b_bar W^H, r = b - A W."""
    y1 = np.linalg.solve(np.conj(np.swapaxes(R, -1, -2)), g[..., None])[..., 0]
    z = np.linalg.solve(R, y1[..., None])[..., 0]
    bbar = np.einsum("...ki,...i->...k", Q, y1)
    res = b - np.einsum("...ki,...i->...k", A, W)
    Abar = res[..., :, None] * z.conj()[..., None, :] - bbar[..., :, None] * W.conj()[..., None, :]
    return bbar, Abar


def bdim_pullback(op, quad, Sbar, Dbar, rowptr, colind, *, green_multiplier, parameters=None):
    """Gradient of J with dJ = Re sum conj(Sbar) d(dS) + conj(Dbar) d(dD),
    (dS, dD) = bdim_correction(op, quad, quad, SL, DL; green_multiplier) in
    the self pattern, Sbar / Dbar given on the corrections' CSR pattern
    (rowptr, colind). Returns {"nodes": (3, n), "normals": (3, n),
    "weights": (n,), "k": float}."""
    if quad.dim != 3:
        raise NotImplementedError("bdim_pullback: 3D quadratures")
    if getattr(op, "ki", 0.0) != 0:
        raise NotImplementedError("bdim_pullback: real wavenumbers only")
    if len(quad.etype_qtags) != 1:
        raise NotImplementedError("bdim_pullback: one element type")
    params = parameters or DimParameters()
    k = op.k if isinstance(op, Helmholtz) else 0.0
    y, nrm, w = quad.coords, quad.normals, quad.weights
    n = len(quad)
    qtags = quad.qtags
    nq, ne = qtags.shape
    ns = int(np.ceil(params.sources_oversample_factor * nq))
    mu = np.asarray(green_multiplier, dtype=np.float64)
    z, u, lo, hi, imin, imax, R = monopoles_3d(y, ns, params.sources_radius_multiplier)
    B0, B1, (r, d, G, Gp, Gpp, Gk, Gpk, rn) = traces_3d(k, y, nrm, z)
    # Theta = mu G(x_i, z_k) + S B1 - D B0 on the GPU (uncorrected sums)
    S = IntegralOperator(SingleLayerKernel(op), quad, quad)
    D = IntegralOperator(DoubleLayerKernel(op), quad, quad)
    cplx = isinstance(op, Helmholtz)
    if not cplx:
        B0, B1, G = B0.real, B1.real, G.real
    Theta = (mu[:, None] * G + S.apply(np.ascontiguousarray(B1)) - D.apply(np.ascontiguousarray(B0))).astype(np.complex128)
    # per element and node: the least squares (by QR, as the build) and
    # their cotangents
    Ae = np.concatenate([B0[qtags.T], B1[qtags.T]], axis=1).transpose(0, 2, 1)  # (ne, ns, 2nq)
    Ae = Ae[:, None]                                  # one factorisation per element, nq right-hand sides
    b = Theta[qtags.T]                                                          # (ne, nq, ns)
    W, Qe, Rq = lsq_solve(Ae, b)
    # cotangent of W from the CSR cotangents: dD[i, jg[c]] = W[c],
    # dS[i, jg[c]] = -W[nq + c]; row i = jg[t] holds the element's nodes in
    # increasing order
    rank = np.argsort(np.argsort(qtags, axis=0), axis=0)                        # (nq, ne)
    P = rowptr[qtags][:, None, :] + rank[None, :, :]                            # (t, c, e)
    if not np.array_equal(colind[P], np.broadcast_to(qtags[None, :, :], P.shape)):
        raise ValueError("bdim_pullback: the corrections are not the self pattern of these elements")
    P = P.transpose(2, 0, 1)                                                    # (e, t, c)
    g = np.concatenate([np.asarray(Dbar)[P], -np.asarray(Sbar)[P]], axis=2)     # (ne, nq, 2nq)
    bbar, Abar = lsq_backward(Ae, Qe, Rq, W, b, g)
    Tbar = np.zeros((n, ns), dtype=np.complex128)
    Tbar[qtags.T] = bbar
    Ab = Abar.sum(axis=1)                             # (ne, ns, 2nq): the element's nq targets summed
    B0bar = np.zeros((n, ns), dtype=np.complex128)
    B1bar = np.zeros((n, ns), dtype=np.complex128)
    np.add.at(B0bar, qtags.T, Ab[:, :, :nq].transpose(0, 2, 1))
    np.add.at(B1bar, qtags.T, Ab[:, :, nq:].transpose(0, 2, 1))
    gy = np.zeros((3, n))
    gn = np.zeros((3, n))
    gw = np.zeros(n)
    gk = 0.0
    # Theta's direct sums, one monopole at a time
    for kk in range(ns):
        tb = Tbar[:, kk] if cplx else Tbar[:, kk].real
        gs = S.pullback(tb, B1[:, kk], wrt=("sigma", "nodes", "source_weights") + (("k",) if cplx else ()))
        gd = D.pullback(-tb, B0[:, kk],
                        wrt=("sigma", "nodes", "source_normals", "source_weights") + (("k",) if cplx else ()))
        B1bar[:, kk] += gs["sigma"]
        B0bar[:, kk] += gd["sigma"]
        gy += gs["nodes"] + gd["nodes"]
        gn += gd["source_normals"]
        gw += gs["source_weights"] + gd["source_weights"]
        if cplx:
            gk += float(gs["k"]) + float(gd["k"])
    # mu G(x_i - z_k): the same radial functions as B0 (targets are the nodes)
    cT = np.conj(Tbar) * mu[:, None]
    gradG = Gp / d * r                                  # d/dy of G(y - z), (3, n, ns)
    gy += np.real(np.einsum("jk,djk->dj", cT, gradG))
    zbar = -np.real(np.einsum("jk,djk->dk", cT, gradG))
    gk += float(np.real(np.sum(cT * Gk)))
    # traces
    c0, c1 = np.conj(B0bar), np.conj(B1bar)
    dB1dy = (rn * (Gpp / d - Gp / (d * d)) / d) * r + (Gp / d)[None] * nrm[:, :, None]
    gy += np.real(np.einsum("jk,djk->dj", c0, gradG) + np.einsum("jk,djk->dj", c1, dB1dy))
    zbar -= np.real(np.einsum("jk,djk->dk", c0, gradG) + np.einsum("jk,djk->dk", c1, dB1dy))
    gn += np.real(np.einsum("jk,djk->dj", c1, gradG))
    gk += float(np.real(np.sum(c0 * Gk) + np.sum(c1 * Gpk * rn / d)))
    # monopoles: z = xc + R u, xc = (lo + hi) / 2, R = rm |hi - lo| / 2
    xcbar = zbar.sum(axis=1)
    Rbar = float(np.sum(zbar * u))
    dh = hi - lo
    dRdh = params.sources_radius_multiplier * dh / (2 * np.linalg.norm(dh))
    lobar = xcbar / 2 - Rbar * dRdh
    hibar = xcbar / 2 + Rbar * dRdh
    for dd in range(3):
        gy[dd, imin[dd]] += lobar[dd]
        gy[dd, imax[dd]] += hibar[dd]
    return {"nodes": gy, "normals": gn, "weights": gw, "k": gk}
