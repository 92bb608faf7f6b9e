"""vdim_correction / volume_potential on the GPU (Inti.jl src/vdim.jl:30-129,
src/api.jl:282-378).

Theta = S[dP/dn] - D[P] - V[p] + mu P(x) for every basis monomial p and its
polynomial solution P (_vdim_auxiliary_quantities, vdim.jl:260-288) is three
multi-right-hand-side GPU applies; the per-element least squares with the
Vandermonde L[n, q] = p_n(y_q) of each element's nodes (vdim.jl:96-117) run as
one batched GPU kernel (inti_lsq_build).
"""
import ctypes as C

import numpy as np

from . import _lib
from .bdim import etype_to_nearest_points
from .kernels import SingleLayerKernel
from .nystrom import IntegralOperator, SparseCorrection, _green_multiplier
from .polynomials import evaluate_basis, normal_derivative, polynomial_solutions_vdim


def _apply(A, X):
    return A.apply(X) if hasattr(A, "apply") else A @ X


def vdim_auxiliary_quantities(op, target, source, boundary, Sop, Dop, Vop, mu, order, center=None):
    """Theta (m x nbasis) and the basis, vdim.jl:260-288."""
    ps, Ps, idx = polynomial_solutions_vdim(op, order)
    b = evaluate_basis(ps, source.coords, center)                    # p at volume nodes
    g0 = evaluate_basis(Ps, boundary.coords, center)                 # P on the boundary
    g1 = normal_derivative(Ps, boundary.coords, boundary.normals, center)
    Pt = evaluate_basis(Ps, target.coords, center)                   # P at the targets
    dt = np.complex128 if op.k != 0 else np.float64
    Theta = _apply(Sop, g1.astype(dt)) - _apply(Dop, g0.astype(dt)) - _apply(Vop, b.astype(dt))
    Theta = Theta + mu[:, None] * Pt
    return Theta, ps, b


def vdim_correction(op, target, source, boundary, Sop, Dop, Vop, *, green_multiplier,
                    interpolation_order=None, maxdist=np.inf, center=None, near=None,
                    diagnostics=None):
    """Correction dV such that V + dV is an accurate volume potential
    (vdim.jl:30-129). Sop/Dop map the boundary to the targets (already
    corrected), Vop maps the volume source to the targets. Returns a
    SparseCorrection."""
    if getattr(op, "ki", 0.0) != 0:
        raise NotImplementedError("vdim_correction: real wavenumbers only")
    mu = np.ascontiguousarray(green_multiplier, dtype=np.float64)
    if mu.shape != (len(target),):
        raise ValueError("green_multiplier must have one entry per target")
    order = interpolation_order if interpolation_order is not None else source.qorder
    Theta, ps, b = vdim_auxiliary_quantities(op, target, source, boundary, Sop, Dop, Vop, mu,
                                             order, center)
    nb = len(ps)
    # L[n, q] = p_n(y_jglob[q]) per element: b[qtags.T] is (ne, nq, nb), i.e.
    # each element's nb x nq matrix in column-major order; element types one
    # after another (vdim.jl:76-121 loops over source.etype2qtags)
    types = source.etype_qtags
    A = np.ascontiguousarray(np.concatenate([b[q.T].reshape(-1) for q in types]))
    nqs, nes, tags = source.etype_flat()
    if near is None:
        near = etype_to_nearest_points(target, source, maxdist)
    if near is None:
        near_ptr = near_idx = None
    else:
        lists = [near] if len(types) == 1 else list(near)
        if len(lists) != len(types):
            raise ValueError("near: one (near_ptr, near_idx) pair per element type")
        near_ptr = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.int64) for p, _ in lists]))
        near_idx = np.ascontiguousarray(np.concatenate([np.asarray(i, dtype=np.int64) for _, i in lists]))
    cplx = op.k != 0
    rhs = np.ascontiguousarray(Theta.T) if cplx else np.ascontiguousarray(Theta.T.real)
    ctx = _lib.default_context()
    h = C.c_void_p()
    diag = np.zeros(2)
    ctx.check(ctx.lib.inti_lsq_build_types(
        ctx.handle, 1 if cplx else 0, nb, len(types), nqs.ctypes.data_as(C.POINTER(C.c_int)), _lib.i64ptr(nes),
        _lib.dptr(A), _lib.i64ptr(tags), len(target), len(source),
        _lib.i64ptr(near_ptr), _lib.i64ptr(near_idx), _lib.dptr(rhs.reshape(-1)), C.byref(h),
        _lib.dptr(diag)))
    if diagnostics is not None:
        diagnostics.update(vander_rdiag_ratio=diag[0], max_weight_norm=diag[1], nbasis=nb)
    return SparseCorrection(ctx, h)


def volume_potential(*, op, target, source, compression=None, correction=None):
    """Volume potential V: source (volume quadrature) -> target with the DIM
    correction (api.jl:282-378). correction = {"method": "dim", "boundary":
    Quadrature, "target_location" or "green_multiplier", "maxdist",
    "interpolation_order", "center", and optionally prebuilt boundary
    operators "S_b2d"/"D_b2d"}; compression {"method": "gpu"} (matrix-free,
    default) or {"method": "none"} (dense)."""
    from .api import (CorrectedOperator, _normalize_compression, _normalize_correction,
                      single_double_layer)
    if getattr(op, "ki", 0.0) != 0:
        raise NotImplementedError("volume_potential: real wavenumbers only")
    compression = _normalize_compression(compression or {"method": "gpu"})
    correction = _normalize_correction(correction or {"method": "none"}, target, source)
    V = IntegralOperator(SingleLayerKernel(op), target, source)
    if correction["method"] == "none":
        if compression["method"] == "none":
            from .nystrom import assemble_matrix
            return np.asarray(assemble_matrix(V))
        return V
    if "green_multiplier" in correction:
        mu = np.asarray(correction["green_multiplier"], dtype=np.float64)
    else:
        loc = "inside" if target is source else correction["target_location"]
        mu = np.full(len(target), _green_multiplier(loc))
    if "boundary" in correction:
        boundary = correction["boundary"]
    elif getattr(source, "boundary_elements", None) is not None:
        # the boundary of the source's own mesh at twice its quadrature order
        from .quadrature import boundary_quadrature
        boundary = boundary_quadrature(source, 2 * source.qorder)
    else:
        raise ValueError("Missing correction.boundary field for :dim method on a volume potential")
    if "S_b2d" in correction and "D_b2d" in correction:
        S, D = correction["S_b2d"], correction["D_b2d"]
    else:
        bcorr = {"method": "dim", "maxdist": correction["maxdist"]}
        if "green_multiplier" in correction:
            bcorr["green_multiplier"] = mu
        else:
            bcorr["target_location"] = loc
        S, D = single_double_layer(op=op, target=target, source=boundary, compression=compression,
                                   correction=bcorr)
    dV = vdim_correction(op, target, source, boundary, S, D, V, green_multiplier=mu,
                         interpolation_order=correction.get("interpolation_order"),
                         maxdist=correction["maxdist"], center=correction.get("center"))
    if compression["method"] == "none":
        from .nystrom import assemble_matrix
        return np.asarray(assemble_matrix(V) + dV.to_scipy())
    V.attach(dV)
    return CorrectedOperator(V)
