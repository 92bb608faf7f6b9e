"""bdim_correction on the GPU (Inti.jl src/bdim.jl:50-196) and the
nearest-element target lists of src/quadrature.jl:249-286."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .kernels import (AdjointDoubleLayerKernel, DoubleLayerKernel, HyperSingularKernel,
                      SingleLayerKernel)
from .nystrom import IntegralOperator, SparseCorrection


@dataclass
class DimParameters:
    """DimParameters (bdim.jl:7-10)."""
    sources_oversample_factor: float = 3.0
    sources_radius_multiplier: float = 1.5


def etype_to_nearest_points(target, source, maxdist=np.inf):
    """Element -> target lists (quadrature.jl:249-286), as CSR over elements,
    one (near_ptr, near_idx) pair per element type.

    target is source: every element lists its own nodes (returns None, the
    self pattern). Otherwise every target goes to the element owning its
    nearest source node (dropped beyond maxdist); within an element, targets
    are grouped by node, then in target order."""
    if target is source:
        return None
    from scipy.spatial import cKDTree
    tree = cKDTree(source.coords.T)
    d, nearest = tree.query(target.coords.T, k=1)
    nearest = np.where(d > maxdist, -1, nearest)
    n = len(source)
    by_node = [[] for _ in range(n)]
    for i, j in enumerate(nearest):
        if j >= 0:
            by_node[j].append(i)
    out = []
    for qtags in source.etype_qtags:
        ne = qtags.shape[1]
        near_ptr = np.zeros(ne + 1, dtype=np.int64)
        near_idx = []
        for e in range(ne):
            for j in qtags[:, e]:
                near_idx.extend(by_node[j])
            near_ptr[e + 1] = len(near_idx)
        out.append((near_ptr, np.asarray(near_idx, dtype=np.int64)))
    return out[0] if len(out) == 1 else out


def element_centers_and_diameters(quad):
    """Per element type: the centres of the elements (the image of the
    reference element's centre, s = 1/2 on P2 lines and (1/3, 1/3) on P2
    triangles, when the quadrature keeps its elements; else the weighted mean
    of the element's nodes) and their diameters (the largest distance between
    two of the element's geometry nodes, or quadrature nodes)."""
    out = []
    els = quad.elements if len(quad.etype_qtags) == 1 else None
    for qtags in quad.etype_qtags:
        X = quad.coords[:, qtags]                       # (dim, nq, ne)
        if els is not None and els.ndim == 3 and els.shape[1] in (3, 6):
            V = np.asarray(els, dtype=np.float64)       # (ne, 3 or 6, dim)
            if V.shape[1] == 3:                         # P2 line: nodes at s = 0, 1/2, 1
                c = V[:, 1, :].T
            else:                                       # P2 triangle at (1/3, 1/3)
                c = (-(V[:, 0] + V[:, 2] + V[:, 5]) / 9 + 4 * (V[:, 1] + V[:, 3] + V[:, 4]) / 9).T
            P = V.transpose(2, 1, 0)                    # (dim, nodes, ne)
        else:
            w = quad.weights[qtags]
            c = np.sum(X * w[None], axis=1) / np.sum(w, axis=0)[None]
            P = X
        diff = P[:, :, None, :] - P[:, None, :, :]
        diam = np.sqrt(np.sum(diff * diff, axis=0)).max(axis=(0, 1))
        out.append((c, diam))
    return out


def near_interaction_list(X, quad, tol):
    """For each element of each type of quad, the points of X (dim x m)
    within tol of the element's centre (mesh.jl:610-642), as one
    (near_ptr, near_idx) CSR over elements per type, indices ascending;
    a stencil for bdim_correction(..., near=...)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.asarray(X, dtype=np.float64).T)
    out = []
    for c, _ in element_centers_and_diameters(quad):
        lists = tree.query_ball_point(c.T, r=tol, workers=-1, return_sorted=True)
        ptr = np.zeros(len(lists) + 1, dtype=np.int64)
        ptr[1:] = np.cumsum([len(l) for l in lists])
        idx = (np.concatenate([np.asarray(l, dtype=np.int64) for l in lists]) if len(lists)
               else np.zeros(0, dtype=np.int64))
        out.append((ptr, idx))
    return out[0] if len(out) == 1 else out


def bdim_correction(op, target, source, Sop, Dop, *, green_multiplier, parameters=None,
                    derivative=False, maxdist=np.inf, near=None, diagnostics=None):
    """Corrections (dS, dD) such that S + dS and D + dD are accurate
    (bdim.jl:50-61). Sop/Dop are the GPU IntegralOperators of SL/DL (or
    ADL/HS if derivative) over (target, source); Theta = mu*G + S*B1 - D*B0 is
    computed matrix-free on the GPU. A source with several element types
    takes one near list per type (etype_to_nearest_points returns them).
    Returns two SparseCorrection objects
    (CSR on the device; .to_scipy() for host copies)."""
    if not isinstance(Sop, IntegralOperator) or not isinstance(Dop, IntegralOperator):
        raise TypeError("Sop and Dop must be GPU IntegralOperators over (target, source)")
    want = (AdjointDoubleLayerKernel, HyperSingularKernel) if derivative else (SingleLayerKernel, DoubleLayerKernel)
    if not (isinstance(Sop.kernel, want[0]) and isinstance(Dop.kernel, want[1])):
        raise TypeError("Sop/Dop kernels do not match `derivative`")
    if Sop.target is not target or Sop.source is not source:
        raise ValueError("Sop must map source to target")
    mu = np.ascontiguousarray(green_multiplier, dtype=np.float64)
    if mu.shape != (len(target),):
        raise ValueError("green_multiplier must have one entry per target")
    params = parameters or DimParameters()
    pv = np.array([params.sources_oversample_factor, params.sources_radius_multiplier])
    if near is None:
        near = etype_to_nearest_points(target, source, maxdist)
    ntypes = len(source.etype_qtags)
    if near is None:
        near_ptr = near_idx = None
    else:
        lists = [near] if ntypes == 1 else list(near)
        if len(lists) != ntypes:
            raise ValueError("near: one (near_ptr, near_idx) pair per element type")
        near_ptr = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.int64) for p, _ in lists]))
        near_idx = np.ascontiguousarray(np.concatenate([np.asarray(i, dtype=np.int64) for _, i in lists]))
    nqs, nes, tags = source.etype_flat()
    ctx = Sop.ctx
    hS, hD = C.c_void_p(), C.c_void_p()
    diag = np.zeros(4)
    ctx.check(ctx.lib.inti_bdim_build_types(
        Sop.handle, Dop.handle, 1 if derivative else 0, ntypes,
        nqs.ctypes.data_as(C.POINTER(C.c_int)), _lib.i64ptr(nes), _lib.i64ptr(tags),
        _lib.i64ptr(near_ptr), _lib.i64ptr(near_idx), _lib.dptr(mu), _lib.dptr(pv),
        C.byref(hS), C.byref(hD), _lib.dptr(diag)))
    if diagnostics is not None:
        diagnostics.update(rdiag_ratio=diag[0], max_weight_norm=diag[2], max_theta_norm=diag[3])
    return SparseCorrection(ctx, hS), SparseCorrection(ctx, hD)
