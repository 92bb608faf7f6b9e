"""single_double_layer / adj_double_layer_hypersingular (Inti.jl
src/api.jl:65-212) with the GPU backend.

compression:
  {"method": "none"}  dense numpy matrices: entries assembled on the GPU and
                      the sparse correction scatter-added (api.jl:171-173);
  {"method": "gpu"}   matrix-free GPU operators with the correction fused
                      into every apply (the role LinearMap(FMM) + LinearMap(dS)
                      plays in api.jl:182-185).
correction: {"method": "none"} or {"method": "dim", ...} (api.jl:96-158);
  the adaptive correction is not part of this backend.
"""
import numpy as np

from .bdim import DimParameters, bdim_correction, etype_to_nearest_points
from .dense import DenseOperator
from .kernels import (AdjointDoubleLayerKernel, CombinedFieldKernel, DoubleLayerKernel,
                      Helmholtz, HyperSingularKernel, SingleLayerKernel)
from .nystrom import (DoubleLayerPotential, IntegralOperator, SingleLayerPotential,
                      _green_multiplier, assemble_matrix)

COMPRESSION_METHODS = ("none", "gpu")
CORRECTION_METHODS = ("none", "dim")


def _normalize_compression(compression):
    """utils.jl:189-197."""
    m = compression.get("method")
    if m not in COMPRESSION_METHODS:
        raise ValueError(f"Unknown compression.method {m}. Available options: {COMPRESSION_METHODS}")
    return dict(compression)


def _normalize_correction(correction, target, source):
    """utils.jl:199-228."""
    m = correction.get("method")
    if m not in CORRECTION_METHODS:
        raise ValueError(f"Unknown correction.method {m}. Available options: {CORRECTION_METHODS}")
    c = dict(correction)
    if m == "dim":
        if target is not source and "target_location" not in c and "green_multiplier" not in c:
            raise ValueError("missing target_location field in correction")
        if "target_location" in c and "green_multiplier" in c:
            raise ValueError("conflicting target_location and green_multiplier fields in correction")
        c.setdefault("maxdist", np.inf)
    return c


def _dim_corrections(op, target, source, Sop, Dop, correction, derivative):
    if "green_multiplier" in correction:
        mu = np.asarray(correction["green_multiplier"], dtype=np.float64)
        if mu.shape != (len(target),):
            raise ValueError("green_multiplier must have one entry per target")
    else:
        loc = "on" if target is source else correction["target_location"]
        mu = np.full(len(target), _green_multiplier(loc))
    # "near": element -> target lists given by the caller (e.g. a stencil of
    # near_interaction_list), in place of the nearest-node assignment
    near = correction["near"] if "near" in correction else etype_to_nearest_points(target, source,
                                                                                    correction["maxdist"])
    return bdim_correction(op, target, source, Sop, Dop, green_multiplier=mu,
                           parameters=correction.get("parameters", DimParameters()),
                           derivative=derivative, near=near)


class CorrectedOperator:
    """Matrix-free A + dA on the GPU (one fused apply)."""

    def __init__(self, iop):
        self.iop = iop

    @property
    def shape(self):
        return self.iop.shape

    @property
    def dtype(self):
        return self.iop.dtype

    def apply(self, x, y=None, alpha=1.0, beta=0.0):
        return self.iop.apply(x, y, alpha, beta)

    def __matmul__(self, x):
        return self.iop.apply(x)


def single_double_layer(*, op, target, source, compression=None, correction=None,
                        derivative=False, precision="fp64"):
    """(S, D) for op (or (ADL, HS) if derivative), api.jl:65-187."""
    compression = _normalize_compression(compression or {"method": "none"})
    correction = _normalize_correction(correction or {"method": "dim"}, target, source)
    G = AdjointDoubleLayerKernel(op) if derivative else SingleLayerKernel(op)
    dG = HyperSingularKernel(op) if derivative else DoubleLayerKernel(op)
    Sop = IntegralOperator(G, target, source, precision=precision)
    Dop = IntegralOperator(dG, target, source, precision=precision)
    dS = dD = None
    if correction["method"] == "dim":
        dS, dD = _dim_corrections(op, target, source, Sop, Dop, correction, derivative)
    if compression["method"] == "none":
        # assemble_matrix and axpy!(true, dS, Smat) (api.jl:84-85,171-173) on
        # the device, one download per matrix
        if dS is None:
            return assemble_matrix(Sop), assemble_matrix(Dop)
        return (DenseOperator(Sop, corrections=[(dS, 1.0)]).to_numpy(),
                DenseOperator(Dop, corrections=[(dD, 1.0)]).to_numpy())
    if dS is not None:
        Sop.attach(dS)
        Dop.attach(dD)
    return CorrectedOperator(Sop), CorrectedOperator(Dop)


def single_double_layer_potential(*, op, source):
    """(single-layer, double-layer) IntegralPotentials (api.jl:220-226)."""
    return SingleLayerPotential(op, source), DoubleLayerPotential(op, source)


def adj_double_layer_hypersingular(*, op, target, source=None, compression=None,
                                   correction=None, precision="fp64"):
    """(ADL, HS) with their DIM corrections (api.jl:197-212)."""
    return single_double_layer(op=op, target=target, source=target if source is None else source,
                               compression=compression, correction=correction, derivative=True,
                               precision=precision)


def combined_field(*, op, quad, c_single, c_double, c_identity=0.0, correction=None,
                   precision="fp64", corrections=None):
    """Corrected c_identity*I + c_double*D + c_single*S on one surface as a
    single fused GPU operator (e.g. I/2 + D - i k S,
    docs/src/examples/helmholtz_scattering.jl:230). Returns
    (operator, (dS, dD)); pass corrections=(dS, dD) to reuse them."""
    correction = _normalize_correction(correction or {"method": "dim"}, quad, quad)
    if corrections is None and correction["method"] == "dim":
        Sop = IntegralOperator(SingleLayerKernel(op), quad, quad)
        Dop = IntegralOperator(DoubleLayerKernel(op), quad, quad)
        corrections = _dim_corrections(op, quad, quad, Sop, Dop, correction, False)
    A = IntegralOperator(CombinedFieldKernel(op, c_single, c_double, "SD"), quad, quad,
                         precision=precision)
    if corrections is not None:
        dS, dD = corrections
        A.attach(dS, c_single)
        A.attach(dD, c_double)
    if c_identity != 0:
        A.set_identity(c_identity)
    return CorrectedOperator(A), corrections


def is_complex_op(op):
    return isinstance(op, Helmholtz)
