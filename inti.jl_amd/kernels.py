"""PDE and kernel descriptors mirroring Inti.jl src/kernels.jl.

Only the Laplace (:118-204) and Helmholtz (:310-413) operators in 2D/3D are
provided; values are computed on the GPU by libinti_amd.
"""
import numpy as np

from . import _lib


class Laplace:
    """Laplace(; dim): -Delta u (note the sign), kernels.jl:124-131."""

    pde = _lib.LAPLACE

    def __init__(self, *, dim):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.k = 0.0

    def __repr__(self):
        return f"Laplace operator in {self.dim} dimensions: -Δu"


class Helmholtz:
    """Helmholtz(; k, dim): -Delta u - k^2 u (kernels.jl:322-335).

    k may be complex (kernels.jl takes any k::Number): k = kr + i ki with
    kr > 0 runs the 3D kernels in FP64 (one-direction sums; no mixed
    precision, far fields, batches or pullbacks). In 2D a complex k needs
    the Hankel function of a complex argument and raises, and so does a
    purely imaginary k (the reference's Yukawa remap). A complex k with zero
    imaginary part is taken as real."""

    pde = _lib.HELMHOLTZ

    def __init__(self, *, k, dim):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        kr, ki = (k.real, k.imag) if isinstance(k, complex) else (k, 0.0)
        if ki != 0 and dim != 3:
            raise NotImplementedError("complex wavenumbers: 3D only (2D needs complex-argument Hankel functions)")
        if not kr > 0:
            raise ValueError("k must have a positive real part")
        if not np.isfinite(ki):
            raise ValueError("k must be finite")
        self.dim = dim
        self.kr = float(kr)
        self.ki = float(ki)
        self.k = complex(self.kr, self.ki) if self.ki != 0 else self.kr

    @property
    def is_complex_k(self):
        return self.ki != 0

    def __repr__(self):
        return f"Helmholtz operator in {self.dim} dimensions: -Δu - k²u"


def default_kernel_eltype(op):
    return float if isinstance(op, Laplace) else complex


default_density_eltype = default_kernel_eltype


class _Kernel:
    kind = None

    def __init__(self, op):
        self.op = op

    @property
    def is_complex(self):
        return isinstance(self.op, Helmholtz)

    def __repr__(self):
        return f"{type(self).__name__}({self.op!r})"


class SingleLayerKernel(_Kernel):
    kind = _lib.SL


class DoubleLayerKernel(_Kernel):
    kind = _lib.DL


class AdjointDoubleLayerKernel(_Kernel):
    kind = _lib.ADL


class HyperSingularKernel(_Kernel):
    kind = _lib.HS


class _FarField(_Kernel):
    def __init__(self, op):
        if not isinstance(op, Helmholtz):
            raise ValueError("far-field kernels are defined for Helmholtz only")
        super().__init__(op)


class FarFieldSingleLayerKernel(_FarField):
    """Far-field pattern of the single-layer potential: targets are unit
    directions xh, K = P e^{-ik xh.y}, P = 1/(4 pi) in 3D and
    e^{i pi/4}/sqrt(8 pi k) in 2D (u ~ e^{ikr} r^{-(dim-1)/2} u_inf)."""
    kind = _lib.FF_SL


class FarFieldDoubleLayerKernel(_FarField):
    """Far-field pattern of the double-layer potential: P e^{-ik xh.y} (-ik xh.n_y)."""
    kind = _lib.FF_DL


class CombinedFieldKernel(_Kernel):
    """c0*SL + c1*DL (pair='SD'), c0*ADL + c1*HS (pair='AH') or the far field
    of c0*SL + c1*DL (pair='FF'), evaluated in one pass (shares the distance
    and e^{ikd} / Hankel evaluation)."""

    def __init__(self, op, c0, c1, pair="SD"):
        super().__init__(op)
        if pair not in ("SD", "AH", "FF"):
            raise ValueError("pair must be 'SD', 'AH' or 'FF'")
        if pair == "FF" and not isinstance(op, Helmholtz):
            raise ValueError("far-field kernels are defined for Helmholtz only")
        self.kind = {"SD": _lib.COMB_SD, "AH": _lib.COMB_AH, "FF": _lib.FF_COMB}[pair]
        self.pair = pair
        self.c0 = complex(c0)
        self.c1 = complex(c1)
