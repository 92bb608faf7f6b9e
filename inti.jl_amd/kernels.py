"""PDE and kernel descriptors mirroring Inti.jl src/kernels.jl.

Only the Laplace (:118-204) and Helmholtz (:310-413) operators in 2D/3D are
provided; values are computed on the GPU by libinti_amd.
"""
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

    Complex k with zero imaginary part is taken as real; a genuinely complex
    or purely imaginary k (the reference's Yukawa remap) is not supported by
    the GPU kernels and raises."""

    pde = _lib.HELMHOLTZ

    def __init__(self, *, k, dim):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        if isinstance(k, complex):
            if k.imag != 0:
                raise NotImplementedError("complex wavenumbers are not supported on the GPU path")
            k = k.real
        if not k > 0:
            raise ValueError("k must be positive")
        self.dim = dim
        self.k = float(k)

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
