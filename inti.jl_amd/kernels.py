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


class CombinedFieldKernel(_Kernel):
    """c0*SL + c1*DL (pair='SD') or c0*ADL + c1*HS (pair='AH'), evaluated in
    one pass (shares the distance and e^{ikd} / Hankel evaluation)."""

    def __init__(self, op, c0, c1, pair="SD"):
        super().__init__(op)
        if pair not in ("SD", "AH"):
            raise ValueError("pair must be 'SD' or 'AH'")
        self.kind = _lib.COMB_SD if pair == "SD" else _lib.COMB_AH
        self.pair = pair
        self.c0 = complex(c0)
        self.c1 = complex(c1)
