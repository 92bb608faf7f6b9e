"""MI355X backend for Inti.jl's corrected Nystrom operators.

Host-side mirror of the Inti.jl interface for this path (Quadrature,
Laplace/Helmholtz kernels, IntegralOperator, assemble_matrix,
bdim_correction, single_double_layer, vdim_correction, volume_potential,
gmres) over the C-ABI library
libinti_amd.so (hand-written gfx950 HIP kernels). The directory name is not a
Python identifier; load it with load_package() from __graft_entry__.
"""
from . import _lib
from .api import (CorrectedOperator, adj_double_layer_hypersingular, combined_field,
                  single_double_layer, single_double_layer_potential)
from .batch import BatchOperator
from .dense import DenseOperator
from .bdim import (DimParameters, bdim_correction, element_centers_and_diameters, etype_to_nearest_points,
                   near_interaction_list)
from .bdim_pullback import bdim_pullback
from .kernels import (AdjointDoubleLayerKernel, CombinedFieldKernel, DoubleLayerKernel,
                      FarFieldDoubleLayerKernel, FarFieldSingleLayerKernel, Helmholtz,
                      HyperSingularKernel, Laplace, SingleLayerKernel, default_density_eltype,
                      default_kernel_eltype)
from .nystrom import (DoubleLayerPotential, IntegralOperator, IntegralPotential,
                      SingleLayerPotential, SparseCorrection, _green_multiplier, assemble_matrix,
                      gmres, gmres_adjoint, refine_solve, rrule, set_default_symmetric)
from .quadrature import (Quadrature, area_quadrature, boundary_quadrature, box_quadratures, circle_quadrature, curve_quadrature,
                         disk_quadratures, ellipse_disk_quadratures, ellipse_quadrature,
                         gauss_legendre, gauss_legendre_order, join_quadratures, points, sphere_quadrature, surface_quadrature,
                         vioreanu_rokhlin_tetrahedron, vioreanu_rokhlin_triangle)
from . import gmsh, polynomials
from .vdim import vdim_auxiliary_quantities, vdim_correction, volume_potential

IntiError = _lib.IntiError
default_context = _lib.default_context
Context = _lib.Context

__all__ = [
    "AdjointDoubleLayerKernel", "CombinedFieldKernel", "Context", "CorrectedOperator",
    "DimParameters", "DoubleLayerKernel", "Helmholtz", "HyperSingularKernel",
    "IntegralOperator", "IntiError", "set_default_symmetric", "Laplace", "Quadrature", "SingleLayerKernel",
    "SparseCorrection", "_green_multiplier", "adj_double_layer_hypersingular",
    "assemble_matrix", "bdim_correction", "circle_quadrature", "combined_field",
    "area_quadrature", "disk_quadratures", "ellipse_disk_quadratures", "boundary_quadrature", "curve_quadrature",
    "polynomials", "gmsh", "vdim_auxiliary_quantities", "vdim_correction",
    "volume_potential", "IntegralPotential", "SingleLayerPotential", "DoubleLayerPotential",
    "single_double_layer_potential", "FarFieldSingleLayerKernel", "FarFieldDoubleLayerKernel",
    "rrule", "BatchOperator", "DenseOperator",
    "default_context", "default_density_eltype", "default_kernel_eltype",
    "ellipse_quadrature", "etype_to_nearest_points", "bdim_pullback", "near_interaction_list",
    "element_centers_and_diameters", "gauss_legendre", "gauss_legendre_order", "gmres", "join_quadratures",
    "gmres_adjoint",
    "refine_solve",
    "points", "single_double_layer", "sphere_quadrature", "surface_quadrature", "vioreanu_rokhlin_triangle",
    "vioreanu_rokhlin_tetrahedron", "box_quadratures",
]
