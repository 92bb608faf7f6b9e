"""
    IntiMI355X

Julia binding of libinti_amd (the MI355X backend of the corrected Nyström
operators). It keeps Inti's API: `IntegralOperator`s are uploaded once and
applied as `LinearMap`s; `bdim_correction` returns `SparseMatrixCSC`s.

Not exercised in this repository (no Julia toolchain in the build image);
the C-ABI it calls is tested from Python with the same argument conventions.
"""
module IntiMI355X

using Inti
using LinearAlgebra
using LinearMaps
using SparseArrays

const libinti = joinpath(@__DIR__, "..", "libinti_amd.so")

const LAPLACE, HELMHOLTZ = Cint(0), Cint(1)
const SL, DL, ADL, HS, COMB_SD, COMB_AH = Cint.(0:5)
const FP64, MIXED = Cint(0), Cint(1)

mutable struct Context
    ptr::Ptr{Cvoid}
end

function check(ctx::Ptr{Cvoid}, code::Cint)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:inti_last_error, libinti), Cstring, (Ptr{Cvoid},), ctx))
    code == 1 && throw(ArgumentError(msg))
    return error("libinti_amd error $code: $msg")
end

const CTX = Ref{Context}()
function __init__()
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:inti_ctx_create, libinti), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, p))
    CTX[] = Context(p[])
    finalizer(c -> ccall((:inti_ctx_destroy, libinti), Cint, (Ptr{Cvoid},), c.ptr), CTX[])
end

# QuadratureNode{N,T} (src/quadrature.jl:7-11) -> SoA, as the FMM ext does
function upload(Q::Inti.Quadrature{N}) where {N}
    n = length(Q)
    X = Matrix{Float64}(undef, n, N)        # column d = coordinate d (SoA)
    Nrm = Matrix{Float64}(undef, n, N)
    w = Vector{Float64}(undef, n)
    for (i, q) in enumerate(Q.qnodes)
        X[i, :] .= Inti.coords(q)
        Nrm[i, :] .= Inti.normal(q)
        w[i] = Inti.weight(q)
    end
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(CTX[].ptr, ccall((:inti_quad_upload, libinti), Cint,
        (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
        CTX[].ptr, N, n, X, Nrm, w, p))
    return p[]
end

kind_id(::Inti.SingleLayerKernel) = SL
kind_id(::Inti.DoubleLayerKernel) = DL
kind_id(::Inti.AdjointDoubleLayerKernel) = ADL
kind_id(::Inti.HyperSingularKernel) = HS
pde_id(::Inti.Laplace) = LAPLACE
pde_id(::Inti.Helmholtz) = HELMHOLTZ
wavenumber(::Inti.Laplace) = 0.0
wavenumber(op::Inti.Helmholtz) = Float64(op.k)

const QUADS = IdDict{Any,Ptr{Cvoid}}()
const HANDLES = IdDict{Any,Ptr{Cvoid}}()

quad_handle(Q) = get!(() -> upload(Q), QUADS, Q)

"""
    op_handle(iop::Inti.IntegralOperator; precision = FP64)

Device operator of `iop` (created once per operator and cached).
"""
function op_handle(iop::Inti.IntegralOperator; precision = FP64)
    return get!(HANDLES, iop) do
        K = Inti.kernel(iop)
        op = Inti.operator(K)
        N = Inti.ambient_dimension(iop.source)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(CTX[].ptr, ccall((:inti_op_create, libinti), Cint,
            (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
             Ref{Ptr{Cvoid}}),
            CTX[].ptr, pde_id(op), kind_id(K), N, wavenumber(op), C_NULL,
            quad_handle(iop.target), quad_handle(iop.source), precision, h))
        h[]
    end
end

"""
    gpu_operator(iop::Inti.IntegralOperator; precision = FP64)

Matrix-free GPU counterpart of `iop` as a `LinearMap` (mul! calls
`inti_op_apply`), to be combined with a sparse correction as
`gpu_operator(Sop) + LinearMap(δS)` (src/api.jl:182-185), or with the
correction attached on the device by `attach!`.
"""
function gpu_operator(iop::Inti.IntegralOperator; precision = FP64)
    h = op_handle(iop; precision)
    T = eltype(iop)
    m, n = size(iop)
    return LinearMap{T}(m, n) do y, x
        α = T[1]; β = T[0]
        check(CTX[].ptr, ccall((:inti_op_apply, libinti), Cint,
            (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}), h, 1, x, y, α, β))
        return y
    end
end

"""
    set_symmetric!(iop, mode)

Symmetric evaluation of an operator whose target is its source (the same
`Quadrature` object, hence the same device quadrature): 0 off, 1 automatic
(the default), 2 always.
"""
function set_symmetric!(iop::Inti.IntegralOperator, mode::Integer)
    check(CTX[].ptr, ccall((:inti_op_set_symmetric, libinti), Cint, (Ptr{Cvoid}, Cint, Cint),
        op_handle(iop), mode, 0))
    return iop
end

"""
    pullback(iop, ȳ, σ; dcoefs = nothing)

Reverse mode of `y = iop * σ` on the GPU: the gradients of Re⟨ȳ, y⟩ with
respect to σ (`iop' * ȳ`), the source nodes, normals and weights, the
target nodes and k. For an operator on one quadrature (`iop.target ===
iop.source`) the node gradient is the sum of the source and target parts.
"""
function pullback(iop::Inti.IntegralOperator, ȳ::Vector{T}, σ::Vector{T};
                  dcoefs = nothing) where {T}
    m, n = size(iop)
    N = Inti.ambient_dimension(iop.source)
    σ̄ = similar(σ)
    # gradients of points are n x N (column d = coordinate d), the layout of
    # the uploaded coordinates
    ys, ns, xs, ws = zeros(n, N), zeros(n, N), zeros(m, N), zeros(n)
    k̄ = Ref(0.0)
    dc = dcoefs === nothing ? C_NULL : Float64[real(dcoefs[1]), imag(dcoefs[1]), real(dcoefs[2]), imag(dcoefs[2])]
    check(CTX[].ptr, ccall((:inti_op_adjoint, libinti), Cint,
        (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
        op_handle(iop), ȳ, σ, σ̄, ys, ns, ws, xs, dc, k̄))
    nodes = iop.target === iop.source ? ys .+ xs : ys
    return (; sigma = σ̄, nodes, target_nodes = xs, normals = ns, weights = ws, k = k̄[])
end

"""
    assemble_matrix(iop) -> Matrix   (src/nystrom.jl:103-124 on the GPU)
"""
function assemble_matrix(iop::Inti.IntegralOperator)
    T = eltype(iop)
    out = Matrix{T}(undef, size(iop)...)
    h = op_handle(iop)
    check(CTX[].ptr, ccall((:inti_op_assemble_dense, libinti), Cint, (Ptr{Cvoid}, Ptr{T}), h, out))
    return out
end

# CSR (0-based, rows = targets) -> SparseMatrixCSC (src/bdim.jl:193-194)
function to_csc(csr::Ptr{Cvoid}, ::Type{T}) where {T}
    m, n, nnz, cplx = Ref{Int64}(), Ref{Int64}(), Ref{Int64}(), Ref{Cint}()
    ccall((:inti_csr_info, libinti), Cint,
          (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ref{Cint}), csr, m, n, nnz, cplx)
    rowptr = Vector{Int64}(undef, m[] + 1)
    colind = Vector{Int64}(undef, nnz[])
    val = Vector{T}(undef, nnz[])
    check(CTX[].ptr, ccall((:inti_csr_export, libinti), Cint,
        (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{T}), csr, rowptr, colind, val))
    I = Int[]
    for i in 1:m[], p in (rowptr[i] + 1):rowptr[i + 1]
        push!(I, i)
    end
    return sparse(I, colind .+ 1, val, m[], n[])
end

"""
    bdim_correction(op, target, source, Sop, Dop; green_multiplier,
                    parameters = Inti.DimParameters(), derivative = false)

GPU counterpart of `Inti.bdim_correction` (src/bdim.jl:50-61) for
`target === source` (self pattern) or explicit near lists.
"""
function bdim_correction(op, target, source::Inti.Quadrature, Sop, Dop;
                         green_multiplier::Vector{<:Real},
                         parameters = Inti.DimParameters(), derivative = false,
                         maxdist = Inf)
    T = eltype(Sop)
    (E, qtags), = collect(source.etype2qtags)          # single element type
    nq, ne = size(qtags)
    qt = Int64.(qtags .- 1)                             # 0-based, column-major
    near_ptr = near_idx = C_NULL
    if target !== source
        lists = Inti.etype_to_nearest_points(target, source; maxdist)[E]
        near_ptr = Int64[0; cumsum(length.(lists))]
        near_idx = Int64.(reduce(vcat, lists; init = Int[]) .- 1)
    end
    params = Float64[parameters.sources_oversample_factor, parameters.sources_radius_multiplier]
    dS, dD = Ref{Ptr{Cvoid}}(C_NULL), Ref{Ptr{Cvoid}}(C_NULL)
    check(CTX[].ptr, ccall((:inti_bdim_build, libinti), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
         Ptr{Cdouble}, Ptr{Cdouble}, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}, Ptr{Cdouble}),
        op_handle(Sop), op_handle(Dop), derivative, nq, ne, qt, near_ptr, near_idx,
        Float64.(green_multiplier), params, dS, dD, C_NULL))
    return to_csc(dS[], T), to_csc(dD[], T)
end

end # module
