"""
    IntiMI355X

Julia binding of libinti_amd (the MI355X backend of the corrected Nyström
operators). It keeps Inti's API: `IntegralOperator`s are uploaded once and
applied as `LinearMap`s (`GPUOperator`); `bdim_correction` returns
`SparseMatrixCSC`s; `single_double_layer` / `adj_double_layer_hypersingular`
return corrected operators (dense matrices with `compression = (method = :none,)`);
`*(::GPUOperator, σ)` has a ChainRulesCore `rrule`.

Not exercised in this repository (no Julia toolchain in the build image);
the C-ABI it calls is tested from Python with the same argument conventions
and call sequences (tests/test_gpu_shim_sequences.py).
"""
module IntiMI355X

using Inti
using LinearAlgebra
using LinearMaps
using SparseArrays
using ChainRulesCore

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

# QuadratureNode{N,T} (src/quadrature.jl:7-11) -> SoA, as the FMM ext does.
# Nodes without normals (volume quadratures, point targets) upload none.
function upload(Q::Inti.Quadrature{N}) where {N}
    n = length(Q)
    X = Matrix{Float64}(undef, n, N)        # column d = coordinate d (SoA)
    has_normals = n > 0 && Q.qnodes[1].normal !== nothing
    Nrm = has_normals ? Matrix{Float64}(undef, n, N) : nothing
    w = Vector{Float64}(undef, n)
    for (i, q) in enumerate(Q.qnodes)
        X[i, :] .= Inti.coords(q)
        has_normals && (Nrm[i, :] .= Inti.normal(q))
        w[i] = Inti.weight(q)
    end
    p = Ref{Ptr{Cvoid}}(C_NULL)
    nrm_ptr = has_normals ? pointer(Nrm) : Ptr{Float64}(C_NULL)
    GC.@preserve Nrm begin
        check(CTX[].ptr, ccall((:inti_quad_upload, libinti), Cint,
            (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
            CTX[].ptr, N, n, X, nrm_ptr, w, p))
    end
    return p[]
end

kind_id(::Inti.SingleLayerKernel) = SL
kind_id(::Inti.DoubleLayerKernel) = DL
kind_id(::Inti.AdjointDoubleLayerKernel) = ADL
kind_id(::Inti.HyperSingularKernel) = HS
pde_id(::Inti.Laplace) = LAPLACE
pde_id(::Inti.Helmholtz) = HELMHOLTZ
wavenumber(::Inti.Laplace) = 0.0
wavenumber(op::Inti.Helmholtz) = Float64(real(op.k))
wavenumber_imag(::Inti.Laplace) = 0.0
wavenumber_imag(op::Inti.Helmholtz) = Float64(imag(op.k))

const QUADS = IdDict{Any,Ptr{Cvoid}}()
# one device operator per (IntegralOperator, precision): a mixed-precision
# handle must never be returned for an FP64 request (the adjoint is FP64 only)
const HANDLES = Dict{Tuple{UInt,Cint},Tuple{Any,Ptr{Cvoid}}}()

quad_handle(Q) = get!(() -> upload(Q), QUADS, Q)

"""
    op_handle(iop::Inti.IntegralOperator; precision = FP64)

Device operator of `iop` at `precision` (created once per pair and cached).
"""
function op_handle(iop::Inti.IntegralOperator; precision = FP64)
    key = (objectid(iop), Cint(precision))
    entry = get(HANDLES, key, nothing)
    entry !== nothing && entry[1] === iop && return entry[2]
    K = Inti.kernel(iop)
    op = Inti.operator(K)
    N = Inti.ambient_dimension(iop.source)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(CTX[].ptr, ccall((:inti_op_create, libinti), Cint,
        (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Ptr{Cdouble}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
         Ref{Ptr{Cvoid}}),
        CTX[].ptr, pde_id(op), kind_id(K), N, wavenumber(op), C_NULL,
        quad_handle(iop.target), quad_handle(iop.source), precision, h))
    ki = wavenumber_imag(op)
    if ki != 0   # Helmholtz(; k) with a complex k (3D, FP64)
        check(CTX[].ptr, ccall((:inti_op_set_wavenumber_imag, libinti), Cint, (Ptr{Cvoid}, Cdouble), h[], ki))
    end
    HANDLES[key] = (iop, h[])   # the operator is kept alive with its handle
    return h[]
end

"""
    GPUOperator{T}

Matrix-free GPU counterpart of an `IntegralOperator` (with any attached
DIM correction and identity term) as a `LinearMap`: `mul!(y, A, x, α, β)`
calls `inti_op_apply`.
"""
struct GPUOperator{T} <: LinearMaps.LinearMap{T}
    handle::Ptr{Cvoid}
    m::Int
    n::Int
    iop::Inti.IntegralOperator
    precision::Cint
end

Base.size(A::GPUOperator) = (A.m, A.n)

function LinearMaps._unsafe_mul!(y::AbstractVecOrMat, A::GPUOperator{T}, x::AbstractVecOrMat,
                                 α::Number = true, β::Number = false) where {T}
    xs, ys = convert(Array{T}, x), convert(Array{T}, y)
    a, b = T[α], T[β]
    check(CTX[].ptr, ccall((:inti_op_apply, libinti), Cint,
        (Ptr{Cvoid}, Cint, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}), A.handle, size(xs, 2), xs, ys, a, b))
    y === ys || copyto!(y, ys)
    return y
end

"""
    gpu_operator(iop::Inti.IntegralOperator; precision = FP64)

Matrix-free GPU counterpart of `iop`, to be combined with a sparse
correction as `gpu_operator(Sop) + LinearMap(δS)` (src/api.jl:182-185), or
with the correction attached on the device by `attach!`.
"""
function gpu_operator(iop::Inti.IntegralOperator; precision = FP64)
    h = op_handle(iop; precision)
    m, n = size(iop)
    return GPUOperator{eltype(iop)}(h, m, n, iop, Cint(precision))
end

"""
    attach!(A::GPUOperator, csr::Ptr{Cvoid}, coef = 1)

Fuse `coef * δA` (a device CSR from the DIM build) into every apply of `A`.
"""
function attach!(A::GPUOperator, csr::Ptr{Cvoid}, coef::Number = 1)
    c = ComplexF64(coef)
    check(CTX[].ptr, ccall((:inti_op_attach_correction, libinti), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble), A.handle, csr, real(c), imag(c)))
    return A
end

"""
    set_symmetric!(A, mode)

Symmetric evaluation of an operator whose target is its source (the same
`Quadrature` object, hence the same device quadrature): 0 off, 1 automatic
(the default), 2 always.
"""
function set_symmetric!(A::GPUOperator, mode::Integer)
    check(CTX[].ptr, ccall((:inti_op_set_symmetric, libinti), Cint, (Ptr{Cvoid}, Cint, Cint),
        A.handle, mode, 0))
    return A
end
set_symmetric!(iop::Inti.IntegralOperator, mode::Integer) = set_symmetric!(gpu_operator(iop), mode)

"""
    pullback(A, ȳ, σ; dcoefs = nothing, geometry = true)

Reverse mode of `y = A * σ` on the GPU: the gradients of Re⟨ȳ, y⟩ with
respect to σ (`A' * ȳ`) and, if `geometry`, the source nodes, normals and
weights, the target nodes and k. For an operator on one quadrature
(`iop.target === iop.source`) the node gradient is the sum of the source
and target parts. With a DIM correction attached only σ̄ is available.
"""
function pullback(A::GPUOperator, ȳ::Vector{T}, σ::Vector{T}; dcoefs = nothing,
                  geometry = true) where {T}
    m, n = size(A)
    iop = A.iop
    N = Inti.ambient_dimension(iop.source)
    σ̄ = similar(σ)
    # gradients of points are n x N (column d = coordinate d), the layout of
    # the uploaded coordinates
    ys, ns, xs, ws = zeros(n, N), zeros(n, N), zeros(m, N), zeros(n)
    k̄ = Ref(0.0)
    dc = dcoefs === nothing ? Ptr{Float64}(C_NULL) :
         Float64[real(dcoefs[1]), imag(dcoefs[1]), real(dcoefs[2]), imag(dcoefs[2])]
    if geometry
        check(CTX[].ptr, ccall((:inti_op_adjoint, libinti), Cint,
            (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
            A.handle, ȳ, σ, σ̄, ys, ns, ws, xs, dc, k̄))
    else
        check(CTX[].ptr, ccall((:inti_op_adjoint, libinti), Cint,
            (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            A.handle, ȳ, σ, σ̄, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL))
    end
    nodes = iop.target === iop.source ? ys .+ xs : ys
    return (; sigma = σ̄, nodes, target_nodes = xs, normals = ns, weights = ws, k = k̄[])
end
pullback(iop::Inti.IntegralOperator, ȳ, σ; kw...) = pullback(gpu_operator(iop), ȳ, σ; kw...)

"""
    rrule(*, A::GPUOperator, σ)

ChainRulesCore rule of the GPU apply: the pullback recomputes kernel values
(nothing N×N is stored). The operator's tangent carries the gradients of
its geometry and wavenumber (`nodes`, `normals`, `weights`, `k`); for an
operator with a DIM correction attached (whose weights depend on the
geometry through the least squares) only σ̄ is returned.
"""
function ChainRulesCore.rrule(::typeof(*), A::GPUOperator{T}, σ::Vector{T}) where {T}
    y = A * σ
    function times_pullback(ȳ)
        ȳv = convert(Vector{T}, unthunk(ȳ))
        geometry = !has_correction(A)
        g = pullback(A, ȳv, σ; geometry)
        dA = geometry ? Tangent{typeof(A)}(; nodes = g.nodes, normals = g.normals,
                                            weights = g.weights, k = g.k) : NoTangent()
        return NoTangent(), dA, g.sigma
    end
    return y, times_pullback
end

const CORRECTED = Set{Ptr{Cvoid}}()
has_correction(A::GPUOperator) = A.handle in CORRECTED

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

# the device CSR pair (δS, δD) of the DIM build
function bdim_build(op, target, source::Inti.Quadrature, Sop, Dop; green_multiplier,
                    parameters = Inti.DimParameters(), derivative = false, maxdist = Inf)
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
    return dS[], dD[]
end

free_csr(c::Ptr{Cvoid}) = ccall((:inti_csr_destroy, libinti), Cint, (Ptr{Cvoid},), c)

"""
    bdim_correction(op, target, source, Sop, Dop; green_multiplier,
                    parameters = Inti.DimParameters(), derivative = false)

GPU counterpart of `Inti.bdim_correction` (src/bdim.jl:50-61) for
`target === source` (self pattern) or explicit near lists.
"""
function bdim_correction(op, target, source::Inti.Quadrature, Sop, Dop; kw...)
    T = eltype(Sop)
    dS, dD = bdim_build(op, target, source, Sop, Dop; kw...)
    out = (to_csc(dS, T), to_csc(dD, T))
    free_csr(dS); free_csr(dD)
    return out
end

"""
    single_double_layer(; op, target, source = target, compression = (method = :gpu,),
                        correction = (method = :dim,), derivative = false, precision = FP64)

`Inti.single_double_layer` (src/api.jl:65-187) on the GPU. With
`compression.method == :gpu` the two operators are matrix-free
`GPUOperator`s with the DIM correction fused into every apply (the
`LinearMap(Smat) + LinearMap(δS)` of src/api.jl:182-185); with `:none`
they are dense matrices (`assemble_matrix` + `axpy!` of the correction,
src/api.jl:171-173). `derivative = true` gives the adjoint double layer and
hypersingular pair.
"""
function single_double_layer(; op, target, source = target, compression = (method = :gpu,),
                             correction = (method = :dim,), derivative = false,
                             precision = FP64)
    compression.method in (:gpu, :none) ||
        throw(ArgumentError("compression.method must be :gpu or :none"))
    correction.method in (:dim, :none) ||
        throw(ArgumentError("correction.method must be :dim or :none"))
    G = derivative ? Inti.AdjointDoubleLayerKernel(op) : Inti.SingleLayerKernel(op)
    dG = derivative ? Inti.HyperSingularKernel(op) : Inti.DoubleLayerKernel(op)
    Sop = Inti.IntegralOperator(G, target, source)
    Dop = Inti.IntegralOperator(dG, target, source)
    dS = dD = C_NULL
    if correction.method == :dim
        μ = if haskey(correction, :green_multiplier)
            correction.green_multiplier
        else
            loc = target === source ? :on : correction.target_location
            fill(Inti._green_multiplier(loc), length(target))
        end
        dS, dD = bdim_build(op, target, source, Sop, Dop; green_multiplier = μ,
                            parameters = get(correction, :parameters, Inti.DimParameters()),
                            derivative, maxdist = get(correction, :maxdist, Inf))
    end
    if compression.method == :none
        S, D = assemble_matrix(Sop), assemble_matrix(Dop)
        if dS != C_NULL
            T = eltype(Sop)
            axpy!(true, to_csc(dS, T), S)
            axpy!(true, to_csc(dD, T), D)
            free_csr(dS); free_csr(dD)
        end
        return S, D
    end
    S = gpu_operator(Sop; precision)
    D = gpu_operator(Dop; precision)
    if dS != C_NULL
        attach!(S, dS); attach!(D, dD)       # the operators keep merged copies
        push!(CORRECTED, S.handle, D.handle)
        free_csr(dS); free_csr(dD)
    end
    return S, D
end

"""
    adj_double_layer_hypersingular(; op, target, source = target, kw...)

`Inti.adj_double_layer_hypersingular` (src/api.jl:197-212) on the GPU.
"""
adj_double_layer_hypersingular(; op, target, source = target, kw...) =
    single_double_layer(; op, target, source, derivative = true, kw...)

end # module
