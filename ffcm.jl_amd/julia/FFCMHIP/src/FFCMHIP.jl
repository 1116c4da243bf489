# MI355X (gfx950) backend of ForceCouplingMethod.jl's `gpu_acceleration` path: a package that
# binds libffcm_hip.so through `ccall` and plugs into the reference's own seams, with no edit to
# the reference. `using ForceCouplingMethod, FFCMHIP` makes `FFCMConfig(...; gpu_acceleration =
# true)` build HIP-backed buffers.
#
# Seams (the reference calls them; this package adds methods):
#   _assemble_gpu_buffers(T, N, num_grid_points, M_G, num_cells_total, k_x, k_y, k_z,
#                         neighbor_map_host)                       src/config.jl:495-502
#   the L1 step barriers of mobility!'s generic body               src/mobility.jl:35-43
#   mul!(f̂_c, forward plan, f_c) / mul!(u_c, inverse plan, f̂_c)    src/stokes_solve.jl:24-40
#   _stage_mobility_io! / _retrieve_mobility_output!               src/mobility.jl:55-56
# and adds a fused `mobility!` for HIP-backed configs with host matrices (one ffcm_mobility
# call: every step on one HIP stream), a position-cached `mul!` for FFCMMobility over a
# HIP-backed config, and the ChainRulesCore rrule of the out-of-place apply.
#
# `_assemble_gpu_buffers` is defined for the argument types the reference passes; loading this
# package next to the reference's CUDA extension replaces that extension's method (load one GPU
# backend per session).
#
# The buffers of a HIP-backed config are `HIPBuffer`s: handles on the context's device buffers
# with `size`, `Array(x)` (device -> host) and `copyto!(x, host)` (host -> device), the
# operations the reference's GPU tests apply to config fields (test/cuda/*.jl). Grid fields
# and the spectrum are StructArrays of per-component handles, so `components(...)` works as in
# src/stokes_solve.jl:22-23.
module FFCMHIP

using ForceCouplingMethod
using ForceCouplingMethod: FFCMConfig, FFCMMobility, CellBuffers, ParticleBuffers, GridBuffers, SolverState,
                           StencilBuffers
using StaticArrays: SVector
using StructArrays: StructArray
import ChainRulesCore
import LinearAlgebra
import LinearAlgebra: mul!

import ForceCouplingMethod:
    _apply_inverse_stokes_kernel!, _assemble_gpu_buffers, _assign_cells_kernel!,
    _build_cell_list_kernel!, _correct_velocities_kernel!, _gather_particles_kernel!,
    _interpolate_velocities_kernel!, _retrieve_mobility_output!, _spread_forces_kernel!,
    _stage_mobility_io!, mobility!, wrap_positions!

export mobility_vjp!, apply_mobility

const libffcm = get(ENV, "FFCM_HIP_LIB", joinpath(@__DIR__, "..", "..", "..", "libffcm_hip.so"))

# ffcm_params (include/ffcm_hip.h)
struct FfcmParams
    dtype::Int32
    pad0::Int32
    N::Int64
    M::NTuple{3, Int32}
    M_G::Int32
    num_cells::NTuple{3, Int32}
    pad1::Int32
    L::NTuple{3, Float64}
    R_c::Float64
    a::Float64
    sigma::Float64
    Sigma::Float64
    h::Float64
    inv_h::Float64
    mu::Float64
    cell_size::NTuple{3, Float64}
    inv_cell_size::NTuple{3, Float64}
    k_x::Ptr{Cvoid}
    k_y::Ptr{Cvoid}
    k_z::Ptr{Cvoid}
end

const BUF = (Y_input = 0, F_input = 1, V_output = 2, Y_wrapped = 3, Y_sorted = 4, F_sorted = 5,
             cell_hash = 6, original_index = 7, cell_start = 8, cell_end = 9, force_density = 10,
             fluid_velocity = 11, fluid_hat = 12, neighbor_map = 13, counting_sort_scratch = 14)

mutable struct HIPContext
    ptr::Ptr{Cvoid}
    scalars_set::Bool
    cached_Y::Union{Nothing, Matrix}  # positions whose cell list the context holds (FFCMMobility)
end

_dtype(::Type{Float32}) = Int32(0)
_dtype(::Type{Float64}) = Int32(1)

function _check(ctx::Ptr{Cvoid}, rc::Cint, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:ffcm_last_error, libffcm), Cstring, (Ptr{Cvoid},), ctx))
    rc == 1 && throw(ArgumentError("$what: $msg"))
    rc == 2 && throw(DimensionMismatch("$what: $msg"))
    rc == 4 && throw(OutOfMemoryError())
    error("$what: $msg")
end

# Handle on one context buffer; indexing goes through a host copy (scalar device indexing is
# not a hot-path operation, as with CuArray's allowscalar(false)).
struct HIPBuffer{T, N} <: AbstractArray{T, N}
    ctx::HIPContext
    id::Int32
    dims::NTuple{N, Int}
end
Base.size(b::HIPBuffer) = b.dims
Base.IndexStyle(::Type{<:HIPBuffer}) = IndexLinear()
function Base.Array(b::HIPBuffer{T, N}) where {T, N}
    out = Array{T, N}(undef, b.dims)
    _check(b.ctx.ptr, ccall((:ffcm_get_buffer, libffcm), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t),
                            b.ctx.ptr, b.id, out, sizeof(out)), "Array")
    return out
end
Base.getindex(b::HIPBuffer, i::Int) = Array(b)[i]
function Base.copyto!(b::HIPBuffer{T}, src::AbstractArray{T}) where {T}
    host = Array(src)
    _check(b.ctx.ptr, ccall((:ffcm_set_buffer, libffcm), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t),
                            b.ctx.ptr, b.id, host, sizeof(host)), "copyto!")
    return b
end
Base.copyto!(dst::AbstractArray{T}, b::HIPBuffer{T}) where {T} = copyto!(dst, Array(b))

# One component of a grid field (force_density, fluid_velocity: real, (Mx, My, Mz)) or of the
# spectrum (fluid_hat: complex, (Mx÷2+1, My, Mz)). The context's buffer holds the three
# components of a field; the buffer copies convert to the reference's SoA layout.
struct HIPComponent{T, E} <: AbstractArray{E, 3}
    field::HIPBuffer{E, 4}  # (3, dims...) in the reference layout
    c::Int
end
Base.size(x::HIPComponent) = x.field.dims[2:4]
Base.Array(x::HIPComponent) = Array(x.field)[x.c, :, :, :]
Base.getindex(x::HIPComponent, I::Vararg{Int, 3}) = Array(x)[I...]
_ctxof(x::HIPComponent) = x.field.ctx.ptr

const HIPReal{T} = HIPComponent{T, T}
const HIPSpectral{T} = HIPComponent{T, Complex{T}}

# The two transform plans of SolverState (src/config.jl:102-109): handles of the context's FFT.
struct HIPForwardPlan
    ctx::HIPContext
end
struct HIPInversePlan
    ctx::HIPContext
end

# _assemble_gpu_buffers (src/config.jl:461-464, called at :495-502): the nine arguments the
# reference passes, the five sub-structs in the reference's field order.
function _assemble_gpu_buffers(::Type{T}, N::Integer, num_grid_points::NTuple{3, Int32}, M_G::Int32,
                               num_cells_total::Integer, k_x::Vector{T}, k_y::Vector{T}, k_z::Vector{T},
                               neighbor_map_host::Vector{Int32}) where {T <: Union{Float32, Float64}}
    out = Ref{Ptr{Cvoid}}(C_NULL)
    M = Int32[num_grid_points...]
    rc = GC.@preserve M k_x k_y k_z neighbor_map_host ccall(
        (:ffcm_assemble, libffcm), Cint,
        (Ref{Ptr{Cvoid}}, Int32, Int64, Ptr{Int32}, Int32, Int64, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{Int32}),
        out, _dtype(T), N, M, M_G, num_cells_total, k_x, k_y, k_z, neighbor_map_host)
    _check(C_NULL, rc, "FFCMConfig")
    ctx = HIPContext(out[], false, nothing)
    finalizer(c -> ccall((:ffcm_destroy, libffcm), Cint, (Ptr{Cvoid},), c.ptr), ctx)
    Mx, My, Mz = Int.(num_grid_points)
    C = Int(num_cells_total)
    ivec(id, n) = HIPBuffer{Int32, 1}(ctx, id, (n,))
    mat(id) = HIPBuffer{T, 2}(ctx, id, (3, Int(N)))
    cells = CellBuffers(ivec(BUF.cell_hash, Int(N)), ivec(BUF.original_index, Int(N)), ivec(BUF.cell_start, C),
                        ivec(BUF.cell_end, C), ivec(BUF.counting_sort_scratch, C),
                        ivec(BUF.neighbor_map, 13C))
    particles = ParticleBuffers(mat(BUF.Y_wrapped), mat(BUF.Y_sorted), mat(BUF.F_sorted),
                                mat(BUF.Y_input), mat(BUF.F_input), mat(BUF.V_output))
    fd = HIPBuffer{T, 4}(ctx, BUF.force_density, (3, Mx, My, Mz))
    fv = HIPBuffer{T, 4}(ctx, BUF.fluid_velocity, (3, Mx, My, Mz))
    real3(f) = StructArray{SVector{3, T}}((HIPComponent{T, T}(f, 1), HIPComponent{T, T}(f, 2), HIPComponent{T, T}(f, 3)))
    grid = GridBuffers(real3(fd), real3(fv))
    fh = HIPBuffer{Complex{T}, 4}(ctx, BUF.fluid_hat, (3, Mx ÷ 2 + 1, My, Mz))
    fluid_hat = StructArray{SVector{3, Complex{T}}}((HIPComponent{T, Complex{T}}(fh, 1),
                                                     HIPComponent{T, Complex{T}}(fh, 2),
                                                     HIPComponent{T, Complex{T}}(fh, 3)))
    solver = SolverState(fluid_hat, k_x, k_y, k_z, HIPForwardPlan(ctx), HIPInversePlan(ctx))
    stencil = StencilBuffers(nothing, nothing, nothing)  # the kernels build their weights in registers
    return (cells, particles, grid, solver, stencil)
end

const HIPConfig{T} = FFCMConfig{T, <:CellBuffers{<:HIPBuffer}}
_context(config::HIPConfig) = config.cells.cell_hash.ctx
_ctx(config::HIPConfig) = _context(config).ptr

# ffcm_set_scalars: every derived scalar of the config (computed by Julia in T), once.
function _ensure_scalars!(config::HIPConfig{T}) where {T}
    c = _context(config)
    c.scalars_set && return nothing
    k_x, k_y, k_z = config.solver.k_x, config.solver.k_y, config.solver.k_z
    p = FfcmParams(_dtype(T), 0, size(config.particles.Y_sorted, 2), config.num_grid_points, config.M_G,
                   config.num_cells, 0, Float64.(config.L), config.R_c, config.a, config.σ, config.Σ,
                   config.h, config.inv_h, config.μ, Float64.(config.cell_size), Float64.(config.inv_cell_size),
                   pointer(k_x), pointer(k_y), pointer(k_z))
    GC.@preserve k_x k_y k_z _check(c.ptr, ccall((:ffcm_set_scalars, libffcm), Cint, (Ptr{Cvoid}, Ref{FfcmParams}),
                                                 c.ptr, p), "FFCMConfig")
    c.scalars_set = true
    return nothing
end

function _shape_guard(N, arrays...)
    all(size(x) == (3, N) for x in arrays) || throw(DimensionMismatch(
        "V, Y, and F must each be 3xN for the config's N = $(N); got sizes $(map(size, arrays))"))
end

# mobility! (src/mobility.jl:23-45) for host matrices: one call runs the staged pipeline on the
# context stream (page-locked staging, one synchronisation at the end).
function mobility!(V::Matrix{T}, config::HIPConfig{T}, Y::Matrix{T}, F::Matrix{T}) where {T}
    N = size(config.particles.Y_sorted, 2)
    _shape_guard(N, V, Y, F)
    _ensure_scalars!(config)
    _context(config).cached_Y = nothing
    _check(_ctx(config), ccall((:ffcm_mobility, libffcm), Cint,
                               (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Int64), _ctx(config), V, Y, F, N), "mobility!")
    return V
end

# Device-resident mobility! on raw device pointers (3xN column-major in T): the zero-copy
# branch of _stage_mobility_io! (ext/ForceCouplingMethodCUDAExt/mobility.jl:1-8). Enqueued on
# the context stream; FFCMHIPAMDGPUExt wraps it for ROCArrays.
function mobility_device!(V::Ptr{T}, config::HIPConfig{T}, Y::Ptr{T}, F::Ptr{T}) where {T}
    _ensure_scalars!(config)
    _context(config).cached_Y = nothing
    N = size(config.particles.Y_sorted, 2)
    _check(_ctx(config), ccall((:ffcm_mobility_device, libffcm), Cint,
                               (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Int64), _ctx(config), V, Y, F, N), "mobility!")
    return V
end
synchronize(config::HIPConfig) = _check(_ctx(config), ccall((:ffcm_synchronize, libffcm), Cint, (Ptr{Cvoid},),
                                                             _ctx(config)), "synchronize")

# FFCMMobility mul! (src/mobility.jl:115-147) over a HIP-backed config: the cell list of M.Y
# stays in the context between applies (ffcm_set_positions once, ffcm_apply_forces per mul!);
# it is rebuilt when M.Y differs from the positions it was built for, or another call replaced
# it, so the result is always mobility!(v, config, M.Y, f).
function _apply_cached!(V::Matrix{T}, M::FFCMMobility{T, <:HIPConfig{T}}, F::Matrix{T}) where {T}
    config = M.config
    _ensure_scalars!(config)
    c = _context(config)
    N = size(M.Y, 2)
    if c.cached_Y === nothing || c.cached_Y != M.Y
        Yh = Matrix{T}(M.Y)
        _check(c.ptr, ccall((:ffcm_set_positions, libffcm), Cint, (Ptr{Cvoid}, Ptr{T}, Int64), c.ptr, Yh, N),
               "mul!")
        c.cached_Y = Yh
    end
    _check(c.ptr, ccall((:ffcm_apply_forces, libffcm), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Int64), c.ptr, V, F, N),
           "mul!")
    return V
end

function mul!(v::AbstractVector{T}, M::FFCMMobility{T, <:HIPConfig{T}}, f::AbstractVector{T}) where {T}
    copyto!(M.F_scratch, f)
    _apply_cached!(M.V_scratch, M, M.F_scratch)
    copyto!(v, M.V_scratch)
    return v
end

function mul!(v::AbstractVector{T}, M::FFCMMobility{T, <:HIPConfig{T}}, f::AbstractVector{T}, α::Number,
              β::Number) where {T}
    copyto!(M.F_scratch, f)
    _apply_cached!(M.V_scratch, M, M.F_scratch)
    αT, βT = convert(T, α), convert(T, β)
    if iszero(βT)
        @inbounds for k in eachindex(v)
            v[k] = αT * M.V_scratch[k]
        end
    else
        @inbounds for k in eachindex(v)
            v[k] = αT * M.V_scratch[k] + βT * v[k]
        end
    end
    return v
end

# ---------------------------------------------------------------------------------------------
# The L1 barriers of mobility!'s generic body, for HIP buffers: each passes the scalars the
# reference gives it to the matching per-barrier entry point, which runs the step on the
# context's buffers (the arguments name those same buffers).
_c(b::HIPBuffer) = b.ctx.ptr
_i3(t) = Int32[t...]
_d3(t) = Float64[t...]

function wrap_positions!(dst::HIPBuffer{T, 2}, src::HIPBuffer{T, 2}, L::NTuple{3, T}) where {T}
    _check(_c(dst), ccall((:ffcm_wrap_positions_kernel, libffcm), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}),
                          _c(dst), dst.id, src.id, _d3(L)), "wrap_positions!")
    dst.ctx.cached_Y = nothing
    return dst
end

function _assign_cells_kernel!(cell_hash::HIPBuffer{Int32, 1}, Y::HIPBuffer{T, 2}, inv_cell_size::NTuple{3, T},
                               num_cells::NTuple{3, Int32}) where {T}
    _check(_c(Y), ccall((:ffcm_assign_cells_kernel, libffcm), Cint,
                        (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ptr{Int32}),
                        _c(Y), cell_hash.id, Y.id, _d3(inv_cell_size), _i3(num_cells)), "_assign_cells_kernel!")
    return cell_hash
end

function _build_cell_list_kernel!(original_index::HIPBuffer{Int32, 1}, cell_start::HIPBuffer{Int32, 1},
                                  cell_end::HIPBuffer{Int32, 1}, counting_sort_scratch::HIPBuffer{Int32, 1},
                                  cell_hash::HIPBuffer{Int32, 1})
    _check(_c(cell_hash), ccall((:ffcm_build_cell_list_kernel, libffcm), Cint, (Ptr{Cvoid},), _c(cell_hash)),
           "_build_cell_list_kernel!")
    return original_index
end

function _gather_particles_kernel!(Y_sorted::HIPBuffer{T, 2}, F_sorted::HIPBuffer{T, 2}, Y::HIPBuffer{T, 2},
                                   F::HIPBuffer{T, 2}, original_index::HIPBuffer{Int32, 1}) where {T}
    _check(_c(Y), ccall((:ffcm_gather_particles_kernel, libffcm), Cint, (Ptr{Cvoid}, Cint, Cint), _c(Y), Y.id, F.id),
           "_gather_particles_kernel!")
    return Y_sorted
end

function _spread_forces_kernel!(force_density, Y_sorted::HIPBuffer{T, 2}, F_sorted::HIPBuffer{T, 2}, σ::T, Σ::T,
                                h::T, inv_h::T, num_grid_points::NTuple{3, Int32}, M_G::Int32, stencil_gaussian,
                                stencil_r², stencil_index) where {T}
    _check(_c(Y_sorted), ccall((:ffcm_spread_forces_kernel, libffcm), Cint,
                               (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Ptr{Int32}, Int32),
                               _c(Y_sorted), σ, Σ, h, inv_h, _i3(num_grid_points), M_G), "_spread_forces_kernel!")
    return force_density
end

# stokes_solve! (src/stokes_solve.jl:21-42): mul! with the two plans, and the projection.
function mul!(out::HIPSpectral{T}, p::HIPForwardPlan, x::HIPReal{T}) where {T}
    out.c == x.c || throw(ArgumentError("forward transform between different components"))
    _check(p.ctx.ptr, ccall((:ffcm_forward_transform, libffcm), Cint, (Ptr{Cvoid}, Cint), p.ctx.ptr, x.c - 1),
           "mul!(fluid_hat, forward_fourier_transform, force_density)")
    return out
end

function mul!(out::HIPReal{T}, p::HIPInversePlan, x::HIPSpectral{T}) where {T}
    out.c == x.c || throw(ArgumentError("inverse transform between different components"))
    _check(p.ctx.ptr, ccall((:ffcm_inverse_transform, libffcm), Cint, (Ptr{Cvoid}, Cint), p.ctx.ptr, out.c - 1),
           "mul!(fluid_velocity, inverse_fourier_transform, fluid_hat)")
    return out
end

function _apply_inverse_stokes_kernel!(fx̂::HIPSpectral{T}, fŷ::HIPSpectral{T}, fẑ::HIPSpectral{T}, k_x::Vector{T},
                                       k_y::Vector{T}, k_z::Vector{T}, μ::T, inv_M::T) where {T}
    _check(_ctxof(fx̂), ccall((:ffcm_apply_inverse_stokes_kernel, libffcm), Cint, (Ptr{Cvoid}, Float64, Float64),
                             _ctxof(fx̂), μ, inv_M), "_apply_inverse_stokes_kernel!")
    return nothing
end

function _interpolate_velocities_kernel!(V::HIPBuffer{T, 2}, fluid_velocity, Y_sorted::HIPBuffer{T, 2},
                                         original_index::HIPBuffer{Int32, 1}, σ::T, Σ::T, h::T, inv_h::T,
                                         num_grid_points::NTuple{3, Int32}, M_G::Int32, stencil_gaussian,
                                         stencil_r², stencil_index) where {T}
    _check(_c(V), ccall((:ffcm_interpolate_velocities_kernel, libffcm), Cint,
                        (Ptr{Cvoid}, Cint, Float64, Float64, Float64, Float64, Ptr{Int32}, Int32),
                        _c(V), V.id, σ, Σ, h, inv_h, _i3(num_grid_points), M_G), "_interpolate_velocities_kernel!")
    return V
end

function _correct_velocities_kernel!(V::HIPBuffer{T, 2}, Y_sorted::HIPBuffer{T, 2}, F_sorted::HIPBuffer{T, 2},
                                     cell_start::HIPBuffer{Int32, 1}, cell_end::HIPBuffer{Int32, 1},
                                     original_index::HIPBuffer{Int32, 1}, neighbor_map::HIPBuffer{Int32, 1},
                                     num_cells::NTuple{3, Int32}, inv_cell_size::NTuple{3, T}, L::NTuple{3, T}, σ::T,
                                     Σ::T, a::T, μ::T, R_c::T) where {T}
    _check(_c(V), ccall((:ffcm_correct_velocities_kernel, libffcm), Cint,
                        (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64,
                         Float64),
                        _c(V), V.id, _i3(num_cells), _d3(inv_cell_size), _d3(L), σ, Σ, a, μ, R_c),
           "_correct_velocities_kernel!")
    return V
end

# The staging seam (src/mobility.jl:55-56): host arrays are copied into the device staging
# buffers; the generic mobility! body then runs the barriers above.
function _stage_mobility_io!(particles::ParticleBuffers{<:HIPBuffer, <:HIPBuffer}, Y, F, V)
    copyto!(particles.Y_input, Y)
    copyto!(particles.F_input, F)
    particles.Y_input.ctx.cached_Y = nothing
    return (particles.Y_input, particles.F_input, particles.V_output)
end
function _retrieve_mobility_output!(V, ::ParticleBuffers{<:HIPBuffer, <:HIPBuffer}, V_staged)
    copyto!(V, Array(V_staged))
    return V
end

# ---------------------------------------------------------------------------------------------
# Pullback of mobility! (no reference counterpart): for L = <G, M(Y) F>, dF = M(Y) G and
# dY = dL/dY (gradient-of-Gaussian weights plus the derivative of the pair correction).
function mobility_vjp!(dF::Matrix{T}, dY::Matrix{T}, config::HIPConfig{T}, Y::Matrix{T}, F::Matrix{T},
                       G::Matrix{T}) where {T}
    N = size(config.particles.Y_sorted, 2)
    all(size(x) == (3, N) for x in (dF, dY, Y, F, G)) || throw(DimensionMismatch(
        "dF, dY, Y, F and G must each be 3xN for the config's N = $(N)"))
    _ensure_scalars!(config)
    _context(config).cached_Y = nothing
    _check(_ctx(config), ccall((:ffcm_mobility_vjp, libffcm), Cint,
                               (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Int64),
                               _ctx(config), dF, dY, Y, F, G, N), "mobility_vjp!")
    return dF, dY
end

apply_mobility(config::HIPConfig{T}, Y::Matrix{T}, F::Matrix{T}) where {T} = mobility!(similar(F), config, Y, F)

# The ChainRulesCore rrule of the out-of-place apply V = M(Y) F: the pullback recomputes the
# cell list and weights from Y (nothing from the forward call is kept but Y and F).
function ChainRulesCore.rrule(::typeof(apply_mobility), config::HIPConfig{T}, Y::Matrix{T},
                              F::Matrix{T}) where {T}
    V = mobility!(similar(F), config, Y, F)
    function apply_mobility_pullback(V̄)
        G = Matrix{T}(ChainRulesCore.unthunk(V̄))
        dF, dY = mobility_vjp!(similar(F), similar(Y), config, Y, F, G)
        return ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent(), dY, dF
    end
    return V, apply_mobility_pullback
end

end # module FFCMHIP
