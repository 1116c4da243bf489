# MI355X (gfx950) backend of ForceCouplingMethod.jl's `gpu_acceleration` path, as a package
# extension that binds libffcm_hip.so through `ccall`.
#
# It plugs into the reference's seams the same way ext/ForceCouplingMethodCUDAExt does:
#   _assemble_gpu_buffers            (src/config.jl:461-464, called at :495-502)
#   the L1 step barriers             (imports at ext/ForceCouplingMethodCUDAExt/ForceCouplingMethodCUDAExt.jl:18-33)
#   _stage_mobility_io! / _retrieve_mobility_output!   (src/mobility.jl:55-56)
# and overrides `mobility!` for HIP-backed configs with the single `ffcm_mobility` call, which
# runs every step on one HIP stream and copies V back.
#
# The buffers of a HIP-backed config are `HIPBuffer`s: handles on the context's device
# buffers with `size`, `Array(x)` (device -> host) and `copyto!(x, host)` (host -> device), the
# operations the reference's GPU tests apply to config fields (test/cuda/*.jl).
module ForceCouplingMethodHIPExt

using ForceCouplingMethod
using StructArrays: StructArray
import ChainRulesCore

import ForceCouplingMethod:
    _apply_inverse_stokes_kernel!, _assemble_gpu_buffers, _assign_cells_kernel!,
    _build_cell_list_kernel!, _correct_velocities_kernel!, _gather_particles_kernel!,
    _interpolate_velocities_kernel!, _retrieve_mobility_output!, _spread_forces_kernel!,
    _stage_mobility_io!, mobility!, wrap_positions!

const libffcm = get(ENV, "FFCM_HIP_LIB", joinpath(@__DIR__, "..", "libffcm_hip.so"))

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
             fluid_velocity = 11, fluid_hat = 12, neighbor_map = 13)

mutable struct HIPContext
    ptr::Ptr{Cvoid}
end

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

# Grid fields are StructArrays of three component handles in the reference; the backend keeps
# them as one interleaved device buffer, so each component reads back through the whole field.
struct HIPComponent{T} <: AbstractArray{T, 3}
    field::HIPBuffer{T, 4}
    c::Int
end
Base.size(x::HIPComponent) = x.field.dims[2:4]
Base.Array(x::HIPComponent) = Array(x.field)[x.c, :, :, :]
Base.getindex(x::HIPComponent, I::Vararg{Int, 3}) = Array(x)[I...]

function _assemble_gpu_buffers(::Type{T}, N::Integer, num_grid_points::NTuple{3, Int32}, M_G::Int32,
                               num_cells_total::Integer, k_x::Vector{T}, k_y::Vector{T}, k_z::Vector{T},
                               neighbor_map_host::Vector{Int32}, scalars) where {T}
    p = FfcmParams(T === Float32 ? 0 : 1, 0, N, num_grid_points, M_G, scalars.num_cells, 0,
                   Float64.(scalars.L), scalars.R_c, scalars.a, scalars.σ, scalars.Σ, scalars.h,
                   scalars.inv_h, scalars.μ, Float64.(scalars.cell_size), Float64.(scalars.inv_cell_size),
                   pointer(k_x), pointer(k_y), pointer(k_z))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = GC.@preserve k_x k_y k_z ccall((:ffcm_create, libffcm), Cint, (Ref{Ptr{Cvoid}}, Ref{FfcmParams}), out, p)
    _check(C_NULL, rc, "FFCMConfig")
    ctx = HIPContext(out[])
    finalizer(c -> ccall((:ffcm_destroy, libffcm), Cint, (Ptr{Cvoid},), c.ptr), ctx)
    Mx, My, Mz = Int.(num_grid_points)
    C = Int(num_cells_total)
    vec(id, n) = HIPBuffer{Int32, 1}(ctx, id, (n,))
    mat(id) = HIPBuffer{T, 2}(ctx, id, (3, N))
    cells = ForceCouplingMethod.CellBuffers(vec(BUF.cell_hash, N), vec(BUF.original_index, N),
                                            vec(BUF.cell_start, C), vec(BUF.cell_end, C), vec(BUF.cell_end, C),
                                            vec(BUF.neighbor_map, 13C))
    particles = ForceCouplingMethod.ParticleBuffers(mat(BUF.Y_wrapped), mat(BUF.Y_sorted), mat(BUF.F_sorted),
                                                    mat(BUF.Y_input), mat(BUF.F_input), mat(BUF.V_output))
    field(id, dims) = HIPBuffer{T, 4}(ctx, id, (3, dims...))
    fd = field(BUF.force_density, (Mx, My, Mz))
    fv = field(BUF.fluid_velocity, (Mx, My, Mz))
    grid = ForceCouplingMethod.GridBuffers(
        StructArray{NTuple{3, T}}((HIPComponent(fd, 1), HIPComponent(fd, 2), HIPComponent(fd, 3))),
        StructArray{NTuple{3, T}}((HIPComponent(fv, 1), HIPComponent(fv, 2), HIPComponent(fv, 3))))
    fh = HIPBuffer{Complex{T}, 4}(ctx, BUF.fluid_hat, (3, Mx ÷ 2 + 1, My, Mz))
    solver = ForceCouplingMethod.SolverState(fh, nothing, nothing, k_x, k_y, k_z)
    stencil = ForceCouplingMethod.StencilBuffers(nothing, nothing, nothing)
    return (cells, particles, grid, solver, stencil)
end

const HIPConfig{T} = FFCMConfig{T, <:ForceCouplingMethod.CellBuffers{<:HIPBuffer}}
_ctx(config::HIPConfig) = config.cells.cell_hash.ctx.ptr

# mobility! (src/mobility.jl:23-45): one call runs the staged pipeline on the context stream.
function mobility!(V::Matrix{T}, config::HIPConfig{T}, Y::Matrix{T}, F::Matrix{T}) where {T}
    N = size(config.particles.Y_sorted, 2)
    (size(V) == (3, N) && size(Y) == (3, N) && size(F) == (3, N)) || throw(DimensionMismatch(
        "V, Y, and F must each be 3xN for the config's N = $(N)"))
    _check(_ctx(config), ccall((:ffcm_mobility, libffcm), Cint,
                               (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Int64), _ctx(config), V, Y, F, N), "mobility!")
    return V
end

# The L1 barriers for HIP buffers: each runs the matching per-step entry point on the
# context's own buffers (the arguments name those same buffers).
for (barrier, entry) in ((:_assign_cells_kernel!, :ffcm_assign_cells),
                         (:_build_cell_list_kernel!, :ffcm_sort_particles_by_cell),
                         (:_spread_forces_kernel!, :ffcm_spread_forces),
                         (:_interpolate_velocities_kernel!, :ffcm_interpolate_velocities),
                         (:_correct_velocities_kernel!, :ffcm_correct_velocities))
    @eval function $barrier(first::HIPBuffer, args...)
        _check(first.ctx.ptr, ccall(($(QuoteNode(entry)), libffcm), Cint, (Ptr{Cvoid},), first.ctx.ptr),
               $(string(barrier)))
        return first
    end
end
function wrap_positions!(dst::HIPBuffer, src::HIPBuffer, L)
    _check(dst.ctx.ptr, ccall((:ffcm_wrap_positions, libffcm), Cint, (Ptr{Cvoid},), dst.ctx.ptr), "wrap_positions!")
    return dst
end
# the gather is fused into ffcm_sort_particles_by_cell
_gather_particles_kernel!(Ys::HIPBuffer, Fs::HIPBuffer, Y, F, original_index) = Ys
# the FFT passes and the projection run inside ffcm_stokes_solve (stokes_solve! calls the
# forward transform, the projection and the inverse transform; the first runs the whole solve)
function _apply_inverse_stokes_kernel!(fx̂::HIPBuffer, args...)
    return fx̂
end

# Pullback of mobility! (no reference counterpart): for L = <G, M(Y) F>, dF = M(Y) G and
# dY = dL/dY (gradient-of-Gaussian weights plus the derivative of the pair correction).
function mobility_vjp!(dF::Matrix{T}, dY::Matrix{T}, config::HIPConfig{T}, Y::Matrix{T}, F::Matrix{T},
                       G::Matrix{T}) where {T}
    N = size(config.particles.Y_sorted, 2)
    all(size(x) == (3, N) for x in (dF, dY, Y, F, G)) || throw(DimensionMismatch(
        "dF, dY, Y, F and G must each be 3xN for the config's N = $(N)"))
    _check(_ctx(config), ccall((:ffcm_mobility_vjp, libffcm), Cint,
                               (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Ptr{T}, Int64),
                               _ctx(config), dF, dY, Y, F, G, N), "mobility_vjp!")
    return dF, dY
end

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
apply_mobility(config::HIPConfig{T}, Y::Matrix{T}, F::Matrix{T}) where {T} = mobility!(similar(F), config, Y, F)

_stage_mobility_io!(particles::ForceCouplingMethod.ParticleBuffers{<:HIPBuffer, <:HIPBuffer}, Y, F, V) =
    (copyto!(particles.Y_input, Y); copyto!(particles.F_input, F);
     (particles.Y_input, particles.F_input, particles.V_output))
_retrieve_mobility_output!(V, ::ForceCouplingMethod.ParticleBuffers{<:HIPBuffer, <:HIPBuffer}, V_staged) =
    copyto!(V, Array(V_staged))

end # module ForceCouplingMethodHIPExt
