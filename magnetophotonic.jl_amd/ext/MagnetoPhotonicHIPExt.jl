module MagnetoPhotonicHIPExt

# HIP backend for MagnetoPhotonic.jl on AMD Instinct MI355X (gfx950).
#
# Like the CUDA extension (ext/MagnetoPhotonicCUDAExt.jl) every method here is MORE
# SPECIFIC than the core's AbstractBackend fallbacks: it adds `::HIPBackend` / `::HIPArray`
# methods for the backend hooks of src/core/Backend.jl and for the `_ka_*` launchers of
# src/fdtd/Kernels.jl, which it forwards to libmphip.so through `ccall`.  No
# KernelAbstractions kernel runs on this backend.  The C structs below mirror
# include/mphip.h field for field.
#
# Device arrays are `HIPArray`s: a pointer from mp_malloc plus dims, freed by a
# finalizer (or eagerly by free_device!).  Integer index vectors adapted to the device
# become `HIPIndexArray`s: Int32 and 0-based on the device (cell lists, position maps with
# 0 -> -1), the library's convention, and 1-based again when `to_host` downloads them.
#
# The step: `step!` reaches `_step_typed!`, whose `::HIPBackend` method runs the whole
# step sequence as one `mp_run_steps` call (fused film pass, side-stream film update);
# `run!(state, steps)` without a callback issues all the steps in one call.  The
# per-launcher methods below serve callers that drive the launchers themselves.

using MagnetoPhotonic
import MagnetoPhotonic: HIPBackend, FieldState, Grid3D, FDTDParams, DLPole, ADEState,
                        MagnetoOpticADEState, ThermalState, MagnetizationState,
                        AbsorptionState, MagnetoOpticModel, FerrimagnetParameters, FDTDState,
                        PlaneSource, ModeSource, GaussianPulse, CarrierGaussianPulse, ContinuousSource
import Adapt

const libmphip = get(ENV, "MPHIP_LIBRARY", joinpath(@__DIR__, "..", "mphip", "libmphip.so"))

# ---- errors --------------------------------------------------------------------------------
function _check(rc::Cint, what::AbstractString)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:mp_last_error, libmphip), Cstring, ()))
    rc == -1 ? throw(ArgumentError("$what: $msg")) : error("$what: $msg")
end

_dtype(::Type{Float32}) = Int32(0)
_dtype(::Type{Float64}) = Int32(1)

# ---- device arrays ---------------------------------------------------------------------------
mutable struct HIPArray{T,N} <: AbstractArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    function HIPArray{T,N}(dims::NTuple{N,Int}) where {T,N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:mp_malloc, libmphip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, prod(dims) * sizeof(T)),
               "mp_malloc")
        a = new{T,N}(Ptr{T}(p[]), dims)
        finalizer(_free!, a)
        return a
    end
end

function _free!(a::HIPArray)
    a.ptr == C_NULL && return nothing
    ccall((:mp_free, libmphip), Cint, (Ptr{Cvoid},), a.ptr)
    a.ptr = C_NULL
    return nothing
end

Base.size(a::HIPArray) = a.dims
Base.IndexStyle(::Type{<:HIPArray}) = IndexLinear()
Base.getindex(::HIPArray, ::Int) = error("scalar indexing of a HIPArray; copy it with to_host first")
Base.pointer(a::HIPArray) = a.ptr
Base.unsafe_convert(::Type{Ptr{Cvoid}}, a::HIPArray) = Ptr{Cvoid}(a.ptr)

function Base.copyto!(dst::HIPArray{T}, src::Array{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto! sizes differ"))
    _check(ccall((:mp_memcpy_h2d, libmphip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                 dst.ptr, src, sizeof(src)), "mp_memcpy_h2d")
    return dst
end

function Base.Array(a::HIPArray{T,N}) where {T,N}
    out = Array{T,N}(undef, a.dims)
    _check(ccall((:mp_memcpy_d2h, libmphip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                 out, a.ptr, sizeof(out)), "mp_memcpy_d2h")
    return out
end

function Base.copyto!(dst::Array{T}, src::HIPArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto! sizes differ"))
    isempty(src) || _check(ccall((:mp_memcpy_d2h, libmphip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                                 dst, src.ptr, sizeof(dst)), "mp_memcpy_d2h")
    return dst
end

Base.similar(::HIPArray, ::Type{S}, dims::Dims{N}) where {S,N} = HIPArray{S,N}(dims)

function Base.fill!(a::HIPArray{T}, v) where {T<:Union{Float32,Float64}}
    _check(ccall((:mp_fill, libmphip), Cint, (Ptr{Cvoid}, Int32, Int64, Cdouble),
                 a.ptr, _dtype(T), length(a), Float64(v)), "mp_fill")
    return a
end

function Base.copyto!(dst::HIPArray{T}, src::HIPArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto! sizes differ"))
    isempty(src) || _check(ccall((:mp_memcpy_d2d, libmphip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                                 dst.ptr, src.ptr, length(src) * sizeof(T)), "mp_memcpy_d2d")
    return dst
end

function _upload(x::Array{T,N}) where {T,N}
    d = HIPArray{T,N}(size(x))
    isempty(x) || copyto!(d, x)
    return d
end

# Index arrays: Int32 and 0-based on the device (0 = "absent" becomes -1); to_host gives
# the reference's 1-based Int values back.
struct HIPIndexArray{N} <: AbstractArray{Int,N}
    data::HIPArray{Int32,N}
end
Base.size(a::HIPIndexArray) = size(a.data)
Base.IndexStyle(::Type{<:HIPIndexArray}) = IndexLinear()
Base.getindex(::HIPIndexArray, ::Int) = error("scalar indexing of a HIPIndexArray; copy it with to_host first")
_upload_index(x::AbstractArray{<:Integer}) = HIPIndexArray(_upload(Int32.(x) .- Int32(1)))

# dest .= src (a HIPArray or a host Array of the same element type) and dest .= scalar:
# the broadcasts the reference applies to device arrays (apply_magnetization!,
# Solver.jl:474-483; fill!-style resets).  Anything else is refused.
struct HIPStyle <: Broadcast.AbstractArrayStyle{Any} end
HIPStyle(::Val) = HIPStyle()
Base.BroadcastStyle(::Type{<:HIPArray}) = HIPStyle()
Base.BroadcastStyle(::HIPStyle, ::Broadcast.DefaultArrayStyle) = HIPStyle()
function Base.copyto!(dest::HIPArray{T}, bc::Broadcast.Broadcasted{HIPStyle}) where {T}
    if bc.f === identity && length(bc.args) == 1
        src = bc.args[1]
        (src isa HIPArray{T} || src isa Array{T}) && return copyto!(dest, src)
        src isa Number && return fill!(dest, src)
    end
    if bc.f === abs && length(bc.args) == 1 && dest isa HIPArray{Float32,2}
        ax = _plane_view(bc.args[1])
        if ax !== nothing
            f, a, i = ax
            nx, ny, nz = size(f)
            _check(ccall((:mp_plane_abs, libmphip), Cint,
                         (Ptr{Cvoid}, Int32, Int64, Int64, Int64, Int32, Int64, Ptr{Cvoid}),
                         _p(f), _dtype(eltype(f)), nx, ny, nz, Int32(a), Int64(i) - 1, _p(dest)), "mp_plane_abs")
            return dest
        end
    end
    throw(ArgumentError("HIPArray broadcasting supports dest .= src (a HIPArray or Array of the same " *
                        "element type), dest .= scalar and dest .= abs.(view(field, plane)) into Float32"))
end

# (field, axis 0/1/2, 1-based index) of a whole-plane view of a 3-D HIPArray
# (view(f, i, :, :), view(f, :, j, :), view(f, :, :, k)): the planes whose |values| the HDF5
# frame writers stage with dev .= abs.(view) (ext/MagnetoPhotonicHDF5Ext.jl:619-624).
function _plane_view(v)
    v isa SubArray || return nothing
    f = parent(v)
    (f isa HIPArray && ndims(f) == 3) || return nothing
    I = parentindices(v)
    whole(i, n) = i isa Base.Slice || (i isa AbstractUnitRange && first(i) == 1 && length(i) == n)
    nx, ny, nz = size(f)
    I[1] isa Integer && whole(I[2], ny) && whole(I[3], nz) && return (f, 0, Int(I[1]))
    whole(I[1], nx) && I[2] isa Integer && whole(I[3], nz) && return (f, 1, Int(I[2]))
    whole(I[1], nx) && whole(I[2], ny) && I[3] isa Integer && return (f, 2, Int(I[3]))
    return nothing
end

# ---- backend hooks (src/core/Backend.jl) --------------------------------------------------------
MagnetoPhotonic.zeros_backend(::HIPBackend, ::Type{T}, dims::Integer...) where {T} =
    fill!(HIPArray{T,length(dims)}(Int.(dims)), zero(T))
MagnetoPhotonic.array_type(::HIPBackend) = HIPArray
MagnetoPhotonic.adapt_backend(::HIPBackend, x::Array{<:AbstractFloat}) = _upload(x)
MagnetoPhotonic.adapt_backend(::HIPBackend, x::Array{<:DLPole}) = _upload(x)
MagnetoPhotonic.adapt_backend(::HIPBackend, x::AbstractArray{<:Integer}) = _upload_index(collect(x))
MagnetoPhotonic.adapt_backend(::HIPBackend, x) = Adapt.adapt(HIPArray, x)
MagnetoPhotonic.to_host(x::HIPArray) = Array(x)
MagnetoPhotonic.to_host(x::HIPIndexArray) = Int.(Array(x.data)) .+ 1
MagnetoPhotonic.synchronize(::HIPBackend) = _check(ccall((:mp_synchronize, libmphip), Cint, ()), "mp_synchronize")
MagnetoPhotonic.free_device!(x::HIPArray) = _free!(x)
MagnetoPhotonic.free_device!(x::HIPIndexArray) = _free!(x.data)
MagnetoPhotonic.reclaim_device_memory!(::HIPBackend) = nothing

function MagnetoPhotonic.device_memory_info(::HIPBackend)
    f = Ref{Csize_t}(0); t = Ref{Csize_t}(0)
    _check(ccall((:mp_mem_info, libmphip), Cint, (Ref{Csize_t}, Ref{Csize_t}), f, t), "mp_mem_info")
    return (free_GiB=f[] / 2^30, total_GiB=t[] / 2^30)
end

# ---- C structs (include/mphip.h) ----------------------------------------------------------------
struct MpYee
    nx::Int64; ny::Int64; nz::Int64
    field_dtype::Int32; compute_dtype::Int32
    ex::Ptr{Cvoid}; ey::Ptr{Cvoid}; ez::Ptr{Cvoid}
    hx::Ptr{Cvoid}; hy::Ptr{Cvoid}; hz::Ptr{Cvoid}
    dx::Ptr{Cvoid}; dy::Ptr{Cvoid}; dz::Ptr{Cvoid}
    inv_eps_x::Ptr{Cvoid}; inv_eps_y::Ptr{Cvoid}; inv_eps_z::Ptr{Cvoid}
    inv_d_cell_x::Ptr{Cvoid}; inv_d_cell_y::Ptr{Cvoid}; inv_d_cell_z::Ptr{Cvoid}
    inv_d_dual_x::Ptr{Cvoid}; inv_d_dual_y::Ptr{Cvoid}; inv_d_dual_z::Ptr{Cvoid}
    eps_index::Ptr{UInt8}; eps_table::Ptr{Cfloat}; n_eps::Int32; kernel_form::Int32
end

struct MpCpml
    npml_x::Int32; npml_y::Int32; npml_z::Int32; psi_dtype::Int32
    inv_kappa::NTuple{3,Ptr{Cvoid}}; a::NTuple{3,Ptr{Cvoid}}; b::NTuple{3,Ptr{Cvoid}}
    psi_h::NTuple{12,Ptr{Cvoid}}; psi_d::NTuple{12,Ptr{Cvoid}}
    psi_scale::Cdouble
end

struct MpAde
    n::Int64; npoles::Int32; dtype::Int32; state_dtype::Int32; _pad::Int32
    state_scale::Cdouble
    P::Ptr{Cvoid}; J::Ptr{Cvoid}; E_old::Ptr{Cvoid}
    idx::Ptr{Cvoid}; fill::Ptr{Cvoid}; inv_eps::Ptr{Cvoid}; poles::Ptr{Cvoid}
end

struct MpMo
    n::Int64; npoles::Int32; dtype::Int32
    Pa_from_b::Ptr{Cvoid}; Ja_from_b::Ptr{Cvoid}; Pb_from_a::Ptr{Cvoid}; Jb_from_a::Ptr{Cvoid}
    E_old_a::Ptr{Cvoid}; E_old_b::Ptr{Cvoid}
    idx::Ptr{Cvoid}; fill::Ptr{Cvoid}; mpos::Ptr{Cvoid}; poles::Ptr{Cvoid}
end

struct MpFilm
    n::Int64; dtype::Int32; _pad::Int32
    cells::Ptr{Cvoid}; pos_x::Ptr{Cvoid}; pos_y::Ptr{Cvoid}; pos_z::Ptr{Cvoid}
    Te::Ptr{Cvoid}; Tl::Ptr{Cvoid}; Ts_TM::Ptr{Cvoid}; Ts_RE::Ptr{Cvoid}
    m_TM_x::Ptr{Cvoid}; m_TM_y::Ptr{Cvoid}; m_TM_z::Ptr{Cvoid}
    m_RE_x::Ptr{Cvoid}; m_RE_y::Ptr{Cvoid}; m_RE_z::Ptr{Cvoid}
    U_abs::Ptr{Cvoid}; window_sum::Ptr{Cvoid}; peak::Ptr{Cvoid}
end

struct MpLut
    tm::Ptr{Cvoid}; re::Ptr{Cvoid}
    T_min::Cdouble; inv_dT::Cdouble
    n::Int32; dtype::Int32
end

struct MpMlp
    params::Ptr{Cfloat}; packed::Ptr{Cvoid}; norm::Ptr{Cfloat}
end

struct MpPlane
    axis::Int32; _pad::Int32; index::Int64
    dft_ea::Ptr{Cvoid}; dft_eb::Ptr{Cvoid}; dft_ha::Ptr{Cvoid}; dft_hb::Ptr{Cvoid}
    energy::Ptr{Cvoid}
end

struct MpPulse
    kind::Int32; _pad::Int32
    amplitude::Cdouble; t0::Cdouble; tau::Cdouble; omega::Cdouble; phase::Cdouble
end

struct MpSource
    kind::Int32; component::Int32; axis::Int32; _pad::Int32
    i::Int64; j::Int64; k::Int64
    i0::Int64; i1::Int64
    w0::Cdouble; w1::Cdouble
    profile::Ptr{Cvoid}
    hcorr::Cdouble
    pulse::MpPulse
end

struct MpState
    yee::Ptr{MpYee}; cpml::Ptr{MpCpml}; ade::NTuple{3,Ptr{MpAde}}; mo::Ptr{MpMo}
    mo_ea::Ptr{Cvoid}; mo_eb::Ptr{Cvoid}; mo_iea::Ptr{Cvoid}; mo_ieb::Ptr{Cvoid}
    film::Ptr{MpFilm}; params::Ptr{FerrimagnetParameters{Float64}}; lut::Ptr{MpLut}
    source::MpSource
    planes::Ptr{MpPlane}; nplanes::Int32; nb::Int32
    omega_bins::Ptr{Cdouble}; phase::Ptr{Cdouble}
    dt::Cdouble; mu0::Cdouble; pabs_factor::Cdouble; Q_TM::Cdouble; Q_RE::Cdouble; t::Cdouble
    n::Int64; last_mp_n::Int64
    multiphysics_every::Int32; subcycles::Int32; brillouin_iters::Int32; multiphysics_enabled::Int32
    has_ade_work::Int32; record_dft::Int32
    surrogate::Ptr{MpMlp}; surrogate_tol::Cdouble; surrogate_fallbacks::Ptr{Int64}
    film_lists_coincide::Int32; film_serial::Int32
    yee_alt::NTuple{6,Ptr{Cvoid}}
end

_p(x::HIPArray) = Ptr{Cvoid}(x.ptr)
_p(x::HIPIndexArray) = Ptr{Cvoid}(x.data.ptr)
_p(::Nothing) = C_NULL

function _yee(f::FieldState{T}, ie_x, ie_y, ie_z, dcx, dcy, dcz, ddx, ddy, ddz, ::Type{CT}) where {T,CT}
    nx, ny, nz = size(f.Ex)
    return MpYee(nx, ny, nz, _dtype(T), _dtype(CT),
                 _p(f.Ex), _p(f.Ey), _p(f.Ez), _p(f.Hx), _p(f.Hy), _p(f.Hz), _p(f.Dx), _p(f.Dy), _p(f.Dz),
                 _p(ie_x), _p(ie_y), _p(ie_z), _p(dcx), _p(dcy), _p(dcz), _p(ddx), _p(ddy), _p(ddz),
                 Ptr{UInt8}(C_NULL), Ptr{Cfloat}(C_NULL), Int32(0), Int32(0))
end

_psi_dtype(c) = _dtype(eltype(c.psi_Hxy_lo))

function _cpml(c)
    names_h = (:psi_Hxy_lo, :psi_Hxy_hi, :psi_Hxz_lo, :psi_Hxz_hi, :psi_Hyz_lo, :psi_Hyz_hi,
               :psi_Hyx_lo, :psi_Hyx_hi, :psi_Hzx_lo, :psi_Hzx_hi, :psi_Hzy_lo, :psi_Hzy_hi)
    names_d = (:psi_Dxy_lo, :psi_Dxy_hi, :psi_Dxz_lo, :psi_Dxz_hi, :psi_Dyz_lo, :psi_Dyz_hi,
               :psi_Dyx_lo, :psi_Dyx_hi, :psi_Dzx_lo, :psi_Dzx_hi, :psi_Dzy_lo, :psi_Dzy_hi)
    ph = ntuple(i -> _p(getfield(c, names_h[i])), 12)
    pd = ntuple(i -> _p(getfield(c, names_d[i])), 12)
    return MpCpml(c.Npml_x, c.Npml_y, c.Npml_z, _psi_dtype(c),
                  (_p(c.x.inv_kappa), _p(c.y.inv_kappa), _p(c.z.inv_kappa)),
                  (_p(c.x.a), _p(c.y.a), _p(c.z.a)), (_p(c.x.b), _p(c.y.b), _p(c.z.b)), ph, pd, 1.0)
end

# ---- reductions, scalar and plane downloads (src/core/Backend.jl:49-83) --------------------------
const MP_RED_SUM, MP_RED_COUNT_NEGATIVE, MP_RED_NORM_SUM = Int32(0), Int32(1), Int32(2)

function _reduce(op::Int32, x::HIPArray{T}, y=nothing, z=nothing) where {T<:Union{Float32,Float64}}
    out = HIPArray{Float64,1}((1,))
    _check(ccall((:mp_reduce, libmphip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int64, Int32, Ptr{Cvoid}),
                 _p(x), _p(y), _p(z), _dtype(T), Int64(length(x)), op, _p(out)), "mp_reduce")
    r = Array(out)[1]
    _free!(out)
    return r
end

MagnetoPhotonic.reduce_sum(x::HIPArray) = isempty(x) ? 0.0 : _reduce(MP_RED_SUM, x)
MagnetoPhotonic.reduce_mean(x::HIPArray) = isempty(x) ? NaN : _reduce(MP_RED_SUM, x) / length(x)
MagnetoPhotonic.reduce_count_negative(x::HIPArray) = isempty(x) ? 0 : Int(_reduce(MP_RED_COUNT_NEGATIVE, x))
function MagnetoPhotonic.reduce_norm_mean(x::HIPArray{T}, y::HIPArray{T}, z::HIPArray{T}) where {T}
    isempty(x) && return NaN
    (length(y) == length(x) && length(z) == length(x)) || throw(DimensionMismatch("reduce_norm_mean lengths differ"))
    return _reduce(MP_RED_NORM_SUM, x, y, z) / length(x)
end

function MagnetoPhotonic.any_nonfinite(x::HIPArray{T}) where {T<:Union{Float32,Float64}}
    isempty(x) && return false
    flag = HIPArray{Int32,1}((1,))
    _check(ccall((:mp_memset, libmphip), Cint, (Ptr{Cvoid}, Cint, Csize_t), _p(flag), 0, sizeof(Int32)), "mp_memset")
    _check(ccall((:mp_any_nonfinite, libmphip), Cint, (Ptr{Cvoid}, Int32, Int64, Ptr{Cvoid}),
                 _p(x), _dtype(T), Int64(length(x)), _p(flag)), "mp_any_nonfinite")
    r = Array(flag)[1] != 0
    _free!(flag)
    return r
end

function MagnetoPhotonic.scalar_to_host(x::HIPArray{T}, inds::Integer...) where {T}
    li = LinearIndices(x)[Int.(inds)...]
    out = Ref{T}()
    _check(ccall((:mp_memcpy_d2h, libmphip), Cint, (Ptr{T}, Ptr{Cvoid}, Csize_t),
                 out, Ptr{Cvoid}(x.ptr + (li - 1) * sizeof(T)), sizeof(T)), "mp_memcpy_d2h")
    return out[]
end

# plane_to_host: the plane in the order of the reference's views (x -> field[i,:,:],
# y -> field[:,i,:], z -> field[:,:,i]); a 2-D field is a 3-D one with one z plane.
function MagnetoPhotonic.plane_to_host(field::HIPArray{T,N}, axis::Symbol, idx::Integer) where {T,N}
    N == 1 && (axis === :x || throw(ArgumentError("1-D field plane axis must be :x")))
    N == 1 && return [MagnetoPhotonic.scalar_to_host(field, idx)]
    N in (2, 3) || throw(ArgumentError("field planes require a 1-D, 2-D, or 3-D array"))
    nx, ny = size(field, 1), size(field, 2)
    nz = N == 3 ? size(field, 3) : 1
    a = axis === :x ? Int32(0) : axis === :y ? Int32(1) : (axis === :z && N == 3) ? Int32(2) :
        throw(ArgumentError(N == 3 ? "3-D field plane axis must be :x, :y, or :z" : "2-D field plane axis must be :x or :y"))
    dims = a == 0 ? (ny, nz) : a == 1 ? (nx, nz) : (nx, ny)
    out = HIPArray{T,2}(dims)
    _check(ccall((:mp_plane_gather, libmphip), Cint, (Ptr{Cvoid}, Int32, Int64, Int64, Int64, Int32, Int64, Ptr{Cvoid}),
                 _p(field), _dtype(T), nx, ny, nz, a, Int64(idx) - 1, _p(out)), "mp_plane_gather")
    r = Array(out)
    _free!(out)
    return N == 3 ? r : vec(r)
end

# sum(work_trace; dims=(1, 2)) of the probe trace buffer (ProbeReadout.jl:291): the sums
# over the leading dimensions, returned as a device array of size (1, 1, n3).
function Base.sum(a::HIPArray{Float64,3}; dims)
    Tuple(dims) == (1, 2) || throw(ArgumentError("sum of a HIPArray supports dims=(1, 2)"))
    out = HIPArray{Float64,3}((1, 1, size(a, 3)))
    _check(ccall((:mp_sum_blocks, libmphip), Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                 _p(a), Int64(size(a, 1) * size(a, 2)), Int64(size(a, 3)), _p(out)), "mp_sum_blocks")
    return out
end

# ---- launchers (src/fdtd/Kernels.jl:522-699) ---------------------------------------------------
function MagnetoPhotonic._ka_update_H_3d!(::HIPBackend, fields::FieldState, grid::Grid3D, p::FDTDParams,
                                          dt::Real, cpml, ::Type{CT};
                                          inv_d_cell_x=grid.x.inv_d_cell, inv_d_cell_y=grid.y.inv_d_cell,
                                          inv_d_cell_z=grid.z.inv_d_cell) where {CT}
    y = _yee(fields, nothing, nothing, nothing, inv_d_cell_x, inv_d_cell_y, inv_d_cell_z,
             nothing, nothing, nothing, CT)
    _check(ccall((:mp_update_H_3d, libmphip), Cint, (Ref{MpYee}, Ref{MpCpml}, Cdouble, Cdouble),
                 y, _cpml(cpml), Float64(dt), Float64(p.mu0)), "mp_update_H_3d")
    return fields
end

function MagnetoPhotonic._ka_update_E_3d!(::HIPBackend, fields::FieldState, grid::Grid3D, p::FDTDParams,
                                          dt::Real, inv_eps_x, inv_eps_y, inv_eps_z, cpml, ::Type{CT};
                                          inv_d_dual_x=grid.x.inv_d_dual, inv_d_dual_y=grid.y.inv_d_dual,
                                          inv_d_dual_z=grid.z.inv_d_dual) where {CT}
    y = _yee(fields, inv_eps_x, inv_eps_y, inv_eps_z, nothing, nothing, nothing,
             inv_d_dual_x, inv_d_dual_y, inv_d_dual_z, CT)
    _check(ccall((:mp_update_E_3d, libmphip), Cint, (Ref{MpYee}, Ref{MpCpml}, Cdouble),
                 y, _cpml(cpml), Float64(dt)), "mp_update_E_3d")
    return fields
end

function MagnetoPhotonic._ka_inject_soft_3d!(::HIPBackend, fields::FieldState, component::Symbol,
                                             index::NTuple{3,Int}, value::Real, inv_eps, ::Type{CT}) where {CT}
    ie = (inv_eps, inv_eps, inv_eps)
    y = _yee(fields, ie..., nothing, nothing, nothing, nothing, nothing, nothing, CT)
    _check(ccall((:mp_inject_soft_3d, libmphip), Cint, (Ref{MpYee}, Int32, Int64, Int64, Int64, Cdouble),
                 y, MagnetoPhotonic._component_code(component), index[1] - 1, index[2] - 1, index[3] - 1,
                 Float64(value)), "mp_inject_soft_3d")
    return fields
end

function MagnetoPhotonic._ka_inject_mode_x_3d!(::HIPBackend, fields::FieldState, src, grid::Grid3D,
                                               value::Real, inv_eps, ::Type{CT}) where {CT}
    src.axis === :x || throw(ArgumentError("HIP ModeSource injection currently supports axis=:x"))
    i = MagnetoPhotonic._source_index(grid.x, src.position)
    eta0 = sqrt(FDTDParams().mu0 / FDTDParams().eps0)
    y = _yee(fields, inv_eps, inv_eps, inv_eps, nothing, nothing, nothing, nothing, nothing, nothing, CT)
    _check(ccall((:mp_inject_mode_x_3d, libmphip), Cint, (Ref{MpYee}, Int32, Int64, Ptr{Cvoid}, Cdouble, Cdouble),
                 y, MagnetoPhotonic._component_code(src.component), i - 1, _p(src.profile), Float64(value),
                 src.neff / eta0), "mp_inject_mode_x_3d")
    return fields
end

function _ade(state::ADEState, active_idx, fill, inv_eps, poles, ::Type{CT}) where {CT}
    return MpAde(length(active_idx), length(poles), _dtype(CT), _dtype(CT), 0, 1.0,
                 _p(state.P), _p(state.J), _p(state.E_old), _p(active_idx), _p(fill), _p(inv_eps), _p(poles))
end

function MagnetoPhotonic._ka_patch_E_dispersive!(::HIPBackend, E_arr::HIPArray{T}, state::ADEState, active_idx,
                                                 fill, inv_eps, poles, dt::Real, ::Type{CT}) where {T,CT}
    a = Ref(_ade(state, active_idx, fill, inv_eps, poles, CT))
    _check(ccall((:mp_patch_E_dispersive, libmphip), Cint, (Ptr{Cvoid}, Ref{MpAde}, Int32, Cdouble),
                 _p(E_arr), a, _dtype(T), Float64(dt)), "mp_patch_E_dispersive")
    return E_arr
end

function MagnetoPhotonic._ka_patch_E_mo_gyration!(::HIPBackend, Ey::HIPArray{T}, Ez, state::MagnetoOpticADEState,
                                                  active_idx, fill, material_pos, inv_eps_y, inv_eps_z,
                                                  m_TM_x, m_RE_x, poles, Q_voigt_TM::Real, Q_voigt_RE::Real,
                                                  dt::Real, ::Type{CT}) where {T,CT}
    m = Ref(MpMo(length(active_idx), length(poles), _dtype(CT), _p(state.Py_from_z), _p(state.Jy_from_z),
                 _p(state.Pz_from_y), _p(state.Jz_from_y), _p(state.E_old_y), _p(state.E_old_z),
                 _p(active_idx), _p(fill), _p(material_pos), _p(poles)))
    _check(ccall((:mp_patch_E_mo_gyration, libmphip), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ref{MpMo}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ptr{Cvoid},
                  Cdouble, Cdouble, Cdouble),
                 _p(Ey), _p(Ez), m, _p(inv_eps_y), _p(inv_eps_z), _dtype(T), _p(m_TM_x), _p(m_RE_x),
                 Float64(Q_voigt_TM), Float64(Q_voigt_RE), Float64(dt)), "mp_patch_E_mo_gyration")
    return Ey, Ez
end

function _film(th::ThermalState, mag, absorption, region, ::Type{CT}) where {CT}
    return MpFilm(length(region.material_cells), _dtype(CT), 0, _p(region.material_cells),
                  _p(region.pos_x), _p(region.pos_y), _p(region.pos_z),
                  _p(th.Te), _p(th.Tl), _p(th.Ts_TM), _p(th.Ts_RE),
                  mag === nothing ? C_NULL : _p(mag.m_TM_x), mag === nothing ? C_NULL : _p(mag.m_TM_y),
                  mag === nothing ? C_NULL : _p(mag.m_TM_z), mag === nothing ? C_NULL : _p(mag.m_RE_x),
                  mag === nothing ? C_NULL : _p(mag.m_RE_y), mag === nothing ? C_NULL : _p(mag.m_RE_z),
                  absorption === nothing ? C_NULL : _p(absorption.U_abs),
                  absorption === nothing ? C_NULL : _p(absorption.window_sum),
                  absorption === nothing ? C_NULL : _p(absorption.peak))
end

function MagnetoPhotonic._ka_pabs_accumulate!(::HIPBackend, absorption::AbsorptionState, fields, region,
                                              model::MagnetoOpticModel, dt::Real;
                                              absorption_model::Symbol=:cycle_average,
                                              eps0::Real=FDTDParams().eps0, ade_x=nothing, ade_y=nothing,
                                              ade_z=nothing)
    gd = model.params
    CT = eltype(region.material_fill)
    pabs_factor = 0.5 * Float64(gd.omega_pump) * Float64(eps0) * Float64(gd.eps_imag_pump)
    hw = absorption_model === :ade_work && ade_x !== nothing && length(ade_x.J) > 0
    # only the cell list, position maps and the FP64 accumulators are read here
    f = Ref(MpFilm(length(region.material_cells), _dtype(CT), 0, _p(region.material_cells),
                   _p(region.pos_x), _p(region.pos_y), _p(region.pos_z), C_NULL, C_NULL, C_NULL, C_NULL,
                   C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL,
                   _p(absorption.U_abs), _p(absorption.window_sum), _p(absorption.peak)))
    y = Ref(_yee(fields, nothing, nothing, nothing, nothing, nothing, nothing, nothing, nothing, nothing, CT))
    np_ = hw ? size(ade_x.J, 2) : 0
    mk(s, idx) = s === nothing ? nothing : Ref(MpAde(length(idx), np_, _dtype(CT), _dtype(CT), 0, 1.0,
                                                      _p(s.P), _p(s.J), _p(s.E_old), _p(idx), C_NULL, C_NULL, C_NULL))
    ax, ay, az = hw ? (mk(ade_x, region.ade_idx_x), mk(ade_y, region.ade_idx_y), mk(ade_z, region.ade_idx_z)) :
                      (nothing, nothing, nothing)
    _check(ccall((:mp_pabs_accumulate, libmphip), Cint,
                 (Ref{MpFilm}, Ref{MpYee}, Ptr{MpAde}, Ptr{MpAde}, Ptr{MpAde}, Int32, Cdouble, Cdouble),
                 f, y, something(ax, C_NULL), something(ay, C_NULL), something(az, C_NULL), Int32(hw),
                 pabs_factor, Float64(dt)), "mp_pabs_accumulate")
    return absorption
end

function MagnetoPhotonic._ka_multiphysics_step!(::HIPBackend, thermal::ThermalState, mag::MagnetizationState,
                                                fields, region, model::MagnetoOpticModel, lut, dt::Real;
                                                subcycles::Integer=1, absorption=nothing,
                                                brillouin_iters::Integer=2, compute_T::Type=Float32,
                                                _ignored...)
    tm_lut, re_lut, T_min, inv_dT, lut_N = lut
    CT = eltype(thermal.Te)
    f = Ref(_film(thermal, mag, absorption, region, CT))
    gd = Ref(FerrimagnetParameters{Float64}(ntuple(i -> Float64(getfield(model.params, i)),
                                                   fieldcount(typeof(model.params)))...))
    l = Ref(MpLut(_p(tm_lut), _p(re_lut), Float64(T_min), Float64(inv_dT), Int32(lut_N), _dtype(CT)))
    sur = SURROGATE[]
    if sur === nothing
        _check(ccall((:mp_multiphysics_step, libmphip), Cint,
                     (Ref{MpFilm}, Ref{FerrimagnetParameters{Float64}}, Ref{MpLut}, Cdouble, Int32, Int32, Int32),
                     f, gd, l, Float64(dt), Int32(subcycles), Int32(brillouin_iters), Int32(absorption !== nothing)),
               "mp_multiphysics_step")
    else
        mlp, tol, fallbacks = sur
        _check(ccall((:mp_surrogate_step, libmphip), Cint,
                     (Ref{MpFilm}, Ref{FerrimagnetParameters{Float64}}, Ref{MpLut}, Ref{MpMlp}, Cdouble, Int32, Int32,
                      Int32, Cdouble, Ptr{Int64}),
                     f, gd, l, Ref(mlp), Float64(dt), Int32(subcycles), Int32(brillouin_iters),
                     Int32(absorption !== nothing), tol, fallbacks),
               "mp_surrogate_step")
    end
    return thermal
end

# Optional learned film update (no counterpart in the reference): after
# use_surrogate!(params, packed, norm; tol) every film update of this backend goes through
# mp_surrogate_step, with the plain integrator as its checked fallback.  params (Float32,
# MP_MLP_NPARAM), packed (MP_MLP_PACKED_BYTES, filled by mp_mlp_pack) and norm (Float32,
# MP_MLP_NNORM) are device arrays; fallbacks, if given, is a device Int64 counter.
const SURROGATE = Ref{Any}(nothing)
function use_surrogate!(params::HIPArray, packed::HIPArray, norm::HIPArray; tol::Real=1e-4,
                        fallbacks::Union{Nothing,HIPArray}=nothing)
    mlp = MpMlp(Ptr{Cfloat}(params.ptr), Ptr{Cvoid}(packed.ptr), Ptr{Cfloat}(norm.ptr))
    _check(ccall((:mp_mlp_pack, libmphip), Cint, (Ref{MpMlp},), Ref(mlp)), "mp_mlp_pack")
    SURROGATE[] = (mlp, Float64(tol), fallbacks === nothing ? Ptr{Int64}(C_NULL) : Ptr{Int64}(fallbacks.ptr))
    return nothing
end
clear_surrogate!() = (SURROGATE[] = nothing; nothing)

function MagnetoPhotonic._ka_probe_dft_accumulate!(::HIPBackend, fields,
                                                   dft_Ey_pre, dft_Ez_pre, dft_Hy_pre, dft_Hz_pre,
                                                   dft_Ey_post, dft_Ez_post, dft_Hy_post, dft_Hz_post,
                                                   energy_pre, energy_post, omega_bins, phase_factors,
                                                   ix_pre::Integer, ix_post::Integer, t::Real, dt::Real)
    T = eltype(fields.Ex)
    y = Ref(_yee(fields, nothing, nothing, nothing, nothing, nothing, nothing, nothing, nothing, nothing, T))
    planes = [MpPlane(0, 0, ix_pre - 1, _p(dft_Ey_pre), _p(dft_Ez_pre), _p(dft_Hy_pre), _p(dft_Hz_pre), _p(energy_pre)),
              MpPlane(0, 0, ix_post - 1, _p(dft_Ey_post), _p(dft_Ez_post), _p(dft_Hy_post), _p(dft_Hz_post), _p(energy_post))]
    _check(ccall((:mp_probe_dft_accumulate, libmphip), Cint,
                 (Ref{MpYee}, Ptr{MpPlane}, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Cdouble, Cdouble),
                 y, planes, Int32(2), _p(omega_bins), _p(phase_factors), Int32(length(omega_bins)),
                 Float64(t), Float64(dt)), "mp_probe_dft_accumulate")
    return nothing
end

# probe_shot's plane reductions on the device (ProbeReadout.jl:322-351): one call, 4nb+10
# doubles downloaded instead of the ten plane buffers.  Returns the values probe_shot
# derives its T/R/A, spectra and Jones angles from.
function probe_shot_reductions(m::MagnetoPhotonic.ProbeReadout)
    nb = length(m.omega_bins)
    pre = MpPlane(0, 0, 0, _p(m.dft_Ey_pre), _p(m.dft_Ez_pre), _p(m.dft_Hy_pre), _p(m.dft_Hz_pre), _p(m.energy_pre))
    post = MpPlane(0, 0, 0, _p(m.dft_Ey_post), _p(m.dft_Ez_post), _p(m.dft_Hy_post), _p(m.dft_Hz_post),
                   _p(m.energy_post))
    out = HIPArray{Float64,1}((4nb + 10,))
    _check(ccall((:mp_probe_shot_reduce, libmphip), Cint,
                 (Ref{MpPlane}, Ref{MpPlane}, Int64, Int64, Int32, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 Ref(pre), Ref(post), Int64(m.Ny), Int64(m.Nz), Int32(nb), Int32(m.center_bin - 1),
                 _p(m.mode_w_dev), _p(m.area_yz_dev), _p(out)), "mp_probe_shot_reduce")
    r = Array(out)
    _free!(out)
    Sx_pre = ComplexF64[complex(r[2q - 1], r[2q]) for q in 1:nb]
    Sx_post = ComplexF64[complex(r[2nb + 2q - 1], r[2nb + 2q]) for q in 1:nb]
    inv = 1.0 / m.mode_norm
    j(i) = complex(r[4nb + 2 + 2i - 1], r[4nb + 2 + 2i]) * inv
    return (; Sx_pre, Sx_post, net_pre=r[4nb + 1], net_post=r[4nb + 2],
            jEy_far=j(1), jEz_far=j(2), jEy_kerr=j(3), jEz_kerr=j(4))
end

# probe_trace_plane_kernel! (Kernels.jl:467-478): the four mode-weighted plane samples into
# work_trace (Ny, Nz, 4); ProbeReadout.jl:291 then sums them with sum(; dims=(1, 2)) above.
function MagnetoPhotonic._ka_probe_trace_plane!(::HIPBackend, work_trace::HIPArray{Float64,3}, fields,
                                                mode_w, area_yz, ix_pre::Integer, ix_post::Integer)
    (size(work_trace, 1) > 0 && size(work_trace, 2) > 0) || return nothing
    y = Ref(_yee(fields, nothing, nothing, nothing, nothing, nothing, nothing, nothing, nothing, nothing,
                 eltype(fields.Ex)))
    _check(ccall((:mp_probe_trace_samples, libmphip), Cint,
                 (Ref{MpYee}, Int64, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 y, Int64(ix_pre) - 1, Int64(ix_post) - 1, _p(mode_w), _p(area_yz), _p(work_trace)),
           "mp_probe_trace_samples")
    return nothing
end

# inject!(::FieldState, ::Grid3D, ::PlaneSource) (Source.jl:188-236) is a host loop over
# the plane's cells in the reference; on device fields it is one mp_inject_plane_3d call
# with the same _axis_weights (E += w v and D += w v / inv_eps formed in FP64).
function MagnetoPhotonic.inject!(fields::FieldState{T,<:HIPArray}, grid::Grid3D, src::PlaneSource, t::Real,
                                 p::FDTDParams, inv_eps) where {T}
    ax = src.axis === :x ? Int32(0) : src.axis === :y ? Int32(1) : src.axis === :z ? Int32(2) :
         throw(ArgumentError("3-D PlaneSource axis must be :x, :y or :z"))
    w = MagnetoPhotonic._axis_weights(getfield(grid, src.axis), src.position)
    i1, w1 = length(w) > 1 ? (Int64(w[2][1]) - 1, Float64(w[2][2])) : (Int64(-1), 0.0)
    y = Ref(_yee(fields, inv_eps.inv_eps_x, inv_eps.inv_eps_y, inv_eps.inv_eps_z, nothing, nothing, nothing,
                 nothing, nothing, nothing, T))
    _check(ccall((:mp_inject_plane_3d, libmphip), Cint,
                 (Ref{MpYee}, Int32, Int32, Int64, Cdouble, Int64, Cdouble, Cdouble),
                 y, MagnetoPhotonic._component_code(src.component), ax, Int64(w[1][1]) - 1, Float64(w[1][2]), i1, w1,
                 Float64(MagnetoPhotonic.source_value(src.pulse, t))), "mp_inject_plane_3d")
    return fields
end

# ---- the whole step as one mp_run_steps call (Solver.jl:277-383, 557-563) ----------------------
_pulse(p::GaussianPulse) = MpPulse(1, 0, p.amplitude, p.t0, p.tau, p.omega, p.phase)
_pulse(p::CarrierGaussianPulse) = MpPulse(2, 0, p.amplitude, p.t0, p.sigma, p.omega, 0.0)
_pulse(p::ContinuousSource) = MpPulse(3, 0, p.amplitude, 0.0, 1.0, p.omega, p.phase)
_pulse(::Any) = nothing

const _NO_PULSE = MpPulse(0, 0, 0.0, 0.0, 1.0, 0.0, 0.0)

# The source of _step_typed! (Solver.jl:296-311) as an mp_source; nothing when the step
# driver has no such source (the launcher path then runs the step).
function _mp_source(state::FDTDState, source)
    source === nothing && return MpSource(0, 0, 0, 0, 0, 0, 0, 0, -1, 0.0, 0.0, C_NULL, 0.0, _NO_PULSE)
    if source isa Tuple
        pulse, component, index = source
        pl = _pulse(pulse)
        pl === nothing && return nothing
        return MpSource(1, MagnetoPhotonic._component_code(component), 0, 0, index[1] - 1, index[2] - 1,
                        index[3] - 1, 0, -1, 0.0, 0.0, C_NULL, 0.0, pl)
    elseif source isa PlaneSource
        pl = _pulse(source.pulse)
        (pl === nothing || !(source.axis in (:x, :y, :z))) && return nothing
        ax = source.axis === :x ? 0 : source.axis === :y ? 1 : 2
        w = MagnetoPhotonic._axis_weights(getfield(state.grid, source.axis), source.position)
        i1, w1 = length(w) > 1 ? (w[2][1] - 1, Float64(w[2][2])) : (-1, 0.0)
        return MpSource(2, MagnetoPhotonic._component_code(source.component), ax, 0, 0, 0, 0, w[1][1] - 1, i1,
                        Float64(w[1][2]), w1, C_NULL, 0.0, pl)
    elseif source isa ModeSource
        pl = _pulse(source.pulse)
        (pl === nothing || source.axis !== :x || !(source.component in (:Ey, :Ez))) && return nothing
        i = MagnetoPhotonic._source_index(state.grid.x, source.position)
        eta0 = sqrt(FDTDParams().mu0 / FDTDParams().eps0)   # as _ka_inject_mode_x_3d! (Kernels.jl:584)
        return MpSource(3, MagnetoPhotonic._component_code(source.component), 0, 0, i - 1, 0, 0, 0, -1, 0.0, 0.0,
                        _p(source.profile), source.neff / eta0, pl)
    end
    return nothing
end

# Whether every ADE list and the MO list is the film list (entry n is film cell n) and the
# ADE state is in the compute type: the step then runs its list work as one fused pass.
# Decided once per state from host copies of the index maps.
const _COINCIDE = WeakKeyDict{FDTDState,Bool}()
function _lists_coincide(state::FDTDState)
    get!(_COINCIDE, state) do
        r = state.region
        (state.ade_x === nothing || state.thermal === nothing || r === nothing) && return false
        eltype(state.ade_x.P) == state.compute_T || return false
        cells = MagnetoPhotonic.to_host(r.material_cells)
        ident = collect(1:length(cells))
        for (idx, pos) in ((r.ade_idx_x, r.pos_x), (r.ade_idx_y, r.pos_y), (r.ade_idx_z, r.pos_z))
            (MagnetoPhotonic.to_host(idx) == cells && MagnetoPhotonic.to_host(pos) == ident) || return false
        end
        if state.mo !== nothing
            (MagnetoPhotonic.to_host(r.mo_idx) == cells && MagnetoPhotonic.to_host(state.mo_pos) == ident) ||
                return false
        end
        return true
    end
end

# One mp_run_steps call of `steps` steps on the state; false when the state holds a part
# the step driver does not take (then nothing ran).
function _run_steps!(state::FDTDState, steps::Integer)
    CT = state.compute_T
    src = _mp_source(state, state.source)
    src === nothing && return false
    r = state.region
    yee = Ref(_yee(state.fields, state.inv_eps_x, state.inv_eps_y, state.inv_eps_z, state.inv_d_cell_x,
                   state.inv_d_cell_y, state.inv_d_cell_z, state.inv_d_dual_x, state.inv_d_dual_y,
                   state.inv_d_dual_z, CT))
    cpml = Ref(_cpml(state.cpml))
    ade = state.ade_x === nothing ? nothing :
          (Ref(_ade(state.ade_x, r.ade_idx_x, r.ade_f_x, r.ade_inv_x, state.diag_poles, CT)),
           Ref(_ade(state.ade_y, r.ade_idx_y, r.ade_f_y, r.ade_inv_y, state.diag_poles, CT)),
           Ref(_ade(state.ade_z, r.ade_idx_z, r.ade_f_z, r.ade_inv_z, state.diag_poles, CT)))
    film_on = state.thermal !== nothing && state.mag !== nothing && state.absorption !== nothing
    mo_on = state.mo !== nothing && state.mag !== nothing
    mo = mo_on ? Ref(MpMo(length(r.mo_idx), length(state.mo_poles), _dtype(CT), _p(state.mo.Py_from_z),
                          _p(state.mo.Jy_from_z), _p(state.mo.Pz_from_y), _p(state.mo.Jz_from_y),
                          _p(state.mo.E_old_y), _p(state.mo.E_old_z), _p(r.mo_idx), _p(r.mo_fill),
                          _p(state.mo_pos), _p(state.mo_poles))) : nothing
    film = film_on ? Ref(_film(state.thermal, state.mag, state.absorption, r, CT)) : nothing
    gd = state.model isa MagnetoOpticModel ? state.model.params : nothing
    params = gd === nothing ? nothing :
             Ref(FerrimagnetParameters{Float64}(ntuple(i -> Float64(getfield(gd, i)), fieldcount(typeof(gd)))...))
    lut = nothing
    if film_on
        tm_lut, re_lut, T_min, inv_dT, lut_N = state.lut
        lut = Ref(MpLut(_p(tm_lut), _p(re_lut), Float64(T_min), Float64(inv_dT), Int32(lut_N), _dtype(CT)))
    end
    sur = SURROGATE[]
    mlp = sur === nothing ? nothing : Ref(sur[1])
    hw = state.absorption_model === :ade_work && state.ade_x !== nothing && length(state.ade_x.J) > 0
    pabs_factor = gd === nothing ? 0.0 :
                  0.5 * Float64(gd.omega_pump) * Float64(state.params.eps0) * Float64(gd.eps_imag_pump)
    ptr(x) = x === nothing ? C_NULL : Base.unsafe_convert(Ptr{eltype(x)}, x)
    GC.@preserve yee cpml ade mo film params lut mlp state begin
        st = Ref(MpState(ptr(yee), ptr(cpml),
                         ade === nothing ? (Ptr{MpAde}(C_NULL), Ptr{MpAde}(C_NULL), Ptr{MpAde}(C_NULL)) :
                                           (ptr(ade[1]), ptr(ade[2]), ptr(ade[3])),
                         mo_on ? ptr(mo) : Ptr{MpMo}(C_NULL),
                         mo_on ? _p(state.fields.Ey) : C_NULL, mo_on ? _p(state.fields.Ez) : C_NULL,
                         mo_on ? _p(state.inv_eps_y) : C_NULL, mo_on ? _p(state.inv_eps_z) : C_NULL,
                         film_on ? ptr(film) : Ptr{MpFilm}(C_NULL),
                         params === nothing ? Ptr{FerrimagnetParameters{Float64}}(C_NULL) : ptr(params),
                         film_on ? ptr(lut) : Ptr{MpLut}(C_NULL),
                         src, Ptr{MpPlane}(C_NULL), Int32(0), Int32(0), Ptr{Cdouble}(C_NULL), Ptr{Cdouble}(C_NULL),
                         state.dt, Float64(state.params.mu0), pabs_factor,
                         gd === nothing ? 0.0 : Float64(gd.Q_voigt_TM), gd === nothing ? 0.0 : Float64(gd.Q_voigt_RE),
                         state.t, Int64(state.n), Int64(state.last_mp_n),
                         Int32(state.multiphysics_every), Int32(state.subcycles), Int32(state.brillouin_iters),
                         Int32(state.multiphysics_enabled && film_on), Int32(hw), Int32(0),
                         mlp === nothing ? Ptr{MpMlp}(C_NULL) : ptr(mlp), sur === nothing ? 0.0 : sur[2],
                         sur === nothing ? Ptr{Int64}(C_NULL) : sur[3],
                         Int32(film_on && _lists_coincide(state)), Int32(0), ntuple(_ -> C_NULL, 6)))
        _check(ccall((:mp_run_steps, libmphip), Cint, (Ref{MpState}, Int64), st, Int64(steps)), "mp_run_steps")
        state.n = Int(st[].n)
        state.t = st[].t
        state.last_mp_n = Int(st[].last_mp_n)
    end
    return true
end

# step! on this backend (Solver.jl:277-369): the fixed sequence in one mp_run_steps call
# (H, source, E, the fused film pass or the list kernels, absorbed power, the film update
# when its window closes).  A state with a source the step driver does not take (a point
# source with trilinear weights, a non-x mode source) runs the reference's launcher
# sequence instead, every launch of which is a method of this extension.
function MagnetoPhotonic._step_typed!(state::FDTDState, backend::HIPBackend, ::Type{CT},
                                      inv_eps_x, inv_eps_y, inv_eps_z, inv_d_cell_x, inv_d_cell_y, inv_d_cell_z,
                                      inv_d_dual_x, inv_d_dual_y, inv_d_dual_z,
                                      source, region, diag_poles, mo_poles, mo_pos, model) where {CT}
    _run_steps!(state, 1) && return state
    return invoke(MagnetoPhotonic._step_typed!,
                  Tuple{FDTDState,Any,Type{CT},Any,Any,Any,Any,Any,Any,Any,Any,Any,Any,Any,Any,Any,Any,Any},
                  state, backend, CT, inv_eps_x, inv_eps_y, inv_eps_z, inv_d_cell_x, inv_d_cell_y, inv_d_cell_z,
                  inv_d_dual_x, inv_d_dual_y, inv_d_dual_z, source, region, diag_poles, mo_poles, mo_pos, model)
end

# run!(state, steps) (Solver.jl:557-563) without a callback: every step in one
# mp_run_steps call, so each step's film update overlaps the next step's Yee updates.
# Other backends, or a callback, take the reference's loop.
function MagnetoPhotonic.run!(state::FDTDState, steps::Int; callback=nothing)
    if callback === nothing && state.backend isa HIPBackend && steps > 0 && _run_steps!(state, steps)
        return state
    end
    return invoke(MagnetoPhotonic.run!, Tuple{FDTDState,Integer}, state, steps; callback=callback)
end

function __init__()
    ok = Libc.Libdl.dlopen(libmphip; throw_error=false) !== nothing &&
         ccall((:mp_init, libmphip), Cint, (Cint,), 0) == 0
    MagnetoPhotonic._GPU_FUNCTIONAL[] = ok
    return nothing
end

end # module
