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
# become Int32 and 0-based (cell lists, position maps with 0 -> -1), the library's
# convention.

using MagnetoPhotonic
import MagnetoPhotonic: HIPBackend, FieldState, Grid3D, FDTDParams, DLPole, ADEState,
                        MagnetoOpticADEState, ThermalState, MagnetizationState,
                        AbsorptionState, MagnetoOpticModel, FerrimagnetParameters
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

function Base.fill!(a::HIPArray{T}, v) where {T<:Union{Float32,Float64}}
    _check(ccall((:mp_fill, libmphip), Cint, (Ptr{Cvoid}, Int32, Int64, Cdouble),
                 a.ptr, _dtype(T), length(a), Float64(v)), "mp_fill")
    return a
end

function _upload(x::Array{T,N}) where {T,N}
    d = HIPArray{T,N}(size(x))
    isempty(x) || copyto!(d, x)
    return d
end

# Index lists: Int32, 0-based (0 = "absent" becomes -1).
_upload_index(x::AbstractArray{<:Integer}) = _upload(Int32.(x) .- Int32(1))

# ---- backend hooks (src/core/Backend.jl) --------------------------------------------------------
MagnetoPhotonic.zeros_backend(::HIPBackend, ::Type{T}, dims::Integer...) where {T} =
    fill!(HIPArray{T,length(dims)}(Int.(dims)), zero(T))
MagnetoPhotonic.array_type(::HIPBackend) = HIPArray
MagnetoPhotonic.adapt_backend(::HIPBackend, x::Array{<:AbstractFloat}) = _upload(x)
MagnetoPhotonic.adapt_backend(::HIPBackend, x::Array{<:DLPole}) = _upload(x)
MagnetoPhotonic.adapt_backend(::HIPBackend, x::AbstractArray{<:Integer}) = _upload_index(collect(x))
MagnetoPhotonic.adapt_backend(::HIPBackend, x) = Adapt.adapt(HIPArray, x)
MagnetoPhotonic.to_host(x::HIPArray) = Array(x)
MagnetoPhotonic.synchronize(::HIPBackend) = _check(ccall((:mp_synchronize, libmphip), Cint, ()), "mp_synchronize")
MagnetoPhotonic.free_device!(x::HIPArray) = _free!(x)
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

_p(x::HIPArray) = Ptr{Cvoid}(x.ptr)
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

function __init__()
    ok = Libc.Libdl.dlopen(libmphip; throw_error=false) !== nothing &&
         ccall((:mp_init, libmphip), Cint, (Cint,), 0) == 0
    MagnetoPhotonic._GPU_FUNCTIONAL[] = ok
    return nothing
end

end # module
