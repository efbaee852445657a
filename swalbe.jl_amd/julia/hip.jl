# hip.jl -- binds Swalbe.jl's 2D thin-film step to libswalbe_hip (MI355X).
#
# Included inside `module Swalbe` (after the other includes of src/Swalbe.jl),
# together with one branch in Sys (src/initialize.jl:491):
#
#     device == "HIP" && return HipState{T}(sysc.Lx, sysc.Ly; kind = kind, g = sysc.param.g)
#
# Then
#     sys   = Swalbe.SysConst(Lx = 512, Ly = 512, param = Swalbe.Taumucs(Tmax = 10_000))
#     state = Swalbe.Sys(sys, "HIP"; T = Float64)        # a HipState on the GPU
#     Swalbe.equilibrium!(state, sys)
#     Swalbe.time_loop(sys, state)                       # one fused kernel per step
#     h = state.height                                   # a live view (HipField)
#
# Every function below is a more specific method of a Swalbe function for
# HipState; the call sites in src/simulate.jl (time_loop, run_*) and user
# scripts stay as they are.  `state.height` is a HipField, an AbstractArray
# view of the device array: reads see the device's values and in-place writes
# (`state.height .= 1.0`, `state.height[i, j] = v`, randinterface! and
# singledroplet writing into it, `state.ftemp .= state.fout`) reach the device
# before the next call on the state.

# path of libswalbe_hip.so (swalbe.jl_amd/lib/ after its build)
const libswalbe = get(ENV, "SWALBE_HIP_LIB", "libswalbe_hip")

const HIP_OK, HIP_EDOMAIN = Cint(0), Cint(1)
const HIP_F64, HIP_F32 = Int32(0), Int32(1)
const HIP_SIMPLE, HIP_THERMAL = Int32(0), Int32(1)
const HIP_FIELDS = Dict(:fout => 0, :ftemp => 1, :feq => 2, :height => 3, :velx => 4, :vely => 5,
                    :vsq => 6, :pressure => 7, :Fx => 8, :Fy => 9, :slipx => 10, :slipy => 11,
                    :h∇px => 12, :h∇py => 13, :kbtx => 14, :kbty => 15, :θ => 16)
const HIP_DISTS = (:fout, :ftemp, :feq)

# swalbe_params = Taumucs{T} (src/initialize.jl:43-60)
struct HipParams
    Tmax::Int64; tdump::Int64
    tau::Float64; cs::Float64; mu::Float64; delta::Float64; kbt::Float64; gamma::Float64
    n::Int32; m::Int32
    hmin::Float64; hcrit::Float64; theta::Float64; g::Float64
end
HipParams(p) = HipParams(p.Tmax, p.tdump, p.τ, p.cₛ, p.μ, p.δ, p.kbt, p.γ, p.n, p.m, p.hmin,
                     p.hcrit, p.θ, p.g)

# swalbe_step_opts
struct HipStepOpts
    use_theta_field::Int32; thermal::Int32; body::Int32; materialize::Int32
    seed::UInt64; t0::Int64
    body_ax::Float64; body_ay::Float64; body_fac::Float64; wet_thresh::Float64
    dh_series::Ptr{Float64}; wet_series::Ptr{Int64}; mass_series::Ptr{Float64}
    tile_y::Int32; variant::Int32; nontemporal::Int32; no_fp64_guard::Int32
    snap_every::Int64; snap_series::Ptr{Float64}
    maxh_series::Ptr{Float64}; area_series::Ptr{Float64}; energy_series::Ptr{Float64}
    fluid_series::Ptr{UInt8}; surf_thresh::Float64; surf_theta_field::Int32
    final_outputs_only::Int32
end
HipStepOpts(; use_theta_field = 0, thermal = 0, body = 0, materialize = 1, seed = 0, t0 = 0,
         body_ax = 0.0, body_ay = 0.0, body_fac = 0.0, wet_thresh = 0.055,
         dh_series = C_NULL, wet_series = C_NULL, mass_series = C_NULL,
         maxh_series = C_NULL, area_series = C_NULL, energy_series = C_NULL,
         fluid_series = C_NULL, surf_thresh = 0.055, surf_theta_field = 0,
         final_outputs_only = 0, nontemporal = -1) =
    HipStepOpts(use_theta_field, thermal, body, materialize, seed, t0, body_ax, body_ay, body_fac,
             wet_thresh, dh_series, wet_series, mass_series, 0, 0, nontemporal, 0, 0, C_NULL,
             maxh_series, area_series, energy_series, fluid_series, surf_thresh,
             surf_theta_field, final_outputs_only)

# The store mode (nontemporal; every mode gives the same bits) measured fastest
# for a film on MI355X, as swalbe.store_mode of the Python mirror: FP64 films 2;
# FP32 films with more than 32 MB of inputs 1 when every step writes its
# outputs, 0 when only the last step does; smaller FP32 films the other way.
function store_mode(Lx, Ly, ::Type{T}, final_outputs_only) where {T}
    T === Float64 && return Int32(2)
    last = final_outputs_only != 0
    return Int32(Lx * Ly * 12 > (32 << 20) ? (last ? 0 : 1) : (last ? 1 : 0))
end
with_store_mode(o::HipStepOpts, mode) =
    HipStepOpts((f === :nontemporal ? Int32(mode) : getfield(o, f) for f in fieldnames(HipStepOpts))...)

function hip_check(rc::Cint, what)
    rc == HIP_OK && return nothing
    msg = unsafe_string(ccall((:swalbe_last_error, libswalbe), Cstring, ()))
    rc == HIP_EDOMAIN && throw(DomainError(what, msg))
    error("$what: $msg (status $rc)")
end

# ---- state fields that read and write through ---------------------------------
# The reference's drivers and scripts write into the state's arrays in place
# (src/simulate.jl:241, 289, 350-353, 394; scripts/Rivulet_stability.jl:94-95).
# A state keeps a Float64 host mirror of each field read or written element by
# element: downloaded on first use, edited in place, marked dirty.  Every call
# into the library goes through `state.handle`, which uploads the dirty
# mirrors and marks all mirrors stale (the call may change any field), so the
# next read downloads again.  Whole-field writes (fill!, copyto!, `.=`) go to
# the device directly.  FP32 states decode to and re-encode from Float64
# exactly, so untouched nodes of an uploaded mirror keep their bits.
mutable struct HipMirror
    A::Array{Float64}
    version::Int
    dirty::Bool
end

struct HipField{T,N,S} <: AbstractArray{T,N}
    s::S
    name::Symbol
    dims::NTuple{N,Int}
end
Base.size(f::HipField) = f.dims
Base.IndexStyle(::Type{<:HipField}) = IndexLinear()

function hip_mirror(f::HipField)
    s = f.s
    ms = getfield(s, :mirrors)
    v = getfield(s, :version)
    m = get(ms, f.name, nothing)
    if m === nothing || (m.version != v && !m.dirty)
        A = Array{Float64}(undef, f.dims)
        hip_download!(s, f.name, A)
        m = HipMirror(A, v, false)
        ms[f.name] = m
    end
    return m
end

Base.getindex(f::HipField{T}, i::Int) where {T} = T(hip_mirror(f).A[i])
function Base.setindex!(f::HipField, x, i::Int)
    m = hip_mirror(f)
    m.A[i] = x
    m.dirty = true
    return f
end

# upload the fields written on the host, mark every mirror stale and return
# the library handle: the path of every call that may read or change the state
function hip_flush!(s)
    h = getfield(s, :handle)
    for (name, m) in getfield(s, :mirrors)
        if m.dirty
            hip_upload!(s, name, m.A)
            m.dirty = false
        end
    end
    setfield!(s, :version, getfield(s, :version) + 1)
    return h
end

# whole-field writes go straight to the device
function hip_set!(f::HipField, A::AbstractArray)
    length(A) == prod(f.dims) || throw(DimensionMismatch("$(size(A)) into $(f.dims)"))
    B = Array{Float64}(undef, f.dims)
    copyto!(B, A)  # read first: A may be a field of the same state
    hip_flush!(f.s)
    delete!(getfield(f.s, :mirrors), f.name)
    hip_upload!(f.s, f.name, B)
    return f
end
function Base.fill!(f::HipField, x::Number)
    hip_flush!(f.s)
    delete!(getfield(f.s, :mirrors), f.name)
    hip_fill!(f.s, f.name, Float64(x))
    return f
end
Base.copyto!(dest::HipField, src::AbstractArray) = hip_set!(dest, src)
# `state.X .= expr`: a scalar fills on the device, anything else is evaluated
# on the host (reading fields through their mirrors) and uploaded whole
function Base.copyto!(dest::HipField, bc::Base.Broadcast.Broadcasted{Nothing})
    axes(dest) == axes(bc) || throw(DimensionMismatch("broadcast into state.$(dest.name)"))
    if bc.f === identity && length(bc.args) == 1 && bc.args[1] isa Number
        return fill!(dest, bc.args[1])
    end
    B = Array{Float64}(undef, dest.dims)
    copyto!(B, bc)
    return hip_set!(dest, B)
end
Base.Array(f::HipField{T}) where {T} = T.(hip_mirror(f).A)
Base.copy(f::HipField) = Array(f)
# whole-array reductions on one host copy instead of one mirror lookup per element
Base.sum(f::HipField; kw...) = sum(Array(f); kw...)
Base.maximum(f::HipField; kw...) = maximum(Array(f); kw...)
Base.minimum(f::HipField; kw...) = minimum(Array(f); kw...)

"""GPU-resident State{T,3} / State_thermal{T,3}: the library owns the arrays;
`state.<field>` is a HipField."""
mutable struct HipState{T} <: LBM_state_2D
    handle::Ptr{Cvoid}
    Lx::Int
    Ly::Int
    t::Int
    thermal::Bool
    version::Int
    mirrors::Dict{Symbol,HipMirror}
end

hip_download!(s::HipState, name::Symbol, A::Array{Float64}) =
    hip_check(ccall((:swalbe_download, libswalbe), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t),
                getfield(s, :handle), HIP_FIELDS[name], A, sizeof(A)), "download $name")
hip_upload!(s::HipState, name::Symbol, A::Array{Float64}) =
    hip_check(ccall((:swalbe_upload, libswalbe), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t),
                getfield(s, :handle), HIP_FIELDS[name], A, sizeof(A)), "upload $name")
hip_fill!(s::HipState, name::Symbol, x::Float64) =
    hip_check(ccall((:swalbe_fill, libswalbe), Cint, (Ptr{Cvoid}, Int32, Float64),
                getfield(s, :handle), HIP_FIELDS[name], x), "fill $name")

function HipState{T}(Lx, Ly; kind = "simple", href = 1.0, g = 0.0) where {T}
    out = Ref{Ptr{Cvoid}}(C_NULL)
    hip_check(ccall((:swalbe_create, libswalbe), Cint,
                (Int32, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
                Lx, Ly, T === Float32 ? HIP_F32 : HIP_F64, kind == "thermal" ? HIP_THERMAL : HIP_SIMPLE, out),
          "swalbe_create")
    s = HipState{T}(out[], Lx, Ly, 0, kind == "thermal", 0, Dict{Symbol,HipMirror}())
    T === Float32 && hip_check(ccall((:swalbe_set_reference_height, libswalbe), Cint,
                                 (Ptr{Cvoid}, Float64, Float64), s.handle, href, g),
                           "set_reference_height")
    finalizer(s) do x
        h = getfield(x, :handle)
        h == C_NULL || ccall((:swalbe_destroy, libswalbe), Cint, (Ptr{Cvoid},), h)
        setfield!(x, :handle, C_NULL)
    end
    return s
end

function Base.getproperty(s::HipState{T}, name::Symbol) where {T}
    name === :handle && return hip_flush!(s)
    name in fieldnames(HipState) && return getfield(s, name)
    haskey(HIP_FIELDS, name) || error("HipState has no field $name")
    dims = name in HIP_DISTS ? (getfield(s, :Lx), getfield(s, :Ly), 9) :
           (getfield(s, :Lx), getfield(s, :Ly))
    return HipField{T,length(dims),HipState{T}}(s, name, dims)
end

function Base.setproperty!(s::HipState{T}, name::Symbol, A) where {T}
    name in fieldnames(HipState) && return setfield!(s, name, A)
    f = getproperty(s, name)
    A isa HipField && A.s === s && A.name === name && return A  # after an in-place op
    return A isa Number ? fill!(f, A) : hip_set!(f, A)
end

# ---- ops (src/equilibrium.jl:118, src/collide.jl:151, src/moments.jl:64,
#      src/pressure.jl:119, src/forcing.jl:168, :42, :297, :363)
equilibrium!(s::HipState, sys::SysConst) =
    hip_check(ccall((:swalbe_equilibrium, libswalbe), Cint, (Ptr{Cvoid}, Float64), s.handle,
                sys.param.g), "equilibrium!")
BGKandStream!(s::HipState, sys::SysConst) =
    hip_check(ccall((:swalbe_bgk_and_stream, libswalbe), Cint, (Ptr{Cvoid}, Float64), s.handle,
                sys.param.τ), "BGKandStream!")
moments!(s::HipState) =
    hip_check(ccall((:swalbe_moments, libswalbe), Cint, (Ptr{Cvoid},), s.handle), "moments!")

function filmpressure!(s::HipState, sys::SysConst; θ = sys.param.θ, γ = sys.param.γ,
                       n = sys.param.n, m = sys.param.m, hmin = sys.param.hmin,
                       hcrit = sys.param.hcrit)
    field = !(θ isa Number)
    field && (s.θ = θ)
    hip_check(ccall((:swalbe_filmpressure, libswalbe), Cint,
                (Ptr{Cvoid}, Float64, Float64, Int32, Int32, Float64, Float64, Int32, Int32),
                s.handle, γ, field ? 0.0 : θ, n, m, hmin, hcrit, field, 0), "filmpressure!")
end
h∇p!(s::HipState) =
    hip_check(ccall((:swalbe_hgradp, libswalbe), Cint, (Ptr{Cvoid},), s.handle), "h∇p!")
slippage!(s::HipState, sys::SysConst) =
    hip_check(ccall((:swalbe_slippage, libswalbe), Cint, (Ptr{Cvoid}, Float64, Float64), s.handle,
                sys.param.δ, sys.param.μ), "slippage!")
thermal!(s::HipState, sys::SysConst; seed = 0, step = s.t) =
    hip_check(ccall((:swalbe_thermal, libswalbe), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, UInt64, UInt64), s.handle,
                sys.param.kbt, sys.param.μ, sys.param.δ, seed, step), "thermal!")
function inclination!(α::Vector, s::HipState; t = 1000, tstart = 0, tsmooth = 1)
    fac = 0.5 + 0.5 * tanh((t - tstart) / tsmooth)
    hip_check(ccall((:swalbe_inclination, libswalbe), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64), s.handle, α[1], α[2], fac),
          "inclination!")
end

# ---- time_loop(sys, state[, θ | Δh | f, measure]) (src/simulate.jl:6-96)
function fused!(sys, s::HipState{T}, o::HipStepOpts; theta = sys.param.θ) where {T}
    o.nontemporal < 0 && (o = with_store_mode(o, store_mode(sys.Lx, sys.Ly, T, o.final_outputs_only)))
    p = HipParams(sys.param)
    p = HipParams(p.Tmax, p.tdump, p.tau, p.cs, p.mu, p.delta, p.kbt, p.gamma, p.n, p.m, p.hmin,
                p.hcrit, theta, p.g)
    hip_check(ccall((:swalbe_time_loop, libswalbe), Cint,
                (Ptr{Cvoid}, Ref{HipParams}, Int64, Ref{HipStepOpts}), s.handle, p,
                sys.param.Tmax, o), "time_loop")
    s.t += sys.param.Tmax
    return s
end

function time_loop(sys::SysConst, s::HipState; verbose = false, thermal = false, seed = 0)
    mass = zeros(sys.param.Tmax ÷ sys.param.tdump + 2)
    GC.@preserve mass fused!(sys, s, HipStepOpts(thermal = thermal, seed = seed, t0 = s.t,
                                              mass_series = pointer(mass)))
    if verbose
        k = 1
        for t in (s.t - sys.param.Tmax + 1):s.t
            t % sys.param.tdump == 0 || continue
            println("Time step $t mass is $(round(mass[k], digits=3))")
            k += 1
        end
    end
    return s
end

function time_loop(sys::SysConst, s::HipState, θ; verbose = false)
    θ isa Number && return fused!(sys, s, HipStepOpts(t0 = s.t); theta = θ)
    s.θ = θ
    return fused!(sys, s, HipStepOpts(use_theta_field = 1, t0 = s.t))
end

function time_loop(sys::SysConst, s::HipState, Δh::Vector; verbose = false)
    dh = zeros(sys.param.Tmax)
    GC.@preserve dh fused!(sys, s, HipStepOpts(t0 = s.t, dh_series = pointer(dh)))
    append!(Δh, dh)
    return s
end

# wetted! as the callback runs fused: the count of every step is taken on the device
function time_loop(sys::SysConst, s::HipState, f::typeof(wetted!), measure::Vector;
                   verbose = false)
    wet = zeros(Int64, sys.param.Tmax)
    GC.@preserve wet fused!(sys, s, HipStepOpts(t0 = s.t, wet_series = pointer(wet)))
    append!(measure, wet)
    return s, measure
end

# any other measurement callback sees the state between the force sum and
# equilibrium!, so it runs the per-op kernels (src/simulate.jl:69-96)
function time_loop(sys::SysConst, s::HipState, f::Function, measure::Vector; verbose = false)
    for t in 1:sys.param.Tmax
        filmpressure!(s, sys)
        h∇p!(s)
        slippage!(s, sys)
        hip_check(ccall((:swalbe_force_sum, libswalbe), Cint, (Ptr{Cvoid}, Int32), s.handle, 0),
              "force sum")
        f(measure, s)
        equilibrium!(s, sys)
        BGKandStream!(s, sys)
        moments!(s)
        s.t += 1
    end
    return s, measure
end


# ---- measurements on the device (src/measures.jl) ----------------------------
function hip_measure(s::HipState, thresh; θ = 0.0, surface = true, mask = C_NULL, nmask = 0)
    out = zeros(4)
    usefield = 0
    th = θ
    if !(θ isa Number)
        s.θ = θ
        usefield, th = 1, 0.0
    end
    hip_check(ccall((:swalbe_measure, libswalbe), Cint,
                (Ptr{Cvoid}, Float64, Int32, Float64, Int32, Ptr{Float64}, Ptr{UInt8}, Csize_t),
                s.handle, thresh, usefield, th, surface, out, mask, nmask), "measure")
    return out  # count, maximum, liquid-vapour area, reduced energy
end

# wetted!(area_size, maxheight, height, t; hthresh) (src/measures.jl:6-11)
function wetted!(area_size, maxheight, s::HipState, t; hthresh = 0.055)
    m = hip_measure(s, hthresh; surface = 0)
    area_size[t] = m[1]
    maxheight[t] = m[2]
    return nothing
end
# wetted!(area_size, state; hthresh) (:13-17)
function wetted!(area_size, s::HipState; hthresh = 0.055)
    push!(area_size, Int(hip_measure(s, hthresh; surface = 0)[1]))
    return nothing
end
# fluid_dry!(fluid, dummy, height, t; hthresh) (:23-28)
function fluid_dry!(fluid, dummy, s::HipState, t; hthresh = 0.055)
    m = zeros(UInt8, s.Lx * s.Ly)
    GC.@preserve m hip_measure(s, hthresh; surface = 0, mask = pointer(m), nmask = length(m))
    dummy .= reshape(m .== 1, s.Lx, s.Ly)
    fluid[t, :] .= vec(dummy)
    return nothing
end
# surfacearea!(area_lv, red_energy, height, θ, ∇hx, ∇hy, dgrad, surface, t; htresh) (:126-168)
function surfacearea!(area_lv, red_energy, s::HipState, θ, ∇hx, ∇hy, dgrad, surface, t;
                      htresh = 0.055)
    m = hip_measure(s, htresh; θ = θ)
    area_lv[t] = m[3]
    red_energy[t] = m[4]
    return nothing
end

# ---- one-dimensional films (D1Q3) ---------------------------------------------
# Sys(sysc::Consts_1D; T, kind) gains a `device = "HIP"` keyword branch:
#     device == "HIP" && return HipState1D{T}(sysc.L; kind = kind)
const HIP1D_FIELDS = Dict(:fout => 0, :ftemp => 1, :feq => 2, :height => 3, :vel => 4,
                      :pressure => 5, :F => 6, :slip => 7, :h∇p => 8, :kbt => 9, :θ => 10,
                      :γ => 11, :∇γ => 12, :border1 => 13, :border2 => 14, :interior => 15)

struct HipStepOpts1D
    use_theta_field::Int32; use_gamma_field::Int32; grad_gamma::Int32; thermal::Int32
    body::Int32; bound::Int32; materialize::Int32; variant::Int32
    seed::UInt64; t0::Int64
    alpha::Float64; body_fac::Float64; wet_thresh::Float64
    dh_series::Ptr{Float64}; wet_series::Ptr{Int64}; mass_series::Ptr{Float64}
    snap_every::Int64; snap_series::Ptr{Float64}
    threads::Int32; pad_::Int32
end
HipStepOpts1D(; use_theta_field = 0, use_gamma_field = 0, grad_gamma = 0, thermal = 0, body = 0,
           bound = 0, materialize = 1, variant = 0, seed = 0, t0 = 0, alpha = 0.0,
           body_fac = 0.0, wet_thresh = 0.055, dh_series = C_NULL, wet_series = C_NULL,
           mass_series = C_NULL, snap_every = 0, snap_series = C_NULL) =
    HipStepOpts1D(use_theta_field, use_gamma_field, grad_gamma, thermal, body, bound, materialize,
               variant, seed, t0, alpha, body_fac, wet_thresh, dh_series, wet_series,
               mass_series, snap_every, snap_series, 0, 0)

"""GPU-resident State_1D (and its thermal / gamma kinds): the library owns the
arrays; `s.basestate` is `s` itself and `s.<field>` is a HipField."""
mutable struct HipState1D{T} <: LBM_state_1D
    handle::Ptr{Cvoid}
    L::Int
    t::Int
    version::Int
    mirrors::Dict{Symbol,HipMirror}
end
function HipState1D{T}(L; kind = "simple") where {T}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    hip_check(ccall((:swalbe1d_create, libswalbe), Cint, (Int32, Int32, Int32, Ref{Ptr{Cvoid}}),
                L, T === Float32 ? HIP_F32 : HIP_F64,
                kind == "thermal" ? HIP_THERMAL : HIP_SIMPLE, h), "Sys (1D)")
    x = HipState1D{T}(h[], L, 0, 0, Dict{Symbol,HipMirror}())
    finalizer(x) do x
        h = getfield(x, :handle)
        h == C_NULL || ccall((:swalbe1d_destroy, libswalbe), Cint, (Ptr{Cvoid},), h)
        setfield!(x, :handle, C_NULL)
    end
    return x
end
hip_download!(s::HipState1D, name::Symbol, A::Array{Float64}) =
    hip_check(ccall((:swalbe1d_download, libswalbe), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t),
                getfield(s, :handle), HIP1D_FIELDS[name], A, sizeof(A)), "download $name")
hip_upload!(s::HipState1D, name::Symbol, A::Array{Float64}) =
    hip_check(ccall((:swalbe1d_upload, libswalbe), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t),
                getfield(s, :handle), HIP1D_FIELDS[name], A, sizeof(A)), "upload $name")
hip_fill!(s::HipState1D, name::Symbol, x::Float64) =
    hip_check(ccall((:swalbe1d_fill, libswalbe), Cint, (Ptr{Cvoid}, Int32, Float64),
                getfield(s, :handle), HIP1D_FIELDS[name], x), "fill $name")
function Base.getproperty(s::HipState1D{T}, name::Symbol) where {T}
    name === :handle && return hip_flush!(s)
    name in fieldnames(HipState1D) && return getfield(s, name)
    name === :basestate && return s
    haskey(HIP1D_FIELDS, name) || error("HipState1D has no field $name")
    L = getfield(s, :L)
    dims = name in (:fout, :ftemp, :feq) ? (L, 3) : (L,)
    return HipField{T,length(dims),HipState1D{T}}(s, name, dims)
end
function Base.setproperty!(s::HipState1D, name::Symbol, A)
    name in fieldnames(HipState1D) && return setfield!(s, name, A)
    f = getproperty(s, name)
    A isa HipField && A.s === s && A.name === name && return A  # after an in-place op
    return A isa Number ? fill!(f, A) : hip_set!(f, A)
end

function hip_walls!(s::HipState1D, sys::SysConstWithBound_1D)
    s.border1 = sys.border[1]
    s.border2 = sys.border[2]
    s.interior = sys.interior
end

equilibrium!(s::HipState1D, sys::Consts_1D; g = sys.param.g) =
    hip_check(ccall((:swalbe1d_equilibrium, libswalbe), Cint, (Ptr{Cvoid}, Float64), s.handle, g),
          "equilibrium!")
BGKandStream!(s::HipState1D, sys::SysConst_1D) =
    hip_check(ccall((:swalbe1d_bgk_and_stream, libswalbe), Cint, (Ptr{Cvoid}, Float64, Int32),
                s.handle, sys.param.τ, 0), "BGKandStream!")
function BGKandStream!(s::HipState1D, sys::SysConstWithBound_1D)
    hip_walls!(s, sys)
    hip_check(ccall((:swalbe1d_bgk_and_stream, libswalbe), Cint, (Ptr{Cvoid}, Float64, Int32),
                s.handle, sys.param.τ, 1), "BGKandStream!")
end
moments!(s::HipState1D) =
    hip_check(ccall((:swalbe1d_moments, libswalbe), Cint, (Ptr{Cvoid},), s.handle), "moments!")
function filmpressure!(s::HipState1D, sys::Consts_1D; θ = sys.param.θ, n = sys.param.n,
                       m = sys.param.m, hmin = sys.param.hmin, hcrit = sys.param.hcrit,
                       γ = sys.param.γ)
    usθ, usγ = 0, 0
    if !(θ isa Number); s.θ = θ; usθ, θ = 1, 0.0; end
    if !(γ isa Number); s.γ = γ; usγ, γ = 1, 0.0; end
    hip_check(ccall((:swalbe1d_filmpressure, libswalbe), Cint,
                (Ptr{Cvoid}, Float64, Float64, Int32, Int32, Float64, Float64, Int32, Int32, Int32),
                s.handle, γ, θ, n, m, hmin, hcrit, usθ, usγ, 0), "filmpressure!")
end
h∇p!(s::HipState1D) =
    hip_check(ccall((:swalbe1d_hgradp, libswalbe), Cint, (Ptr{Cvoid},), s.handle), "h∇p!")
slippage!(s::HipState1D, sys::Consts_1D) =
    hip_check(ccall((:swalbe1d_slippage, libswalbe), Cint, (Ptr{Cvoid}, Float64, Float64),
                s.handle, sys.param.δ, sys.param.μ), "slippage!")
thermal!(s::HipState1D, sys::Consts_1D; seed = 0) =
    hip_check(ccall((:swalbe1d_thermal, libswalbe), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, UInt64, UInt64), s.handle,
                sys.param.kbt, sys.param.μ, sys.param.δ, seed, s.t), "thermal!")
inclination!(α::Float64, s::HipState1D; t = 1000, tstart = 0, tsmooth = 1) =
    hip_check(ccall((:swalbe1d_inclination, libswalbe), Cint, (Ptr{Cvoid}, Float64, Float64),
                s.handle, α, 0.5 + 0.5 * tanh((t - tstart) / tsmooth)), "inclination!")
∇γ!(s::HipState1D) =
    hip_check(ccall((:swalbe1d_grad_gamma, libswalbe), Cint, (Ptr{Cvoid}, Int32, Float64),
                s.handle, 0, 1.0), "∇γ!")
∇γ!(s::HipState1D, sys::SysConst_1D) =
    hip_check(ccall((:swalbe1d_grad_gamma, libswalbe), Cint, (Ptr{Cvoid}, Int32, Float64),
                s.handle, 1, sys.param.δ), "∇γ!")

function fused1d!(sys, s::HipState1D, o::HipStepOpts1D; theta = sys.param.θ, gamma = sys.param.γ)
    p = HipParams(sys.param)
    p = HipParams(p.Tmax, p.tdump, p.tau, p.cs, p.mu, p.delta, p.kbt, gamma, p.n, p.m, p.hmin,
                p.hcrit, theta, p.g)
    hip_check(ccall((:swalbe1d_time_loop, libswalbe), Cint,
                (Ptr{Cvoid}, Ref{HipParams}, Int64, Ref{HipStepOpts1D}), s.handle, p,
                sys.param.Tmax, o), "time_loop")
    s.t += sys.param.Tmax
    return s
end

# time_loop(sys::SysConst_1D, state[, θ | Δh | f, measure]) (src/simulate.jl:98-178)
time_loop(sys::SysConst_1D, s::HipState1D; verbose = false) =
    fused1d!(sys, s, HipStepOpts1D(t0 = s.t))
function time_loop(sys::SysConst_1D, s::HipState1D, θ; verbose = false)
    θ isa Number && return fused1d!(sys, s, HipStepOpts1D(t0 = s.t); theta = θ)
    s.θ = θ
    return fused1d!(sys, s, HipStepOpts1D(use_theta_field = 1, t0 = s.t))
end
function time_loop(sys::SysConst_1D, s::HipState1D, Δh::Vector; verbose = false)
    dh = zeros(sys.param.Tmax)
    GC.@preserve dh fused1d!(sys, s, HipStepOpts1D(t0 = s.t, dh_series = pointer(dh)))
    append!(Δh, dh)
    return s
end
function time_loop(sys::SysConst_1D, s::HipState1D, f::typeof(inclination!), α; verbose = false)
    fused1d!(sys, s, HipStepOpts1D(t0 = s.t, body = 1, alpha = α, body_fac = 1.0))
    return s, α
end
function time_loop(sys::SysConst_1D, s::HipState1D, f::Function, measure; verbose = false)
    for t in 1:sys.param.Tmax
        filmpressure!(s, sys); h∇p!(s); slippage!(s, sys)
        hip_check(ccall((:swalbe1d_force_sum, libswalbe), Cint, (Ptr{Cvoid}, Int32, Int32),
                    s.handle, 0, 0), "force sum")
        f(measure, s)
        equilibrium!(s, sys); BGKandStream!(s, sys); moments!(s)
        s.t += 1
    end
    return s, measure
end
# the walled loop (src/simulate.jl:181-205)
function time_loop(sys::SysConstWithBound_1D, s::HipState1D; verbose = false, diff = 0.05)
    hip_walls!(s, sys)
    return fused1d!(sys, s, HipStepOpts1D(t0 = s.t, bound = 1))
end

# ---- raw-array methods on caller-owned device arrays --------------------------
# The reference's positional methods (src/equilibrium.jl:63, src/collide.jl:70,
# src/moments.jl:43, src/pressure.jl:72, src/differences.jl:57-206,
# src/forcing.jl:42, :297), for AMDGPU.jl ROCArrays in Julia's layout.  Each
# method is one swalbe_*_ptr call, enqueued on AMDGPU's task-local stream, so it
# orders with the caller's own broadcasts on the same arrays.
using AMDGPU: ROCArray, ROCMatrix
hipptr(A) = A === nothing ? C_NULL : reinterpret(Ptr{Cvoid}, pointer(A))
hipstream() = AMDGPU.stream().stream
hipdt(::Type{Float64}) = HIP_F64
hipdt(::Type{Float32}) = HIP_F32

function equilibrium!(feq::ROCArray{T,3}, height::ROCMatrix{T}, velx::ROCMatrix{T},
                      vely::ROCMatrix{T}, vsq::ROCMatrix{T}, gravity) where {T}
    Lx, Ly = size(height)
    hip_check(ccall((:swalbe_equilibrium_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64,
                 Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(feq), hipptr(height), hipptr(velx), hipptr(vely), hipptr(vsq),
                gravity, Lx, Ly, hipstream()), "equilibrium!")
    return nothing
end
function BGKandStream!(fout::ROCArray{T,3}, feq::ROCArray{T,3}, ftemp::ROCArray{T,3},
                       Fx::ROCMatrix{T}, Fy::ROCMatrix{T}, τ) where {T}
    Lx, Ly = size(Fx)
    hip_check(ccall((:swalbe_bgk_and_stream_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64,
                 Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(fout), hipptr(feq), hipptr(ftemp), hipptr(Fx), hipptr(Fy), τ,
                Lx, Ly, hipstream()), "BGKandStream!")
    return nothing
end
function moments!(height::ROCMatrix{T}, velx::ROCMatrix{T}, vely::ROCMatrix{T},
                  fout::ROCArray{T,3}) where {T}
    Lx, Ly = size(height)
    hip_check(ccall((:swalbe_moments_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(height), hipptr(velx), hipptr(vely), hipptr(fout), Lx, Ly,
                hipstream()), "moments!")
    return nothing
end
# the raw form: (n, m) ∈ {(9, 3), (3, 2)} only, DomainError otherwise (src/pressure.jl:102-107)
function filmpressure!(output::ROCMatrix{T}, f::ROCMatrix{T}, dgrad, γ, θ, n, m, hmin,
                       hcrit) where {T}
    Lx, Ly = size(f)
    field = θ isa ROCMatrix ? hipptr(θ) : C_NULL
    θs = θ isa Number ? Float64(θ) : 0.0
    hip_check(ccall((:swalbe_filmpressure_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Ptr{Cvoid}, Int32, Int32,
                 Float64, Float64, Int32, Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(output), hipptr(f), γ, θs, field, n, m, hmin, hcrit, 1, Lx, Ly,
                hipstream()), "filmpressure!")
    return nothing
end
function hip_grad!(outputx::ROCMatrix{T}, outputy::ROCMatrix{T}, f::ROCMatrix{T}, a) where {T}
    Lx, Ly = size(f)
    hip_check(ccall((:swalbe_grad_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(outputx), hipptr(outputy), hipptr(f), hipptr(a), Lx, Ly,
                hipstream()), "∇f!")
    return nothing
end
∇f!(outputx::ROCMatrix, outputy::ROCMatrix, f::ROCMatrix) = hip_grad!(outputx, outputy, f, nothing)
∇f!(outputx::ROCMatrix, outputy::ROCMatrix, f::ROCMatrix, a::ROCMatrix) =
    hip_grad!(outputx, outputy, f, a)
∇f!(outputx::ROCMatrix, outputy::ROCMatrix, f::ROCMatrix, dgrad, a::ROCMatrix) =
    hip_grad!(outputx, outputy, f, a)
# a scalar `a` scales the unweighted stencil: one product either way, the same
# bits in FP64 (in FP32 the stencil is rounded before the product)
function ∇f!(outputx::ROCMatrix, outputy::ROCMatrix, f::ROCMatrix, a::Number)
    hip_grad!(outputx, outputy, f, nothing)
    outputx .*= a
    outputy .*= a
    return nothing
end
∇f!(outputx::ROCMatrix, outputy::ROCMatrix, f::ROCMatrix, dgrad, a::Number) =
    ∇f!(outputx, outputy, f, a)
function ∇²f!(output::ROCMatrix{T}, f::ROCMatrix{T}, γ) where {T}
    Lx, Ly = size(f)
    hip_check(ccall((:swalbe_laplacian_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(output), hipptr(f), γ, Lx, Ly, hipstream()), "∇²f!")
    return nothing
end
function slippage!(slipx::ROCMatrix{T}, slipy::ROCMatrix{T}, height::ROCMatrix{T},
                   velx::ROCMatrix{T}, vely::ROCMatrix{T}, δ, μ) where {T}
    Lx, Ly = size(height)
    hip_check(ccall((:swalbe_slippage_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64,
                 Float64, Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(slipx), hipptr(slipy), hipptr(height), hipptr(velx),
                hipptr(vely), δ, μ, Lx, Ly, hipstream()), "slippage!")
    return nothing
end
# the normals come from the counter-based stream (seed, step, node), not Julia's randn!
function thermal!(fluc_x::ROCMatrix{T}, fluc_y::ROCMatrix{T}, height::ROCMatrix{T}, kᵦT, μ, δ;
                  seed = 0, step = 0) where {T}
    Lx, Ly = size(height)
    hip_check(ccall((:swalbe_thermal_ptr, libswalbe), Cint,
                (Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Float64, UInt64,
                 UInt64, Int32, Int32, Ptr{Cvoid}),
                hipdt(T), hipptr(fluc_x), hipptr(fluc_y), hipptr(height), kᵦT, μ, δ, seed, step,
                Lx, Ly, hipstream()), "thermal!")
    return nothing
end
synchronize(s::HipState) =
    hip_check(ccall((:swalbe_synchronize, libswalbe), Cint, (Ptr{Cvoid},), s.handle), "synchronize")

# ---- batches and the server (no reference counterpart: the reference runs a
# parameter sweep as a serial loop of run_* calls, scripts/Rivulet_stability.jl:141-190)
# swalbe_film_init
struct HipFilmInit
    kind::Int32; thermal::Int32; body::Int32; clamp::Int32
    seed::UInt64
    base::Float64; amp::Float64; precursor::Float64; href::Float64
    r1::Float64; theta1::Float64; c1x::Float64; c1y::Float64
    r2::Float64; theta2::Float64; c2x::Float64; c2y::Float64
    body_ax::Float64; body_ay::Float64; body_fac::Float64
end
# kind: 0 flat, 1 base + amp N(0,1), 2 one cap, 3 two caps (1-based centres, θ in units of π)
HipFilmInit(; kind = 0, thermal = 0, body = 0, clamp = 0, seed = 0, base = 1.0, amp = 0.0,
         precursor = 0.05, href = 0.0, r1 = 0.0, theta1 = 0.0, c1x = 0.0, c1y = 0.0, r2 = 0.0,
         theta2 = 0.0, c2x = 0.0, c2y = 0.0, body_ax = 0.0, body_ay = 0.0, body_fac = 0.0) =
    HipFilmInit(kind, thermal, body, clamp, seed, base, amp, precursor, href, r1, theta1, c1x,
             c1y, r2, theta2, c2x, c2y, body_ax, body_ay, body_fac)

# FP32 films of any sizes in one arena, stepped together by one launch per step
mutable struct HipBatch
    handle::Ptr{Cvoid}
    function HipBatch(capacity_bytes; max_films = 4096, graph_steps = 0, autotune = true)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        hip_check(ccall((:swalbe_batch_create, libswalbe), Cint,
                    (Int32, Csize_t, Int32, Ref{Ptr{Cvoid}}), HIP_F32, capacity_bytes, max_films,
                    out), "batch create")
        b = finalizer(new(out[])) do x
            x.handle == C_NULL || ccall((:swalbe_batch_destroy, libswalbe), Cint, (Ptr{Cvoid},),
                                        x.handle)
        end
        autotune && hip_check(ccall((:swalbe_batch_set_autotune, libswalbe), Cint,
                                    (Ptr{Cvoid}, Int32), b.handle, 1), "autotune")
        graph_steps > 0 && hip_check(ccall((:swalbe_batch_set_graph, libswalbe), Cint,
                                           (Ptr{Cvoid}, Int32), b.handle, graph_steps), "graph")
        return b
    end
end
function Base.push!(b::HipBatch, sys::SysConst, init::HipFilmInit, nsteps)
    slot = Ref{Int32}(0)
    hip_check(ccall((:swalbe_batch_add, libswalbe), Cint,
                (Ptr{Cvoid}, Ref{HipParams}, Int32, Int32, Ref{HipFilmInit}, Int64, Ref{Int32}),
                b.handle, HipParams(sys.param), sys.Lx, sys.Ly, init, nsteps, slot), "batch add")
    return slot[]
end
Base.delete!(b::HipBatch, slot) =
    hip_check(ccall((:swalbe_batch_remove, libswalbe), Cint, (Ptr{Cvoid}, Int32), b.handle, slot),
          "batch remove")
step!(b::HipBatch, nsteps) =
    hip_check(ccall((:swalbe_batch_step, libswalbe), Cint, (Ptr{Cvoid}, Int64), b.handle, nsteps),
          "batch step")
function batch_height(b::HipBatch, slot, Lx, Ly)
    h = Matrix{Float64}(undef, Lx, Ly)
    hip_check(ccall((:swalbe_batch_download, libswalbe), Cint,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}, Csize_t), b.handle, slot,
                HIP_FIELDS[:height], h, sizeof(h)), "batch download")
    return h
end

# continuous batching of solve requests: submit, then poll until each is done
mutable struct HipServer
    handle::Ptr{Cvoid}
    function HipServer(capacity_bytes; max_films = 4096, tick = 500, graph_steps = 0)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        hip_check(ccall((:swalbe_serve_create, libswalbe), Cint,
                    (Csize_t, Int32, Int32, Int32, Ref{Ptr{Cvoid}}), capacity_bytes, max_films,
                    tick, graph_steps, out), "serve create")
        return finalizer(new(out[])) do x
            x.handle == C_NULL || ccall((:swalbe_serve_destroy, libswalbe), Cint, (Ptr{Cvoid},),
                                        x.handle)
        end
    end
end
function submit!(s::HipServer, sys::SysConst, init::HipFilmInit, nsteps)
    id = Ref{Int64}(0)
    hip_check(ccall((:swalbe_serve_submit, libswalbe), Cint,
                (Ptr{Cvoid}, Ref{HipParams}, Int32, Int32, Ref{HipFilmInit}, Int64, Ref{Int64}),
                s.handle, HipParams(sys.param), sys.Lx, sys.Ly, init, nsteps, id), "submit")
    return id[]
end
# one tick: admit, enqueue, hand back what finished; returns [(id, height), ...]
function poll!(s::HipServer)
    done, resident, queued = Ref{Int32}(0), Ref{Int32}(0), Ref{Int32}(0)
    hip_check(ccall((:swalbe_serve_poll, libswalbe), Cint,
                (Ptr{Cvoid}, Ref{Int32}, Ref{Int32}, Ref{Int32}), s.handle, done, resident, queued),
          "poll")
    out = Tuple{Int64,Matrix{Float64}}[]
    for _ in 1:done[]
        id, Lx, Ly = Ref{Int64}(0), Ref{Int32}(0), Ref{Int32}(0)
        hip_check(ccall((:swalbe_serve_peek, libswalbe), Cint,
                    (Ptr{Cvoid}, Ref{Int64}, Ref{Int32}, Ref{Int32}), s.handle, id, Lx, Ly), "peek")
        h = Matrix{Float64}(undef, Lx[], Ly[])
        hip_check(ccall((:swalbe_serve_take, libswalbe), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                    s.handle, h, sizeof(h)), "take")
        push!(out, (id[], h))
    end
    return out
end
