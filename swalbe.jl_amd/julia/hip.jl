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
#     h = state.height                                   # downloads on access
#
# Every function below is a more specific method of a Swalbe function for
# HipState; the call sites in src/simulate.jl (time_loop, run_*) and user
# scripts stay as they are.  `state.height` downloads a copy; assignments
# `state.height = A` or `state.height = 1.0` upload or fill on the device.

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
    tile_y::Int32; variant::Int32; nontemporal::Int32; pad_::Int32
    snap_every::Int64; snap_series::Ptr{Float64}
end
HipStepOpts(; use_theta_field = 0, thermal = 0, body = 0, materialize = 1, seed = 0, t0 = 0,
         body_ax = 0.0, body_ay = 0.0, body_fac = 0.0, wet_thresh = 0.055,
         dh_series = C_NULL, wet_series = C_NULL, mass_series = C_NULL) =
    HipStepOpts(use_theta_field, thermal, body, materialize, seed, t0, body_ax, body_ay, body_fac,
             wet_thresh, dh_series, wet_series, mass_series, 0, 0, 0, 0, 0, C_NULL)

function hip_check(rc::Cint, what)
    rc == HIP_OK && return nothing
    msg = unsafe_string(ccall((:swalbe_last_error, libswalbe), Cstring, ()))
    rc == HIP_EDOMAIN && throw(DomainError(what, msg))
    error("$what: $msg (status $rc)")
end

"""GPU-resident State{T,3} / State_thermal{T,3}: the library owns the arrays."""
mutable struct HipState{T} <: LBM_state_2D
    handle::Ptr{Cvoid}
    Lx::Int
    Ly::Int
    t::Int
    thermal::Bool
end

function HipState{T}(Lx, Ly; kind = "simple", href = 1.0, g = 0.0) where {T}
    out = Ref{Ptr{Cvoid}}(C_NULL)
    hip_check(ccall((:swalbe_create, libswalbe), Cint,
                (Int32, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
                Lx, Ly, T === Float32 ? HIP_F32 : HIP_F64, kind == "thermal" ? HIP_THERMAL : HIP_SIMPLE, out),
          "swalbe_create")
    s = HipState{T}(out[], Lx, Ly, 0, kind == "thermal")
    T === Float32 && hip_check(ccall((:swalbe_set_reference_height, libswalbe), Cint,
                                 (Ptr{Cvoid}, Float64, Float64), s.handle, href, g),
                           "set_reference_height")
    finalizer(s) do x
        x.handle == C_NULL || ccall((:swalbe_destroy, libswalbe), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return s
end

function Base.getproperty(s::HipState{T}, name::Symbol) where {T}
    name in fieldnames(HipState) && return getfield(s, name)
    haskey(HIP_FIELDS, name) || error("HipState has no field $name")
    dims = name in HIP_DISTS ? (s.Lx, s.Ly, 9) : (s.Lx, s.Ly)
    A = Array{T}(undef, dims)
    hip_check(ccall((:swalbe_download, libswalbe), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t),
                getfield(s, :handle), HIP_FIELDS[name], A, sizeof(A)), "download $name")
    return A
end

function Base.setproperty!(s::HipState{T}, name::Symbol, A) where {T}
    name in fieldnames(HipState) && return setfield!(s, name, A)
    if A isa Number
        return hip_check(ccall((:swalbe_fill, libswalbe), Cint, (Ptr{Cvoid}, Int32, Float64),
                           s.handle, HIP_FIELDS[name], A), "fill $name")
    end
    B = convert(Array{T}, A)
    hip_check(ccall((:swalbe_upload, libswalbe), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Csize_t),
                s.handle, HIP_FIELDS[name], B, sizeof(B)), "upload $name")
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
function fused!(sys, s::HipState, o::HipStepOpts; theta = sys.param.θ)
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

# a measurement callback sees the state between the force sum and
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

