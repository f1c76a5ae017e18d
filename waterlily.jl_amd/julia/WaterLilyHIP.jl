"""
    WaterLilyHIP

Drop-in gfx950 (MI355X) time step for WaterLily.jl through its own plugin seam
(`Simulation(...; flow_ctor, pois_ctor)`, src/WaterLily.jl:96-97,103,105):

    using WaterLily, WaterLilyHIP
    sim = Simulation((256,128,128), (1,0,0), 32; ν=32/3700,
                     body=AutoBody((x,t)->√sum(abs2,x .- 64)-16),
                     flow_ctor=HipFlow, pois_ctor=HipPoisson)
    sim_step!(sim, 10; remeasure=false)

`HipFlow <: AbstractFlow` owns the device state behind the C ABI of `include/wl.h`;
`mom_step!(::HipFlow, ::HipPoisson)` runs the whole predictor/projection/corrector/CFL step on
the GPU and keeps `Δt` and `pois.n` on the host exactly as `Flow`/`MultiLevelPoisson` do.
Bodies are measured by WaterLily's own `measure!` on a host `Flow` (any `AbstractBody`) and the
BDIM fields are uploaded; `Sphere`/`Ellipsoid` descriptors can be measured on the device.
Fields read through `flow.u`, `flow.p`, ... are host copies (downloads).
"""
module WaterLilyHIP

using WaterLily
import WaterLily: mom_step!, measure!, update!, AbstractFlow, AbstractPoisson, AbstractBody

const libwl = get(ENV, "WATERLILY_HIP_LIB", joinpath(@__DIR__, "..", "libwaterlily_hip.so"))

# --- mirrors of the C structs (include/wl.h) ---
struct WlFlowDesc
    D::Int32; n::NTuple{3,Int32}; dtype::Int32; scheme::Int32; perdir_mask::Int32; exit_bc::Int32
    has_g::Int32; has_uBC_rate::Int32
    uBC::NTuple{3,Float64}; uBC_rate::NTuple{3,Float64}; g0::NTuple{3,Float64}; g1::NTuple{3,Float64}
    nu::Float64; dt0::Float64
end
struct WlBody
    kind::Int32; reserved::Int32; center::NTuple{3,Float64}; radius::Float64; axes::NTuple{3,Float64}
end

const FIELD = (u=0, u⁰=1, f=2, p=3, σ=4, V=5, μ₀=6, μ₁=7)

check(rc) = rc == 0 ? nothing :
    rc == -1 ? throw(AssertionError(unsafe_string(ccall((:wl_last_error, libwl), Cstring, ())))) :
    error("libwaterlily_hip error $rc: " * unsafe_string(ccall((:wl_last_error, libwl), Cstring, ())))

scheme_id(λ) = λ === WaterLily.quick ? 0 : λ === WaterLily.vanLeer ? 1 : λ === WaterLily.cds ? 2 :
    throw(ArgumentError("convective scheme $λ has no device kernel (quick, vanLeer, cds)"))

pad3(t, z) = ntuple(i -> i <= length(t) ? Float64(t[i]) : z, 3)

mutable struct HipFlow{D,T} <: AbstractFlow{D,T}
    h::Ptr{Cvoid}
    N::NTuple{D,Int}
    Δt::Vector{T}
    ν::T
    uBC::Union{NTuple{D,Number},Function}
    g::Union{Function,Nothing}
    perdir::NTuple
    exitBC::Bool
    host::Flow{D,T}    # host-side Flow used by measure! for arbitrary bodies
end

"""
    HipFlow(N, uBC; Δt, ν, g, λ, T, perdir, exitBC, u0, mem)

`flow_ctor` for `Simulation`: same keywords as `Flow` (src/Flow.jl:133-134).  `uBC` must be a
tuple (the device applies uniform boundary values); a time-varying uniform `uBC` and a uniform
`g` are passed as (value, rate) pairs through the C descriptor.
"""
function HipFlow(N::NTuple{D}, uBC; Δt=0.25, ν=0., g=nothing, λ=WaterLily.quick, T=Float32,
                 perdir=(), exitBC=false, u0=nothing, mem=Array, kw...) where D
    isa(uBC, Tuple) || throw(ArgumentError("HipFlow: uBC must be a tuple of boundary velocities"))
    isnothing(g) || throw(ArgumentError("HipFlow: pass body forces as a uniform (g0,g1) descriptor"))
    isnothing(u0) || throw(ArgumentError("HipFlow: the device path starts from u = uBC"))
    desc = WlFlowDesc(Int32(D), ntuple(i -> i <= D ? Int32(N[i]) : Int32(0), 3), T == Float64 ? 1 : 0,
                      scheme_id(λ), Int32(sum(1 << (j - 1) for j in perdir; init=0)), Int32(exitBC), 0, 0,
                      pad3(uBC, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), Float64(ν), Float64(Δt))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wl_flow_create, libwl), Cint, (Ptr{Ptr{Cvoid}}, Ref{WlFlowDesc}), h, desc))
    host = Flow(N, uBC; Δt, ν, λ, T, perdir, exitBC)
    f = HipFlow{D,T}(h[], Tuple(N), T[Δt], T(ν), uBC, g, perdir, exitBC, host)
    finalizer(f -> ccall((:wl_flow_destroy, libwl), Cint, (Ptr{Cvoid},), f.h), f)
    f
end

function download(f::HipFlow{D,T}, name::Symbol) where {D,T}
    Ng = f.N .+ 2
    a = name in (:p, :σ) ? zeros(T, Ng) : name == :μ₁ ? zeros(T, Ng..., D, D) : zeros(T, Ng..., D)
    check(ccall((:wl_flow_download, libwl), Cint, (Ptr{Cvoid}, Cint, Ptr{T}, Csize_t),
                f.h, FIELD[name], a, sizeof(a)))
    a
end
function upload!(f::HipFlow{D,T}, name::Symbol, a::Array{T}) where {D,T}
    check(ccall((:wl_flow_upload, libwl), Cint, (Ptr{Cvoid}, Cint, Ptr{T}, Csize_t),
                f.h, FIELD[name], a, sizeof(a)))
end
Base.getproperty(f::HipFlow, s::Symbol) = s in keys(FIELD) ? download(f, s) : getfield(f, s)

"""
    measure!(flow::HipFlow, body; t, ϵ)

WaterLily's own `measure!` (src/Body.jl:28-51) on the host `Flow`, then μ₀, μ₁, V and σ are
uploaded (any `AbstractBody`, including AD-differentiated AutoBody closures).
"""
function measure!(a::HipFlow, body::AbstractBody; t=zero(eltype(a.Δt)), ϵ=1)
    measure!(a.host, body; t, ϵ)
    for s in (:V, :μ₀, :μ₁, :σ)
        upload!(a, s, getfield(a.host, s))
    end
end

mutable struct HipPoisson{T} <: AbstractPoisson{T,Array{T},Array{T}}
    flow::HipFlow
    n::Vector{Int16}
end
"`pois_ctor` for `Simulation`: the MultiLevelPoisson(flow.p, flow.μ₀, flow.σ) of the device flow."
HipPoisson(flow::HipFlow{D,T}) where {D,T} = (p = HipPoisson{T}(flow, Int16[]); update!(p); p)
update!(p::HipPoisson) = check(ccall((:wl_flow_pois_update, libwl), Cint, (Ptr{Cvoid},), p.flow.h))

"""
    mom_step!(a::HipFlow, b::HipPoisson; udf=nothing)

src/Flow.jl:156-167 on the device: t₁ = sum(Δt), t₀ = t₁-Δt[end]; the CFL Δt is pushed onto
a.Δt and the two V-cycle counts onto b.n.
"""
function mom_step!(a::HipFlow{D,T}, b::HipPoisson; udf=nothing, kw...) where {D,T}
    isnothing(udf) || throw(ArgumentError("HipFlow: user-defined functions cannot run on the device"))
    t₁ = sum(a.Δt); t₀ = t₁ - a.Δt[end]
    dt = Ref{Float64}(0); n = zeros(Int16, 2)
    check(ccall((:wl_mom_step, libwl), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Ptr{Float64}, Ptr{Int16}),
                a.h, a.Δt[end], t₀, t₁, dt, n))
    push!(a.Δt, T(dt[])); append!(b.n, n)
end

WaterLily.time(a::HipFlow) = sum(@view(a.Δt[1:end-1]))

export HipFlow, HipPoisson
end # module
