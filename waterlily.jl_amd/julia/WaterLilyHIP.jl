"""
    WaterLilyHIP

Drop-in gfx950 (MI355X) time step for WaterLily.jl through its own plugin seam
(`Simulation(...; flow_ctor, pois_ctor)`, src/WaterLily.jl:96-97,103,105):

    using WaterLily, WaterLilyHIP
    sim = Simulation((256,128,128), (1,0,0), 32; ν=32/3700,
                     body=DeviceSphere((64,64,64), 16),
                     flow_ctor=HipFlow, pois_ctor=HipPoisson)
    sim_step!(sim, 10; remeasure=false)
    pressure_force(sim)          # device reduction (DeviceBody), or WaterLily's host loop otherwise

`HipFlow <: AbstractFlow` owns the device state behind the C ABI of `include/wl.h`;
`mom_step!(::HipFlow, ::HipPoisson)` runs the whole predictor/projection/corrector/CFL step on
the GPU and keeps `Δt` and `pois.n` on the host exactly as `Flow`/`MultiLevelPoisson` do.

- Bodies: any `AbstractBody` is measured by WaterLily's own `measure!` on a host `Flow` and the
  BDIM fields are uploaded; `DeviceSphere`/`DeviceEllipsoid` (and their unions with `+`) are
  measured, and their forces and moments integrated, on the device.
- Boundary values and forcing: tuple `uBC` and `g = nothing` run entirely on the device;
  function-valued `uBC(i,x,t)` and `g(i,x,t)` are evaluated on the host each step by WaterLily's
  own `apply!`/`accelerate!` and uploaded (src/Flow.jl:69-73,138-141).
- `udf`: `sim_step!(sim; udf, kw...)` runs the user function through `wl_flow_set_udf` between
  `conv_diff!` and `accelerate!` of both phases, on the host `Flow` mirror whose fields are
  downloaded before and uploaded after the call (src/Flow.jl:193,207,255-257).
- `u0` (and its deprecated alias `uλ`): the initial velocity, evaluated with WaterLily's `apply!`.
- Fields read through `flow.u`, `flow.p`, ... are host copies (downloads); `flow.p = A` uploads.
  In-place writes into such a copy (`copyto!(flow.p, A)`) do not reach the device.
- `HipMeanFlow`, `HipBatch`, `solver_mode!` and `CFL` mirror `MeanFlow` (src/Metrics.jl:199-261),
  batched stepping, the solver modes and `CFL` (src/Flow.jl:234-237).
"""
module WaterLilyHIP

using WaterLily
import WaterLily: mom_step!, measure!, measure, update!, AbstractFlow, AbstractPoisson, AbstractBody,
                  pressure_force, viscous_force, pressure_moment, viscous_moment, sim_step!, CFL, reset!, uu

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
const FN_UBC, FN_ACC0, FN_ACC1 = 0, 1, 2
const MFIELD = (P=0, U=1, UU=2, τ=3)
const SOLVER = (plain=0, mixed=1)

check(rc) = rc == 0 ? nothing :
    rc == -1 ? throw(AssertionError(unsafe_string(ccall((:wl_last_error, libwl), Cstring, ())))) :
    error("libwaterlily_hip error $rc: " * unsafe_string(ccall((:wl_last_error, libwl), Cstring, ())))

scheme_id(λ) = λ === WaterLily.quick ? 0 : λ === WaterLily.vanLeer ? 1 : λ === WaterLily.cds ? 2 :
    throw(ArgumentError("convective scheme $λ has no device kernel (quick, vanLeer, cds)"))

pad3(t, z) = ntuple(i -> i <= length(t) ? Float64(t[i]) : z, 3)

# ---------------------------------------------------------------------------------------------
# Flow
# ---------------------------------------------------------------------------------------------
mutable struct HipFlow{D,T} <: AbstractFlow{D,T}
    h::Ptr{Cvoid}
    N::NTuple{D,Int}
    Δt::Vector{T}
    ν::T
    uBC::Union{NTuple{D,Number},Function}
    g::Union{Function,Nothing}
    perdir::NTuple
    exitBC::Bool
    host::Flow{D,T}    # host-side Flow: measure! of arbitrary bodies, apply!/accelerate! of functions
end

"""
    HipFlow(N, uBC; Δt, ν, g, λ, T, perdir, exitBC, u0, uλ, mem)

`flow_ctor` for `Simulation`: same keywords as `Flow` (src/Flow.jl:133-134).  The initial velocity
is `u0` (a function `u0(i,x)` or a tuple; `uλ` is its deprecated alias), by default `uBC` at t = 0,
followed by `BC!` and `exitBC!` (src/Flow.jl:135-141).
"""
function HipFlow(N::NTuple{D}, uBC; Δt=0.25, ν=0., g=nothing, λ=WaterLily.quick, T=Float32,
                 perdir=(), exitBC=false, u0=nothing, uλ=nothing, mem=Array, kw...) where D
    u0 = WaterLily.ic_kwarg(u0, uλ)
    isnothing(g) || g isa Function || throw(ArgumentError("HipFlow: g must be a function g(i,x,t) or nothing"))
    Ut = uBC isa Function ? ntuple(_ -> 0.0, D) : uBC
    desc = WlFlowDesc(Int32(D), ntuple(i -> i <= D ? Int32(N[i]) : Int32(0), 3), T == Float64 ? 1 : 0,
                      scheme_id(λ), Int32(sum(1 << (j - 1) for j in perdir; init=0)), Int32(exitBC), 0, 0,
                      pad3(Ut, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), Float64(ν), Float64(Δt))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wl_flow_create, libwl), Cint, (Ptr{Ptr{Cvoid}}, Ref{WlFlowDesc}), h, desc))
    host = Flow(N, uBC; Δt, ν, g, λ, T, perdir, exitBC)
    f = HipFlow{D,T}(h[], Tuple(N), T[Δt], T(ν), uBC, g, perdir, exitBC, host)
    finalizer(f -> ccall((:wl_flow_destroy, libwl), Cint, (Ptr{Cvoid},), f.h), f)
    if uBC isa Function || !isnothing(u0)  # u = u0 (default uBC(i,x,0)), then BC! and exitBC!
        ubc0 = uBC isa Function ? field_of((i, x) -> uBC(i, x, zero(T)), f) : nothing
        ic = isnothing(u0) ? ubc0 : u0 isa Function ? field_of(u0, f) : field_of((i, x) -> u0[i], f)
        upload!(f, :u, ic)
        isnothing(ubc0) || set_field!(f, FN_UBC, ubc0)
        check(ccall((:wl_flow_apply_ic, libwl), Cint, (Ptr{Cvoid},), f.h))
    end
    f
end

function download(f::HipFlow{D,T}, name::Symbol) where {D,T}
    Ng = f.N .+ 2
    a = name in (:p, :σ) ? zeros(T, Ng) : name == :μ₁ ? zeros(T, Ng..., D, D) : zeros(T, Ng..., D)
    check(ccall((:wl_flow_download, libwl), Cint, (Ptr{Cvoid}, Cint, Ptr{T}, Csize_t),
                f.h, FIELD[name], a, sizeof(a)))
    a
end
function upload!(f::HipFlow{D,T}, name::Symbol, a::AbstractArray) where {D,T}
    b = convert(Array{T}, a)
    check(ccall((:wl_flow_upload, libwl), Cint, (Ptr{Cvoid}, Cint, Ptr{T}, Csize_t),
                f.h, FIELD[name], b, sizeof(b)))
end
Base.getproperty(f::HipFlow, s::Symbol) = s in keys(FIELD) ? download(f, s) : getfield(f, s)
Base.setproperty!(f::HipFlow, s::Symbol, v) = s in keys(FIELD) ? upload!(f, s, v) : setfield!(f, s, v)

# a vector field c[I,i] = fn(i, loc(i,I)) on the ghosted grid (apply!, src/Flow.jl:81)
function field_of(fn, f::HipFlow{D,T}) where {D,T}
    c = zeros(T, (f.N .+ 2)..., D)
    WaterLily.apply!(fn, c)
    c
end
set_field!(f::HipFlow{D,T}, which, a::Array{T}) where {D,T} =
    check(ccall((:wl_flow_set_field_fn, libwl), Cint, (Ptr{Cvoid}, Cint, Ptr{T}, Csize_t), f.h, which, a, sizeof(a)))

"`CFL(a)` (src/Flow.jl:234-237) of the device velocity; Δt_max is the reference's default 10."
function CFL(a::HipFlow; Δt_max=10)
    Δt_max == 10 || throw(ArgumentError("HipFlow: CFL is computed with Δt_max = 10"))
    dt = Ref{Float64}(0)
    check(ccall((:wl_flow_cfl, libwl), Cint, (Ptr{Cvoid}, Ptr{Float64}), a.h, dt))
    eltype(a.Δt)(dt[])
end

# ---------------------------------------------------------------------------------------------
# Bodies
# ---------------------------------------------------------------------------------------------
"""
    DeviceSphere(center, radius), DeviceEllipsoid(center, axes), a + b

Analytic bodies the device measures itself (src/Body.jl:28-51 with the distance functions
√sum(abs2,x-c)-r and (√sum(abs2,(x-c)./a)-1)*minimum(a), unions as SetBody(min)).  Their host
geometry is the equivalent `AutoBody`, so WaterLily's own functions see the same body.
"""
struct DeviceBody <: AbstractBody
    parts::Vector{WlBody}
    host::AbstractBody
end
sphere_sdf(c, r) = (x, t) -> √sum(abs2, x .- c) - r
ellipsoid_sdf(c, a) = (x, t) -> (√sum(abs2, (x .- c) ./ a) - 1) * minimum(a)
DeviceSphere(center, radius) = DeviceBody([WlBody(1, 0, pad3(center, 0.0), Float64(radius), (1.0, 1.0, 1.0))],
                                          AutoBody(sphere_sdf(collect(center), radius)))
DeviceEllipsoid(center, axes) = DeviceBody([WlBody(2, 0, pad3(center, 0.0), 0.0, pad3(axes, 1.0))],
                                           AutoBody(ellipsoid_sdf(collect(center), collect(axes))))
Base.:+(a::DeviceBody, b::DeviceBody) = DeviceBody(vcat(a.parts, b.parts), a.host + b.host)
measure(b::DeviceBody, x, t; kwargs...) = measure(b.host, x, t; kwargs...)

"""
    measure!(flow::HipFlow, body; t, ϵ)

`DeviceBody`: measured on the device.  Any other body: WaterLily's own `measure!`
(src/Body.jl:28-51) on the host `Flow`, then μ₀, μ₁, V and σ are uploaded (AD-differentiated
AutoBody closures included).
"""
function measure!(a::HipFlow, body::DeviceBody; t=zero(eltype(a.Δt)), ϵ=1)
    check(ccall((:wl_flow_measure, libwl), Cint, (Ptr{Cvoid}, Ptr{WlBody}, Cint, Float64, Float64),
                a.h, body.parts, length(body.parts), Float64(t), Float64(ϵ)))
end
function measure!(a::HipFlow, body::AbstractBody; t=zero(eltype(a.Δt)), ϵ=1)
    measure!(a.host, body; t, ϵ)
    for s in (:V, :μ₀, :μ₁, :σ)
        upload!(a, s, getfield(a.host, s))
    end
end

# pressure/viscous force and moment (src/Metrics.jl:126-197) of a DeviceBody: one device reduction
function metrics(a::HipFlow{D}, body::DeviceBody, x₀) where D
    Fp, Fv, Mp, Mv = zeros(D), zeros(D), zeros(D == 2 ? 1 : 3), zeros(D == 2 ? 1 : 3)
    x = Float64[pad3(x₀, 0.0)...]
    check(ccall((:wl_flow_metrics, libwl), Cint,
                (Ptr{Cvoid}, Ptr{WlBody}, Cint, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                a.h, body.parts, length(body.parts), Float64(WaterLily.time(a)), x, Fp, Fv, Mp, Mv))
    moment(M) = D == 2 ? fill(M[1], 2) : M   # the 2D cross product broadcast into both entries
    Fp, Fv, moment(Mp), moment(Mv)
end
pressure_force(a::HipFlow, body::DeviceBody) = metrics(a, body, zeros(3))[1]
viscous_force(a::HipFlow, body::DeviceBody) = metrics(a, body, zeros(3))[2]
pressure_moment(x₀, a::HipFlow, body::DeviceBody) = metrics(a, body, x₀)[3]
viscous_moment(x₀, a::HipFlow, body::DeviceBody) = metrics(a, body, x₀)[4]

# ---------------------------------------------------------------------------------------------
# Poisson and the step
# ---------------------------------------------------------------------------------------------
mutable struct HipPoisson{T} <: AbstractPoisson{T,Array{T},Array{T}}
    flow::HipFlow
    n::Vector{Int16}
end
"`pois_ctor` for `Simulation`: the MultiLevelPoisson(flow.p, flow.μ₀, flow.σ) of the device flow."
HipPoisson(flow::HipFlow{D,T}) where {D,T} = (p = HipPoisson{T}(flow, Int16[]); update!(p); p)
update!(p::HipPoisson) = check(ccall((:wl_flow_pois_update, libwl), Cint, (Ptr{Cvoid},), p.flow.h))

"""
    solver_mode!(pois::HipPoisson; mode=:plain, tol=1e-4, itmx=32)

`:plain` is `solver!(::MultiLevelPoisson)` as the reference runs it; `:mixed` wraps the FP32
V-cycles in an FP64 defect-correction loop to the same Σr² tolerance.
"""
solver_mode!(p::HipPoisson; mode=:plain, tol=1e-4, itmx=32) =
    check(ccall((:wl_flow_set_solver, libwl), Cint, (Ptr{Cvoid}, Cint, Float64, Cint), p.flow.h, SOLVER[mode], tol, itmx))

# the host-evaluated function fields of the coming step: uBC at t₁; g + ∂uBC/∂t at t₀ and t₁
function functions_for_step!(a::HipFlow{D,T}, t₀, t₁) where {D,T}
    a.uBC isa Function && set_field!(a, FN_UBC, field_of((i, x) -> a.uBC(i, x, T(t₁)), a))
    if a.g isa Function || a.uBC isa Function
        for (which, t) in ((FN_ACC0, t₀), (FN_ACC1, t₁))
            acc = zeros(T, (a.N .+ 2)..., D)
            WaterLily.accelerate!(acc, T(t), a.g, a.uBC)
            set_field!(a, which, acc)
        end
    end
end

# udf! through the C ABI: the library calls back (wl_udf_fn) between conv_diff! and accelerate!;
# the trampoline runs WaterLily's own udf! (src/Flow.jl:255-257) on the host Flow mirror, with the
# device fields copied in before and out after, and carries a Julia exception back out.
const UDF_CTX = Ref{Any}(nothing)
const UDF_FIELDS = (:u, :u⁰, :f, :p, :σ)
function udf_trampoline(user::Ptr{Cvoid}, h::Ptr{Cvoid}, phase::Cint, t::Float64)::Cint
    a, udf, kw = UDF_CTX[]
    try
        host = a.host
        for s in UDF_FIELDS
            copyto!(getfield(host, s), download(a, s))
        end
        WaterLily.udf!(host, udf, phase == 0 ? host.u⁰ : host.u, eltype(a.Δt)(t); kw...)
        for s in UDF_FIELDS
            upload!(a, s, getfield(host, s))
        end
        return Cint(0)
    catch e
        UDF_CTX[] = (a, udf, kw, e)
        return Cint(-1)
    end
end
const UDF_C = Ref{Ptr{Cvoid}}(C_NULL)
__init__() = (UDF_C[] = @cfunction(udf_trampoline, Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Float64)))

"""
    mom_step!(a::HipFlow, b::HipPoisson; udf=nothing, kw...)

src/Flow.jl:156-167 on the device: t₁ = sum(Δt), t₀ = t₁-Δt[end]; the CFL Δt is pushed onto
a.Δt and the two V-cycle counts onto b.n; `udf(flow, [u,] t; kw...)` runs in both phases.
"""
function mom_step!(a::HipFlow{D,T}, b::HipPoisson; udf=nothing, kw...) where {D,T}
    t₁ = sum(a.Δt); t₀ = t₁ - a.Δt[end]
    functions_for_step!(a, t₀, t₁)
    dt = Ref{Float64}(0); n = zeros(Int16, 2)
    isnothing(udf) || (UDF_CTX[] = (a, udf, kw);
                       check(ccall((:wl_flow_set_udf, libwl), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                                   a.h, UDF_C[], C_NULL)))
    rc = try
        ccall((:wl_mom_step, libwl), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Ptr{Float64}, Ptr{Int16}),
              a.h, a.Δt[end], t₀, t₁, dt, n)
    finally
        isnothing(udf) || ccall((:wl_flow_set_udf, libwl), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                                a.h, C_NULL, C_NULL)
    end
    ctx = UDF_CTX[]; UDF_CTX[] = nothing
    ctx isa Tuple && length(ctx) == 4 && throw(ctx[4])
    check(rc)
    push!(a.Δt, T(dt[])); append!(b.n, n)
end

WaterLily.time(a::HipFlow) = sum(@view(a.Δt[1:end-1]))

# ---------------------------------------------------------------------------------------------
# MeanFlow on the device (src/Metrics.jl:199-261)
# ---------------------------------------------------------------------------------------------
mutable struct HipMeanFlow{D,T}
    h::Ptr{Cvoid}
    flow::HipFlow{D,T}
    t::Vector{T}
    uu_stats::Bool
end
function HipMeanFlow(flow::HipFlow{D,T}; t_init=WaterLily.time(flow), uu_stats=false) where {D,T}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wl_meanflow_create, libwl), Cint, (Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Cint), h, flow.h, uu_stats))
    m = HipMeanFlow{D,T}(h[], flow, T[t_init], uu_stats)
    finalizer(m -> ccall((:wl_meanflow_destroy, libwl), Cint, (Ptr{Cvoid},), m.h), m)
    m
end
function mdownload(m::HipMeanFlow{D,T}, name::Symbol) where {D,T}
    Ng = m.flow.N .+ 2
    a = name == :P ? zeros(T, Ng) : name == :U ? zeros(T, Ng..., D) : zeros(T, Ng..., D, D)
    check(ccall((:wl_meanflow_download, libwl), Cint, (Ptr{Cvoid}, Cint, Ptr{T}, Csize_t), m.h, MFIELD[name], a, sizeof(a)))
    a
end
Base.getproperty(m::HipMeanFlow, s::Symbol) =
    s in (:P, :U) ? mdownload(m, s) : s == :UU ? (getfield(m, :uu_stats) ? mdownload(m, :UU) : nothing) : getfield(m, s)
WaterLily.time(m::HipMeanFlow) = m.t[end] - m.t[1]
function update!(m::HipMeanFlow{D,T}, flow::HipFlow) where {D,T}
    dt = WaterLily.time(flow) - m.t[end]
    ε = dt / (dt + WaterLily.time(m) + eps(T))
    length(m.t) == 1 && (ε = one(T))
    check(ccall((:wl_meanflow_update, libwl), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64), m.h, flow.h, Float64(ε)))
    push!(m.t, m.t[end] + dt)
end
function reset!(m::HipMeanFlow; t_init=0.0)
    check(ccall((:wl_meanflow_reset, libwl), Cint, (Ptr{Cvoid},), m.h))
    empty!(m.t); push!(m.t, t_init)
end
uu(m::HipMeanFlow) = mdownload(m, :τ)
Base.copy!(a::HipFlow, m::HipMeanFlow) = check(ccall((:wl_meanflow_copy_to_flow, libwl), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), m.h, a.h))

# ---------------------------------------------------------------------------------------------
# Batches: many simulations of different sizes, one launch per stage of the step for all of them
# ---------------------------------------------------------------------------------------------
mutable struct HipBatch
    h::Ptr{Cvoid}
    sims::Vector
end
function HipBatch(sims::AbstractVector)
    flows = Ptr{Cvoid}[s.flow.h for s in sims]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wl_batch_create, libwl), Cint, (Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint), h, flows, length(flows)))
    b = HipBatch(h[], collect(sims))
    finalizer(b -> ccall((:wl_batch_destroy, libwl), Cint, (Ptr{Cvoid},), b.h), b)
    b
end
"`sim_step!(b::HipBatch)`: one `sim_step!(sim; remeasure=false)` of every member."
function sim_step!(b::HipBatch)
    n = length(b.sims)
    dt = [Float64(s.flow.Δt[end]) for s in b.sims]
    t1 = [Float64(sum(s.flow.Δt)) for s in b.sims]
    t0 = [Float64(sum(s.flow.Δt) - s.flow.Δt[end]) for s in b.sims]
    out = zeros(n); nit = zeros(Int16, 2n)
    check(ccall((:wl_batch_step, libwl), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int16}),
                b.h, dt, t0, t1, out, nit))
    for (k, s) in enumerate(b.sims)
        push!(s.flow.Δt, eltype(s.flow.Δt)(out[k]))
        append!(s.pois.n, nit[2k-1:2k])
    end
end

export HipFlow, HipPoisson, DeviceSphere, DeviceEllipsoid, HipMeanFlow, HipBatch, solver_mode!
end # module
