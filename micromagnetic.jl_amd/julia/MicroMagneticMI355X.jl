# MicroMagneticMI355X.jl -- routes MicroMagnetic.jl's finite-difference hot path
# (effective field, LLG right-hand side, RK4 / Heun / Dormand-Prince / BS23
# stepping) through libmmx, the HIP library for MI355X (gfx950), via ccall.
#
# Load it after MicroMagnetic:
#
#     using MicroMagnetic
#     include("MicroMagneticMI355X.jl"); using .MicroMagneticMI355X
#     MicroMagneticMI355X.attach!(sim)        # after the add_* calls
#
# From then on the reference's own entry points -- effective_field(sim, spin, t),
# llg_call_back, advance_step(sim, integrator), run_until, run_sim, relax -- run
# on the GPU; sim.spin is kept in sync at the boundary.  The C ABI it calls is
# declared in include/mmx.h, where each entry point names the reference method
# it replaces.
module MicroMagneticMI355X

using MicroMagnetic
import MicroMagnetic: effective_field, advance_step, llg_call_back

const libmmx = get(ENV, "MMX_LIBRARY", joinpath(@__DIR__, "..", "libmmx.so"))

struct MeshDesc
    nx::Int32
    ny::Int32
    nz::Int32
    pbc::Int32
    dx::Float64
    dy::Float64
    dz::Float64
    precision::Int32
    device::Int32
end

const contexts = IdDict{Any,Ptr{Cvoid}}()

function check(rc::Integer)
    if rc < 0
        msg = unsafe_string(ccall((:mmx_last_error, libmmx), Cstring, ()))
        rc == -3 && (@warn msg; return false)   # NaN error norm (adaptive_rk.jl:112-117)
        rc == -4 && (@warn msg; return false)   # too many tries (adaptive_rk.jl:119-122)
        error("libmmx: " * msg)
    end
    return true
end

host(a) = Array{Float64}(Array(a))

"""
    attach!(sim; device=0)

Create the GPU context for `sim` (an FD `MicroSim`) and copy its mesh, Ms, pins,
m and interaction terms into it, in insertion order.
"""
function attach!(sim::MicroMagnetic.MicroSim; device::Integer=0)
    mesh = sim.mesh
    pbc = Int32(mesh.xperiodic) | (Int32(mesh.yperiodic) << 1) | (Int32(mesh.zperiodic) << 2)
    prec = MicroMagnetic.Float[] == Float32 ? Int32(32) : Int32(64)
    desc = Ref(MeshDesc(mesh.nx, mesh.ny, mesh.nz, pbc, mesh.dx, mesh.dy, mesh.dz, prec, device))
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:mmx_create, libmmx), Cint, (Ref{Ptr{Cvoid}}, Ref{MeshDesc}), ctx, desc))
    h = ctx[]
    contexts[sim] = h
    Ms = host(sim.mu0_Ms) ./ MicroMagnetic.mu_0
    check(ccall((:mmx_set_Ms, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, Ms))
    pins = UInt8.(Array(sim.pins))
    check(ccall((:mmx_set_pinning, libmmx), Cint, (Ptr{Cvoid}, Ptr{UInt8}), h, pins))
    check(ccall((:mmx_init_m0, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), h, host(sim.spin), 0))
    for term in sim.interactions
        add_term!(h, term)
    end
    drv = sim.driver
    if drv isa MicroMagnetic.SpatialLLG  # per-cell damping, src/llg/llg.jl:62-97
        check(ccall((:mmx_set_llg, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cint), h,
                    0.1, Float64(drv.gamma), Cint(drv.precession)))
        check(ccall((:mmx_set_alpha, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, host(drv.alpha)))
    else
        check(ccall((:mmx_set_llg, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cint), h,
                    Float64(drv.alpha), Float64(drv.gamma), Cint(drv.precession)))
    end
    set_integrator!(sim)
    return h
end

function add_term!(h, t::MicroMagnetic.UniformExchange)
    ccall((:mmx_add_exch, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, C_NULL,
          t.Ax, t.Ay, t.Az) |> check
end
function add_term!(h, t::MicroMagnetic.SpatialExchange)
    ccall((:mmx_add_exch, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, host(t.A),
          0, 0, 0) |> check
end
function add_term!(h, t::MicroMagnetic.BulkDMI)
    ccall((:mmx_add_dmi, libmmx), Cint,
          (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, 0, C_NULL,
          C_NULL, C_NULL, t.Dx, t.Dy, t.Dz) |> check
end
function add_term!(h, t::MicroMagnetic.SpatialVectorBulkDMI)
    ccall((:mmx_add_dmi, libmmx), Cint,
          (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, 0,
          host(t.Dx), host(t.Dy), host(t.Dz), 0, 0, 0) |> check
end
function add_term!(h, t::MicroMagnetic.TimeBulkDMI)
    # D scaled by time_D(t) at every evaluation time: src/micro/field.jl:145-163
    idx = ccall((:mmx_add_dmi, libmmx), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, 0,
                C_NULL, C_NULL, C_NULL, t.Dx, t.Dy, t.Dz)
    check(idx)
    set_time_fn!(h, idx, t.time_D)
    return idx
end
function add_term!(h, t::MicroMagnetic.TimeSpatialBulkDMI)
    # one per-cell D for all three components: src/micro/field.jl:180-193
    D = host(t.D)
    idx = ccall((:mmx_add_dmi, libmmx), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, 0,
                D, D, D, 0, 0, 0)
    check(idx)
    set_time_fn!(h, idx, t.time_D)
    return idx
end
function add_term!(h, t::MicroMagnetic.InterfacialDMI)
    ccall((:mmx_add_dmi, libmmx), Cint,
          (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, 1,
          host(t.D), C_NULL, C_NULL, 0, 0, 0) |> check
end
function add_term!(h, t::MicroMagnetic.Anisotropy)
    ccall((:mmx_add_anis, libmmx), Cint,
          (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, host(t.Ku), 0, C_NULL,
          t.axis[1], t.axis[2], t.axis[3]) |> check
end
function add_term!(h, t::MicroMagnetic.SpatialAnisotropy)
    ccall((:mmx_add_anis, libmmx), Cint,
          (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h, host(t.Ku), 0,
          host(t.axes), 0, 0, 0) |> check
end
function add_term!(h, t::MicroMagnetic.CubicAnisotropy)
    # the reference stores axis3 = axis1 x axis2 of the inputs; libmmx rebuilds it the same way
    a1 = collect(Float64, t.axis1)
    a2 = collect(Float64, t.axis2)
    ccall((:mmx_add_cubic_anis, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Ptr{Float64}, Ptr{Float64}),
          h, host(t.Kc), 0, a1, a2) |> check
end
function add_term!(h, t::MicroMagnetic.InterlayerExchange)
    ccall((:mmx_add_exch_int, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint), h, host(t.Js), t.k1,
          t.k2) |> check
end
function add_term!(h, t::MicroMagnetic.InterlayerDMI)
    ccall((:mmx_add_dmi_int, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Cint), h, t.Dx, t.Dy,
          t.Dz, t.k1, t.k2) |> check
end
function add_term!(h, t::MicroMagnetic.SAHETorqueField)
    # the struct keeps only the products sigma x a1, while libmmx takes the inputs of
    # add_sahe_torque (add_field.jl:892-936)
    error("attach the spin anomalous Hall torque with attach_sahe!(sim, sigma_s, sigma_sa1, " *
          "sigma_sa2, a1, a2, a3, beta)")
end
"""
    attach_sahe!(sim, sigma_s, sigma_sa1, sigma_sa2, a1, a2, a3, beta)

add_sahe_torque(sim, sigma_s; ...) on the MI355X path: 3-tuples or 3N AoS vectors.
"""
function attach_sahe!(sim, sigma_s, sigma_sa1, sigma_sa2, a1, a2, a3, beta)
    h = contexts[sim]
    n = sim.n_total
    vec(v) = v isa Tuple ? repeat(collect(Float64, v), n) : collect(Float64, v)
    ccall((:mmx_add_sahe_torque, libmmx), Cint,
          (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
           Cdouble), h, vec(sigma_s), vec(sigma_sa1), vec(sigma_sa2), vec(a1), vec(a2), vec(a3), beta) |> check
end
function add_term!(h, t::MicroMagnetic.TorqueField)
    # torque_fun(y, m, t) on host arrays wrapping libmmx's buffers (3N AoS each)
    fun = t.torque_fun
    n = length(t.field)  # 3N
    f = (s::Cdouble, m::Ptr{Cdouble}, y::Ptr{Cdouble}, user::Ptr{Cvoid}) -> begin
        fun(unsafe_wrap(Array, y, n), unsafe_wrap(Array, m, n), s)
        nothing
    end
    cf = @cfunction($f, Cvoid, (Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cvoid}))
    idx = ccall((:mmx_add_torque, libmmx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), h, cf, C_NULL)
    check(idx)
    timefns[(h, idx)] = cf
    return idx
end
function add_term!(h, t::MicroMagnetic.Magnetoelastic)
    # field_data: the 6N Voigt values, cell by cell (add_field.jl:1413-1435)
    model = t.model == :tensor ? Cint(0) : Cint(1)
    ccall((:mmx_add_magnetoelastic, libmmx), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Cdouble, Cdouble, Cdouble),
          h, model, host(t.field_data), t.lambda_s, t.B1, t.B2) |> check
end
function add_term!(h, t::MicroMagnetic.TwinMonoclinicAnisotropy)
    K = Float64[t.Ka, t.Kb, t.Kaa, t.Kbb, t.Kab, t.Ku]
    ccall((:mmx_add_twin_mono_anis, libmmx), Cint,
          (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), h, K, host(t.axis_a),
          host(t.axis_b), host(t.axis_u111)) |> check
end
function add_term!(h, t::MicroMagnetic.HexagonalAnisotropy)
    ccall((:mmx_add_hex_anis, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble), h, t.K1, t.K2,
          t.K3) |> check
end
function add_term!(h, t::MicroMagnetic.Zeeman)
    ccall((:mmx_add_zeeman, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble), h,
          host(t.field), 0, 0, 0) |> check
end
function add_term!(h, t::MicroMagnetic.TimeZeeman)
    # time_fun(t) returns one factor or (fx, fy, fz): src/micro/field.jl:16-26
    fun, scalar = t.time_fun, t.is_scalar
    f = (s::Cdouble, out::Ptr{Cdouble}, user::Ptr{Cvoid}) -> begin
        v = fun(s)
        fx, fy, fz = scalar ? (v, v, v) : v
        unsafe_store!(out, fx, 1)
        unsafe_store!(out, fy, 2)
        unsafe_store!(out, fz, 3)
        nothing
    end
    cf = @cfunction($f, Cvoid, (Cdouble, Ptr{Cdouble}, Ptr{Cvoid}))
    idx = ccall((:mmx_add_zeeman_time, libmmx), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}), h,
                host(t.init_field), 0, 0, 0, cf, C_NULL)
    check(idx)
    timefns[(h, idx)] = cf
    return idx
end
function add_term!(h, t::MicroMagnetic.Demag)
    error("attach! needs the macro-PBC image counts of the demag term: call attach_demag!(sim, Nx, Ny, Nz)")
end
function add_term!(h, t::MicroMagnetic.DFTorqueField)
    ccall((:mmx_add_sot, libmmx), Cint,
          (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble), h, host(t.aj), 0, t.bj,
          t.px, t.py, t.pz) |> check
end
function add_term!(h, t::MicroMagnetic.SlonczewskiTorque)
    idx = ccall((:mmx_add_stt_slonczewski, libmmx), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble),
                h, host(t.J), 0, t.beta, t.Lambda, t.xi, t.P, t.px, t.py, t.pz)
    check(idx)  # the return code is the term index
    set_time_fn!(h, idx, t.ufun)
    return idx
end
function add_term!(h, t::MicroMagnetic.ZhangLiTorque)
    # the struct already holds bJ = b*J (with the 1/(1+xi^2) factor): pass it as J with b = 1
    idx = ccall((:mmx_add_stt_zhangli, libmmx), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cdouble, Ptr{Float64}, Cdouble, Cint),
                h, host(t.bJ), 0, 0, 0, 1.0, host(t.xi), 0, 0)
    check(idx)
    set_time_fn!(h, idx, t.ufun)
    return idx
end
add_term!(h, t) = error("term $(typeof(t)) is not on the MI355X path")

# ufun(t) of a torque term, called by libmmx at every evaluation time; the
# C-callable closures are kept alive here for as long as the context lives
const timefns = IdDict{Any,Any}()
function set_time_fn!(h, idx, ufun)
    f = (t::Cdouble, user::Ptr{Cvoid}) -> Cdouble(ufun(t))
    cf = @cfunction($f, Cdouble, (Cdouble, Ptr{Cvoid}))
    timefns[(h, idx)] = cf
    ccall((:mmx_set_term_time_fn, libmmx), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}), h, idx, cf,
          C_NULL) |> check
end

"""
    set_integrator!(sim)

Mirror create_driver (src/driver.jl:126-145): RungeKutta/Heun with their step,
DormandPrince/BS23 with their tolerance.
"""
function set_integrator!(sim)
    h = contexts[sim]
    I = sim.driver.integrator
    kind, step, tol = if I isa MicroMagnetic.RungeKutta
        0, I.step, 0.0
    elseif I isa MicroMagnetic.ModifiedEuler
        1, I.step, 0.0
    elseif I isa MicroMagnetic.AdaptiveRK && I.method isa MicroMagnetic.DOPRI54Method
        2, 0.0, I.tol
    elseif I isa MicroMagnetic.AdaptiveRK && I.method isa MicroMagnetic.BS23Method
        3, 0.0, I.tol
    else
        error("integrator $(typeof(I)) is not on the MI355X path")
    end
    check(ccall((:mmx_set_integrator, libmmx), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble), h, kind, step, tol))
end

function pull_spin!(sim)
    m = Vector{Float64}(undef, 3 * sim.n_total)
    check(ccall((:mmx_get_m, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}), contexts[sim], m))
    copyto!(sim.spin, m)
end

function push_spin!(sim, spin)
    check(ccall((:mmx_init_m0, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), contexts[sim], host(spin), 0))
end

# effective_field(sim, spin, t): src/micro/field.jl:496-505
function effective_field(sim::MicroMagnetic.MicroSim, spin, t::Float64=0.0)
    haskey(contexts, sim) || return invoke(effective_field, Tuple{MicroMagnetic.AbstractSim,Any,Float64}, sim, spin, t)
    push_spin!(sim, spin)
    H = Vector{Float64}(undef, 3 * sim.n_total)
    check(ccall((:mmx_effective_field, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Float64}), contexts[sim], t, H))
    copyto!(sim.field, H)
    return nothing
end

# llg_call_back(sim, dm_dt, spin, t): src/llg/llg.jl:45-58
function llg_call_back(sim::MicroMagnetic.MicroSim, dm_dt::AbstractArray{T,1}, spin::AbstractArray{T,1},
                       t::Float64) where {T<:AbstractFloat}
    haskey(contexts, sim) || return invoke(llg_call_back, Tuple{MicroMagnetic.AbstractSim,AbstractArray{T,1},
                                                                AbstractArray{T,1},Float64}, sim, dm_dt, spin, t)
    out = Vector{Float64}(undef, 3 * sim.n_total)
    check(ccall((:mmx_llg_rhs, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Float64}, Ptr{Float64}), contexts[sim], t,
                host(spin), out))
    copyto!(dm_dt, out)
    return nothing
end

# advance_step for the fixed-step and adaptive integrators (rk.jl:23, heun.jl:28, adaptive_rk.jl:93)
for IT in (:RungeKutta, :ModifiedEuler, :AdaptiveRK)
    @eval function advance_step(sim::MicroMagnetic.MicroSim, integrator::MicroMagnetic.$IT)
        haskey(contexts, sim) || return invoke(advance_step, Tuple{MicroMagnetic.AbstractSim,MicroMagnetic.$IT},
                                               sim, integrator)
        ok = check(ccall((:mmx_advance_step, libmmx), Cint, (Ptr{Cvoid}, Int64), contexts[sim], 1))
        t = Ref{Float64}(0.0)
        step = Ref{Float64}(0.0)
        step_next = Ref{Float64}(0.0)
        nsteps = Ref{Int64}(0)
        nfe = Ref{Int64}(0)
        ccall((:mmx_get_state, libmmx), Cint,
              (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Float64}, Ref{Int64}, Ref{Int64}), contexts[sim], t,
              step, step_next, nsteps, nfe)
        integrator.t = t[]
        integrator.nsteps = nsteps[]
        if integrator isa MicroMagnetic.AdaptiveRK
            integrator.step = step[]
            integrator.step_next = step_next[]
            integrator.nfevals = nfe[]
        end
        copyto!(sim.prespin, sim.spin)
        pull_spin!(sim)
        integrator isa MicroMagnetic.AdaptiveRK && copyto!(view(integrator.y_current, 1:(3 * sim.n_total)), sim.spin)
        return ok
    end
end

"""
    attach_demag!(sim, Nx, Ny, Nz)

The reference's Demag struct does not keep its image counts; pass the values
given to add_demag(sim; Nx, Ny, Nz) here (in insertion order with the other
terms: call attach! with the demag term removed, then this, or build the
context term by term).
"""
function attach_demag!(sim, Nx::Integer=0, Ny::Integer=0, Nz::Integer=0)
    check(ccall((:mmx_add_demag, libmmx), Cint, (Ptr{Cvoid}, Cint, Cint, Cint), contexts[sim], Nx, Ny, Nz))
end

function detach!(sim)
    h = pop!(contexts, sim, C_NULL)
    h == C_NULL || ccall((:mmx_destroy, libmmx), Cint, (Ptr{Cvoid},), h)
    return nothing
end

# ---- differentiation form: a ChainRules rrule over the RK4 trajectory ------------
"""
    evolve_rk4(sim, nsteps; every=50) -> m(T)

Advance an attached simulation `nsteps` RK4 steps on the GPU.  With
ChainRulesCore loaded, `rrule(evolve_rk4, sim, nsteps)` returns m(T) and a
pullback mapping dL/dm(T) to (dL/dm0, per-term parameter gradients) through
the hand-written adjoint kernels (mmx_adjoint_record / _backward).
"""
function evolve_rk4(sim, nsteps::Integer; every::Integer=50)
    h = contexts[sim]
    check(ccall((:mmx_adjoint_record, libmmx), Cint, (Ptr{Cvoid}, Int64, Int64), h, nsteps, every))
    pull_spin!(sim)
    return copy(sim.spin)
end

function evolve_rk4_pullback(sim, dL_dmT::AbstractVector)
    h = contexts[sim]
    n = sim.n_total
    g0 = Vector{Float64}(undef, 3n)
    check(ccall((:mmx_adjoint_backward, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), h,
                Vector{Float64}(dL_dmT), g0))
    grads = Dict{Any,Vector{Float64}}()
    nterms = ccall((:mmx_num_terms, libmmx), Cint, (Ptr{Cvoid},), h)
    for t in 0:(nterms - 1)
        # bulk DMI terms give (dL/dDx, dL/dDy, dL/dDz) per cell (3n values), the others n
        g = Vector{Float64}(undef, 3n)
        rc = ccall((:mmx_adjoint_param_grad, libmmx), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), h, t, g)
        term = sim.interactions[t + 1]
        bulk = term isa MicroMagnetic.BulkDMI || term isa MicroMagnetic.SpatialVectorBulkDMI
        rc >= 0 && (grads[t + 1] = bulk ? g : g[1:n])
    end
    gms = Vector{Float64}(undef, n)
    check(ccall((:mmx_adjoint_ms_grad, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}), h, gms))
    grads[:Ms] = gms
    return g0, grads
end

"""
    llg_rhs_jvp(sim, v; spin=nothing, t=0.0), llg_rhs_vjp(sim, a; spin=nothing, t=0.0)

The Jacobian of one llg_call_back evaluation at `spin` (the current state when
nothing) applied to v, or transposed and applied to a (3N vectors).
"""
function llg_rhs_jvp(sim, v::AbstractVector; spin=nothing, t=0.0)
    out = Vector{Float64}(undef, 3 * sim.n_total)
    m = spin === nothing ? Ptr{Float64}(C_NULL) : Vector{Float64}(spin)
    check(ccall((:mmx_llg_rhs_jvp, libmmx), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                contexts[sim], t, m, Vector{Float64}(v), out))
    return out
end

function llg_rhs_vjp(sim, a::AbstractVector; spin=nothing, t=0.0)
    out = Vector{Float64}(undef, 3 * sim.n_total)
    m = spin === nothing ? Ptr{Float64}(C_NULL) : Vector{Float64}(spin)
    check(ccall((:mmx_llg_rhs_vjp, libmmx), Cint, (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                contexts[sim], t, m, Vector{Float64}(a), out))
    return out
end

"""
    dynamic_matrix_gpu(sim; gamma=2.21e5)

ext/EnzymeExt.jl's dynamic_matrix from the library's JVP: the Jacobian of
gamma m x H_eff(m) column by column (the JVP of the LLG right-hand side with
alpha = 0 is -gamma m x h on the transverse directions), rotated into each
cell's frame, transverse 2x2 blocks (2N x 2N).
"""
function dynamic_matrix_gpu(sim; gamma=2.21e5)
    h = contexts[sim]
    n = sim.n_total
    pull_spin!(sim)
    m0 = Array(sim.spin)
    check(ccall((:mmx_set_llg, libmmx), Cint, (Ptr{Cvoid}, Float64, Float64, Cint), h, 0.0, gamma, 1))
    A = zeros(3n, 3n)
    try
        e = zeros(3n)
        for i in 1:(3n)
            e .= 0
            e[i] = 1
            A[:, i] .= .-llg_rhs_jvp(sim, e; spin=m0)
        end
    finally  # the driver's own alpha, gamma and precession again
        drv = sim.driver
        check(ccall((:mmx_set_llg, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cint), h,
                    drv isa MicroMagnetic.SpatialLLG ? 0.1 : Float64(drv.alpha), Float64(drv.gamma),
                    Cint(drv.precession)))
    end
    C = zeros(2n, 2n)
    Rs = [MicroMagnetic.rotation_matrix(m0[3i-2], m0[3i-1], m0[3i]) for i in 1:n]
    for i in 1:n, j in 1:n
        B = Rs[i]' * A[3i-2:3i, 3j-2:3j] * Rs[j]
        C[2i-1:2i, 2j-1:2j] .= B[1:2, 1:2]
    end
    return C
end

if isdefined(Main, :ChainRulesCore)
    import ChainRulesCore
    function ChainRulesCore.rrule(::typeof(evolve_rk4), sim, nsteps::Integer; every::Integer=50)
        mT = evolve_rk4(sim, nsteps; every=every)
        function pullback(ΔmT)
            g0, grads = evolve_rk4_pullback(sim, ChainRulesCore.unthunk(ΔmT))
            return (ChainRulesCore.NoTangent(), ChainRulesCore.Tangent{typeof(sim)}(; spin=g0, params=grads),
                    ChainRulesCore.NoTangent())
        end
        return mT, pullback
    end
end

# ---- batching form ------------------------------------------------------------------
"""
    batch_advance!(sims, nsteps)

Advance several attached simulations (all RungeKutta or all Heun) together:
one launch per stage over those without demag, one launch per demag pass and
stage for each group of those with demag that share a geometry, replayed
from a HIP graph.  Each simulation ends where `run_sim` stepping it alone
would leave it.
"""
function batch_advance!(sims::AbstractVector, nsteps::Integer)
    hs = Ptr{Cvoid}[contexts[s] for s in sims]
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:mmx_batch_create, libmmx), Cint, (Ref{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint), b, hs, length(hs)))
    try
        check(ccall((:mmx_batch_advance, libmmx), Cint, (Ptr{Cvoid}, Int64), b[], nsteps))
    finally
        ccall((:mmx_batch_destroy, libmmx), Cint, (Ptr{Cvoid},), b[])
    end
    foreach(pull_spin!, sims)
    return nothing
end

# ---- nudged elastic band (src/neb/neb.jl) ----------------------------------------------
"""
    neb_relax!(neb; max_steps, stopping_dmdt)

Run relax(neb) for a reference `NEB` built with the LLG driver: its images
(image_l, spin, image_r) go to the GPU band, the band relaxes there, and
neb.spin, neb.energy_cpu and neb.distance are copied back.
"""
function neb_relax!(neb; max_steps=10000, stopping_dmdt=0.01)
    sim = neb.sim
    haskey(contexts, sim) || attach!(sim)
    n, dof = neb.n_images, neb.dof
    imgs = vcat(host(neb.image_l), host(neb.spin), host(neb.image_r))
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:mmx_neb_create, libmmx), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Float64}, Cint, Cdouble, Cint),
                b, contexts[sim], imgs, n, neb.spring_constant, neb.clib_image))
    try
        d = neb.driver
        check(ccall((:mmx_neb_set_driver, libmmx), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cint, Cdouble), b[],
                    d.alpha, d.gamma, Cint(d.precession), d.integrator.tol))
        steps = Ref{Int64}(0)
        md = Ref{Float64}(0)
        check(ccall((:mmx_neb_relax, libmmx), Cint, (Ptr{Cvoid}, Int64, Cdouble, Ref{Int64}, Ref{Float64}), b[],
                    max_steps, stopping_dmdt, steps, md))
        out = Vector{Float64}(undef, n * dof)
        E = Vector{Float64}(undef, n + 2)
        dist = Vector{Float64}(undef, n + 1)
        check(ccall((:mmx_neb_get, libmmx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    b[], out, E, dist, C_NULL))
        copyto!(neb.spin, out)
        neb.energy_cpu .= E
        neb.distance .= dist
    finally
        ccall((:mmx_neb_destroy, libmmx), Cint, (Ptr{Cvoid},), b[])
    end
    return nothing
end

end # module
