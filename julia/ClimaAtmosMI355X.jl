"""
    ClimaAtmosMI355X

Julia side of `libcadycore` (the C ABI in `include/cadycore.h`): the hooks
ClimaTimeSteppers calls on the ClimaAtmos ODE function, forwarded to the
MI355X library, and a drop-in for `args_integrator`
(`src/simulation/integrator.jl:183-236` of ClimaAtmos) that builds the same
`CTS.ClimaODEFunction` with native-backed functions.

Two ways to run, chosen by where ClimaCore keeps `Y`:

  * `:hooks` -- `Y` (and every stage FieldVector ClimaTimeSteppers allocates)
    lives in device memory, i.e. ClimaCore's parent arrays are `ROCArray`s.
    Each hook passes `pointer(parent(Y.c))`, `pointer(parent(Y.f))` and the
    strides of their DataLayout as a `ca_fieldvector`; ClimaTimeSteppers owns
    the ARS343 stage loop, exactly as in the reference.
  * `:fused` -- `Y` stays a CPU FieldVector (`ClimaComms.CPUSingleThreaded`,
    the only non-CUDA device ClimaComms offers today).  The state is uploaded
    once into a native device buffer, `ca_step` runs whole steps on it, and
    the CPU FieldVector is refreshed (`ca_unpack_state` + copy) only when a
    callback needs it.  `solve_atmos!` keeps its loop: the integrator's step
    is replaced, not its callbacks.

Device memory and streams come from AMDGPU.jl.  Nothing here is exercised by
the repository's tests (there is no Julia in its build image); the Python
mirror `climaatmos.jl_amd/cadycore/lib.py` binds the same symbols and is.
"""
module ClimaAtmosMI355X

import LinearAlgebra
import AMDGPU
import ClimaCore: Fields, DataLayouts
import ClimaTimeSteppers as CTS
import ClimaAtmos as CA

const lib = get(ENV, "CADYCORE_LIB", "libcadycore.so")

# ------------------------------------------------------------------ structs
# include/cadycore.h, field for field
struct CaGeometry
    nelem::Int32; ncol::Int32; nlev::Int32; radius::Float64
    z_face::Ptr{Float64}; J_h::Ptr{Float64}; gup::Ptr{Float64}; gdn::Ptr{Float64}
    f3_num::Ptr{Float64}; f12::Ptr{Float64}; basis::Ptr{Float64}
    n_dss::Int32; dss_ptr::Ptr{Int32}; dss_col::Ptr{Int32}; dss_w::Ptr{Float64}
    gll_w::Ptr{Float64}; gll_D::Ptr{Float64}
end
struct CaParams
    R_d::Float64; cp_d::Float64; cv_d::Float64; T_0::Float64; p_ref_theta::Float64
    grav::Float64; T_surf_ref::Float64; T_min_ref::Float64; T_min_sgs::Float64
end
struct CaConfig
    hyperdiff::Int32; nu4_vort::Float64; nu4_scal::Float64; div_damp::Float64
    upwinding::Int32; rayleigh::Int32; zd_rayleigh::Float64; alpha_uh::Float64; alpha_w::Float64
    alpha_w_col::Ptr{Float64}; vert_diff::Int32; D0_col::Ptr{Float64}; dt::Float64
    sfc_flux::Int32; Cd::Float64; Ch::Float64; gustiness::Float64; T_sfc_col::Ptr{Float64}
    implicit_diffusion::Int32; solve_iters::Int32
    vert_diff_model::Int32; vert_diff_H::Float64; vert_diff_C_E::Float64
end
struct CaLayout
    offset::Int64; s_v::Int64; s_i::Int64; s_j::Int64; s_h::Int64
end
struct CaFieldVector
    c::Ptr{Float32}; f::Ptr{Float32}
    lay::NTuple{5, CaLayout}
end

# ------------------------------------------------------------------- errors
struct CaError <: Exception
    status::Cint
    msg::String
end
Base.showerror(io::IO, e::CaError) = print(io, "cadycore (status ", e.status, "): ", e.msg)

last_error(h) = unsafe_string(ccall((:ca_last_error, lib), Cstring, (Ptr{Cvoid},), h))
check(h, s::Cint) = s == 0 ? nothing : throw(CaError(s, last_error(h)))

stream() = AMDGPU.stream().stream   # the task-local HIP stream

# ------------------------------------------------------------------- handle
"""
    NativeDycore(geometry, params, config)

Owns one `ca_handle`.  `geometry`, `params` and `config` are filled from
ClimaCore's `LocalGeometry` (per stored column 16 h + 4 j + i: J_h, g^ij, g_ij
and the Cartesian basis at one level divided by the deep-atmosphere scale
s(z)^2, the Coriolis components, the shared-node groups with weights WJ / sum
WJ, and the face heights of one column), the ClimaParams set and the
`AtmosModel`; the arrays they point into are kept alive in `keep`.
"""
mutable struct NativeDycore
    h::Ptr{Cvoid}
    keep::Any
    function NativeDycore(g::CaGeometry, p::CaParams, c::CaConfig, keep)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        s = ccall((:ca_create, lib), Cint,
                  (Ref{CaGeometry}, Ref{CaParams}, Ref{CaConfig}, Ref{Ptr{Cvoid}}), g, p, c, out)
        s == 0 || throw(CaError(s, last_error(C_NULL)))
        d = new(out[], keep)
        finalizer(d -> ccall((:ca_destroy, lib), Cint, (Ptr{Cvoid},), d.h), d)
        return d
    end
end
state_size(d::NativeDycore) = ccall((:ca_state_size, lib), Int64, (Ptr{Cvoid},), d.h)

# ----------------------------------------------------- FieldVector layouts
"""
    layouts(Y) -> NTuple{5, CaLayout}

Offsets and strides (in Float32 elements) of ρ, uₕ₁, uₕ₂, ρe_tot in
`parent(Y.c)` and of u₃ in `parent(Y.f)`, read from the DataLayout types at
run time (VIJFH and VIJHF; Nq = 4).
"""
function layouts(Y)
    vc = Fields.field_values(Y.c)
    vf = Fields.field_values(Y.f)
    return (ntuple(f -> component_layout(vc, f), 4)..., component_layout(vf, 1))
end
function component_layout(v::DataLayouts.VIJFH, f)
    Nv, Ni, Nj, F, Nh = size(parent(v))
    return CaLayout((f - 1) * Nv * Ni * Nj, 1, Nv, Nv * Ni, Nv * Ni * Nj * F)
end
function component_layout(v::DataLayouts.VIJHF, f)
    Nv, Ni, Nj, Nh, F = size(parent(v))
    return CaLayout((f - 1) * Nv * Ni * Nj * Nh, 1, Nv, Nv * Ni, Nv * Ni * Nj)
end
function component_layout(v::DataLayouts.VF, f)   # a single column
    Nv, F = size(parent(v))
    return CaLayout((f - 1) * Nv, 1, 0, 0, 0)
end

fieldvector(Y) = CaFieldVector(pointer(parent(Y.c)), pointer(parent(Y.f)), layouts(Y))

# -------------------------------------------------------------- the hooks
# Same arguments, order and meaning as the reference hooks; `p` carries the
# NativeDycore in `p.native` (set by `args_integrator`).
function remaining_tendency!(Yₜ, Yₜ_lim, Y, p, t)
    a, b, c = Ref(fieldvector(Yₜ)), Ref(fieldvector(Yₜ_lim)), Ref(fieldvector(Y))
    check(p.native.h, ccall((:ca_remaining_tendency_fv, lib), Cint,
        (Ptr{Cvoid}, Ref{CaFieldVector}, Ref{CaFieldVector}, Ref{CaFieldVector}, Float64, Ptr{Cvoid}),
        p.native.h, a, b, c, Float64(t), stream()))
    return Yₜ
end
function implicit_tendency!(Yₜ, Y, p, t)
    check(p.native.h, ccall((:ca_implicit_tendency_fv, lib), Cint,
        (Ptr{Cvoid}, Ref{CaFieldVector}, Ref{CaFieldVector}, Float64, Ptr{Cvoid}),
        p.native.h, Ref(fieldvector(Yₜ)), Ref(fieldvector(Y)), Float64(t), stream()))
end
function correct_implicit_advection_tendency!(Yₜ, Y, p, t)
    check(p.native.h, ccall((:ca_correct_implicit_advection_fv, lib), Cint,
        (Ptr{Cvoid}, Ref{CaFieldVector}, Ref{CaFieldVector}, Float64, Ptr{Cvoid}),
        p.native.h, Ref(fieldvector(Yₜ)), Ref(fieldvector(Y)), Float64(t), stream()))
end
for (jl, c) in ((:set_precomputed_quantities!, :ca_set_precomputed_quantities_fv),
                (:set_implicit_precomputed_quantities!, :ca_set_implicit_precomputed_quantities_fv),
                (:dss!, :ca_dss_fv))
    @eval $jl(Y, p, t) = check(p.native.h, ccall(($(QuoteNode(c)), lib), Cint,
        (Ptr{Cvoid}, Ref{CaFieldVector}, Float64, Ptr{Cvoid}),
        p.native.h, Ref(fieldvector(Y)), Float64(t), stream()))
end

"""
    NativeJacobian(d)

Stands where the reference's `Jacobian{A, C}` (`implicit/jacobian.jl:54-62`)
does: the band storage lives in the handle, `update_jacobian!` refreshes it,
`ldiv!` solves with it.
"""
struct NativeJacobian
    d::NativeDycore
end
Base.zero(j::NativeJacobian) = j   # as jacobian.jl:71: no second Jacobian
function update_jacobian!(j::NativeJacobian, Y, p, dtγ, t)
    check(j.d.h, ccall((:ca_update_jacobian_fv, lib), Cint,
        (Ptr{Cvoid}, Ref{CaFieldVector}, Float64, Float64, Ptr{Cvoid}),
        j.d.h, Ref(fieldvector(Y)), Float64(dtγ), Float64(t), stream()))
end
function LinearAlgebra.ldiv!(ΔY::Fields.FieldVector, j::NativeJacobian, R::Fields.FieldVector)
    check(j.d.h, ccall((:ca_ldiv_fv, lib), Cint,
        (Ptr{Cvoid}, Ref{CaFieldVector}, Ref{CaFieldVector}, Ptr{Cvoid}),
        j.d.h, Ref(fieldvector(ΔY)), Ref(fieldvector(R)), stream()))
end
# the Krylov fallback (jacobian.jl:87-96) is the same call on the wrapped FieldVectors
LinearAlgebra.ldiv!(ΔY::AbstractVector, j::NativeJacobian, R::AbstractVector) =
    LinearAlgebra.ldiv!(Fields.FieldVector(ΔY), j, Fields.FieldVector(R))

"""
    args_integrator(Y, p, tspan, ode_algo, callback, native; mode = :hooks)

`CA.args_integrator` with the hooks forwarded to the library.  `p` must carry
the `NativeDycore` as `p.native`.  `mode = :hooks` keeps ClimaTimeSteppers'
stage loop (`Y` in device memory); `mode = :fused` is used through
[`FusedIntegrator`](@ref) instead.
"""
function args_integrator(Y, p, tspan, ode_algo, callback, native::NativeDycore, dt;
                         update_cache_every = "stage", update_constrain_state_every = "step")
    T_imp! = CTS.ODEFunction(implicit_tendency!; jac_prototype = NativeJacobian(native),
                             Wfact = update_jacobian!)
    tendency_function = CTS.ClimaODEFunction(;
        T_exp_T_lim! = remaining_tendency!, T_imp!,
        T_post_imp! = correct_implicit_advection_tendency!,
        cache! = set_precomputed_quantities!,
        cache_imp! = set_implicit_precomputed_quantities!,
        lim! = CA.limiters_func!, dss!, constrain_state! = CA.constrain_state!,
        update_cache = CA.update_cache_signal_handler(update_cache_every),
        update_constrain_state = CA.update_constrain_state_signal_handler(update_constrain_state_every),
        initialize_imp! = CA.initialize_implicit_stage_problem!,
    )
    problem = CTS.ODEProblem(tendency_function, Y, tspan, p)
    t_begin, t_end, dt = promote(tspan[1], tspan[2], dt)
    return (problem, ode_algo), (; saveat = [t_begin, t_end], callback, dt)
end

# --------------------------------------------------------------- fused mode
"""
    FusedIntegrator(native, Y_cpu, dt)

Keeps the state in a native device buffer and steps it with `ca_step`
(one `IMEXAlgorithm(ARS343(), NewtonsMethod(max_iters = 1))` step per call,
as `CTS.step!` does).  `sync!` writes it back into the CPU FieldVector
`Y_cpu` for callbacks, diagnostics and output.
"""
mutable struct FusedIntegrator
    native::NativeDycore
    Y::AMDGPU.ROCVector{Float32}     # native layout
    cf::AMDGPU.ROCVector{Float32}    # parent(Y.c) followed by parent(Y.f), on the device
    nc::Int                          # length(parent(Y.c))
    lays::Vector{CaLayout}
    Y_cpu::Any
    dt::Float64
    t::Float64
end
function FusedIntegrator(native::NativeDycore, Y_cpu, dt, t0 = 0.0)
    pc, pf = parent(Y_cpu.c), parent(Y_cpu.f)
    nc = length(pc)
    # ca_pack_state reads every component relative to one base pointer: the
    # two parent arrays go into one buffer, u_3's offset shifted past Y.c's
    cf = AMDGPU.ROCVector{Float32}(vcat(vec(pc), vec(pf)))
    lays = collect(layouts(Y_cpu))
    l5 = lays[5]
    lays[5] = CaLayout(l5.offset + nc, l5.s_v, l5.s_i, l5.s_j, l5.s_h)
    Y = AMDGPU.ROCVector{Float32}(undef, state_size(native))
    check(native.h, ccall((:ca_pack_state, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Ptr{CaLayout}, Ptr{Float32}, Ptr{Cvoid}),
        native.h, pointer(cf), lays, pointer(Y), stream()))
    return FusedIntegrator(native, Y, cf, nc, lays, Y_cpu, Float64(dt), Float64(t0))
end
function step!(fi::FusedIntegrator, nsteps = 1)
    check(fi.native.h, ccall((:ca_step, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Float64, Int32, Ptr{Cvoid}),
        fi.native.h, pointer(fi.Y), fi.dt, Int32(nsteps), stream()))
    fi.t += nsteps * fi.dt
    return fi
end
function sync!(fi::FusedIntegrator)
    check(fi.native.h, ccall((:ca_unpack_state, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Ptr{CaLayout}, Ptr{Float32}, Ptr{Cvoid}),
        fi.native.h, pointer(fi.Y), fi.lays, pointer(fi.cf), stream()))
    host = Array(fi.cf)
    pc, pf = parent(fi.Y_cpu.c), parent(fi.Y_cpu.f)
    copyto!(pc, reshape(view(host, 1:fi.nc), size(pc)))
    copyto!(pf, reshape(view(host, (fi.nc + 1):length(host)), size(pf)))
    return fi.Y_cpu
end

end # module
