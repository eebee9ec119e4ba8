# HyQMOMAMD: the Julia side of libhyqmom_hip.so (MI355X, gfx950), bound with `ccall`.
#
# It keeps the reference's API for the 3D step -- simulation_runner(params) and
# run_simulation_with_snapshots(params; snapshot_interval) of HyQMOM.jl
# (src/simulation_runner.jl:63-759, 994-1011) -- and replaces the step loop
# (src/simulation_runner.jl:346-560) by the device step.  The initial condition is built
# by HyQMOM.jl's own functions (initialize_moment_field, src/initial_conditions.jl:103), so
# cell membership and Maxwellian moments are the reference's.
#
# The C declarations are in include/hyqmom.h.  This file is source only: the build that
# produced libhyqmom_hip.so has no Julia toolchain to run it.
module HyQMOMAMD

export simulation_runner, run_simulation_with_snapshots, AmdSolver, step!, run!

const libhyqmom = get(ENV, "HYQMOM_AMD_LIB",
                      joinpath(@__DIR__, "..", "build", "libhyqmom_hip.so"))

const PRECISION = Dict(:fp64 => Int32(0), :mixed => Int32(1), :mixed_learned => Int32(2),
                       :mixed_newton => Int32(3))

# Layout of hyq_params (include/hyqmom.h): 8 x Int32, then 11 x Float64.
struct HyqParams
    Nx::Int32; Ny::Int32; Nz::Int32; Nmom::Int32
    flag2D::Int32; positivity::Int32; precision::Int32; reserved::Int32
    tmax::Float64; Kn::Float64; Ma::Float64; CFL::Float64; dtmax::Float64
    xmin::Float64; xmax::Float64; ymin::Float64; ymax::Float64; zmin::Float64; zmax::Float64
end

last_error() = unsafe_string(ccall((:hyq_last_error, libhyqmom), Cstring, ()))

function check(rc::Integer, what::AbstractString)
    rc == 0 && return nothing
    # HYQ_ENONFINITE is where the reference's eigvals throws (eigenvalues6z_hyperbolic_3D.jl:108,121)
    error("$what failed (code $rc): $(last_error())")
end

getp(params, k, default) = hasproperty(params, k) ? getproperty(params, k) : default

function hyq_params(params)
    HyqParams(Int32(params.Nx), Int32(params.Ny), Int32(params.Nz), Int32(params.Nmom),
              Int32(params.flag2D), Int32(getp(params, :positivity, true) ? 1 : 0),
              PRECISION[getp(params, :precision, :fp64)], Int32(0),
              Float64(params.tmax), Float64(params.Kn), Float64(params.Ma), Float64(params.CFL),
              Float64(params.dtmax),
              Float64(getp(params, :xmin, -0.5)), Float64(getp(params, :xmax, 0.5)),
              Float64(getp(params, :ymin, -0.5)), Float64(getp(params, :ymax, 0.5)),
              Float64(getp(params, :zmin, -0.5)), Float64(getp(params, :zmax, 0.5)))
end

mutable struct AmdSolver
    ptr::Ptr{Cvoid}
    p::HyqParams
    t::Float64
    nn::Int64
    function AmdSolver(params)
        p = hyq_params(params)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:hyq_create, libhyqmom), Cint, (Ref{HyqParams}, Ref{Ptr{Cvoid}}), p, out),
              "hyq_create")
        s = new(out[], p, 0.0, 0)
        finalizer(s) do x
            x.ptr != C_NULL && ccall((:hyq_destroy, libhyqmom), Cint, (Ptr{Cvoid},), x.ptr)
            x.ptr = C_NULL
        end
        return s
    end
end

# State transfer in the reference's own layout: (Nx, Ny, Nz, 35), column-major, interior only.
function set_state!(s::AmdSolver, M::Array{Float64,4})
    size(M) == (s.p.Nx, s.p.Ny, s.p.Nz, 35) || throw(DimensionMismatch("state shape"))
    GC.@preserve M check(ccall((:hyq_set_state, libhyqmom), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                               s.ptr, M), "hyq_set_state")
end
function get_state(s::AmdSolver)
    M = Array{Float64,4}(undef, s.p.Nx, s.p.Ny, s.p.Nz, 35)
    GC.@preserve M check(ccall((:hyq_get_state, libhyqmom), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                               s.ptr, M), "hyq_get_state")
    return M
end

# One pass of the loop body of simulation_runner.jl:346-560; returns dt.
function step!(s::AmdSolver)
    t = Ref(s.t); dt = Ref(0.0); nn = Ref(s.nn)
    check(ccall((:hyq_step, libhyqmom), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Int64}),
                s.ptr, t, dt, nn), "hyq_step")
    s.t, s.nn = t[], nn[]
    return dt[]
end

# while t < tmax && nn < nnmax: step (simulation_runner.jl:346), on the device stream.
function run!(s::AmdSolver, tmax::Real, nnmax::Integer)
    t = Ref(s.t); nn = Ref(s.nn)
    check(ccall((:hyq_run, libhyqmom), Cint, (Ptr{Cvoid}, Cdouble, Int64, Ref{Float64}, Ref{Int64}),
                s.ptr, tmax, nnmax, t, nn), "hyq_run")
    s.t, s.nn = t[], nn[]
    return s
end

# Central and standardized moment fields of the current state (compute_central_field,
# compute_standardized_field), computed on the device.
function snapshot_fields(s::AmdSolver)
    C = Array{Float64,4}(undef, s.p.Nx, s.p.Ny, s.p.Nz, 35); S = similar(C)
    GC.@preserve C S check(ccall((:hyq_snapshot_fields, libhyqmom), Cint,
                                 (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), s.ptr, C, S),
                           "hyq_snapshot_fields")
    return C, S
end

# The runner's symmetry check MaxDiff (simulation_runner.jl:565-628), a device reduction.
function symmetry_maxdiff(s::AmdSolver)
    m = Ref(0.0); per = zeros(5)
    check(ccall((:hyq_symmetry_maxdiff, libhyqmom), Cint, (Ptr{Cvoid}, Ref{Float64}, Ptr{Float64}),
                s.ptr, m, per), "hyq_symmetry_maxdiff")
    return m[], per
end

# The first-step vmax breakdown (simulation_runner.jl:404-461): (vx_max, vy_max, vz_max) and
# the 1-based (i, j, k) of the fastest cell in the reference's search order.
function vmax_breakdown(s::AmdSolver)
    v = zeros(3); w = zeros(Int64, 3)
    check(ccall((:hyq_vmax_breakdown, libhyqmom), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}),
                s.ptr, v, w), "hyq_vmax_breakdown")
    return (v[1], v[2], v[3]), (w[1] + 1, w[2] + 1, w[3] + 1)
end

# grid_out of simulation_runner.jl:720-730
function grid_out(params)
    xmin, xmax = getp(params, :xmin, -0.5), getp(params, :xmax, 0.5)
    ymin, ymax = getp(params, :ymin, -0.5), getp(params, :ymax, 0.5)
    zmin, zmax = getp(params, :zmin, -0.5), getp(params, :zmax, 0.5)
    Nx, Ny, Nz = params.Nx, params.Ny, params.Nz
    dx, dy, dz = (xmax - xmin) / Nx, (ymax - ymin) / Ny, (zmax - zmin) / Nz
    return (x = collect(range(xmin, xmax, length=Nx + 1)), y = collect(range(ymin, ymax, length=Ny + 1)),
            z = collect(range(zmin, zmax, length=Nz + 1)), dx = dx, dy = dy, dz = dz,
            xm = collect(range(xmin + dx / 2, step=dx, length=Nx)),
            ym = collect(range(ymin + dy / 2, step=dy, length=Ny)),
            zm = collect(range(zmin + dz / 2, step=dz, length=Nz)))
end

# The initial state, built by HyQMOM.jl (custom ICs: initialize_moment_field,
# src/initial_conditions.jl:103; otherwise the caller passes M0).
function initial_state(params, HyQMOM)
    g = grid_out(params)
    gp = (Nx = params.Nx, Ny = params.Ny, Nz = params.Nz, xm = g.xm, ym = g.ym, zm = g.zm)
    return HyQMOM.initialize_moment_field(gp, params.ic_background, params.ic_jets;
                                          r110 = params.r110, r101 = params.r101, r011 = params.r011)
end

"""
    simulation_runner(params; M0 = nothing, HyQMOM = nothing)

Same contract as HyQMOM.simulation_runner for one device: returns
`(M_final, final_time, time_steps, grid_out)`.  `params` may also carry
`precision` in (:fp64, :mixed, :mixed_newton, :mixed_learned).
"""
function simulation_runner(params; M0 = nothing, HyQMOM = nothing)
    M = M0 !== nothing ? M0 : initial_state(params, HyQMOM)
    s = AmdSolver(params)
    set_state!(s, Array{Float64,4}(M))
    run!(s, params.tmax, params.nnmax)
    return get_state(s), s.t, s.nn, grid_out(params)
end

"""
    run_simulation_with_snapshots(params; snapshot_interval = 2, M0 = nothing, HyQMOM = nothing)

Returns `(snapshots, grid_out)` with `snapshots` a vector of `(M, t, step)` taken every
`snapshot_interval` steps and at the end (the reference writes the same content to JLD2,
src/simulation_runner.jl:291-318, 676-688).
"""
function run_simulation_with_snapshots(params; snapshot_interval = 2, M0 = nothing, HyQMOM = nothing)
    M = M0 !== nothing ? M0 : initial_state(params, HyQMOM)
    s = AmdSolver(params)
    set_state!(s, Array{Float64,4}(M))
    snaps = [(copy(M), 0.0, 0)]
    while s.t < params.tmax && s.nn < params.nnmax
        run!(s, params.tmax, min(params.nnmax, s.nn + snapshot_interval))
        push!(snaps, (get_state(s), s.t, s.nn))
    end
    return snaps, grid_out(params)
end

end # module
