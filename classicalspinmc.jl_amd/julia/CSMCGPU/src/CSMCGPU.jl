"""
    CSMCGPU

Julia binding of libcsmc (the MI355X sweep engine, `include/csmc.h`) for
ClassicalSpinMC.jl.  Every function is a thin `ccall`; the library owns the
device buffers and borrows Julia arrays only for the duration of a call.

Layouts line up with the reference without conversion:
  * `lat.spins` (3xN Float64, column-major, lattice.jl:6) is the library's
    `double[N][3]`;
  * `InteractionMatrix` (interaction_matrix.jl:2-12, nine Float64 in
    row order) is the library's row-major 3x3;
  * cubic / quartic tensors are passed as Julia stores them (column-major).

Usage (drop-in for the sweep path):

    using ClassicalSpinMC, CSMCGPU
    mc = MonteCarlo(T, lat, params)
    CSMCGPU.simulated_annealing!(mc, x -> 0.9^x, 1.0; alg=CSMCGPU.METROPOLIS)
    CSMCGPU.deterministic_updates!(mc)
    mcs = [MonteCarlo(T, lat, params) for T in temps]      # one per replica
    CSMCGPU.parallel_tempering!(mcs)                        # all replicas on one GPU
"""
module CSMCGPU

using ClassicalSpinMC
using ClassicalSpinMC: UnitCell, Lattice, MonteCarlo, InteractionMatrix

const libcsmc = get(ENV, "CSMC_LIB", joinpath(@__DIR__, "..", "csmc", "libcsmc.so"))

const OVERRELAX = Int32(0)       # overrelaxation!        monte_carlo.jl:126-139
const METROPOLIS = Int32(1)      # metropolis!            metropolis.jl:65-82
const CONE = Int32(2)            # metropolis_fixed_cone! metropolis.jl:136-153
const CONE_ADAPTIVE = Int32(3)   # metropolis_adaptive!   metropolis.jl:113-133
const DETERMINISTIC = Int32(4)   # deterministic_updates! monte_carlo.jl:201-213

struct CsmcError <: Exception
    code::Int32
    msg::String
end

function check(rc::Int32)
    rc == 0 && return nothing
    throw(CsmcError(rc, unsafe_string(ccall((:csmc_last_error, libcsmc), Cstring, ()))))
end

# csmc_unit_cell: field order and types follow include/csmc.h
struct CUnitCell
    dim::Int32
    n_basis::Int32
    lattice_vectors::Ptr{Float64}
    basis::Ptr{Float64}
    n_field::Int32
    field_b::Ptr{Int32}
    field_h::Ptr{Float64}
    n_onsite::Int32
    onsite_b::Ptr{Int32}
    onsite_m::Ptr{Float64}
    n_bilinear::Int32
    bilinear_b::Ptr{Int32}
    bilinear_m::Ptr{Float64}
    bilinear_off::Ptr{Int32}
    n_cubic::Int32
    cubic_b::Ptr{Int32}
    cubic_t::Ptr{Float64}
    cubic_off::Ptr{Int32}
    n_quartic::Int32
    quartic_b::Ptr{Int32}
    quartic_t::Ptr{Float64}
    quartic_off::Ptr{Int32}
end

f64(v) = isempty(v) ? zeros(Float64, 1) : collect(Float64, v)
i32(v) = isempty(v) ? zeros(Int32, 1) : collect(Int32, v)
imvec(ms) = Float64[getfield(m, k) for m in ms for k in fieldnames(InteractionMatrix)]

"""Flatten a UnitCell (unit_cell.jl:3-18); returns the C struct and the
arrays it points into (keep them alive with GC.@preserve)."""
function pack(uc::UnitCell{D}) where {D}
    keep = (
        f64([x for v in uc.lattice_vectors for x in v]),
        f64([x for v in uc.basis for x in v]),
        i32([t[1] for t in uc.field]),
        f64([x for t in uc.field for x in t[2]]),
        i32([t[1] for t in uc.onsite]),
        f64(imvec([t[2] for t in uc.onsite])),
        i32([x for t in uc.bilinear for x in (t[1], t[2])]),
        f64(imvec([t[3] for t in uc.bilinear])),
        i32([x for t in uc.bilinear for x in t[4]]),
        i32([x for t in uc.cubic for x in (t[1], t[2], t[3])]),
        f64([x for t in uc.cubic for x in vec(t[4])]),
        i32([x for t in uc.cubic for o in (t[5], t[6]) for x in o]),
        i32([x for t in uc.quartic for x in (t[1], t[2], t[3], t[4])]),
        f64([x for t in uc.quartic for x in vec(t[5])]),
        i32([x for t in uc.quartic for o in (t[6], t[7], t[8]) for x in o]),
    )
    p = map(pointer, keep)
    c = CUnitCell(D, length(uc.basis), p[1], p[2],
                  length(uc.field), p[3], p[4], length(uc.onsite), p[5], p[6],
                  length(uc.bilinear), p[7], p[8], p[9],
                  length(uc.cubic), p[10], p[11], p[12],
                  length(uc.quartic), p[13], p[14], p[15])
    return c, keep
end

mutable struct Context
    h::Ptr{Cvoid}
    lattices::IdDict{Any,Int32}
    function Context(device::Integer=0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:csmc_ctx_create, libcsmc), Int32, (Int32, Ptr{Ptr{Cvoid}}), device, r))
        ctx = new(r[], IdDict{Any,Int32}())
        finalizer(c -> ccall((:csmc_ctx_destroy, libcsmc), Int32, (Ptr{Cvoid},), c.h), ctx)
    end
end

const _default = Ref{Union{Nothing,Context}}(nothing)
function default_context()
    _default[] === nothing && (_default[] = Context(0))
    return _default[]
end

"""Lattice(...) tables on the device (csmc_lattice_create); cached per
unit cell, shape, S and boundary condition."""
function device_lattice(ctx::Context, lat::Lattice)
    key = (lat.unit_cell, lat.shape, lat.S, lat.bc)
    get!(ctx.lattices, key) do
        c, keep = pack(lat.unit_cell)
        shape = Int32[lat.shape...]
        id = Ref{Int32}(0)
        periodic = lat.bc == "periodic" ? Int32(1) : lat.bc == "open" ? Int32(0) : Int32(2)
        GC.@preserve keep shape begin
            check(ccall((:csmc_lattice_create, libcsmc), Int32,
                        (Ptr{Cvoid}, Ref{CUnitCell}, Ptr{Int32}, Int32, Float64, Ref{Int32}),
                        ctx.h, c, shape, periodic, Float64(lat.S), id))
        end
        id[]
    end
end

"""A batch of systems on one GPU (csmc_batch_create): replicas or
independent lattices of uneven size."""
mutable struct Batch
    ctx::Context
    id::Int32
    lattices::Vector{Lattice}
    function Batch(ctx::Context, lats::Vector{<:Lattice}; precision::Integer=32)
        ids = Int32[device_lattice(ctx, l) for l in lats]
        id = Ref{Int32}(0)
        check(ccall((:csmc_batch_create, libcsmc), Int32,
                    (Ptr{Cvoid}, Int32, Ptr{Int32}, Int32, Ref{Int32}),
                    ctx.h, length(ids), ids, precision, id))
        b = new(ctx, id[], collect(Lattice, lats))
        finalizer(b -> ccall((:csmc_batch_destroy, libcsmc), Int32, (Ptr{Cvoid}, Int32), b.ctx.h, b.id), b)
    end
end

set_spins!(b::Batch, i::Integer, spins::Matrix{Float64}) =
    check(ccall((:csmc_set_spins, libcsmc), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                b.ctx.h, b.id, i - 1, spins))
get_spins!(b::Batch, i::Integer, spins::Matrix{Float64}) =
    check(ccall((:csmc_get_spins, libcsmc), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                b.ctx.h, b.id, i - 1, spins))

"""n coloured sweeps of `kind` on every system; returns accepted counts."""
function sweep!(b::Batch, kind::Int32, T::Vector{Float64}, sigma::Vector{Float64}; n::Integer=1,
                seed::UInt64=UInt64(0))
    acc = zeros(Float64, length(b.lattices))
    check(ccall((:csmc_sweep, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, UInt64, Ptr{Float64}),
                b.ctx.h, b.id, kind, T, sigma, n, seed, acc))
    return acc
end

function total_energy(b::Batch)
    E = zeros(Float64, length(b.lattices))
    check(ccall((:csmc_total_energy, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                b.ctx.h, b.id, E, C_NULL))
    return E
end

# ---- drivers with the reference's signatures ---------------------------------

"""simulated_annealing! (monte_carlo.jl:157-190) on the GPU: one launch per
temperature runs its t_thermalization-1 sweeps."""
function simulated_annealing!(mc::MonteCarlo, schedule::Function, T0::Float64=1.0;
                              alg::Int32=METROPOLIS, ctx::Context=default_context(),
                              precision::Integer=64, seed::UInt64=UInt64(0))
    b = Batch(ctx, [mc.lattice]; precision=precision)
    set_spins!(b, 1, mc.lattice.spins)
    p = mc.parameters
    accept_total = p.overrelaxation_rate != 0 ?
                   p.t_thermalization * mc.lattice.size / p.overrelaxation_rate :
                   p.t_thermalization * mc.lattice.size
    T, time = T0, 1
    while T > mc.T
        acc = zeros(Float64, 1)
        check(ccall((:csmc_anneal, libcsmc), Int32,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Int64, Int64, Int32, Float64, UInt64, Ptr{Float64}),
                    ctx.h, b.id, [T], 1, p.t_thermalization, p.overrelaxation_rate, alg,
                    Float64(mc.sigma0), seed, acc))
        println("Acceptance rate at T=$T: $(round(acc[1] / accept_total * 100, digits=5)) % ")
        mc.sigma = mc.sigma0
        T = schedule(time)
        time += 1
    end
    get_spins!(b, 1, mc.lattice.spins)
    return nothing
end

"""deterministic_updates! (monte_carlo.jl:201-213) as coloured sweeps."""
function deterministic_updates!(mc::MonteCarlo; ctx::Context=default_context())
    mc.parameters.t_deterministic > 1 || return nothing
    n = max(1, round(Int, (mc.parameters.t_deterministic - 1) / mc.lattice.size))
    b = Batch(ctx, [mc.lattice]; precision=64)
    set_spins!(b, 1, mc.lattice.spins)
    sweep!(b, DETERMINISTIC, [0.0], [0.0]; n=n)
    get_spins!(b, 1, mc.lattice.spins)
    return nothing
end

# csmc_pt_params: field order and types follow include/csmc.h
struct CPTParams
    t_thermalization::Int64
    t_measurement::Int64
    probe_rate::Int64
    swap_rate::Int64
    overrelaxation_rate::Int64
    report_interval::Int64
    checkpoint_rate::Int64
    alg::Int32
    measure::Int32
end

"""parallel_tempering! (monte_carlo.jl:235-398) with every replica of
`mcs` (ordered by temperature, one per former MPI rank) in one batch."""
function parallel_tempering!(mcs::Vector{MonteCarlo}; alg::Int32=METROPOLIS,
                             ctx::Context=default_context(), precision::Integer=32,
                             seed::UInt64=UInt64(0))
    b = Batch(ctx, [mc.lattice for mc in mcs]; precision=precision)
    for (i, mc) in enumerate(mcs)
        set_spins!(b, i, mc.lattice.spins)
    end
    p = mcs[1].parameters
    cp = CPTParams(p.t_thermalization, p.t_measurement, p.probe_rate, p.swap_rate,
                   p.overrelaxation_rate, p.report_interval, p.checkpoint_rate, alg, 1)
    temps = Float64[mc.T for mc in mcs]
    check(ccall((:csmc_pt_init, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ref{CPTParams}, Float64, UInt64),
                ctx.h, b.id, temps, cp, Float64(mcs[1].sigma0), seed))
    # corr = true (monte_carlo.jl:371-375): S(k) at every probe, on the device;
    # the momenta go as double[n_k][D], i.e. permutedims of the D x n_k matrix
    ks = mcs[1].momentum_vectors
    corr = mcs[1].corr && size(ks, 2) > 0
    if corr
        check(ccall((:csmc_pt_set_correlations, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32),
                    ctx.h, b.id, Matrix{Float64}(permutedims(ks)), size(ks, 2)))
    end
    stopped, event = Ref{Int64}(0), Ref{Int32}(0)
    while event[] != 3
        check(ccall((:csmc_pt_advance, libcsmc), Int32,
                    (Ptr{Cvoid}, Int32, Int64, Ref{Int64}, Ref{Int32}),
                    ctx.h, b.id, -1, stopped, event))
        # event 1: checkpoint due (write_MC_checkpoint! after get_spins!),
        # event 2: report due (print_runtime_statistics!)
    end
    R = length(mcs)
    sys = zeros(Int32, R)
    check(ccall((:csmc_pt_state, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx.h, b.id, sys, C_NULL, C_NULL, C_NULL, C_NULL))
    obs = zeros(Float64, 8, R)
    check(ccall((:csmc_pt_observables, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                ctx.h, b.id, obs))
    for (r, mc) in enumerate(mcs)
        get_spins!(b, sys[r] + 1, mc.lattice.spins)
    end
    if corr
        # mean(mc.observables.correlations) of each slot: 9 x n_k
        nk = size(ks, 2)
        ssf = zeros(Float64, 9, nk, R)
        count = Ref{Int64}(0)
        check(ccall((:csmc_pt_correlations, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ref{Int64}),
                    ctx.h, b.id, ssf, count))
        return obs, ssf
    end
    return obs  # rows: C, dC, chi, dchi, <E>, dE, <|M|>, d|M| per temperature slot
end

"""compute_equal_time_correlations (spin_correlations.jl:6-43) of every
system of `b` (one lattice): 9 x n_k x n_sys."""
function compute_equal_time_correlations(b::Batch, ks::Matrix{Float64})
    nk = size(ks, 2)
    out = zeros(Float64, 9, nk, length(b.lattices))
    check(ccall((:csmc_spin_correlations, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Ptr{Float64}),
                b.ctx.h, b.id, Matrix{Float64}(permutedims(ks)), nk, out))
    return out
end

# ---- resume (no reference counterpart) ---------------------------------------------

"""The whole state of a PT run on batch `b` between two csmc_pt_advance calls."""
function pt_save(b::Batch)
    n = Ref{Int64}(0)
    check(ccall((:csmc_pt_save, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ref{Int64}),
                b.ctx.h, b.id, C_NULL, n))
    buf = Vector{UInt8}(undef, n[])
    check(ccall((:csmc_pt_save, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{UInt8}, Ref{Int64}),
                b.ctx.h, b.id, buf, n))
    return buf
end

"""Continue a saved run, after csmc_pt_init with the saved run's settings."""
pt_load!(b::Batch, state::Vector{UInt8}) =
    check(ccall((:csmc_pt_load, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{UInt8}, Int64),
                b.ctx.h, b.id, state, length(state)))

# ---- launch forms (no reference counterpart) ----------------------------------------

"""Workgroups per system of the PT kernel chosen at csmc_pt_init (1: one; more:
the split kernel's strips, or the parts of a system whose spins stay in device
memory)."""
function pt_split_parts(b::Batch)
    g = Ref{Int32}(0)
    check(ccall((:csmc_pt_split_parts, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ref{Int32}), b.ctx.h, b.id, g))
    return Int(g[])
end

"""Time each sweep form on the batch's own state and keep the fastest
(csmc_autotune); returns the per-form times in ms (NaN: skipped) and the
chosen form's index."""
function autotune!(b::Batch, kind::Int32, T::Vector{Float64}; n_sweeps::Integer=3)
    times = zeros(Float64, 5)
    choice = Ref{Int32}(0)
    check(ccall((:csmc_autotune, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int32, Ptr{Float64}, Ref{Int32}),
                b.ctx.h, b.id, kind, T, n_sweeps, times, choice))
    return times, choice[]
end

"""n_sweeps plain-form sweeps captured as one hipGraph; replay with
graph_launch!(b, g), free with graph_destroy!(g)."""
function graph_sweeps(b::Batch, kind::Int32, T::Vector{Float64}, sigma::Vector{Float64}; n_sweeps::Integer=1,
                      seed::UInt64=UInt64(0))
    g = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:csmc_graph_sweeps, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, UInt64, Ref{Ptr{Cvoid}}),
                b.ctx.h, b.id, kind, T, sigma, n_sweeps, seed, g))
    return g[]
end
graph_launch!(b::Batch, g::Ptr{Cvoid}) =
    check(ccall((:csmc_graph_launch, libcsmc), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), b.ctx.h, g))
graph_destroy!(g::Ptr{Cvoid}) = check(ccall((:csmc_graph_destroy, libcsmc), Int32, (Ptr{Cvoid},), g))

# ---- differentiation ------------------------------------------------------------

"""dE/dS for every site of every system of `b` (3 x N per system)."""
function energy_grad_spins(b::Batch)
    n = sum(l.size for l in b.lattices)
    g = zeros(Float64, 3, n)
    check(ccall((:csmc_energy_grad_spins, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                b.ctx.h, b.id, g))
    return g
end

"""ChainRulesCore.rrule for total_energy(lat) with respect to lat.spins,
defined when ChainRulesCore is loaded:

    function ChainRulesCore.rrule(::typeof(ClassicalSpinMC.total_energy), lat::Lattice)
        b = Batch(default_context(), [lat]; precision=64)
        set_spins!(b, 1, lat.spins)
        E = total_energy(b)[1]
        pullback(dE) = (NoTangent(), Tangent{Lattice}(; spins=dE .* energy_grad_spins(b)))
        return E, pullback
    end
"""
function rrule_doc end

end # module
