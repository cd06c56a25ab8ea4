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

Usage (drop-in for the sweep path; `alg` takes the reference's own plugin
wrappers, metropolis.jl:181-199):

    using ClassicalSpinMC, CSMCGPU
    mc = MonteCarlo(T, lat, params)
    CSMCGPU.simulated_annealing!(mc, x -> 0.9^x, 1.0; alg=Metropolis())
    CSMCGPU.deterministic_updates!(mc)
    mcs = [MonteCarlo(T, lat, params; outpath=out) for T in temps]  # one per replica
    CSMCGPU.parallel_tempering!(mcs, [0, 3]; alg=MetropolisAdaptive())  # saveIC ranks 0 and 3

The plugin form: `CSMCGPU.device_alg(Metropolis())` is an
`alg::FunctionWrapper{Float64,Tuple{MonteCarlo,Float64}}` that runs one
Metropolis sweep of `mc.lattice` on the GPU per call, so the reference's own
drivers can take it: `ClassicalSpinMC.simulated_annealing!(mc, s, T0;
alg=CSMCGPU.device_alg(Metropolis()))`.

With ChainRulesCore loaded (package extension), `total_energy(lat)` and
`coupling_energy(lat, theta)` have rrules backed by the adjoint kernels.
"""
module CSMCGPU

using ClassicalSpinMC
using ClassicalSpinMC: UnitCell, Lattice, MonteCarlo, InteractionMatrix
using FunctionWrappers: FunctionWrapper
using HDF5
using Printf
using Dates

const libcsmc = get(ENV, "CSMC_LIB", joinpath(@__DIR__, "..", "..", "..", "csmc", "libcsmc.so"))

const OVERRELAX = Int32(0)       # overrelaxation!        monte_carlo.jl:126-139
const METROPOLIS = Int32(1)      # metropolis!            metropolis.jl:65-82
const CONE = Int32(2)            # metropolis_fixed_cone! metropolis.jl:136-153
const CONE_ADAPTIVE = Int32(3)   # metropolis_adaptive!   metropolis.jl:113-133
const DETERMINISTIC = Int32(4)   # deterministic_updates! monte_carlo.jl:201-213
const CONSTRAINT_ADAPTIVE = Int32(5)  # metropolis_constraint_adaptive! metropolis.jl:155-179

"""The device form of mc.constraint for metropolis_constraint_adaptive!
(metropolis.jl:155-179): c(lattice) = sum_i (a[:, b] . s_i + s_i' Q[:, :, b] s_i) - c0
over sites i of basis b.  Callable on a Lattice like the reference's closure
(mc.constraint(mc.lattice)), so MonteCarlo(...; constraint=SiteSumConstraint(...))
works with both the reference's drivers and these."""
struct SiteSumConstraint <: Function
    a::Matrix{Float64}      # 3 x n_basis
    Q::Array{Float64,3}     # 3 x 3 x n_basis
    c0::Float64
end
function (k::SiteSumConstraint)(lat::Lattice)::Float64
    nb = length(lat.unit_cell.basis)
    ncell = lat.size ÷ nb
    c = 0.0
    for i in 1:lat.size
        b = (i - 1) ÷ ncell + 1
        b > size(k.a, 2) && continue
        s = lat.spins[:, i]
        c += k.a[:, b]' * s + s' * k.Q[:, :, b] * s
    end
    return c - k.c0
end

const AlgWrapper = FunctionWrapper{Float64,Tuple{MonteCarlo,Float64}}

"""`alg(mc, T)` running one coloured Metropolis sweep of kind K of
mc.lattice on the GPU (the plugin contract of metropolis.jl:181-199: it
mutates mc.lattice.spins and mc.sigma and returns the accepted count)."""
struct DevicePlugin{K} end

"""The device kind of an `alg` plugin: the reference's wrappers
Metropolis(), MetropolisAdaptive(), MetropolisFixedCone()
(metropolis.jl:181-199), a device_alg wrapper, or a kind constant.  Any
other wrapped function has no device form and is refused."""
alg_kind(k::Integer) = Int32(k)
function alg_kind(alg::AlgWrapper)
    alg.objT <: DevicePlugin && return Int32(alg.objT.parameters[1])
    f = alg.obj[]
    f === ClassicalSpinMC.metropolis! && return METROPOLIS
    f === ClassicalSpinMC.metropolis_adaptive! && return CONE_ADAPTIVE
    f === ClassicalSpinMC.metropolis_fixed_cone! && return CONE
    f === ClassicalSpinMC.metropolis_constraint! &&
        throw(ArgumentError("MetropolisConstraint() calls e_diff, which the reference never defines " *
                            "(metropolis.jl:44); it has no device form"))
    f === ClassicalSpinMC.metropolis_constraint_adaptive! && return CONSTRAINT_ADAPTIVE
    throw(ArgumentError("alg wraps $(f), which has no device form: pass Metropolis(), " *
                        "MetropolisAdaptive() or MetropolisFixedCone() (metropolis.jl:181-199)"))
end

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

"""Hand mc.constraint (a SiteSumConstraint), mc.weight, mc.sigma0 and
mc.lambda to system 1 of `b` (csmc_constraint_set / csmc_constraint_lambda);
any other constraint function has no device form."""
function set_constraint!(b::Batch, mc::MonteCarlo)
    k = mc.constraint
    k isa SiteSumConstraint ||
        throw(ArgumentError("MetropolisConstraintAdaptive on the GPU needs mc.constraint to be a " *
                            "CSMCGPU.SiteSumConstraint: the reference calls an arbitrary Julia function " *
                            "per proposal (metropolis.jl:164)"))
    nb = size(k.a, 2)
    Qr = Float64[k.Q[r, c, q] for q in 1:nb for r in 1:3 for c in 1:3]  # row-major per basis
    check(ccall((:csmc_constraint_set, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64),
                b.ctx.h, b.id, nb, vec(k.a), Qr, k.c0, Float64(mc.weight), Float64(mc.sigma0)))
    lam = fill(Float64(mc.lambda), length(b.lattices))
    check(ccall((:csmc_constraint_lambda, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                b.ctx.h, b.id, lam, C_NULL))
    return nothing
end

function get_lambda(b::Batch)
    lam = zeros(Float64, length(b.lattices))
    check(ccall((:csmc_constraint_lambda, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                b.ctx.h, b.id, C_NULL, lam))
    return lam
end

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
                              alg::Union{AlgWrapper,Integer}=Metropolis(), ctx::Context=default_context(),
                              precision::Integer=64, seed::UInt64=UInt64(0))
    kind = alg_kind(alg)
    b = Batch(ctx, [mc.lattice]; precision=precision)
    set_spins!(b, 1, mc.lattice.spins)
    kind == CONSTRAINT_ADAPTIVE && set_constraint!(b, mc)
    p = mc.parameters
    accept_total = p.overrelaxation_rate != 0 ?
                   p.t_thermalization * mc.lattice.size / p.overrelaxation_rate :
                   p.t_thermalization * mc.lattice.size
    T, time = T0, 1
    while T > mc.T
        acc = zeros(Float64, 1)
        check(ccall((:csmc_anneal, libcsmc), Int32,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Int64, Int64, Int32, Float64, UInt64, Ptr{Float64}),
                    ctx.h, b.id, [T], 1, p.t_thermalization, p.overrelaxation_rate, kind,
                    Float64(mc.sigma0), seed, acc))
        println("Acceptance rate at T=$T: $(round(acc[1] / accept_total * 100, digits=5)) % ")
        mc.sigma = mc.sigma0
        kind == CONSTRAINT_ADAPTIVE && (mc.lambda = get_lambda(b)[1])
        if length(mc.outpath) > 0  # checkpoint per temperature (monte_carlo.jl:185-186)
            get_spins!(b, 1, mc.lattice.spins)
            ClassicalSpinMC.write_MC_checkpoint(mc)
        end
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
`mcs` (ordered by temperature; slot r plays MPI rank r - 1) in one batch.
saveIC lists the ranks whose configurations are written to
IC_<rank>/IC_<k>.h5 at every checkpoint (monte_carlo.jl:353-365,
hdf5.jl:185-194); replicas with an outpath get their spins dataset
rewritten at every checkpoint (hdf5.jl:176-180) and the observables and
correlations written at the end (hdf5.jl:204-239).  Returns, per
temperature slot, the rows C, dC, chi, dchi, <E>, dE, <|M|>, d|M| (and the
9 x n_k x R mean correlations when corr = true)."""
function parallel_tempering!(mcs::Vector{MonteCarlo}, saveIC::Vector{Int64}=Int64[];
                             alg::Union{AlgWrapper,Integer}=Metropolis(),
                             ctx::Context=default_context(), precision::Integer=32,
                             seed::UInt64=UInt64(0), verbose::Bool=true)
    kind = alg_kind(alg)
    kind == CONSTRAINT_ADAPTIVE &&
        throw(ArgumentError("MetropolisConstraintAdaptive() in parallel tempering: the device PT loop has " *
                            "no per-proposal multiplier chain; use simulated_annealing! or the plugin form"))
    R = length(mcs)
    R == 1 && @warn "MPI commSize of 1; no replica exchanges will occur!"
    b = Batch(ctx, [mc.lattice for mc in mcs]; precision=precision)
    for (i, mc) in enumerate(mcs)
        set_spins!(b, i, mc.lattice.spins)
    end
    p = mcs[1].parameters
    cp = CPTParams(p.t_thermalization, p.t_measurement, p.probe_rate, p.swap_rate,
                   p.overrelaxation_rate, p.report_interval, p.checkpoint_rate, kind, 1)
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
    # IC directories of the saveIC ranks (monte_carlo.jl:276-281)
    for rank in saveIC
        0 <= rank < R || continue
        path = dirname(mcs[rank + 1].outpath)
        verbose && println("Initializing IC collection on rank $rank")
        isdir(string(path, "/IC_$rank")) || mkdir(string(path, "/IC_$rank"))
    end
    verbose && @printf("Running sweeps on %s.\n", Dates.format(Dates.now(), "dd u yyyy HH:MM:SS"))
    sys = zeros(Int32, R)
    slot_state() = check(ccall((:csmc_pt_state, libcsmc), Int32,
                               (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                Ptr{Float64}), ctx.h, b.id, sys, C_NULL, C_NULL, C_NULL, C_NULL))
    stopped, event = Ref{Int64}(0), Ref{Int32}(0)
    while event[] != 3
        check(ccall((:csmc_pt_advance, libcsmc), Int32,
                    (Ptr{Cvoid}, Int32, Int64, Ref{Int64}, Ref{Int32}),
                    ctx.h, b.id, -1, stopped, event))
        if event[] == 1 && (any(mc -> length(mc.outpath) > 0, mcs) || !isempty(saveIC))
            # checkpoint at sweep `stopped`, after its exchange and before its
            # probe (monte_carlo.jl:353-365)
            slot_state()
            k = (stopped[] - p.t_thermalization) ÷ p.checkpoint_rate
            for (r, mc) in enumerate(mcs)
                get_spins!(b, sys[r] + 1, mc.lattice.spins)
                length(mc.outpath) > 0 && ClassicalSpinMC.write_MC_checkpoint(mc)
                if (r - 1) in saveIC
                    ClassicalSpinMC.write_initial_configuration(
                        string(dirname(mc.outpath), "/IC_$(r - 1)/IC_$k.h5"), mc)
                end
            end
        end
        # event 2: report due (print_runtime_statistics!, helper.jl:25-78)
        event[] == 2 && verbose && println("Sweep $(stopped[]) / $(p.t_thermalization + p.t_measurement)")
    end
    slot_state()
    sig = zeros(Float64, R)
    check(ccall((:csmc_pt_state, libcsmc), Int32,
                (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx.h, b.id, C_NULL, C_NULL, C_NULL, C_NULL, sig))
    obs = zeros(Float64, 8, R)
    check(ccall((:csmc_pt_observables, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                ctx.h, b.id, obs))
    for (r, mc) in enumerate(mcs)
        get_spins!(b, sys[r] + 1, mc.lattice.spins)
        mc.sigma = sig[r]
    end
    ssf = nothing
    if corr
        # mean(mc.observables.correlations) of each slot: 9 x n_k
        nk = size(ks, 2)
        ssf = zeros(Float64, 9, nk, R)
        count = Ref{Int64}(0)
        check(ccall((:csmc_pt_correlations, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ref{Int64}),
                    ctx.h, b.id, ssf, count))
    end
    for (r, mc) in enumerate(mcs)
        length(mc.outpath) > 0 && write_final_observables(mc, view(obs, :, r),
                                                          corr ? ssf[:, :, r] : nothing)
    end
    verbose && @printf("Simulation finished on %s.\n", Dates.format(Dates.now(), "dd u yyyy HH:MM:SS"))
    return corr ? (obs, ssf) : obs
end

"""One replica: parallel_tempering!(mc, saveIC; alg) as the reference calls
it per rank (monte_carlo.jl:235-236), here without exchanges."""
parallel_tempering!(mc::MonteCarlo, saveIC::Vector{Int64}=Int64[]; kw...) =
    parallel_tempering!([mc], saveIC; kw...)

"""write_final_observables (hdf5.jl:204-239) from the device's per-slot
observables: the spins dataset, the observables group (C and chi errors
relative, E and |M| errors absolute, as the reference writes them) and the
spin_correlations group when corr = true."""
function write_final_observables(mc::MonteCarlo, o::AbstractVector{Float64}, ssf)
    h5open(mc.outpath, "r+") do file
        file["spins"][:, :] = mc.lattice.spins
        if haskey(file, "observables")
            g = file["observables"]
            for key in keys(g)
                delete_object(g, key)
            end
        else
            g = create_group(file, "observables")
        end
        g["specific_heat"] = o[1]
        g["specific_heat_err"] = abs(o[2] / o[1])
        g["susceptibility"] = o[3]
        g["susceptibility_err"] = abs(o[4] / o[3])
        g["magnetization"] = o[7]
        g["magnetization_err"] = o[8]
        g["energy"] = o[5]
        g["energy_err"] = o[6]
        if ssf !== nothing
            d = haskey(file, "spin_correlations") ? file["spin_correlations"] :
                create_group(file, "spin_correlations")
            ClassicalSpinMC.overwrite_keys!(d, Dict("SSF" => Matrix{Float64}(ssf),
                                                    "SSF_momentum" => mc.momentum_vectors))
        end
    end
    return nothing
end

# ---- the plugin form: one device sweep per alg(mc, T) call ----------------------

const _plugin_batches = IdDict{MonteCarlo,Any}()
const plugin_seed = Ref{UInt64}(0)  # Philox key of the plugin sweeps

"""One coloured Metropolis sweep of kind K of mc.lattice at T on the GPU; the
spins go up and back every call, since the reference's drivers change them
on the host between plugin calls (overrelaxation!, monte_carlo.jl:126-139)."""
function (::DevicePlugin{K})(mc::MonteCarlo, T::Float64)::Float64 where {K}
    b = get!(() -> Batch(default_context(), [mc.lattice]; precision=64), _plugin_batches, mc)
    set_spins!(b, 1, mc.lattice.spins)
    K == CONSTRAINT_ADAPTIVE && set_constraint!(b, mc)
    sigma = [Float64(mc.sigma)]
    acc = sweep!(b, Int32(K), [T], sigma; n=1, seed=plugin_seed[])
    get_spins!(b, 1, mc.lattice.spins)
    K in (CONE_ADAPTIVE, CONSTRAINT_ADAPTIVE) && (mc.sigma = sigma[1])
    K == CONSTRAINT_ADAPTIVE && (mc.lambda = get_lambda(b)[1])
    return acc[1]
end

"""The device form of a reference plugin wrapper, for the reference's own
drivers: device_alg(Metropolis()) etc."""
device_alg(alg) = AlgWrapper(DevicePlugin{Int(alg_kind(alg))}())

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

"""Tiles per system of the batch-wide sweeps (csmc_batch_tiles): > 0 when
every lattice of the batch has a tile plan, so that a sweep is one launch of
LDS tiles; 0 when the sweeps run as colour passes."""
function batch_tiles(b::Batch)
    n = Ref{Int32}(0)
    check(ccall((:csmc_batch_tiles, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ref{Int32}, Ptr{Int32}), b.ctx.h, b.id, n,
                C_NULL))
    return Int(n[])
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

"""dE/d(coupling) per unit-cell term for every system of `b`:
n_sys x (3 n_field + 9 n_onsite + 9 n2 + 27 n3 + 81 n4), the layout of
`coupling_vector` (unit-cell term order, tensors column-major)."""
function energy_grad_couplings(b::Batch)
    uc = b.lattices[1].unit_cell
    n = 3 * length(uc.field) + 9 * length(uc.onsite) + 9 * length(uc.bilinear) +
        27 * length(uc.cubic) + 81 * length(uc.quartic)
    g = zeros(Float64, n, length(b.lattices))
    check(ccall((:csmc_energy_grad_couplings, libcsmc), Int32, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                b.ctx.h, b.id, g))
    return permutedims(g)
end

"""Every coupling of a unit cell as one vector, in the layout of the
coupling gradient: fields, onsite matrices (row-major), bilinear matrices
(row-major), cubic and quartic tensors (column-major), in term order."""
function coupling_vector(uc::UnitCell)
    v = Float64[]
    for t in uc.field; append!(v, collect(t[2])); end
    for t in uc.onsite; append!(v, imvec([t[2]])); end
    for t in uc.bilinear; append!(v, imvec([t[3]])); end
    for t in uc.cubic; append!(v, vec(t[4])); end
    for t in uc.quartic; append!(v, vec(t[5])); end
    return v
end

"""Replace the couplings of a device lattice whose unit cell keeps its terms
(csmc_lattice_update_couplings): theta in the layout of coupling_vector."""
function update_couplings!(ctx::Context, lat::Lattice, theta::AbstractVector{Float64})
    uc = lat.unit_cell
    k = 0
    take(n) = (k += n; theta[k-n+1:k])
    im(v) = InteractionMatrix(permutedims(reshape(v, 3, 3)))
    u = deepcopy(uc)  # UnitCell is immutable; its term vectors are not
    for (i, t) in enumerate(uc.field); u.field[i] = (t[1], collect(take(3))); end
    for (i, t) in enumerate(uc.onsite); u.onsite[i] = (t[1], im(take(9))); end
    for (i, t) in enumerate(uc.bilinear); u.bilinear[i] = (t[1], t[2], im(take(9)), t[4]); end
    for (i, t) in enumerate(uc.cubic)
        u.cubic[i] = (t[1], t[2], t[3], reshape(take(27), 3, 3, 3), t[5], t[6])
    end
    for (i, t) in enumerate(uc.quartic)
        u.quartic[i] = (t[1], t[2], t[3], t[4], reshape(take(81), 3, 3, 3, 3), t[6], t[7], t[8])
    end
    c, keep = pack(u)
    GC.@preserve keep check(ccall((:csmc_lattice_update_couplings, libcsmc), Int32,
                                  (Ptr{Cvoid}, Int32, Ref{CUnitCell}), ctx.h, device_lattice(ctx, lat), c))
    return nothing
end

"""total_energy(lat) (hamiltonian.jl:70-132) with the unit cell's couplings
replaced by theta (layout of coupling_vector), evaluated on the GPU in FP64;
differentiable in theta through the ChainRulesCore extension."""
function coupling_energy(lat::Lattice, theta::AbstractVector{Float64}; ctx::Context=default_context())
    update_couplings!(ctx, lat, theta)
    b = Batch(ctx, [lat]; precision=64)
    set_spins!(b, 1, lat.spins)
    return total_energy(b)[1]
end

end # module
