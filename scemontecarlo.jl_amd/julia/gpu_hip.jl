# The device Metropolis sweep of SCEMonteCarlo.jl on an AMD MI355X through
# libsce_hip.so (C ABI: include/sce_hip.h).
#
# A maintainer includes this file inside `module SCEMonteCarlo` in place of the
# KernelAbstractions device files (src/SCEMonteCarlo.jl:49-55 and :59 include
# gpu/philox.jl … gpu/gpu_gradient.jl and gpu/gpu_pt.jl):
#
#     include("gpu_hip.jl")        # after state.jl / updates.jl / pt.jl
#
# It defines the same exported names with the same signatures
# (src/SCEMonteCarlo.jl:86-88): GPUTiledHamiltonian(backend, H),
# GPUChainState(gH, st; seed), gpu_metropolis_sweep!, gpu_run_sweeps!,
# to_host!, gpu_run_pt, and the gradient pair gpu_energy_gradient!.  The
# reference's host machinery (ChainState, _renormalize!, the observable
# accumulators, PTResult) is used unchanged.  The library path comes from
# ENV["SCE_HIP_LIB"], else deps/libsce_hip.so beside the package.
#
# `backend` is accepted as the reference accepts a KernelAbstractions backend
# (e.g. AMDGPU's ROCBackend()); the library drives its own HIP device, so the
# argument only selects nothing.  `workgroupsize` keeps its meaning — the
# number of lanes that walk one site, part of the determinism scope — with the
# library's lane counts: 16, 32 or 64 walk with that many lanes; a larger power
# of two walks with one full wavefront (64), so the reference's default 128 runs
# unchanged; anything else is an ArgumentError naming the allowed values.

using StaticArrays: SVector
using Random: Xoshiro

const SCE_HIP_LIB = get(ENV, "SCE_HIP_LIB", joinpath(@__DIR__, "..", "deps", "libsce_hip.so"))

# ---- errors ---------------------------------------------------------------------

function _hip_check(rc::Int32)
    rc == 0 && return nothing
    buf = Vector{UInt8}(undef, 1024)
    ccall((:sce_last_error, SCE_HIP_LIB), Int32, (Ptr{UInt8}, Csize_t), buf, length(buf))
    msg = unsafe_string(pointer(buf))
    rc == -1 && throw(ArgumentError(msg))            # SCE_EINVAL
    rc == -2 && throw(DimensionMismatch(msg))        # SCE_EDIM
    rc == -3 && throw(ArgumentError(msg))            # SCE_ERANGE_COUNTER (gpu_sweep.jl:163-164)
    rc == -5 && throw(OutOfMemoryError())            # SCE_ENOMEM
    error("libsce_hip (status $rc): $msg")           # SCE_EHIP, SCE_ELAYOUT, SCE_ENODEV
end

"""
    hip_lanes(workgroupsize) -> Int32

Lanes per site for a reference `workgroupsize`: 16, 32 and 64 as given, any
larger power of two → 64 (one wavefront per site), else an `ArgumentError`.
"""
function hip_lanes(workgroupsize::Integer)::Int32
    ws = Int(workgroupsize)
    ispow2(ws) || throw(ArgumentError("workgroupsize must be a power of two (got $ws)"))
    ws >= 16 || throw(ArgumentError(
        "workgroupsize $ws is below the device walk's smallest lane count; use 16, 32, 64 " *
        "or a larger power of two (walked by 64 lanes)"))
    return Int32(min(ws, 64))
end

# ---- tables: GPUTiledHamiltonian(backend, H) — src/gpu/gpu_hamiltonian.jl:66-88 --

# Mirrors struct sce_tables_desc (include/sce_hip.h): counts, then pointers into
# the TiledHamiltonian's own arrays (1-based Int32 as the Julia struct holds
# them; Bool vectors are one byte per element).
struct TablesDesc
    n_cell_atoms::Int32
    dims::NTuple{3,Int32}
    n_sites::Int32
    lmax::Int32
    nlm::Int32
    n_terms::Int32
    n_inst::Int32
    n_active::Int32
    n_colors::Int32
    n_prog::Int32
    n_inst_sites::Int64
    n_sent::Int64
    n_sfac::Int64
    n_eent::Int64
    n_efac::Int64
    inst_term::Ptr{Int32}
    inst_ptr::Ptr{Int32}
    inst_sites::Ptr{Int32}
    site_ptr::Ptr{Int32}
    site_inst::Ptr{Int32}
    site_slot::Ptr{Int8}
    site_has_l1::Ptr{UInt8}
    site_active::Ptr{UInt8}
    site_prog::Ptr{Int32}
    sprog_ptr::Ptr{Int32}
    sent_w::Ptr{Float64}
    sent_tgt::Ptr{Int32}
    sfac_ptr::Ptr{Int32}
    sfac_row::Ptr{Int32}
    sfac_slot::Ptr{Int8}
    site_col::Ptr{Int32}
    site_col2::Ptr{Int32}
    pent_row::Ptr{Int32}
    pent_row2::Ptr{Int32}
    eprog_ptr::Ptr{Int32}
    eent_w::Ptr{Float64}
    efac_ptr::Ptr{Int32}
    efac_row::Ptr{Int32}
    term_coef::Ptr{Float64}
    color_ptr::Ptr{Int32}
    color_sites::Ptr{Int32}
end

"""
    tables_desc(H::TiledHamiltonian) -> TablesDesc

The flat tables of `H` (src/hamiltonian.jl:234-266, `_ContractionPrograms`
:62-91) as the C ABI reads them.  The descriptor points into `H`'s arrays:
keep `H` alive (`GC.@preserve H`) while the library reads it.
"""
function tables_desc(H::TiledHamiltonian)::TablesDesc
    pr = H.progs
    return TablesDesc(Int32(H.n_cell_atoms), (Int32(H.dims[1]), Int32(H.dims[2]), Int32(H.dims[3])),
                      Int32(H.n_sites), Int32(H.lmax), Int32(H.nlm), Int32(length(H.terms)),
                      Int32(length(H.inst_term)), Int32(H.n_active), Int32(H.n_colors),
                      Int32(length(pr.sprog_ptr) - 1),
                      Int64(length(H.site_inst)), Int64(length(pr.sent_w)), Int64(length(pr.sfac_row)),
                      Int64(length(pr.eent_w)), Int64(length(pr.efac_row)),
                      pointer(H.inst_term), pointer(H.inst_ptr), pointer(H.inst_sites),
                      pointer(H.site_ptr), pointer(H.site_inst), pointer(H.site_slot),
                      Ptr{UInt8}(pointer(H.site_has_l1)), Ptr{UInt8}(pointer(H.site_active)),
                      pointer(pr.site_prog), pointer(pr.sprog_ptr), pointer(pr.sent_w),
                      pointer(pr.sent_tgt), pointer(pr.sfac_ptr), pointer(pr.sfac_row),
                      pointer(pr.sfac_slot), pointer(pr.site_col), pointer(pr.site_col2),
                      pointer(pr.pent_row), pointer(pr.pent_row2),
                      pointer(pr.eprog_ptr), pointer(pr.eent_w), pointer(pr.efac_ptr),
                      pointer(pr.efac_row), pointer(pr.term_coef),
                      pointer(H.color_ptr), pointer(H.color_sites))
end

"""
    GPUTiledHamiltonian(backend, H::TiledHamiltonian)

Upload the tables of `H` and derive the device walk layout (sce_tables_create):
rejects lmax > 6 with an `ArgumentError` (gpu_hamiltonian.jl:67-68).  Keeps `H`
alongside for host-side bookkeeping, as the reference does.
"""
mutable struct GPUTiledHamiltonian
    ptr::Ptr{Cvoid}
    backend::Any
    host::TiledHamiltonian
end

function GPUTiledHamiltonian(backend, H::TiledHamiltonian)
    H.lmax <= 6 || throw(ArgumentError(
        "lmax = $(H.lmax) unsupported on the device path (gated ≤ 6)"))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve H begin
        d = Ref(tables_desc(H))
        _hip_check(ccall((:sce_tables_create, SCE_HIP_LIB), Int32,
                         (Ref{TablesDesc}, Ref{Ptr{Cvoid}}), d, out))
    end
    gH = GPUTiledHamiltonian(out[], backend, H)
    finalizer(gH) do g
        g.ptr == C_NULL || ccall((:sce_tables_destroy, SCE_HIP_LIB), Int32, (Ptr{Cvoid},), g.ptr)
        g.ptr = C_NULL
    end
    return gH
end

n_sites(gH::GPUTiledHamiltonian) = gH.host.n_sites

# ---- chain: GPUChainState(gH, st; seed) — src/gpu/gpu_state.jl:49-61 -----------

# Mirrors struct sce_chain_info (include/sce_hip.h).
struct ChainInfo
    energy::Float64
    step::Float64
    max_drift::Float64
    last_drift::Float64
    sweep_index::Int64
    acc_metro::Int64
    att_metro::Int64
    acc_or::Int64
    att_or::Int64
    seed::UInt64
    acc_stage1::Int64
end

mutable struct GPUChainState
    ptr::Ptr{Cvoid}
    gH::GPUTiledHamiltonian
    seed::UInt64
end

function GPUChainState(gH::GPUTiledHamiltonian, st::ChainState; seed::Integer = rand(UInt64))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve st _hip_check(ccall((:sce_chain_create, SCE_HIP_LIB), Int32,
                                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, UInt64, Float64,
                                      Ref{Ptr{Cvoid}}),
                                     gH.ptr, Ptr{Float64}(pointer(st.config)), pointer(st.zrows),
                                     st.energy, UInt64(seed), st.step, out))
    gst = GPUChainState(out[], gH, UInt64(seed))
    finalizer(gst) do g
        g.ptr == C_NULL || ccall((:sce_chain_destroy, SCE_HIP_LIB), Int32, (Ptr{Cvoid},), g.ptr)
        g.ptr = C_NULL
    end
    return gst
end

function chain_info(gst::GPUChainState)::ChainInfo
    r = Ref{ChainInfo}()
    _hip_check(ccall((:sce_chain_info_get, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Ref{ChainInfo}), gst.ptr, r))
    return r[]
end

# The reference's GPUChainState fields, read from the device-resident chain.
function Base.getproperty(gst::GPUChainState, f::Symbol)
    f in (:ptr, :gH, :seed) && return getfield(gst, f)
    f in (:energy, :step, :sweep_index, :acc_metro, :att_metro, :acc_or, :att_or, :max_drift) &&
        return getfield(chain_info(gst), f)
    return getfield(gst, f)
end

function Base.setproperty!(gst::GPUChainState, f::Symbol, v)
    if f === :step
        _hip_check(ccall((:sce_chain_set_step, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Float64), gst.ptr, Float64(v)))
        return v
    end
    return setfield!(gst, f, v)
end

# to_host!(st, gst) / _from_host!(gst, st) — src/gpu/gpu_state.jl:71-86
function to_host!(st::ChainState, gst::GPUChainState)::ChainState
    e = Ref{Float64}(0.0)
    GC.@preserve st _hip_check(ccall((:sce_download, SCE_HIP_LIB), Int32,
                                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                                     gst.ptr, Ptr{Float64}(pointer(st.config)), pointer(st.zrows), e))
    st.energy = e[]
    return st
end

function _from_host!(gst::GPUChainState, st::ChainState)::GPUChainState
    GC.@preserve st _hip_check(ccall((:sce_upload, SCE_HIP_LIB), Int32,
                                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64),
                                     gst.ptr, Ptr{Float64}(pointer(st.config)), pointer(st.zrows),
                                     st.energy))
    return gst
end

# _swap_payload!(a, b) — src/gpu/gpu_pt.jl:46-51: config, rows and energy move
# (device pointers and their ownership), the keyed-RNG state stays.
_swap_payload!(a::GPUChainState, b::GPUChainState) =
    _hip_check(ccall((:sce_swap_payload, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), a.ptr, b.ptr))

# Lane-owned counters and the drift record of a restored chain (checkpoint resume,
# src/checkpoint.jl:584-597).
function _restore_lane!(gst::GPUChainState, sweep_index, acc_metro, att_metro, max_drift)
    _hip_check(ccall((:sce_chain_set_counters, SCE_HIP_LIB), Int32,
                     (Ptr{Cvoid}, Int64, Int64, Int64, Int64, Int64),
                     gst.ptr, sweep_index, acc_metro, att_metro, 0, 0))
    _hip_check(ccall((:sce_chain_set_drift, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Float64), gst.ptr, max_drift))
    return gst
end

# ---- sweeps — src/gpu/gpu_sweep.jl:158-218 --------------------------------------

"""
    gpu_metropolis_sweep!(gst, gH, β; workgroupsize = 128) -> Int

One Metropolis sweep on the device; returns the number of accepted moves.
"""
function gpu_metropolis_sweep!(gst::GPUChainState, gH::GPUTiledHamiltonian, β::Float64;
                               workgroupsize::Integer = 128)::Int
    lanes = hip_lanes(workgroupsize)
    n = Ref{Int64}(0)
    _hip_check(ccall((:sce_metropolis_sweep, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Float64, Int32, Ref{Int64}),
                     gst.ptr, β, lanes, n))
    return Int(n[])
end

"""
    gpu_run_sweeps!(gst, gH, st, β, nsweeps; renorm_interval = 1_000,
                    workgroupsize = 128, or_per_metropolis = 0) -> GPUChainState

`nsweeps` device sweeps (each one Metropolis sweep plus `or_per_metropolis`
overrelaxation passes), renormalising on the device every `renorm_interval`
sweeps (≤ 0 disables); downloads the final state into `st`.
"""
function gpu_run_sweeps!(gst::GPUChainState, gH::GPUTiledHamiltonian, st::ChainState,
                         β::Float64, nsweeps::Integer; renorm_interval::Integer = 1_000,
                         workgroupsize::Integer = 128, or_per_metropolis::Integer = 0)::GPUChainState
    lanes = hip_lanes(workgroupsize)
    obs = zeros(Float64, 4)
    n = Ref{Int64}(0)
    _hip_check(ccall((:sce_run_sweeps, SCE_HIP_LIB), Int32,
                     (Ptr{Cvoid}, Float64, Int64, Int32, Int64, Int64, Int32, Ptr{Float64}, Int64, Ref{Int64}),
                     gst.ptr, β, Int64(nsweeps), Int32(or_per_metropolis), Int64(renorm_interval),
                     Int64(0), lanes, obs, Int64(1), n))
    to_host!(st, gst)
    return gst
end

# ---- the all-site gradient — src/gpu/gpu_gradient.jl:187-226 --------------------

function gpu_energy_gradient!(G::Vector{SVector{3,Float64}}, gH::GPUTiledHamiltonian,
                              config::Vector{SVector{3,Float64}}; workgroupsize::Integer = 128)
    length(G) == length(config) == n_sites(gH) ||
        throw(DimensionMismatch("gradient and configuration need $(n_sites(gH)) sites"))
    _hip_check(ccall((:sce_energy_gradient, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int32),
                     gH.ptr, Ptr{Float64}(pointer(config)), Ptr{Float64}(pointer(G)), Int32(workgroupsize)))
    return G
end

function gpu_energy_gradient!(G::Vector{SVector{3,Float64}}, gst::GPUChainState, gH::GPUTiledHamiltonian;
                              workgroupsize::Integer = 128)
    _hip_check(ccall((:sce_chain_energy_gradient, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int32),
                     gst.ptr, Ptr{Float64}(pointer(G)), Int32(workgroupsize)))
    return G
end

# ---- replica exchange: gpu_run_pt — src/gpu/gpu_pt.jl:261-319 --------------------

# the batch handle (sce_batch_*): every rung sweeps in one launch per colour,
# exchanges run on the device
mutable struct _HIPBatch
    ptr::Ptr{Cvoid}
    R::Int
end

function _HIPBatch(gH::GPUTiledHamiltonian, states::Vector{<:ChainState}, seeds::Vector{UInt64},
                   kts::Vector{Float64}, exchange_seed::UInt64)
    R = length(states)
    tabs = fill(gH.ptr, R)
    cfgs = [Ptr{Float64}(pointer(st.config)) for st in states]
    zrs = [pointer(st.zrows) for st in states]
    en = [st.energy for st in states]
    steps = [st.step for st in states]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve states _hip_check(ccall((:sce_batch_create, SCE_HIP_LIB), Int32,
                                         (Int32, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Float64},
                                          Ptr{UInt64}, Ptr{Float64}, Ptr{Float64}, UInt64, Ref{Ptr{Cvoid}}),
                                         Int32(R), tabs, cfgs, zrs, en, seeds, steps, kts, exchange_seed, out))
    b = _HIPBatch(out[], R)
    finalizer(b) do x
        x.ptr == C_NULL || ccall((:sce_batch_destroy, SCE_HIP_LIB), Int32, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    return b
end

function _batch_run!(b::_HIPBatch, n::Int, lanes::Int32, renorm::Int, exchange::Int, done::Int, total::Int)
    obs = zeros(Float64, 4 * b.R)
    nobs = Ref{Int64}(0)
    _hip_check(ccall((:sce_batch_run, SCE_HIP_LIB), Int32,
                     (Ptr{Cvoid}, Int64, Int32, Int64, Int64, Int64, Int64, Int64, Ptr{Float64}, Int64, Ref{Int64}),
                     b.ptr, n, lanes, renorm, 0, exchange, done, total, obs, 1, nobs))
    return nothing
end

function _batch_info(b::_HIPBatch, r::Int)::ChainInfo
    x = Ref{ChainInfo}()
    _hip_check(ccall((:sce_batch_info, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Int32, Ref{ChainInfo}),
                     b.ptr, Int32(r - 1), x))
    return x[]
end

function _batch_download!(st::ChainState, b::_HIPBatch, r::Int)
    e = Ref{Float64}(0.0)
    GC.@preserve st _hip_check(ccall((:sce_batch_download, SCE_HIP_LIB), Int32,
                                     (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                                     b.ptr, Int32(r - 1), Ptr{Float64}(pointer(st.config)), pointer(st.zrows), e))
    st.energy = e[]
    return st
end

"""
    gpu_run_pt(gH::GPUTiledHamiltonian; temperature = nothing, kT = nothing,
               exchange_interval = 10, workgroupsize = 128, kwargs...) -> PTResult
    gpu_run_pt(backend, H::TiledHamiltonian; kwargs...) -> PTResult

Replica exchange on the device with the reference's keywords and result.  All
rungs sweep in one launch per colour and exchange on the device
(`sce_batch_run`: the alternating-parity schedule of `_gpu_pt_phase!`, one
keyed uniform per attempted pair); thermalisation adapts each rung's step every
`adapt_interval` sweeps on its counters (`_gpu_adapt_step!`); every
measurement downloads each rung into its host mirror and feeds the
`observables` accumulators, as the reference does, so any `Observable` /
`Evaluable` works.  The master-seed derivation keeps the reference's order
(lane RNG words, exchange words, device seeds); the exchange uniforms are keyed
by the first exchange word instead of drawn from an exchange Xoshiro.
Checkpoints of ladder runs are written by the library's Python driver
(`sce_amd.gpu_run_pt(checkpoint=…)`); here `checkpoint` must be `nothing`.
"""
function gpu_run_pt(gH::GPUTiledHamiltonian; temperature = nothing, kT = nothing,
                    exchange_interval::Integer = 10, sweeps_therm::Integer = 2_000,
                    sweeps_measure::Integer = 10_000, measure_interval::Integer = 1,
                    step::Real = 0.6, adapt_target::Real = 0.5, adapt_interval::Integer = 50,
                    renorm_interval::Integer = 1_000, nbins::Integer = 32,
                    observables::Vector{Observable} = standard_observables(gH.host),
                    evaluables::Vector{Evaluable} = standard_evaluables(),
                    init = nothing, seed::Integer = rand(UInt64), workgroupsize::Integer = 128,
                    checkpoint::Union{Nothing,AbstractString} = nothing,
                    checkpoint_interval::Integer = 0, persistent::Bool = false)::PTResult
    H = gH.host
    kts = resolve_kt(temperature, kT)
    R = length(kts)
    R >= 2 || throw(ArgumentError("parallel tempering needs a ladder of ≥ 2 temperatures; got $R"))
    (all(diff(kts) .> 0) || all(diff(kts) .< 0)) || throw(ArgumentError(
        "the temperature ladder must be strictly monotone; got $kts"))
    exchange_interval >= 1 || throw(ArgumentError("exchange_interval must be ≥ 1; got $exchange_interval"))
    measure_interval >= 1 || throw(ArgumentError("measure_interval must be ≥ 1; got $measure_interval"))
    checkpoint === nothing || throw(ArgumentError(
        "checkpointed ladder runs of the HIP path are written by sce_amd.gpu_run_pt(checkpoint = …)"))
    lanes = hip_lanes(workgroupsize)
    _check_observables(observables)
    _check_evaluables(observables, evaluables)
    # the reference's master-seed derivation order (gpu_pt.jl:17-29)
    master = Xoshiro(UInt64(seed))
    lane_rngs = [Xoshiro(rand(master, UInt64), rand(master, UInt64), rand(master, UInt64), rand(master, UInt64))
                 for _ = 1:R]
    exch_words = (rand(master, UInt64), rand(master, UInt64), rand(master, UInt64), rand(master, UInt64))
    dev_seeds = [rand(master, UInt64) for _ = 1:R]
    mirrors = [ChainState(H, _initial_config(H, init, lane_rngs[r]), lane_rngs[r], Float64(step)) for r = 1:R]
    b = _HIPBatch(gH, mirrors, dev_seeds, Vector{Float64}(kts), exch_words[1])
    # an addition to the reference's keywords: each run of sweeps as one
    # persistent launch of every rung (bit-identical to per-colour launches)
    _hip_check(ccall((:sce_batch_set_persistent, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Int32), b.ptr, Int32(persistent)))
    ex = Int(exchange_interval)
    # thermalisation: exchanges, renormalisation, and step adaptation every
    # adapt_interval sweeps on the rung's counter window
    done = 0
    while done < sweeps_therm
        n = min(adapt_interval - done % adapt_interval, sweeps_therm - done)
        _batch_run!(b, n, lanes, Int(renorm_interval), ex, done, Int(sweeps_therm))
        done += n
        if done % adapt_interval == 0
            for r = 1:R
                inf = _batch_info(b, r)
                s = inf.att_metro > 0 ?
                    clamp(inf.step * exp(0.5 * (inf.acc_metro / inf.att_metro - adapt_target)), 1e-3, Float64(π)) :
                    inf.step
                _hip_check(ccall((:sce_batch_set_step, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Int32, Float64),
                                 b.ptr, Int32(r - 1), s))
            end
            _hip_check(ccall((:sce_batch_reset_counters, SCE_HIP_LIB), Int32, (Ptr{Cvoid},), b.ptr))
        end
    end
    # thermalisation → measurement boundary: renormalise, fresh counters
    _hip_check(ccall((:sce_batch_renormalize, SCE_HIP_LIB), Int32, (Ptr{Cvoid},), b.ptr))
    _hip_check(ccall((:sce_batch_reset_counters, SCE_HIP_LIB), Int32, (Ptr{Cvoid},), b.ptr))
    planned = fld(sweeps_measure, measure_interval)
    accs = [[ObsAccumulator(o, planned, nbins) for o in observables] for _ = 1:R]
    done = 0
    while done < sweeps_measure
        n = min(measure_interval - done % measure_interval, sweeps_measure - done)
        _batch_run!(b, n, lanes, Int(renorm_interval), ex, done, Int(sweeps_measure))
        done += n
        if done % measure_interval == 0
            for r = 1:R
                _batch_download!(mirrors[r], b, r)
                for acc in accs[r]
                    _measure!(acc, mirrors[r].config, mirrors[r].energy, H)
                end
            end
        end
    end
    points = TempResult[]
    for r = 1:R
        inf = _batch_info(b, r)
        acc_m = inf.att_metro == 0 ? NaN : inf.acc_metro / inf.att_metro
        push!(points, TempResult(kts[r], kts[r] / KB_EV, _finalize_stats(accs[r], evaluables, kts[r], H.n_active),
                                 acc_m, NaN, inf.step, inf.max_drift))
    end
    att = zeros(Int64, R - 1)
    acc = zeros(Int64, R - 1)
    _hip_check(ccall((:sce_batch_swap_stats, SCE_HIP_LIB), Int32, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}),
                     b.ptr, att, acc))
    swaps = [att[i] == 0 ? NaN : acc[i] / att[i] for i = 1:(R - 1)]
    finals = [copy(_batch_download!(mirrors[r], b, r).config) for r = 1:R]
    return PTResult(points, swaps, finals, UInt64(seed))
end

gpu_run_pt(backend, H::TiledHamiltonian; kwargs...)::PTResult = gpu_run_pt(GPUTiledHamiltonian(backend, H); kwargs...)
