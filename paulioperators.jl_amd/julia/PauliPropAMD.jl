# PauliPropAMD — Julia binding of libpauliprop.so (the MI355X SparsePauliVector
# engine) for PauliOperators.jl v3.1.0.
#
# A `DeviceSPV{N,W,T}` is a device-resident twin of the reference's
# `SparsePauliVector{N,W,T}` for T = Float64 or ComplexF64
# (src/type_SparsePauliVector.jl:69-84); the
# methods below keep the reference's names, keyword arguments and exception
# types and forward to the C ABI declared in include/pauliprop.h.  Keys pass
# unchanged: a Vector{UInt64} (N <= 64) or Vector{UInt128} (N <= 128) is already
# W little-endian 64-bit words per term, a Vector{Float64} or
# Vector{ComplexF64} is already the ABI's coefficient layout, and `MergeFilter`
# (src/spv_kernels.jl:33-42) has the byte layout of `pp_filter`.  Julia
# computes cos θ, sin θ and the channels' λ^k itself and passes them in, so
# results match the Julia reference to the bit.
module PauliPropAMD

using LinearAlgebra
using Random
using PauliOperators
import PauliOperators: evolve!, expectation_value, truncate!, stochastic_clip!,
                       coeff_clip!, clip!, weight_clip!, weight_damped_clip!, x_weight_clip!,
                       x_weight_damped_clip!, majorana_weight_clip!,
                       depolarizing_channel!, dephasing_channel!, bit_flip_channel!,
                       bit_phase_flip_channel!, pauli_channel!,
                       commutator, anticommutator, inner_product, check_spv, variance,
                       otimes, osum, largest, largest_diag, offdiag
using PauliOperators: MergeFilter, NOFILTER, _compile_filter, _is_compilable, _pack, _spv_mask,
                      TruncationStrategy, NoTruncation, CoeffTruncation, WeightTruncation,
                      WeightDampedTruncation, XWeightTruncation, XWeightDampedTruncation,
                      MajoranaWeightTruncation, AdaptiveTruncation, StochasticCoeffTruncation,
                      StochasticSamplingTruncation, CompositeTruncation,
                      CorrectionAccumulator, NoCorrection, EnergyCorrection, EnergyVarianceCorrection,
                      WindowCounters

export DeviceSPV, DeviceSPVBatch, SurrogateMLP, lookup_noalloc, capacities, set_resident!

const LIB = joinpath(@__DIR__, "..", "_build", "libpauliprop.so")

# ---- errors -------------------------------------------------------------------

const PP_OK, PP_EARG, PP_ENOMEM, PP_EHIP, PP_EOVERFLOW, PP_EUNSUPPORTED = 0, 1, 2, 3, 4, 5

last_error() = unsafe_string(ccall((:pp_last_error, LIB), Cstring, ()))
version() = unsafe_string(ccall((:pp_version, LIB), Cstring, ()))

# PP_EARG is the reference's ArgumentError / DimensionMismatch
function check(rc::Integer)
    rc == PP_OK && return nothing
    msg = last_error()
    rc == PP_EARG && throw(occursin("mismatch", msg) ? DimensionMismatch(msg) : ArgumentError(msg))
    rc == PP_ENOMEM && throw(OutOfMemoryError())
    error("libpauliprop: ", msg, " (code ", rc, ")")
end

# Julia's Float64^Int as the library restates it (for non-Julia callers'
# channel tables); Julia code uses its own `^`.
julia_powi(x::Float64, n::Integer) = ccall((:pp_julia_powi, LIB), Float64, (Float64, Int64), x, n)

# ---- ABI records ----------------------------------------------------------------

struct PPCounters                       # pp_counters: WindowCounters summed over a call
    rotations::Int64
    terms_swept::Int64
    terms_created::Int64
    merge_in::Int64
    merge_out::Int64
    early_merges::Int64
    t_rotate::Float64
    t_merge::Float64
end
PPCounters() = PPCounters(0, 0, 0, 0, 0, 0, 0.0, 0.0)

const PP_OP_FILTER, PP_OP_ADAPTIVE, PP_OP_STOCHASTIC_COEFF, PP_OP_STOCHASTIC_SAMPLING = 0, 1, 2, 3
struct PPStrategyOp                     # pp_strategy_op
    kind::Int32
    reserved::Int32
    filter::MergeFilter
    max_terms::Int64
    value::Float64
    seed::UInt64
end

const PP_ALG_MUL, PP_ALG_COMMUTATOR, PP_ALG_ANTICOMMUTATOR = 0, 1, 2
struct PPTerms                          # pp_terms
    z::Ptr{UInt64}
    x::Ptr{UInt64}
    re::Ptr{Float64}
    im::Ptr{Float64}
    n::Int64
    conj::Int32
end

# A TruncationStrategy as device operations, members of a composite in order
# (_apply_tup!), the way _apply! runs them (spv_evolve.jl:59-119).  The
# stochastic strategies draw their device seed from their own rng.
function _strategy_ops!(ops::Vector{PPStrategyOp}, s::TruncationStrategy)
    if _is_compilable(s)
        push!(ops, PPStrategyOp(PP_OP_FILTER, 0, _compile_filter(s), 0, 0.0, 0))
    elseif s isa AdaptiveTruncation
        push!(ops, PPStrategyOp(PP_OP_ADAPTIVE, 0, NOFILTER, s.max_terms, s.min_thresh, 0))
    elseif s isa StochasticCoeffTruncation
        push!(ops, PPStrategyOp(PP_OP_STOCHASTIC_COEFF, 0, NOFILTER, 0, s.epsilon, rand(s.rng, UInt64)))
    elseif s isa StochasticSamplingTruncation
        push!(ops, PPStrategyOp(PP_OP_STOCHASTIC_SAMPLING, 0, NOFILTER, s.n_keep, 0.0, rand(s.rng, UInt64)))
    elseif s isa CompositeTruncation
        for t in s.strategies
            _strategy_ops!(ops, t)
        end
    else
        throw(ArgumentError("truncation strategy $(typeof(s)) has no device implementation"))
    end
    return ops
end

# EnergyCorrection(ψ) / EnergyVarianceCorrection(ψ) as the ABI's ket words and
# accumulators (the variance accumulator is C_NULL for EnergyCorrection)
_ket_words(::Type{W}, ψ::Ket) where {W} = Ref{W}((ψ.v % UInt128) % W)
function _correction_args(::Type{W}, corr::CorrectionAccumulator) where {W}
    corr isa NoCorrection && return (C_NULL, Ref{Float64}(0.0), C_NULL)
    corr isa EnergyCorrection &&
        return (_ket_words(W, corr.ψ), Ref{Float64}(corr.accumulated_energy), C_NULL)
    corr isa EnergyVarianceCorrection &&
        return (_ket_words(W, corr.ψ), Ref{Float64}(corr.accumulated_energy), Ref{Float64}(corr.accumulated_variance))
    throw(ArgumentError("$(typeof(corr)) has no device implementation"))
end
function _correction_back!(corr::CorrectionAccumulator, acc, vacc)
    if corr isa EnergyCorrection || corr isa EnergyVarianceCorrection
        corr.accumulated_energy = acc[]
    end
    corr isa EnergyVarianceCorrection && (corr.accumulated_variance = vacc[])
    return corr
end

# pp_counters of a call into a WindowCounters: the device keeps call totals,
# which are added to window 1
function _add_counters!(wc::WindowCounters, c::PPCounters)
    wc.terms_created[1] += c.terms_created
    wc.merge_in[1] += c.merge_in
    wc.merge_out[1] += c.merge_out
    wc.early_merges[1] += c.early_merges
    wc.t_rotate[1] += c.t_rotate
    wc.t_merge[1] += c.t_merge
    return wc
end

# ---- one device vector ----------------------------------------------------------

const DeviceCoeff = Union{Float64,ComplexF64}

mutable struct DeviceSPV{N,W,T<:DeviceCoeff}
    h::Ptr{Cvoid}
end

function _destroy!(d::DeviceSPV)
    d.h == C_NULL && return nothing
    ccall((:pp_destroy, LIB), Cint, (Ptr{Cvoid},), d.h)
    d.h = C_NULL
    return nothing
end

# SparsePauliVector(N, T; capacity) on the device, with the reference's
# virtual buffer lengths (type_SparsePauliVector.jl:94-104,138-141)
function DeviceSPV{N,W,T}(live_cap::Integer=16, append_cap::Integer=16) where {N,W,T<:DeviceCoeff}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = T <: Complex ?
        ccall((:pp_create_typed, LIB), Cint, (Cint, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
              N, live_cap, append_cap, 1, h) :
        ccall((:pp_create, LIB), Cint, (Cint, Int64, Int64, Ref{Ptr{Cvoid}}), N, live_cap, append_cap, h)
    check(rc)
    d = DeviceSPV{N,W,T}(h[])
    finalizer(_destroy!, d)
    return d
end
DeviceSPV{N,W}(live_cap::Integer=16, append_cap::Integer=16) where {N,W} = DeviceSPV{N,W,Float64}(live_cap, append_cap)

function DeviceSPV(v::SparsePauliVector{N,W,T}) where {N,W,T<:DeviceCoeff}
    d = DeviceSPV{N,W,T}(length(v.z), length(v.az))
    upload!(d, v)
    return d
end

function upload!(d::DeviceSPV{N,W,T}, v::SparsePauliVector{N,W,T}) where {N,W,T}
    v.an == 0 || throw(ArgumentError("merge the pending appends first (merge_pending!)"))
    check(ccall((:pp_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{W}, Ptr{W}, Ptr{T}, Int64, Cint),
                d.h, v.z, v.x, v.c, v.n, 1))
    return d
end

# the coefficient type the library holds for d (pp_coeff_type)
function coeff_type(d::DeviceSPV)
    c = Ref{Cint}(0)
    check(ccall((:pp_coeff_type, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}), d.h, c))
    return c[] == 1 ? ComplexF64 : Float64
end

function Base.length(d::DeviceSPV)
    n = Ref{Int64}(0)
    check(ccall((:pp_size, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), d.h, n))
    return Int(n[])
end

function capacities(d::DeviceSPV)
    lc = Ref{Int64}(0)
    ac = Ref{Int64}(0)
    check(ccall((:pp_capacities, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), d.h, lc, ac))
    return (Int(lc[]), Int(ac[]))
end

# PauliSum(d) goes through the host copy: PauliSum(SparsePauliVector(d))
function PauliOperators.SparsePauliVector(d::DeviceSPV{N,W,T}) where {N,W,T}
    n = length(d)
    v = SparsePauliVector(N, T; capacity=max(n, 1))
    got = Ref{Int64}(0)
    check(ccall((:pp_download, LIB), Cint, (Ptr{Cvoid}, Ptr{W}, Ptr{W}, Ptr{T}, Int64, Ref{Int64}),
                d.h, v.z, v.x, v.c, length(v.z), got))
    v.n = Int(got[])
    return v
end
PauliOperators.PauliSum(d::DeviceSPV) = PauliSum(SparsePauliVector(d))

function set_resident!(d::DeviceSPV, mode::Integer=1, cap_limit::Integer=0)
    check(ccall((:pp_set_resident, LIB), Cint, (Ptr{Cvoid}, Cint, Cint), d.h, mode, cap_limit))
    return d
end

# test hook: the merge look-back tag base moved forward (reaches the 22-bit wrap)
set_epoch!(d::DeviceSPV, epoch::Integer) =
    (check(ccall((:pp_set_epoch, LIB), Cint, (Ptr{Cvoid}, UInt32), d.h, epoch)); d)

# evolve!(O, generators, angles; window, truncation, local_truncation,
# correction, counters) (spv_evolve.jl:312-387)
function evolve!(d::DeviceSPV{N,W}, generators::Vector{PauliBasis{N}}, angles::Vector{<:Real};
                 window::Int=1,
                 truncation::TruncationStrategy=NoTruncation(),
                 local_truncation::TruncationStrategy=NoTruncation(),
                 correction::CorrectionAccumulator=NoCorrection(),
                 counters::Union{Nothing,WindowCounters}=nothing) where {N,W}
    length(generators) == length(angles) ||
        throw(DimensionMismatch("generators and angles must have same length"))
    window >= 1 || throw(ArgumentError("window must be >= 1"))
    L = length(generators)
    L == 0 && return d
    _is_compilable(local_truncation) ||
        throw(ArgumentError("local_truncation must be a deterministic " *
                            "(weight/coefficient) strategy — it runs per append"))
    gz = Vector{W}(undef, L)
    gx = Vector{W}(undef, L)
    for i in 1:L
        gz[i], gx[i] = _pack(W, generators[i])
    end
    cosv = Float64[cos(Float64(a)) for a in angles]     # spv_evolve.jl:341-342
    sinv = Float64[sin(Float64(a)) for a in angles]
    fl = Ref(_compile_filter(local_truncation))
    cnt = Ref(PPCounters())
    if _is_compilable(truncation) && correction isa NoCorrection
        fs = Ref(_compile_filter(truncation))
        check(ccall((:pp_evolve, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{W}, Ptr{W}, Ptr{Float64}, Ptr{Float64}, Int64, Int64,
                     Ref{MergeFilter}, Ref{MergeFilter}, Ref{PPCounters}),
                    d.h, gz, gx, cosv, sinv, L, window, fs, fl, cnt))
    else
        # the `compiled || _apply!` branch of _boundary! (spv_evolve.jl:258-284)
        compiled = _is_compilable(truncation)
        fs = Ref(compiled ? _compile_filter(truncation) : NOFILTER)
        ops = compiled ? PPStrategyOp[] : _strategy_ops!(PPStrategyOp[], truncation)
        ket, acc, vacc = _correction_args(W, correction)
        check(ccall((:pp_evolve_ex, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{W}, Ptr{W}, Ptr{Float64}, Ptr{Float64}, Int64, Int64,
                     Ref{MergeFilter}, Ptr{PPStrategyOp}, Cint, Ref{MergeFilter}, Ptr{Cvoid},
                     Ref{Float64}, Ptr{Cvoid}, Ref{PPCounters}),
                    d.h, gz, gx, cosv, sinv, L, window, fs, ops, length(ops), fl, ket, acc, vacc, cnt))
        _correction_back!(correction, acc, vacc)
    end
    counters === nothing || _add_counters!(counters, cnt[])
    return d
end

# evolve!(O, G, θ): one rotation, no truncation (spv_evolve.jl:228-236)
function evolve!(d::DeviceSPV{N,W}, G::PauliBasis{N}, θ::Real) where {N,W}
    gz, gx = _pack(W, G)
    check(ccall((:pp_evolve_single, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{W}, Float64, Float64),
                d.h, gz, gx, cos(Float64(θ)), sin(Float64(θ))))
    return d
end

# truncate!(O, strategy, corr) (truncation.jl:330-337)
function truncate!(d::DeviceSPV{N,W}, strategy::TruncationStrategy,
                   corr::CorrectionAccumulator=NoCorrection()) where {N,W}
    if _is_compilable(strategy) && corr isa NoCorrection
        check(ccall((:pp_compact, LIB), Cint, (Ptr{Cvoid}, Ref{MergeFilter}), d.h, Ref(_compile_filter(strategy))))
        return d
    end
    ops = _strategy_ops!(PPStrategyOp[], strategy)
    ket, acc, vacc = _correction_args(W, corr)
    check(ccall((:pp_truncate, LIB), Cint, (Ptr{Cvoid}, Ptr{PPStrategyOp}, Cint, Ptr{Cvoid}, Ref{Float64}, Ptr{Cvoid}),
                d.h, ops, length(ops), ket, acc, vacc))
    _correction_back!(corr, acc, vacc)
    return d
end

# the clip family (spv_evolve.jl:13-26): order-preserving compaction
_compact!(d::DeviceSPV, f::MergeFilter) =
    (check(ccall((:pp_compact, LIB), Cint, (Ptr{Cvoid}, Ref{MergeFilter}), d.h, Ref(f))); d)
coeff_clip!(d::DeviceSPV, thresh::Real) = _compact!(d, _compile_filter(CoeffTruncation(Float64(thresh))))
clip!(d::DeviceSPV; thresh=1e-16) = coeff_clip!(d, thresh)
weight_clip!(d::DeviceSPV, max_weight::Int) = _compact!(d, _compile_filter(WeightTruncation(max_weight)))
weight_damped_clip!(d::DeviceSPV, alpha::Real, thresh::Real) =
    _compact!(d, _compile_filter(WeightDampedTruncation(Float64(alpha), Float64(thresh))))
x_weight_clip!(d::DeviceSPV, max_weight::Int) = _compact!(d, _compile_filter(XWeightTruncation(max_weight)))
x_weight_damped_clip!(d::DeviceSPV, alpha::Real, thresh::Real) =
    _compact!(d, _compile_filter(XWeightDampedTruncation(Float64(alpha), Float64(thresh))))
majorana_weight_clip!(d::DeviceSPV, max_weight::Int) =
    _compact!(d, _compile_filter(MajoranaWeightTruncation(max_weight)))

# stochastic_clip!(v, ε; rng) (spv_evolve.jl:35-52): the device draws from a
# counter-based generator seeded from rng
function stochastic_clip!(d::DeviceSPV{N,W}, ε::Real; rng::AbstractRNG=Random.default_rng()) where {N,W}
    ops = [PPStrategyOp(PP_OP_STOCHASTIC_COEFF, 0, NOFILTER, 0, Float64(ε), rand(rng, UInt64))]
    check(ccall((:pp_truncate, LIB), Cint, (Ptr{Cvoid}, Ptr{PPStrategyOp}, Cint, Ptr{Cvoid}, Ref{Float64}, Ptr{Cvoid}),
                d.h, ops, 1, C_NULL, Ref{Float64}(0.0), C_NULL))
    return d
end

# ---- channels (channels.jl:203-265): λ^k tables from Julia's own ^ --------------

_table(λ::Float64, N::Int) = Float64[λ^k for k in 0:N]

function _scale!(d::DeviceSPV{N,W}, kind::Integer, qubits, t0, t1=t0, t2=t0) where {N,W}
    M = _spv_mask(W, N, qubits)
    check(ccall((:pp_scale_by_popcount, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ref{W}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                d.h, kind, M, t0, t1, t2))
    return d
end

function depolarizing_channel!(d::DeviceSPV{N}, p::Real; qubits=nothing) where {N}
    0 ≤ p ≤ 1 || throw(ArgumentError("depolarizing parameter p=$p must be in [0,1]"))
    return _scale!(d, 0, qubits, _table(1 - 4p/3, N))
end
function dephasing_channel!(d::DeviceSPV{N}, p::Real; qubits=nothing) where {N}
    0 ≤ p ≤ 1 || throw(ArgumentError("dephasing parameter p=$p must be in [0,1]"))
    return _scale!(d, 1, qubits, _table(1 - 2p, N))
end
function bit_flip_channel!(d::DeviceSPV{N}, p::Real; qubits=nothing) where {N}
    0 ≤ p ≤ 1 || throw(ArgumentError("bit_flip parameter p=$p must be in [0,1]"))
    return _scale!(d, 2, qubits, _table(1 - 2p, N))
end
function bit_phase_flip_channel!(d::DeviceSPV{N}, p::Real; qubits=nothing) where {N}
    0 ≤ p ≤ 1 || throw(ArgumentError("bit_phase_flip parameter p=$p must be in [0,1]"))
    return _scale!(d, 3, qubits, _table(1 - 2p, N))
end
function pauli_channel!(d::DeviceSPV{N}, pX::Real, pY::Real, pZ::Real; qubits=nothing) where {N}
    (pX ≥ 0 && pY ≥ 0 && pZ ≥ 0) ||
        throw(ArgumentError("probabilities must be non-negative (got pX=$pX, pY=$pY, pZ=$pZ)"))
    pX + pY + pZ ≤ 1 + 4*eps(Float64) || throw(ArgumentError("pX+pY+pZ must be ≤ 1 (got $(pX+pY+pZ))"))
    λX = 1 - 2*(pY + pZ)
    λY = 1 - 2*(pX + pZ)
    λZ = 1 - 2*(pX + pY)
    return _scale!(d, 4, qubits, _table(λX, N), _table(λY, N), _table(λZ, N))
end

# ---- observables and reductions --------------------------------------------------

# ⟨ψ|O|ψ⟩ for a computational-basis ket (spv_ops.jl:490-497), of type T
function expectation_value(d::DeviceSPV{N,W,T}, ψ::Ket{N}) where {N,W,T}
    out = Ref{T}(zero(T))
    check(ccall((:pp_expectation_ket, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{T}), d.h, _ket_words(W, ψ), out))
    return out[]
end

# variance(v, ψ) (statistics.jl:6-22)
function variance(d::DeviceSPV{N,W}, ψ::Ket{N}) where {N,W}
    out = Ref{Float64}(0.0)
    check(ccall((:pp_variance_ket, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{Float64}), d.h, _ket_words(W, ψ), out))
    return out[]
end

function LinearAlgebra.norm(d::DeviceSPV, p::Real=2)
    out = Ref{Float64}(0.0)
    check(ccall((:pp_norm, LIB), Cint, (Ptr{Cvoid}, Float64, Ref{Float64}), d.h, Float64(p), out))
    return out[]
end

# Σ over shared keys of conj(c_a) c_b (spv_ops.jl:378-396)
function inner_product(a::DeviceSPV{N,W,T}, b::DeviceSPV{N,W,T}) where {N,W,T}
    out = Ref{T}(zero(T))
    check(ccall((:pp_inner_product, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{T}), a.h, b.h, out))
    return out[]
end

# isapprox(v1, v2; atol, rtol) (spv_ops.jl:88-107)
function Base.isapprox(a::DeviceSPV{N,W}, b::DeviceSPV{N,W}; atol=1e-14, rtol=0) where {N,W}
    dist = Ref{Float64}(0.0)
    check(ccall((:pp_diff_norm, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}), a.h, b.h, dist))
    return dist[] <= atol + rtol * max(norm(a), norm(b))
end

function check_spv(d::DeviceSPV)
    ok = Ref{Cint}(0)
    check(ccall((:pp_check, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}), d.h, ok))
    ok[] == 1 || throw(ArgumentError("live keys are not strictly increasing"))
    return true
end

# ---- Dict idiom (type_SparsePauliVector.jl:286-358) -------------------------------

function _lookup(sym::Symbol, d::DeviceSPV{N,W,T}, p::PauliBasis{N}) where {N,W,T}
    z, x = _pack(W, p)
    c = Ref{T}(zero(T))
    f = Ref{Int32}(0)
    rc = sym === :pp_lookup ?
        ccall((:pp_lookup, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{W}, Int64, Ref{T}, Ref{Int32}),
              d.h, z, x, 1, c, f) :
        ccall((:pp_lookup_noalloc, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{W}, Int64, Ref{T}, Ref{Int32}),
              d.h, z, x, 1, c, f)
    check(rc)
    return f[] != 0, c[]
end
function Base.get(d::DeviceSPV{N}, p::PauliBasis{N}, default) where {N}
    found, c = _lookup(:pp_lookup, d, p)
    return found ? c : default
end
function Base.getindex(d::DeviceSPV{N}, p::PauliBasis{N}) where {N}
    found, c = _lookup(:pp_lookup, d, p)
    found || throw(KeyError(p))
    return c
end
Base.haskey(d::DeviceSPV{N}, p::PauliBasis{N}) where {N} = _lookup(:pp_lookup, d, p)[1]
# get without device allocation (callers under graph capture, serving loops)
lookup_noalloc(d::DeviceSPV{N}, p::PauliBasis{N}, default) where {N} =
    (r = _lookup(:pp_lookup_noalloc, d, p); r[1] ? r[2] : default)

function Base.setindex!(d::DeviceSPV{N,W,T}, val, p::PauliBasis{N}) where {N,W,T}
    z, x = _pack(W, p)
    check(ccall((:pp_setindex, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{W}, Ref{T}, Int64),
                d.h, z, x, convert(T, val), 1))
    return d
end

function Base.delete!(d::DeviceSPV{N,W}, p::PauliBasis{N}) where {N,W}
    z, x = _pack(W, p)
    check(ccall((:pp_delete, LIB), Cint, (Ptr{Cvoid}, Ref{W}, Ref{W}, Int64), d.h, z, x, 1))
    return d
end

# sum!(v1, v2) (spv_ops.jl:182-193) and mul!(v, a) (spv_ops.jl:12-17)
function Base.sum!(a::DeviceSPV{N,W}, b::DeviceSPV{N,W}) where {N,W}
    a === b && return LinearAlgebra.mul!(a, 2)
    check(ccall((:pp_sum, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), a.h, b.h))
    return a
end
function LinearAlgebra.mul!(d::DeviceSPV, s::Real)
    check(ccall((:pp_scale, LIB), Cint, (Ptr{Cvoid}, Float64), d.h, Float64(s)))
    return d
end
# a complex scalar: Complex * Complex on ComplexF64 coefficients; a Float64
# vector takes it only when it is real, as convert(Float64, c * s) requires
function LinearAlgebra.mul!(d::DeviceSPV{N,W,T}, s::Complex) where {N,W,T}
    T <: Complex || return LinearAlgebra.mul!(d, convert(Float64, s))
    check(ccall((:pp_scale_complex, LIB), Cint, (Ptr{Cvoid}, Float64, Float64), d.h, real(s), imag(s)))
    return d
end

# ---- copy, tensor products, selections, Matrix (spv_ops.jl:109-115,342-366,520-587) --

# copy(v) (type_SparsePauliVector.jl:249-254), device to device
function Base.copy(d::DeviceSPV{N,W,T}) where {N,W,T}
    lc, ac = capacities(d)
    out = DeviceSPV{N,W,T}(lc, ac)
    check(ccall((:pp_copy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), out.h, d.h))
    return out
end

# otimes(v1, v2) = v1 ⊗ v2: v2 on the high sites, coefficients c1 * c2
function otimes(a::DeviceSPV{N,W1,T}, b::DeviceSPV{M,W2,T}) where {N,M,W1,W2,T}
    N + M <= 128 || throw(ArgumentError("N1 + N2 must be <= 128"))
    Wo = N + M <= 64 ? UInt64 : UInt128
    out = DeviceSPV{N + M,Wo,T}(16, 16)
    check(ccall((:pp_otimes, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), out.h, a.h, b.h))
    return out
end

# osum(v1, v2) = v1 ⊗ I_M + I_N ⊗ v2
function osum(a::DeviceSPV{N,W1,T}, b::DeviceSPV{M,W2,T}) where {N,M,W1,W2,T}
    I_N = DeviceSPV{N,W1,T}(16, 16)
    I_N[PauliBasis{N}(Int128(0), Int128(0))] = one(T)
    I_M = DeviceSPV{M,W2,T}(16, 16)
    I_M[PauliBasis{M}(Int128(0), Int128(0))] = one(T)
    out = otimes(a, I_M)
    return sum!(out, otimes(I_N, b))
end

function _largest(d::DeviceSPV{N,W,T}, diag_only::Bool) where {N,W,T}
    z = Ref{W}(zero(W))
    x = Ref{W}(zero(W))
    c = Ref{T}(zero(T))
    idx = Ref{Int64}(0)
    check(ccall((:pp_largest, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{W}, Ref{W}, Ref{T}, Ref{Int64}),
                d.h, diag_only, z, x, c, idx))
    return PauliBasis{N}(Int128(z[] % UInt128), Int128(x[] % UInt128)), c[]
end
# largest(v): the first term with the largest |c|, as a one-term vector
function largest(d::DeviceSPV{N,W,T}) where {N,W,T}
    p, c = _largest(d, false)
    out = DeviceSPV{N,W,T}(16, 16)
    out[p] = c
    return out
end
# largest_diag(v): PauliBasis => c of the diagonal term with the largest |c|
largest_diag(d::DeviceSPV) = Pair(_largest(d, true)...)

# diag(v) / offdiag(v): copies holding the x == 0 (x != 0) terms
LinearAlgebra.diag(d::DeviceSPV) =
    (out = copy(d); check(ccall((:pp_select_diag, LIB), Cint, (Ptr{Cvoid}, Cint), out.h, 0)); out)
offdiag(d::DeviceSPV) =
    (out = copy(d); check(ccall((:pp_select_diag, LIB), Cint, (Ptr{Cvoid}, Cint), out.h, 1)); out)

# Matrix(v) for N <= 13, Julia's column-major layout
function Base.Matrix(d::DeviceSPV{N,W,T}) where {N,W,T}
    N <= 13 || throw(ArgumentError("Matrix(v) is limited to N <= 13"))
    out = Matrix{T}(undef, 2^N, 2^N)
    check(ccall((:pp_matrix, LIB), Cint, (Ptr{Cvoid}, Ptr{T}), d.h, out))
    return out
end

# ---- products (spv_ops.jl:255-484) on host SparsePauliVectors ---------------------

const _contexts = Dict{Int,DeviceSPV}()   # one device context per word count

function _algebra(op::Integer, A::SparsePauliVector{N,W,T}, ca::Bool,
                  B::SparsePauliVector{N,W,T}, cb::Bool) where {N,W,T}
    ctx = get!(_contexts, sizeof(W) ÷ 8) do
        DeviceSPV{64 * (sizeof(W) ÷ 8), W}(16, 16)
    end
    cplx = T <: Complex
    ar = Float64[real(A.c[i]) for i in 1:A.n]
    ai = Float64[imag(A.c[i]) for i in 1:A.n]
    br = Float64[real(B.c[i]) for i in 1:B.n]
    bi = Float64[imag(B.c[i]) for i in 1:B.n]
    n = Ref{Int64}(0)
    local re, im, out
    GC.@preserve A B ar ai br bi begin
        ta = Ref(PPTerms(Ptr{UInt64}(pointer(A.z)), Ptr{UInt64}(pointer(A.x)), pointer(ar), pointer(ai), A.n, ca))
        tb = Ref(PPTerms(Ptr{UInt64}(pointer(B.z)), Ptr{UInt64}(pointer(B.x)), pointer(br), pointer(bi), B.n, cb))
        rc = ccall((:pp_algebra, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ref{PPTerms}, Ref{PPTerms}, Ptr{W}, Ptr{W}, Ptr{Float64}, Ptr{Float64},
                    Int64, Ref{Int64}),
                   ctx.h, op, cplx, ta, tb, C_NULL, C_NULL, C_NULL, C_NULL, 0, n)   # size query
        rc in (PP_OK, PP_EOVERFLOW) || check(rc)
        out = SparsePauliVector(N, T; capacity=max(n[], 1))
        re = Vector{Float64}(undef, n[])
        im = Vector{Float64}(undef, n[])
        check(ccall((:pp_algebra, LIB), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Ref{PPTerms}, Ref{PPTerms}, Ptr{W}, Ptr{W}, Ptr{Float64}, Ptr{Float64},
                     Int64, Ref{Int64}),
                    ctx.h, op, cplx, ta, tb, out.z, out.x, re, im, n[], n))
    end
    for i in 1:n[]
        out.c[i] = cplx ? complex(re[i], im[i]) : re[i]
    end
    out.n = n[]
    return out
end
Base.:*(A::SparsePauliVector, B::SparsePauliVector) = _algebra(PP_ALG_MUL, A, false, B, false)
Base.:*(A::Adjoint{<:Any,<:SparsePauliVector}, B::SparsePauliVector) = _algebra(PP_ALG_MUL, parent(A), true, B, false)
Base.:*(A::SparsePauliVector, B::Adjoint{<:Any,<:SparsePauliVector}) = _algebra(PP_ALG_MUL, A, false, parent(B), true)
Base.:*(A::Adjoint{<:Any,<:SparsePauliVector}, B::Adjoint{<:Any,<:SparsePauliVector}) =
    _algebra(PP_ALG_MUL, parent(A), true, parent(B), true)
commutator(A::SparsePauliVector, B::SparsePauliVector) = _algebra(PP_ALG_COMMUTATOR, A, false, B, false)
anticommutator(A::SparsePauliVector, B::SparsePauliVector) = _algebra(PP_ALG_ANTICOMMUTATOR, A, false, B, false)

# ---- batches: many evolve! calls at once (no reference counterpart) -------------

mutable struct DeviceSPVBatch{W}
    h::Ptr{Cvoid}
    Ns::Vector{Int}
end

function _destroy!(b::DeviceSPVBatch)
    b.h == C_NULL && return nothing
    ccall((:pp_batch_destroy, LIB), Cint, (Ptr{Cvoid},), b.h)
    b.h = C_NULL
    return nothing
end

_nqubits(::SparsePauliVector{N}) where {N} = N

function DeviceSPVBatch(vs::Vector{<:SparsePauliVector{<:Any,W,Float64}}) where {W}
    B = length(vs)
    Ns = Cint[_nqubits(v) for v in vs]
    lc = Int64[length(v.z) for v in vs]
    ac = Int64[length(v.az) for v in vs]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pp_batch_create, LIB), Cint, (Cint, Ptr{Cint}, Ptr{Int64}, Ptr{Int64}, Ref{Ptr{Cvoid}}),
                B, Ns, lc, ac, h))
    bt = DeviceSPVBatch{W}(h[], Int.(Ns))
    finalizer(_destroy!, bt)
    wd = Ref{Cint}(0)
    check(ccall((:pp_batch_words, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}), bt.h, wd))
    wd[] == sizeof(W) ÷ 8 || throw(ArgumentError("batch word count $(wd[]) differs from the vectors' $(sizeof(W) ÷ 8)"))
    for (b, v) in enumerate(vs)
        check(ccall((:pp_batch_upload, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{W}, Ptr{W}, Ptr{Float64}, Int64, Cint),
                    bt.h, b - 1, v.z, v.x, v.c, v.n, 1))
    end
    return bt
end

function sizes(bt::DeviceSPVBatch)
    n = zeros(Int64, length(bt.Ns))
    check(ccall((:pp_batch_sizes, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}), bt.h, n))
    return Int.(n)
end

function PauliOperators.SparsePauliVector(bt::DeviceSPVBatch{W}, b::Integer) where {W}
    n = sizes(bt)[b]
    v = SparsePauliVector(bt.Ns[b], Float64; capacity=max(n, 1))
    got = Ref{Int64}(0)
    check(ccall((:pp_batch_download, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{W}, Ptr{W}, Ptr{Float64}, Int64, Ref{Int64}),
                bt.h, b - 1, v.z, v.x, v.c, length(v.z), got))
    v.n = Int(got[])
    return v
end

# one evolve!(v_b, generators[b], angles[b]; truncation = strategies[b]) per
# vector (window = 1); shorter sequences are padded with inactive steps
function evolve!(bt::DeviceSPVBatch{W}, generators::Vector{<:Vector{<:PauliBasis}},
                 angles::Vector{<:Vector{<:Real}}; truncation=NoTruncation(),
                 local_truncation=NoTruncation(), counters::Bool=false) where {W}
    B = length(bt.Ns)
    length(generators) == B == length(angles) || throw(DimensionMismatch("one sequence per vector"))
    S = maximum(length, generators; init=0)
    S == 0 && return counters ? fill(PPCounters(), B) : bt
    strict = truncation isa AbstractVector ? truncation : fill(truncation, B)
    loc = local_truncation isa AbstractVector ? local_truncation : fill(local_truncation, B)
    all(_is_compilable, strict) || throw(ArgumentError("batch truncation must compile to a merge filter"))
    all(_is_compilable, loc) || throw(ArgumentError("local_truncation must be a deterministic " *
                                                    "(weight/coefficient) strategy — it runs per append"))
    gz = zeros(W, S * B)
    gx = zeros(W, S * B)
    cosv = ones(Float64, S * B)
    sinv = zeros(Float64, S * B)
    act = zeros(UInt8, S * B)
    for b in 1:B
        length(generators[b]) == length(angles[b]) ||
            throw(DimensionMismatch("generators and angles must have same length"))
        for s in 1:length(generators[b])
            k = (s - 1) * B + b
            gz[k], gx[k] = _pack(W, generators[b][s])
            cosv[k] = cos(Float64(angles[b][s]))
            sinv[k] = sin(Float64(angles[b][s]))
            act[k] = 1
        end
    end
    fs = MergeFilter[_compile_filter(s) for s in strict]
    fl = MergeFilter[_compile_filter(s) for s in loc]
    cnt = fill(PPCounters(), B)
    check(ccall((:pp_batch_evolve, LIB), Cint,
                (Ptr{Cvoid}, Int64, Ptr{W}, Ptr{W}, Ptr{Float64}, Ptr{Float64}, Ptr{UInt8},
                 Ptr{MergeFilter}, Ptr{MergeFilter}, Ptr{PPCounters}),
                bt.h, S, gz, gx, cosv, sinv, act, fs, fl, cnt))
    return counters ? cnt : bt
end

# channels on every vector: kind 0..4 as pp_scale_by_popcount, -1 = none
function _batch_scale!(bt::DeviceSPVBatch{W}, kind::Integer, lams::NTuple{3,Float64}, qubits) where {W}
    B = length(bt.Ns)
    kinds = fill(Cint(kind), B)
    masks = Vector{W}(undef, B)
    tabs = zeros(Float64, 129 * 3 * B)
    for b in 1:B
        N = bt.Ns[b]
        masks[b] = _spv_mask(W, N, qubits)
        for t in 1:3, k in 0:N
            tabs[((b - 1) * 3 + (t - 1)) * 129 + k + 1] = lams[t]^k
        end
    end
    check(ccall((:pp_batch_scale_by_popcount, LIB), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ptr{W}, Ptr{Float64}),
                bt.h, kinds, masks, tabs))
    return bt
end
depolarizing_channel!(bt::DeviceSPVBatch, p::Real; qubits=nothing) =
    (0 ≤ p ≤ 1 || throw(ArgumentError("depolarizing parameter p=$p must be in [0,1]"));
     _batch_scale!(bt, 0, (1 - 4p/3, 1.0, 1.0), qubits))
dephasing_channel!(bt::DeviceSPVBatch, p::Real; qubits=nothing) =
    (0 ≤ p ≤ 1 || throw(ArgumentError("dephasing parameter p=$p must be in [0,1]"));
     _batch_scale!(bt, 1, (1 - 2p, 1.0, 1.0), qubits))
bit_flip_channel!(bt::DeviceSPVBatch, p::Real; qubits=nothing) =
    (0 ≤ p ≤ 1 || throw(ArgumentError("bit_flip parameter p=$p must be in [0,1]"));
     _batch_scale!(bt, 2, (1 - 2p, 1.0, 1.0), qubits))
bit_phase_flip_channel!(bt::DeviceSPVBatch, p::Real; qubits=nothing) =
    (0 ≤ p ≤ 1 || throw(ArgumentError("bit_phase_flip parameter p=$p must be in [0,1]"));
     _batch_scale!(bt, 3, (1 - 2p, 1.0, 1.0), qubits))
function pauli_channel!(bt::DeviceSPVBatch, pX::Real, pY::Real, pZ::Real; qubits=nothing)
    (pX ≥ 0 && pY ≥ 0 && pZ ≥ 0) ||
        throw(ArgumentError("probabilities must be non-negative (got pX=$pX, pY=$pY, pZ=$pZ)"))
    pX + pY + pZ ≤ 1 + 4*eps(Float64) || throw(ArgumentError("pX+pY+pZ must be ≤ 1 (got $(pX+pY+pZ))"))
    return _batch_scale!(bt, 4, (1 - 2*(pY + pZ), 1 - 2*(pX + pZ), 1 - 2*(pX + pY)), qubits)
end

function expectation_value(bt::DeviceSPVBatch{W}, ψs::Vector{<:Ket}) where {W}
    B = length(bt.Ns)
    length(ψs) == B || throw(DimensionMismatch("one ket per vector"))
    kets = W[(ψ.v % UInt128) % W for ψ in ψs]
    out = zeros(Float64, B)
    check(ccall((:pp_batch_expectation_ket, LIB), Cint, (Ptr{Cvoid}, Ptr{W}, Ptr{Float64}), bt.h, kets, out))
    return out
end

reserve!(bt::DeviceSPVBatch, cap::Integer) =
    (check(ccall((:pp_batch_reserve, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}), bt.h, fill(Int64(cap), length(bt.Ns)))); bt)
set_tiles!(bt::DeviceSPVBatch, tile::Integer=0, merge_tile::Integer=0) =
    (check(ccall((:pp_batch_set_tiles, LIB), Cint, (Ptr{Cvoid}, Cint, Cint), bt.h, tile, merge_tile)); bt)
set_graphs!(bt::DeviceSPVBatch, on::Bool) =
    (check(ccall((:pp_batch_set_graphs, LIB), Cint, (Ptr{Cvoid}, Cint), bt.h, on)); bt)
set_timing!(bt::DeviceSPVBatch, on::Bool) =
    (check(ccall((:pp_batch_set_timing, LIB), Cint, (Ptr{Cvoid}, Cint), bt.h, on)); bt)
set_epoch!(bt::DeviceSPVBatch, epoch::Integer) =
    (check(ccall((:pp_batch_set_epoch, LIB), Cint, (Ptr{Cvoid}, UInt32), bt.h, epoch)); bt)
set_resident!(bt::DeviceSPVBatch, mode::Integer=1, cap_limit::Integer=0) =
    (check(ccall((:pp_batch_set_resident, LIB), Cint, (Ptr{Cvoid}, Cint, Cint), bt.h, mode, cap_limit)); bt)

function graph_stats(bt::DeviceSPVBatch)
    c = Ref{Int64}(0)
    l = Ref{Int64}(0)
    check(ccall((:pp_batch_graph_stats, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), bt.h, c, l))
    return (captures=c[], launches=l[])
end
function merge_time(bt::DeviceSPVBatch; reset::Bool=false)
    s = Ref{Float64}(0.0)
    l = Ref{Int64}(0)
    check(ccall((:pp_batch_merge_time, LIB), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}, Cint), bt.h, s, l, reset))
    return (seconds=s[], launches=l[])
end
function resident_stats(bt::DeviceSPVBatch)
    a = Ref{Int64}(0)
    f = Ref{Int64}(0)
    x = Ref{Int64}(0)
    check(ccall((:pp_batch_resident_stats, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                bt.h, a, f, x))
    return (launches=a[], fallbacks=f[], aborts=x[])
end
function resident_time(bt::DeviceSPVBatch; reset::Bool=false)
    s = Ref{Float64}(0.0)
    check(ccall((:pp_batch_resident_time, LIB), Cint, (Ptr{Cvoid}, Ref{Float64}, Cint), bt.h, s, reset))
    return s[]
end
function resident_work(bt::DeviceSPVBatch; reset::Bool=false)
    r = Ref{Int64}(0)
    s = Ref{Int64}(0)
    o = Ref{Int64}(0)
    q = Ref{Int64}(0)
    check(ccall((:pp_batch_resident_work, LIB), Cint,
                (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ref{Int64}, Cint), bt.h, r, s, o, q, reset))
    return (rotations=r[], swept=s[], merge_out=o[], rebalances=q[])
end

# ---- learned surrogate of ⟨O⟩(θ) (no reference counterpart) ----------------------

mutable struct SurrogateMLP
    h::Ptr{Cvoid}
    E::Int
    din::Int
    count::Int
end

function _destroy!(m::SurrogateMLP)
    m.h == C_NULL && return nothing
    ccall((:pp_mlp_destroy, LIB), Cint, (Ptr{Cvoid},), m.h)
    m.h = C_NULL
    return nothing
end

function SurrogateMLP(E::Integer, din::Integer, hidden::Integer=512, layers::Integer=3)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pp_mlp_create, LIB), Cint, (Cint, Cint, Cint, Cint, Ref{Ptr{Cvoid}}), E, din, hidden, layers, h))
    n = Ref{Int64}(0)
    check(ccall((:pp_mlp_param_count, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), h[], n))
    m = SurrogateMLP(h[], E, din, Int(n[]))
    finalizer(_destroy!, m)
    return m
end

set_params!(m::SurrogateMLP, p::Vector{Float32}) =
    (length(p) == m.E * m.count || throw(DimensionMismatch("E x param_count floats expected"));
     check(ccall((:pp_mlp_set_params, LIB), Cint, (Ptr{Cvoid}, Ptr{Float32}), m.h, p)); m)
function get_params(m::SurrogateMLP)
    p = Vector{Float32}(undef, m.E * m.count)
    check(ccall((:pp_mlp_get_params, LIB), Cint, (Ptr{Cvoid}, Ptr{Float32}), m.h, p))
    return p
end
load_data!(m::SurrogateMLP, x::Matrix{Float32}, y::Vector{Float32}) =
    (size(x, 1) == m.din && size(x, 2) == length(y) || throw(DimensionMismatch("x is din x n, y has n entries"));
     check(ccall((:pp_mlp_load_data, LIB), Cint, (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Int64), m.h, x, y, length(y)));
     m)
function train_epoch!(m::SurrogateMLP, perm::Vector{Int32}, batch::Integer, lr::Real)
    loss = zeros(Float32, m.E)
    check(ccall((:pp_mlp_train_epoch, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Cint, Float32, Ptr{Float32}),
                m.h, perm, batch, lr, loss))
    return loss
end
function gradients(m::SurrogateMLP, x::Matrix{Float32}, t::Vector{Float32})
    g = Vector{Float32}(undef, m.E * m.count)
    loss = zeros(Float32, m.E)
    check(ccall((:pp_mlp_gradients, LIB), Cint, (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Cint, Ptr{Float32}, Ptr{Float32}),
                m.h, x, t, length(t), g, loss))
    return g, loss
end
function predict(m::SurrogateMLP, x::Matrix{Float32})
    n = size(x, 2)
    mean = zeros(Float64, n)
    sd = zeros(Float64, n)
    t = Ref{Float64}(0.0)
    check(ccall((:pp_mlp_predict, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{Float32}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                m.h, x, n, C_NULL, mean, sd, t))
    return mean, sd
end

end # module
