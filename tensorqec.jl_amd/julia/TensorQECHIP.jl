"""
    TensorQECHIP

Julia binding of `libtqec_hip.so` for TensorQEC.jl, two drop-in decoders:

  * `TNMMAPHIP`: `compile` builds the reference's factor graph and
    contraction tree (`compile(::TNMMAP, problem)`, src/decoding/tndecoder.jl:97-146
    for CSS codes and :186-219 for detector error models) and hands the tree
    to the HIP library as a flat post-order step list; `decode` /
    `decode_batch` return `DecodingResult`s with the reference's layout
    (src/decoding/interfaces.jl:116-119).  The pure error and the logical
    fix reuse the reference's own `error_pattern`
    (src/decoding/tndecoder.jl:167-174, 255-271), so a result is the
    reference's result whenever the decided class is.
  * `TNMAPHIP`: the most probable error configuration
    (`compile(::TNMAP, ::GeneralDecodingProblem)` / `decode`,
    src/decoding/tndecoder.jl:46-57): the reference's UAI model
    (`stg2uaimodel`, :33-40) with the syndrome variables as sliced labels,
    contracted by the library in the max-product semiring with a per-syndrome
    arg-max backtrace (`tqec_map_decode_batch`).  CSS and classical problems
    reach it through the reference's own reductions (interfaces.jl:123-142).

This file is source only: it needs a Julia environment with TensorQEC and
OMEinsum, which this build does not have.
"""
module TensorQECHIP

using TensorQEC
using TensorQEC.OMEinsum

export TNMMAPHIP, TNMAPHIP, decode_batch

const libtqec = joinpath(@__DIR__, "..", "tensorqec", "libtqec_hip.so")

const MODE_PLAIN_F64 = Int32(0)
const MODE_FUSED_F64 = Int32(1)
const MODE_FUSED_F32 = Int32(2)
const MODE_WIDE_F32 = Int32(3)
const MODE_SPEC_F32 = Int32(4)     # plan-specialised kernel (hiprtc), FP32 with FP64 re-check of low margins
const MODE_SPEC_F64 = Int32(5)
const MODE_SPEC_BF16 = Int32(6)    # GEMM-shaped steps on BF16 MFMA (circuit-level DEMs)

const SEMIRING_SUM = Int32(0)      # marginals (TNMMAP)
const SEMIRING_MAX = Int32(1)      # most probable configuration (TNMAP)

# Mirrors `tqec_plan_desc` in include/tqec_hip.h field by field.
struct PlanDesc
    n_vars::Int32
    n_leaves::Int32
    leaf_rank::Ptr{Int32}
    leaf_labels::Ptr{Int32}
    leaf_data::Ptr{Float64}
    leaf_evidence::Ptr{Int32}
    n_steps::Int32
    step_lhs::Ptr{Int32}
    step_rhs::Ptr{Int32}
    step_out_rank::Ptr{Int32}
    step_out_labels::Ptr{Int32}
    n_out::Int32
    out_labels::Ptr{Int32}
    n_syndrome_bits::Int32
    n_sliced::Int32
    sliced_label::Ptr{Int32}
    sliced_bit::Ptr{Int32}
    tabulate_bits::Int32
    semiring::Int32
end

last_error() = unsafe_string(ccall((:tqec_last_error, libtqec), Cstring, ()))
check(rc) = rc == 0 || error("tqec error $rc: $(last_error())")

mutable struct HipPlan
    handle::Ptr{Cvoid}
    n_out::Int
    n_syn::Int
    function HipPlan(h, n_out, n_syn)
        p = new(h, n_out, n_syn)
        finalizer(p -> ccall((:tqec_plan_destroy, libtqec), Cvoid, (Ptr{Cvoid},), p.handle), p)
        return p
    end
end

"""
    flatten(code::NestedEinsum, tensors; evidence, sliced, ixs) -> (leaf ranks, labels, data, evidence, steps...)

Post-order list of the binary contractions of an OMEinsum `NestedEinsum`
(leaves keep their `tensorindex`; step i becomes node `n_leaves + i`).
`evidence` maps a tensor index to the syndrome bit (0-based) whose value
selects its one-hot vector: the evidence vectors of `update_syndrome!`
(tndecoder.jl:148-158, 240-246).  `sliced` maps a label to the syndrome bit
that fixes it (TNMAP's evidence variables); such labels stay on the leaves
(`ixs`, by default the tree's own leaf labels) and never appear in the
tree.  Labels are shifted to 0-based.
"""
function flatten(code, tensors; evidence::Dict{Int,Int} = Dict{Int,Int}(),
                 sliced::Dict{Int,Int} = Dict{Int,Int}(), ixs = OMEinsum.getixsv(code))
    L = length(tensors)
    lhs, rhs, outr, outl = Int32[], Int32[], Int32[], Int32[]
    function visit(ne)::Int32
        OMEinsum.isleaf(ne) && return Int32(ne.tensorindex - 1)
        length(ne.args) == 2 || error("the contraction tree must be binary")
        a = visit(ne.args[1])
        b = visit(ne.args[2])
        iy = OMEinsum.getiyv(ne.eins)
        push!(lhs, a); push!(rhs, b)
        push!(outr, Int32(length(iy)))
        append!(outl, Int32.(iy .- 1))
        return Int32(L + length(lhs) - 1)
    end
    visit(code)
    ranks = Int32[length(ix) for ix in ixs]
    labels = Int32[l - 1 for ix in ixs for l in ix]
    data = Float64[x for t in tensors for x in vec(t)]
    ev = fill(Int32(-1), L)
    for (t, bit) in evidence
        ev[t] = Int32(bit)
    end
    sl = sort!(collect(sliced))
    sliced_label = Int32[l - 1 for (l, _) in sl]
    sliced_bit = Int32[b for (_, b) in sl]
    iy = Int32.(OMEinsum.getiyv(code) .- 1)
    nvars = maximum(vcat(labels, outl, iy, sliced_label)) + 1
    return (; ranks, labels, data, evidence = ev, lhs, rhs, outr, outl, iy, nvars, sliced_label, sliced_bit)
end

function HipPlan(code, tensors, n_syndrome::Int; tabulate_bits::Int = 4, evidence::Dict{Int,Int} = Dict{Int,Int}(),
                 sliced::Dict{Int,Int} = Dict{Int,Int}(), ixs = OMEinsum.getixsv(code), semiring::Int32 = SEMIRING_SUM)
    f = flatten(code, tensors; evidence, sliced, ixs)
    GC.@preserve f begin
        desc = PlanDesc(Int32(f.nvars), Int32(length(f.ranks)), pointer(f.ranks), pointer(f.labels),
                        pointer(f.data), pointer(f.evidence), Int32(length(f.lhs)), pointer(f.lhs),
                        pointer(f.rhs), pointer(f.outr), pointer(f.outl), Int32(length(f.iy)), pointer(f.iy),
                        Int32(n_syndrome), Int32(length(f.sliced_label)),
                        isempty(f.sliced_label) ? Ptr{Int32}(C_NULL) : pointer(f.sliced_label),
                        isempty(f.sliced_bit) ? Ptr{Int32}(C_NULL) : pointer(f.sliced_bit),
                        Int32(tabulate_bits), semiring)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tqec_plan_create, libtqec), Cint, (Ref{PlanDesc}, Ref{Ptr{Cvoid}}), desc, h))
    end
    return HipPlan(h[], length(f.iy), n_syndrome)
end

# The last n tensors of a compiled TNMMAP network are its evidence vectors,
# set per syndrome bit in order by `update_syndrome!`.
last_as_evidence(L::Int, n::Int) = Dict(L - n + i => i - 1 for i in 1:n)

"""
    TNMMAPHIP(; optimizer=TreeSA(), factorize=true, mode=MODE_SPEC_F32)

`TNMMAP` (src/decoding/tndecoder.jl:77-80) executed by the HIP library.
`factorize` applies to detector error models, as in the reference (its CSS
network is never factorized).
"""
Base.@kwdef struct TNMMAPHIP <: TensorQEC.AbstractGeneralDecoder
    optimizer::OMEinsum.CodeOptimizer = TreeSA()
    factorize::Bool = true
    mode::Int32 = MODE_SPEC_F32
end

struct CompiledTNMMAPHIP <: TensorQEC.CompiledDecoder
    ct::TensorQEC.CompiledTNMMAP
    plan::HipPlan
    mode::Int32
end

tabulation(mode) = mode == MODE_PLAIN_F64 ? 0 : 4

function TensorQEC.compile(decoder::TNMMAPHIP, problem::TensorQEC.IndependentDepolarizingDecodingProblem)
    ct = TensorQEC.compile(TNMMAP(; optimizer = decoder.optimizer, factorize = false), problem)
    ns = length(ct.syndrome_indices)
    plan = HipPlan(ct.code, ct.tensors, ns; tabulate_bits = tabulation(decoder.mode),
                   evidence = last_as_evidence(length(ct.tensors), ns))
    return CompiledTNMMAPHIP(ct, plan, decoder.mode)
end

struct CompiledDEMTNMMAPHIP <: TensorQEC.CompiledDecoder
    ct::TensorQEC.CompiledDEMTNMMAP
    plan::HipPlan
    mode::Int32
end

# compile(::TNMMAP, ::DetectorErrorModel), src/decoding/tndecoder.jl:186-219
function TensorQEC.compile(decoder::TNMMAPHIP, dem::TensorQEC.DetectorErrorModel)
    ct = TensorQEC.compile(TNMMAP(; optimizer = decoder.optimizer, factorize = decoder.factorize), dem)
    ns = length(ct.syndrome_indices)
    plan = HipPlan(ct.code, ct.tensors, ns; tabulate_bits = tabulation(decoder.mode),
                   evidence = last_as_evidence(length(ct.tensors), ns))
    return CompiledDEMTNMMAPHIP(ct, plan, decoder.mode)
end

syndrome_bytes(s::CSSSyndrome) = UInt8[x.x for x in vcat(s.sx, s.sz)]
syndrome_bytes(s::SimpleSyndrome) = UInt8[x.x for x in s.s]

# B syndromes as the library's row-major B x nsyn bytes (column b = syndrome b).
function syndrome_rows(syndromes, nsyn::Int)
    rows = Matrix{UInt8}(undef, nsyn, length(syndromes))
    for (b, s) in enumerate(syndromes)
        rows[:, b] = syndrome_bytes(s)
    end
    return rows
end

# Arg-max classes of a batch (0-based, column-major over the n_out class bits).
function decode_classes(plan::HipPlan, mode::Int32, syndromes)
    B = length(syndromes)
    rows = syndrome_rows(syndromes, plan.n_syn)
    mar = Matrix{Float64}(undef, 1 << plan.n_out, B)
    cls = Vector{Int32}(undef, B)
    tie = Vector{UInt8}(undef, B)
    check(ccall((:tqec_decode_batch, libtqec), Cint,
                (Ptr{Cvoid}, Ptr{UInt8}, Int64, Int32, Ptr{Float64}, Ptr{Int32}, Ptr{UInt8}, Ptr{Cvoid}),
                plan.handle, rows, B, mode, mar, cls, tie, C_NULL))
    return cls
end

"""
    decode_batch(c::CompiledTNMMAPHIP, syndromes::Vector{CSSSyndrome})

One library call for the whole batch; returns `DecodingResult`s.
"""
function decode_batch(c::CompiledTNMMAPHIP, syndromes::Vector{CSSSyndrome})
    cls = decode_classes(c.plan, c.mode, syndromes)
    shape = ntuple(_ -> 2, c.plan.n_out)
    return map(eachindex(syndromes)) do b
        pos = CartesianIndices(shape)[cls[b] + 1]      # Julia findmax position (column-major)
        DecodingResult(true, TensorQEC.error_pattern(pos, c.ct, syndromes[b]))
    end
end

TensorQEC.decode(c::CompiledTNMMAPHIP, s::CSSSyndrome) = decode_batch(c, [s])[1]

"""
    decode_batch(c::CompiledDEMTNMMAPHIP, syndromes::Vector{SimpleSyndrome})

Detector error model: one library call for the batch.  `findmax` of the
reference's marginal is a linear index for one logical and a CartesianIndex
for several (src/decoding/tndecoder.jl:248-253); `error_pattern` (:255-271)
is the reference's.
"""
function decode_batch(c::CompiledDEMTNMMAPHIP, syndromes::Vector{SimpleSyndrome})
    cls = decode_classes(c.plan, c.mode, syndromes)
    nl = c.plan.n_out
    shape = ntuple(_ -> 2, nl)
    return map(eachindex(syndromes)) do b
        pos = nl == 1 ? Int(cls[b]) + 1 : CartesianIndices(shape)[cls[b] + 1]
        DecodingResult(true, TensorQEC.error_pattern(pos, c.ct, syndromes[b]))
    end
end

TensorQEC.decode(c::CompiledDEMTNMMAPHIP, s::SimpleSyndrome) = decode_batch(c, [s])[1]

"""
    TNMAPHIP(; optimizer=TreeSA())

`TNMAP` (src/decoding/tndecoder.jl:9-11) executed by the HIP library.
"""
Base.@kwdef struct TNMAPHIP <: TensorQEC.AbstractGeneralDecoder
    optimizer::OMEinsum.CodeOptimizer = TreeSA()
end

struct CompiledTNMAPHIP <: TensorQEC.CompiledDecoder
    plan::HipPlan
    qubit_num::Int
end

# compile(::TNMAP, ::GeneralDecodingProblem), src/decoding/tndecoder.jl:46-50:
# the same UAI model; the evidence variables nq+1..nq+ns become labels fixed
# by syndrome bits 0..ns-1 instead of TensorInference evidence.
function TensorQEC.compile(decoder::TNMAPHIP, problem::TensorQEC.GeneralDecodingProblem)
    uai = TensorQEC.stg2uaimodel(problem.tanner, problem.ptn)
    nq, ns = problem.tanner.nq, problem.tanner.ns
    ixs = [collect(Int, f.vars) for f in uai.factors]
    tensors = [f.vals for f in uai.factors]
    sliced = Dict(nq + s => s - 1 for s in 1:ns)
    code = OMEinsum.DynamicEinCode([filter(l -> !haskey(sliced, l), ix) for ix in ixs], Int[])
    code = optimize_code(code, uniformsize(code, 2), decoder.optimizer)
    plan = HipPlan(code, tensors, ns; tabulate_bits = 0, sliced, ixs, semiring = SEMIRING_MAX)
    return CompiledTNMAPHIP(plan, nq)
end

"""
    decode_batch(c::CompiledTNMAPHIP, syndromes::Vector{SimpleSyndrome})

Most probable configurations of the error variables for a batch, one
library call (`tqec_map_decode_batch`); `decode` (tndecoder.jl:53-57) returns
`DecodingResult(true, Mod2.(config[1:qubit_num]))`.
"""
function decode_batch(c::CompiledTNMAPHIP, syndromes::Vector{SimpleSyndrome})
    B = length(syndromes)
    rows = syndrome_rows(syndromes, c.plan.n_syn)
    cfg = Matrix{UInt8}(undef, c.qubit_num, B)    # column b = configuration b (row-major B x qubit_num in C)
    logp = Vector{Float64}(undef, B)
    check(ccall((:tqec_map_decode_batch, libtqec), Cint,
                (Ptr{Cvoid}, Ptr{UInt8}, Int64, Int32, Ptr{UInt8}, Ptr{Float64}, Ptr{Cvoid}),
                c.plan.handle, rows, B, Int32(c.qubit_num), cfg, logp, C_NULL))
    return [DecodingResult(true, Mod2.(cfg[:, b] .!= 0)) for b in 1:B]
end

TensorQEC.decode(c::CompiledTNMAPHIP, s::SimpleSyndrome) = decode_batch(c, [s])[1]

end # module
