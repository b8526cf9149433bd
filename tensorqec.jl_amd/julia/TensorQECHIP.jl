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
    # typed, so that it is more specific than the outer HipPlan(code, tensors, n_syndrome::Int)
    function HipPlan(h::Ptr{Cvoid}, n_out::Int, n_syn::Int)
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
(`ixs`, by default the tree's own leaf labels) and are dropped from every
intermediate (an OMEinsum `SlicedEinsum` keeps its sliced labels in the
inner tree).  Labels are shifted to 0-based.
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
        iy = filter(l -> !haskey(sliced, l), OMEinsum.getiyv(ne.eins))
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

"""
    HipPlan(code::SlicedEinsum, tensors, n_syndrome; kwargs...)

A contraction with sliced labels (`optimize_code(...; slicer = TreeSASlicer(...))`,
the reference's slicer hook, src/nonclifford/tensornetwork.jl:174-192): the
j-th sliced label becomes syndrome bit `n_syndrome + j - 1` of the inner
tree's plan, so one library row decodes one slice; `sliced_rows` expands
syndromes into their 2^k slice rows and the marginals of a syndrome are
the sum over its rows.
"""
function HipPlan(code::OMEinsum.SlicedEinsum, tensors, n_syndrome::Int; tabulate_bits::Int = 4,
                 evidence::Dict{Int,Int} = Dict{Int,Int}(), semiring::Int32 = SEMIRING_SUM)
    semiring == SEMIRING_SUM || error("sliced contractions are summed: max-product plans cannot be sliced")
    k = length(code.slicing)
    sliced = Dict(l => n_syndrome + j - 1 for (j, l) in enumerate(code.slicing))
    p = HipPlan(code.eins, tensors, n_syndrome + k; tabulate_bits, evidence, sliced,
                ixs = OMEinsum.getixsv(code.eins))
    # a slice's top-2 margin decides nothing: only rows that underflowed are
    # re-checked by the library; close syndromes are re-decided on the sum
    check(ccall((:tqec_plan_set_tolerances, libtqec), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), p.handle, 1e-12, 0.0))
    return SlicedPlan(p, k, n_syndrome, p.n_out)
end

struct SlicedPlan
    plan::HipPlan
    k::Int          # sliced labels; 2^k rows per syndrome
    n_syn::Int      # syndrome bits of the unsliced problem
    n_out::Int      # class bits, as HipPlan's
end

# Row r = (b - 1) * 2^k + s of a sliced batch: syndrome b followed by bit j of
# the slice index s as syndrome bit n_syn + j.
function sliced_rows(rows::Matrix{UInt8}, k::Int)
    n_syn, B = size(rows)
    m = 1 << k
    out = Matrix{UInt8}(undef, n_syn + k, B * m)
    for b in 1:B, s in 0:m-1
        c = (b - 1) * m + s + 1
        out[1:n_syn, c] = rows[:, b]
        for j in 0:k-1
            out[n_syn + j + 1, c] = UInt8((s >> j) & 1)
        end
    end
    return out
end

# The last n tensors of a compiled TNMMAP network are its evidence vectors,
# set per syndrome bit in order by `update_syndrome!`.
last_as_evidence(L::Int, n::Int) = Dict(L - n + i => i - 1 for i in 1:n)

"""
    TNMMAPHIP(; optimizer=TreeSA(), factorize=true, mode=MODE_SPEC_F32, slicer=nothing)

`TNMMAP` (src/decoding/tndecoder.jl:77-80) executed by the HIP library.
`factorize` applies to detector error models, as in the reference (its CSS
network is never factorized).  `slicer` (e.g.
`TreeSASlicer(score = ScoreFunction(sc_target = 21))`) re-optimizes the
reference's tree with sliced labels for networks too wide for one tensor
per workgroup (toric L=6, multi-round circuit-level models); the compiled
decoder then holds a `SlicedPlan`.
"""
Base.@kwdef struct TNMMAPHIP <: TensorQEC.AbstractGeneralDecoder
    optimizer::OMEinsum.CodeOptimizer = TreeSA()
    factorize::Bool = true
    mode::Int32 = MODE_SPEC_F32
    slicer = nothing
end

struct CompiledTNMMAPHIP{P} <: TensorQEC.CompiledDecoder
    ct::TensorQEC.CompiledTNMMAP
    plan::P
    mode::Int32
end

tabulation(mode) = mode == MODE_PLAIN_F64 ? 0 : 4

# The reference's tree, or, with a slicer, the same network re-optimized
# with sliced labels (an OMEinsum SlicedEinsum).
function hip_plan(decoder::TNMMAPHIP, ct)
    ns = length(ct.syndrome_indices)
    code = ct.code
    if decoder.slicer !== nothing
        code = optimize_code(OMEinsum.flatten(code), uniformsize(code, 2), decoder.optimizer;
                             slicer = decoder.slicer)
    end
    return HipPlan(code, ct.tensors, ns; tabulate_bits = tabulation(decoder.mode),
                   evidence = last_as_evidence(length(ct.tensors), ns))
end

function TensorQEC.compile(decoder::TNMMAPHIP, problem::TensorQEC.IndependentDepolarizingDecodingProblem)
    ct = TensorQEC.compile(TNMMAP(; optimizer = decoder.optimizer, factorize = false), problem)
    return CompiledTNMMAPHIP(ct, hip_plan(decoder, ct), decoder.mode)
end

struct CompiledDEMTNMMAPHIP{P} <: TensorQEC.CompiledDecoder
    ct::TensorQEC.CompiledDEMTNMMAP
    plan::P
    mode::Int32
end

# compile(::TNMMAP, ::DetectorErrorModel), src/decoding/tndecoder.jl:186-219
function TensorQEC.compile(decoder::TNMMAPHIP, dem::TensorQEC.DetectorErrorModel)
    ct = TensorQEC.compile(TNMMAP(; optimizer = decoder.optimizer, factorize = decoder.factorize), dem)
    return CompiledDEMTNMMAPHIP(ct, hip_plan(decoder, ct), decoder.mode)
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

# FP64 counterpart of each reduced-precision mode, for re-deciding close sums.
fp64_mode(mode) = mode == MODE_FUSED_F32 ? MODE_FUSED_F64 : mode == MODE_WIDE_F32 ? MODE_PLAIN_F64 :
                  mode in (MODE_SPEC_F32, MODE_SPEC_BF16) ? MODE_SPEC_F64 : mode

# Summed marginals (2^n_out x B) of syndrome columns `rows` over their slices.
function sliced_marginals(p::SlicedPlan, mode::Int32, rows::Matrix{UInt8})
    B = size(rows, 2)
    m = 1 << p.k
    ext = sliced_rows(rows, p.k)
    nout = 1 << p.n_out
    mar = Matrix{Float64}(undef, nout, B * m)
    cls = Vector{Int32}(undef, B * m)
    tie = Vector{UInt8}(undef, B * m)
    check(ccall((:tqec_decode_batch, libtqec), Cint,
                (Ptr{Cvoid}, Ptr{UInt8}, Int64, Int32, Ptr{Float64}, Ptr{Int32}, Ptr{UInt8}, Ptr{Cvoid}),
                p.plan.handle, ext, B * m, mode, mar, cls, tie, C_NULL))
    return dropdims(sum(reshape(mar, nout, m, B); dims = 2); dims = 2)
end

# Sliced plans: classes from the sums over the slices; a syndrome whose sum
# has no positive entry or a top-2 gap within 1e-3 of its maximum (the
# executors' re-check margin) is decided again in FP64.
function decode_classes(p::SlicedPlan, mode::Int32, syndromes)
    rows = syndrome_rows(syndromes, p.n_syn)
    tot = sliced_marginals(p, mode, rows)
    if fp64_mode(mode) != mode
        near = map(eachcol(tot)) do c
            s = sort(c; rev = true)
            !(s[1] > 0) || (length(s) > 1 && s[1] - s[2] <= 1e-3 * s[1])
        end
        if any(near)
            tot[:, near] = sliced_marginals(p, fp64_mode(mode), rows[:, near])
        end
    end
    return Int32[argmax(c) - 1 for c in eachcol(tot)]    # first maximum, column-major class index
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
