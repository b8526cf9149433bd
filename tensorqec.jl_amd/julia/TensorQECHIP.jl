"""
    TensorQECHIP

Julia binding of `libtqec_hip.so` for TensorQEC.jl: a drop-in `TNMMAPHIP`
decoder whose `compile` builds the reference's factor graph and contraction
tree (`compile(::TNMMAP, problem)`, src/decoding/tndecoder.jl:97-146 and
:186-219) and hands the tree to the HIP library as a flat post-order step
list, and whose `decode` / `decode_batch` return `DecodingResult`s with the
reference's layout (src/decoding/interfaces.jl:116-119).

The pure error and the logical fix reuse the reference's own
`error_pattern` (src/decoding/tndecoder.jl:167-174, 255-271), so a result is
the reference's result whenever the decided class is.

This file is source only: it needs a Julia environment with TensorQEC and
OMEinsum, which this build does not have.
"""
module TensorQECHIP

using TensorQEC
using TensorQEC.OMEinsum

export TNMMAPHIP, decode_batch

const libtqec = joinpath(@__DIR__, "..", "tensorqec", "libtqec_hip.so")

const MODE_PLAIN_F64 = Int32(0)
const MODE_FUSED_F64 = Int32(1)
const MODE_FUSED_F32 = Int32(2)
const MODE_WIDE_F32 = Int32(3)
const MODE_SPEC_F32 = Int32(4)     # plan-specialised kernel (hiprtc), bit-identical to MODE_FUSED_F32
const MODE_SPEC_F64 = Int32(5)
const MODE_SPEC_BF16 = Int32(6)    # GEMM-shaped steps on BF16 MFMA (circuit-level DEMs)

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
    flatten(code::NestedEinsum, tensors, n_syndrome) -> (leaf ranks, labels, data, evidence, steps...)

Post-order list of the binary contractions of an OMEinsum `NestedEinsum`
(leaves keep their `tensorindex`; step i becomes node `n_leaves + i`).  The
last `n_syndrome` tensors are the evidence vectors of `update_syndrome!`
(tndecoder.jl:148-158): they become one-hot leaves selected by syndrome bit
`i` (0-based, `sx` first).  Labels are shifted to 0-based.
"""
function flatten(code, tensors, n_syndrome::Int)
    L = length(tensors)
    ixs = OMEinsum.getixsv(code)
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
    evidence = fill(Int32(-1), L)
    for i in 1:n_syndrome
        evidence[L - n_syndrome + i] = Int32(i - 1)
    end
    iy = Int32.(OMEinsum.getiyv(code) .- 1)
    nvars = maximum(vcat(labels, outl, iy)) + 1
    return (; ranks, labels, data, evidence, lhs, rhs, outr, outl, iy, nvars)
end

function HipPlan(code, tensors, n_syndrome::Int; tabulate_bits::Int = 4)
    f = flatten(code, tensors, n_syndrome)
    GC.@preserve f begin
        desc = PlanDesc(Int32(f.nvars), Int32(length(f.ranks)), pointer(f.ranks), pointer(f.labels),
                        pointer(f.data), pointer(f.evidence), Int32(length(f.lhs)), pointer(f.lhs),
                        pointer(f.rhs), pointer(f.outr), pointer(f.outl), Int32(length(f.iy)), pointer(f.iy),
                        Int32(n_syndrome), Int32(0), C_NULL, C_NULL, Int32(tabulate_bits), Int32(0))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tqec_plan_create, libtqec), Cint, (Ref{PlanDesc}, Ref{Ptr{Cvoid}}), desc, h))
    end
    return HipPlan(h[], length(f.iy), n_syndrome)
end

"""
    TNMMAPHIP(; optimizer=TreeSA(), mode=MODE_SPEC_F32)

`TNMMAP` (src/decoding/tndecoder.jl:77-80) executed by the HIP library.
"""
Base.@kwdef struct TNMMAPHIP <: TensorQEC.AbstractGeneralDecoder
    optimizer::OMEinsum.CodeOptimizer = TreeSA()
    mode::Int32 = MODE_SPEC_F32
end

struct CompiledTNMMAPHIP <: TensorQEC.CompiledDecoder
    ct::TensorQEC.CompiledTNMMAP
    plan::HipPlan
    mode::Int32
end

function TensorQEC.compile(decoder::TNMMAPHIP, problem::TensorQEC.IndependentDepolarizingDecodingProblem)
    ct = TensorQEC.compile(TNMMAP(; optimizer = decoder.optimizer, factorize = false), problem)
    ns = length(ct.syndrome_indices)
    tab = decoder.mode == MODE_PLAIN_F64 ? 0 : 4
    return CompiledTNMMAPHIP(ct, HipPlan(ct.code, ct.tensors, ns; tabulate_bits = tab), decoder.mode)
end

syndrome_bytes(s::CSSSyndrome) = UInt8[x.x for x in vcat(s.sx, s.sz)]

"""
    decode_batch(c::CompiledTNMMAPHIP, syndromes::Vector{CSSSyndrome})

One library call for the whole batch; returns `DecodingResult`s.
"""
function decode_batch(c::CompiledTNMMAPHIP, syndromes::Vector{CSSSyndrome})
    B = length(syndromes)
    nsyn = c.plan.n_syn
    rows = Matrix{UInt8}(undef, nsyn, B)          # column b = syndrome b: row-major B x nsyn in C
    for (b, s) in enumerate(syndromes)
        rows[:, b] = syndrome_bytes(s)
    end
    nout = 1 << c.plan.n_out
    mar = Matrix{Float64}(undef, nout, B)
    cls = Vector{Int32}(undef, B)
    tie = Vector{UInt8}(undef, B)
    check(ccall((:tqec_decode_batch, libtqec), Cint,
                (Ptr{Cvoid}, Ptr{UInt8}, Int64, Int32, Ptr{Float64}, Ptr{Int32}, Ptr{UInt8}, Ptr{Cvoid}),
                c.plan.handle, rows, B, c.mode, mar, cls, tie, C_NULL))
    shape = ntuple(_ -> 2, c.plan.n_out)
    return map(1:B) do b
        pos = CartesianIndices(shape)[cls[b] + 1]      # Julia findmax position (column-major)
        DecodingResult(true, TensorQEC.error_pattern(pos, c.ct, syndromes[b]))
    end
end

TensorQEC.decode(c::CompiledTNMMAPHIP, s::CSSSyndrome) = decode_batch(c, [s])[1]

end # module
