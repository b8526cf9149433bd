// plan.hpp — compiled contraction plan shared by the host compiler and the
// HIP kernels.
//
// A plan step computes
//   C[c] = sum_{k, sa, sb} A[a(c) | sa<<A.sshift | k<<A.kshift] * B[b(c) | sb<<B.sshift | k<<B.kshift]
// where c runs over C's entries in C's memory order, a(c)/b(c) move the bits
// of c that belong to each operand into its low bits (keeping C's order), sa
// and sb run over labels summed on one side only and k over the labels
// contracted between A and B (the top bits of both operands).  Every
// operand's memory order is chosen by its single consumer, so no step ever
// needs a separate transpose pass, and lanes that handle consecutive outputs
// read consecutive (or identical) words.
#pragma once
#include <stdint.h>

namespace tqec {

constexpr int kMaxSeg = 12;   // bit-field moves per operand
constexpr int kMaxEv = 8;     // syndrome bits selecting a sliced-leaf variant
constexpr int kMaxRank = 24;  // widest per-syndrome tensor
constexpr int kMaxOut = 16;   // result rank

enum OperandKind : int32_t {
    kInter = 0,   // per-syndrome intermediate in the workspace
    kConst = 1,   // syndrome-independent tensor in the constant pool
    kSliced = 2,  // leaf sliced by syndrome bits: one variant per bit pattern
};

struct Seg {
    int16_t src;    // first bit in the output index
    int16_t width;  // number of consecutive bits
    int16_t dst;    // first bit in the operand index
    int16_t pad;
};

struct OperandDesc {
    int32_t kind;
    int32_t rank;      // effective rank (after slicing / pre-summing)
    int32_t nsingle;   // labels summed on this side only (bits sshift..)
    int32_t nseg;
    int32_t sshift;    // first bit of the single-side labels
    int32_t kshift;    // first bit of the contracted labels
    int64_t off;       // kInter: entry offset per syndrome; kConst/kSliced: pool offset
    int32_t slot;      // kInter: scale slot; otherwise -1
    int32_t cexp;      // power-of-two exponent folded out of kConst/kSliced data
    int32_t nev;       // kSliced: number of syndrome bits selecting the variant
    int32_t ev_bit[kMaxEv];
    Seg seg[kMaxSeg];
};

struct StepDesc {
    OperandDesc a;
    OperandDesc b;
    int32_t rc;       // output rank
    int32_t nk;       // contracted bits (low bits of both operands)
    int64_t c_off;    // output entry offset per syndrome
    int32_t c_slot;   // output scale slot
    int32_t is_root;
};

// Extract the operand's high index bits from the output index.
__host__ __device__ inline uint32_t gather_bits(uint32_t c, const OperandDesc &op) {
    uint32_t r = 0;
    for (int i = 0; i < op.nseg; ++i) {
        const Seg s = op.seg[i];
        r |= ((c >> s.src) & ((1u << s.width) - 1u)) << s.dst;
    }
    return r;
}

}  // namespace tqec
