// spec_codegen.hpp — plan-specialised FP32 / FP64 kernel source.
//
// The generic fused kernel interprets the step descriptors at run time: per
// output entry it walks the operands' bit-field segments, loops over
// contracted and single-side labels with run-time trip counts and reads
// table operands from global memory.  For a fixed plan all of that is
// known, so this generator writes the same per-syndrome program as straight
// HIP source (constant shifts and masks, unrolled contraction loops, LDS
// offsets as immediates, tables staged in LDS), compiled at run time with
// hiprtc for gfx950.  The arithmetic — order of every sum, fma, scale and
// comparison — is the generic FP32 kernel's, so both give bit-identical
// results, except for the register-resident FP32 kernel by default: it
// replaces the per-step scaling by plan-time static exponents (reg_source).
// The FP64 variant (no scaling) matches the generic FP64 fused kernel and
// serves the FP64 re-check and TQEC_MODE_SPEC_F64.
#pragma once
#include <stddef.h>

#include <string>
#include <vector>

#include "plan_compile.hpp"

namespace tqec {

struct SpecCfg {
    int glog = 0;        // log2 of syndromes per workgroup tile
    int threads = 128;   // threads per workgroup (>= tile)
    bool tab_lds = true; // constant and sliced tables staged in LDS
    int waves = 0;       // >0: ask the compiler for at least this many waves per SIMD (caps VGPRs)
    bool fp64 = false;   // FP64 without scaling (the generic FP64 fused kernel's arithmetic)
    bool slab = false;   // intermediates in a per-workgroup slab of global memory instead of LDS
    int unroll = 1;      // output entries per lane in flight (unroll of the output loop)
    bool mfma = false;   // FP32 tiles of one syndrome: GEMM-shaped steps on BF16 MFMA (FP32 accumulation)
    bool mfma32 = false; // FP32 tiles of one syndrome: GEMM-shaped steps on FP32 MFMA (operands kept in FP32)
    bool reg = false;    // FP32, one syndrome per quad of lanes, every tensor in registers (glog = log2(threads / 4))
    bool dyn = false;    // reg FP32: per-step scaling from operand maxima (bit-identical to the generic kernel)
                         // instead of plan-time static exponents
};

// Register-resident variant (cfg.reg): each syndrome is owned by a quad of
// lanes; a tensor's entries are spread over the quad by two of its index bits
// and over registers by the rest, operands cross lanes with DPP quad
// permutations (and, with cfg.dyn, per-tensor maxima with two DPP steps), so
// the step program touches no LDS and needs no barrier.  Returns false (with the reason) when
// the plan does not suit it: more than 16 classes, or more live tensor entries
// per lane than the register file holds.
bool spec_reg_supported(const CompiledPlan &cp, std::string *why = nullptr);

// The register-resident FP32 kernel re-contracts the syndromes it flags in
// FP64 itself, at the end of each tile (one syndrome at a time over the
// workgroup, in LDS), when that fits a modest LDS budget; otherwise the
// separate FP64 re-check kernel follows it.
bool spec_reg_inline_recheck(const CompiledPlan &cp, const SpecCfg &cfg);

// Peak live tensor registers per lane of the register-resident variant.
int spec_reg_peak(const CompiledPlan &cp);

// Choose the register-resident variant's lane layouts once (a tree search
// over every tensor's lane bits) and keep them in cp.reg_lay; later calls
// for this plan reuse them.
void spec_reg_prepare(CompiledPlan &cp);

// A step runs on MFMA (v_mfma_f32_32x32x16_bf16) when the configuration asks
// for it, the tile holds one syndrome, and per syndrome the step is a GEMM
// with M, N >= 32 (output bits feeding only A / only B) and K >= 16
// (contracted bits), with no labels summed on one side only.
bool spec_step_uses_mfma(const StepDesc &st, const SpecCfg &cfg);

// A step runs on FP32 MFMA when the configuration asks for it (cfg.mfma32,
// not with cfg.mfma), the tile holds one syndrome, and per syndrome the step
// is a GEMM with no labels summed on one side only: M, N >= 32 and K >= 2
// (v_mfma_f32_32x32x2f32 tiles), else M, N >= 16 and K >= 4
// (v_mfma_f32_16x16x4f32).  Unlike the BF16 steps the operands keep
// their FP32 values; only the summation order differs from the scalar steps.
bool spec_step_uses_mfma32(const StepDesc &st, const SpecCfg &cfg);
bool spec_plan_uses_mfma32(const CompiledPlan &cp, const SpecCfg &cfg);

// Static LDS bytes of the specialised kernel for this configuration.
size_t spec_lds_bytes(const CompiledPlan &cp, const SpecCfg &cfg);

// Entries of one workgroup's slab (slab configurations).
size_t spec_slab_entries(const CompiledPlan &cp, const SpecCfg &cfg);

// HIP source of the kernel `tqec_spec_f32` / `tqec_spec_bf16` /
// `tqec_spec_f64` (extern "C", see spec_kernel_name).  Parameters:
//   (const uint8_t *syn, int64_t n, const int32_t *list, const int32_t *list_count,
//    const T *tab, double *mar, int32_t *cls, uint8_t *tie,
//    int32_t *rl, int32_t *rcount, int64_t rcap, double tie_tol, double recheck_tol, T *slab,
//    int nseg, int segcap, const double *tab64)
// `slab` holds spec_slab_entries() entries per workgroup (slab configurations only).
// `tab` is the constant pool followed by the sliced pool, in the kernel's
// precision (float, or double when cfg.fp64), `tab64` the same in double
// (the inline FP64 re-check's); the FP64 variant ignores the re-check
// arguments.
std::string spec_source(const CompiledPlan &cp, const SpecCfg &cfg);

// One plan of the mixed-code kernel: its FP32 tile configuration, the FP64
// (tile 1) configuration of its re-checks and its tolerances.  Both share
// the kernel's thread count.
struct MixedSpecPlan {
    const CompiledPlan *cp = nullptr;
    SpecCfg f32, f64;
    double tie_tol = 1e-12, recheck_tol = 1e-3;
};

// Static LDS of the mixed kernel (the largest tile of any plan) and its
// source: kernel `tqec_mixed_spec(const MixedBatch *, const MixedHeader *,
// MixedCounters *, unsigned long long *t_done, const float *const *tabs32,
// const double *const *tabs64, float *slab32, double *slab64)` with the
// batch table, header and counters of the mixed executor (tqec_hip.hip);
// slabs hold slab32_entries / slab64_entries per workgroup.
size_t spec_mixed_lds_bytes(const std::vector<MixedSpecPlan> &plans);
std::string spec_mixed_source(const std::vector<MixedSpecPlan> &plans, int threads, size_t slab32_entries,
                              size_t slab64_entries);
// Max-product (TNMAP) plan: kernel `tqec_map_spec(const unsigned char *syn,
// long long n, const double *cpool, const double *spool, unsigned char
// *config, int n_config, double *logp)`, G = 2^glog syndromes per
// workgroup, the generic max-product kernel's LDS layout, order of
// comparisons and backtrace with every descriptor value substituted (the
// backtrace's label words in registers).  Bit-identical to it.
std::string spec_map_source(const CompiledPlan &cp, int glog, int threads);

inline const char *spec_kernel_name(const SpecCfg &cfg) {
    return cfg.fp64 ? "tqec_spec_f64" : (cfg.mfma ? "tqec_spec_bf16" : "tqec_spec_f32");
}

}  // namespace tqec
