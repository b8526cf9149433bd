/*
 * tqec_hip.h — C ABI of the MI355X tensor-network decoding library.
 *
 * The library executes a flattened contraction plan of a TNMMAP factor graph
 * over a batch of syndromes on one GPU (gfx950) and returns, per syndrome,
 * the unnormalised joint probability P(syndrome, logical class) for every
 * logical class, the arg-max class and a tie flag.
 *
 * It replaces, for the TNMMAP decoding path of TensorQEC.jl:
 *   - the per-syndrome evidence update  (src/decoding/tndecoder.jl:148-158,
 *     240-246  `update_syndrome!`),
 *   - the contraction `ct.code(ct.tensors...)` evaluated by OMEinsum
 *     (src/decoding/tndecoder.jl:162, 250),
 *   - `findmax(mar)` (src/decoding/tndecoder.jl:163, 251).
 * The host keeps building the factor graph and the contraction tree
 * (src/decoding/tndecoder.jl:97-146, 186-219) and hands the tree over as a
 * post-order list of pairwise steps.
 *
 * Conventions
 *   - Every variable (label) has size 2.  A tensor of rank r is 2^r doubles;
 *     bit j of the flat index is the value of the tensor's j-th label, which
 *     is the column-major layout of a Julia Array{Float64} with all sizes 2.
 *   - Node ids: leaves are 0..n_leaves-1, step i produces node n_leaves+i.
 *   - All functions return TQEC_OK (0) or a negative TQEC_E* code and never
 *     abort; tqec_last_error() returns a thread-local message.
 *   - The contraction a plan describes is fixed at creation.  Only its
 *     execution settings change afterwards (tqec_plan_set_tolerances,
 *     _set_tile, _set_spec, _set_map_kernel, _set_timing, and the
 *     autotuner); each setter takes the plan's lock and applies to calls
 *     made after it returns.  Per-call device workspace is kept per (plan,
 *     stream); calls on different streams may run concurrently, calls on the
 *     same stream are ordered by the stream.
 *   - tqec_decode_batch takes host pointers and synchronises the stream
 *     before returning; tqec_decode_batch_device takes device pointers and
 *     only enqueues work on the stream.
 */
#ifndef TQEC_HIP_H
#define TQEC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQEC_OK 0
#define TQEC_EINVAL (-1)      /* malformed descriptor or argument       */
#define TQEC_ENOMEM (-2)      /* host or device allocation failed        */
#define TQEC_EDEVICE (-3)     /* HIP runtime error                       */
#define TQEC_EUNSUPPORTED (-4)/* plan exceeds a limit of the chosen mode */
#define TQEC_ENULL (-5)       /* null handle or pointer                  */

/* Execution modes of tqec_decode_batch*. */
#define TQEC_MODE_PLAIN_F64 0 /* one batched kernel per plan step, FP64,
                                 intermediates in HBM                     */
#define TQEC_MODE_FUSED_F64 1 /* whole plan per workgroup tile, FP64,
                                 intermediates in LDS                     */
#define TQEC_MODE_WIDE_F32 3  /* plain layout (intermediates in HBM) in FP32
                                 with per-tensor power-of-two scaling;
                                 low-margin syndromes re-contracted by the
                                 plain FP64 executor (for plans too wide
                                 for LDS; synchronises the stream once)  */

#define TQEC_SEMIRING_SUM 0   /* sum-product: marginals (TNMMAP)            */
#define TQEC_SEMIRING_MAX 1   /* max-product with arg-max backtrace (TNMAP)  */

#define TQEC_MODE_FUSED_F32 2 /* fused, FP32 with per-tensor power-of-two
                                 scaling; syndromes whose top-2 class
                                 margin is below the re-check threshold are
                                 re-contracted in FP64                    */
#define TQEC_MODE_SPEC_F32 4  /* FUSED_F32 with a kernel generated for the
                                 plan and compiled at first use (hiprtc):
                                 constant index arithmetic, unrolled
                                 contractions, tables in LDS or registers;
                                 LDS-tile kernels are bit-identical to
                                 FUSED_F32, the register-resident kernel
                                 (default where the plan fits) uses
                                 plan-time static exponents, see
                                 TQEC_SPEC_REG                            */
#define TQEC_MODE_SPEC_F64 5  /* FUSED_F64 with the plan-specialised kernel;
                                 bit-identical to FUSED_F64               */
#define TQEC_MODE_SPEC_BF16 6 /* SPEC_F32 with one syndrome per tile and the
                                 GEMM-shaped steps (M, N >= 32, K >= 16 per
                                 syndrome) on BF16 MFMA with FP32
                                 accumulation; syndromes whose top-2 margin
                                 is below max(recheck_tol, 0.05) are
                                 re-contracted in FP64                    */

/*
 * Flattened contraction plan.  Mirrors the NestedEinsum that
 * `optimize_code(DynamicEinCode(ixs, iy), size_dict, TreeSA())` returns in
 * src/decoding/tndecoder.jl:140-144 and :213-217.
 */
typedef struct tqec_plan_desc {
    int32_t n_vars;                 /* labels are 0..n_vars-1               */
    int32_t n_leaves;
    const int32_t *leaf_rank;       /* [n_leaves]                           */
    const int32_t *leaf_labels;     /* concatenated, sum(leaf_rank) entries */
    const double *leaf_data;        /* concatenated 2^rank values per leaf;
                                       entries of evidence leaves ignored  */
    const int32_t *leaf_evidence;   /* [n_leaves]: syndrome bit whose value
                                       selects the one-hot vector of an
                                       evidence leaf (rank 1), or -1       */
    int32_t n_steps;                /* must be n_leaves - 1                 */
    const int32_t *step_lhs;        /* [n_steps] node ids                   */
    const int32_t *step_rhs;        /* [n_steps] node ids                   */
    const int32_t *step_out_rank;   /* [n_steps]                            */
    const int32_t *step_out_labels; /* concatenated                         */
    int32_t n_out;                  /* rank of the result                   */
    const int32_t *out_labels;      /* result label order (the einsum iy)   */
    int32_t n_syndrome_bits;        /* bytes per syndrome row               */
    int32_t n_sliced;               /* labels fixed by a syndrome bit       */
    const int32_t *sliced_label;    /* [n_sliced] label id                  */
    const int32_t *sliced_bit;      /* [n_sliced] syndrome bit              */
    int32_t tabulate_bits;          /* 0: only syndrome-independent subtrees
                                       are contracted at plan creation;
                                       1..8: subtrees that depend on at most
                                       this many syndrome bits are also
                                       contracted there, once per bit
                                       pattern, and looked up per syndrome */
    int32_t semiring;               /* TQEC_SEMIRING_SUM or _MAX; a max-
                                       product plan must have n_out = 0 and
                                       is decoded by tqec_map_decode_batch */
} tqec_plan_desc;

typedef struct tqec_plan tqec_plan;

/* Mirrors `contraction_complexity(ct)` (src/decoding/tndecoder.jl:273-279)
 * and adds what the executor derived from the plan. */
typedef struct tqec_stats {
    double tc;              /* multiply-adds over all steps               */
    double rw;              /* elements read + written over all steps     */
    int32_t sc;             /* log2 of the widest tensor                  */
    int32_t n_steps;
    int32_t n_dep_steps;    /* steps executed per syndrome                */
    double dep_tc;          /* multiply-adds per syndrome                 */
    double dep_rw;          /* per-syndrome elements read+written         */
    int64_t plain_ws_entries;  /* per-syndrome HBM workspace, in entries  */
    int64_t fused_lds_entries; /* per-syndrome LDS footprint, in entries  */
    int32_t n_out;          /* number of logical classes = 2^n_out        */
    int32_t fused_tile_f64; /* syndromes per workgroup, FUSED_F64         */
    int32_t fused_tile_f32; /* syndromes per workgroup, FUSED_F32         */
    int64_t const_entries;  /* folded syndrome-independent operand data    */
    int64_t table_entries;  /* tabulated operand data (all bit patterns)   */
    int32_t tabulate_bits;  /* subtrees on at most this many syndrome bits
                               were tabulated at plan creation             */
} tqec_stats;

/* Build a plan.  The descriptor's arrays are copied; the caller keeps them. */
int tqec_plan_create(const tqec_plan_desc *desc, tqec_plan **out);

void tqec_plan_destroy(tqec_plan *plan);

int tqec_plan_stats(const tqec_plan *plan, tqec_stats *out);

/*
 * Decode n syndromes.
 *   syndromes  n x n_syndrome_bits bytes, row-major, each 0 or 1
 *   marginals  n x 2^n_out doubles: row b is P(syndrome_b, class) for every
 *              class, classes in column-major order of out_labels (NULL to
 *              skip)
 *   classes    n: index of the first maximum (Julia `findmax`) (NULL to skip)
 *   tie_flags  n: 1 when the top-2 relative gap is <= tie_tol (NULL to skip)
 *   stream     hipStream_t or NULL for the null stream
 */
int tqec_decode_batch(tqec_plan *plan, const uint8_t *syndromes, int64_t n, int32_t mode,
                      double *marginals, int32_t *classes, uint8_t *tie_flags, void *stream);

/* Same with device pointers; enqueues on `stream` and returns. */
int tqec_decode_batch_device(tqec_plan *plan, const uint8_t *d_syndromes, int64_t n, int32_t mode,
                             double *d_marginals, int32_t *d_classes, uint8_t *d_tie_flags,
                             void *stream);

/*
 * tqec_decode_batch_device without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Modes TQEC_MODE_SPEC_F32 and TQEC_MODE_SPEC_F64 only, at most
 * TQEC_NOALLOC_MAX_ROWS syndromes per call (TQEC_EINVAL beyond).  The FP64
 * re-check list and counts live in one static device array per GPU
 * (~4.3 MB, held until the process ends) instead of the per-stream
 * workspace; every call on a GPU shares it, so calls that overlap on one GPU
 * (different streams, or threads) race and must be serialised by the caller.
 * The plan's tables and kernels must already be on the device: make one
 * tqec_decode_batch[_device] call in the same mode there first (TQEC_EINVAL
 * "plan not prepared" otherwise).  Plans whose kernels need a global slab
 * return TQEC_EUNSUPPORTED.  Results equal tqec_decode_batch_device's (same
 * kernels); tqec_last_recheck_count does not see these calls.
 */
#define TQEC_NOALLOC_MAX_ROWS (1 << 20)
int tqec_decode_batch_device_noalloc(tqec_plan *plan, const uint8_t *d_syndromes, int64_t n, int32_t mode,
                                     double *d_marginals, int32_t *d_classes, uint8_t *d_tie_flags,
                                     void *stream);

/*
 * Most probable configuration (TNMAP) for a max-product plan: replaces
 * TensorInference's `most_probable_config` after `update_evidence!`
 * (src/decoding/tndecoder.jl:53-57).  The contraction keeps, per output
 * entry, the maximum over the labels each step eliminates (first maximum in
 * the order: contracted labels, then single-side labels of A, then of B) and
 * records it; a per-syndrome backtrace from the scalar root assigns every
 * label.  Writes the values of labels 0..n_config-1 (n x n_config bytes) and,
 * when logp is not NULL, the natural log of the maximum (n doubles).
 * Host pointers for tqec_map_decode_batch (synchronises `stream`), device
 * pointers for the _device variant (enqueues only).
 */
int tqec_map_decode_batch(tqec_plan *plan, const uint8_t *syndromes, int64_t n, int32_t n_config,
                          uint8_t *config, double *logp, void *stream);
int tqec_map_decode_batch_device(tqec_plan *plan, const uint8_t *d_syndromes, int64_t n, int32_t n_config,
                                 uint8_t *d_config, double *d_logp, void *stream);

/* Decode only the syndromes whose row numbers are listed in d_list; the
 * list length is read on the device from *d_count (at most n_max), so the
 * call can follow a kernel that builds the list without a host round trip.
 * Outputs are written at the listed rows.  Fused modes only. */
int tqec_decode_list_device(tqec_plan *plan, const uint8_t *d_syndromes, const int32_t *d_list,
                            const int32_t *d_count, int64_t n_max, int32_t mode, double *d_marginals,
                            int32_t *d_classes, uint8_t *d_tie_flags, void *stream);

/* ---------------------------------------------------------------------------
 * Learned front-end.  No counterpart in the reference: an MLP from syndrome
 * bits to logical-class logits (3 hidden ReLU layers of `hidden` units, BF16
 * weights and activations, FP32 accumulation on MFMA), trained on class
 * posteriors of the TN decoder, with a confidence gate whose rejected
 * syndromes go to tqec_decode_list_device.
 * Parameters (FP32, flat): W1[kin][hidden] b1 W2[hidden][hidden] b2
 * W3[hidden][hidden] b3 W4[hidden][n_out] b4, kin = n_in rounded up to a
 * multiple of 32 (rows k >= n_in of W1 are kept at zero).
 * All data pointers are device pointers except tqec_mlp_set/get_params and
 * the host loss of tqec_mlp_gradients / tqec_mlp_train_step (NULL to skip;
 * when given, the call synchronises `stream`).
 * ------------------------------------------------------------------------- */
typedef struct tqec_mlp tqec_mlp;
int tqec_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, tqec_mlp **out);
int64_t tqec_mlp_num_params(const tqec_mlp *mlp);
int tqec_mlp_set_params(tqec_mlp *mlp, const float *host_params);
int tqec_mlp_get_params(tqec_mlp *mlp, float *host_params);
int tqec_mlp_forward(tqec_mlp *mlp, const uint8_t *d_syndromes, int64_t n, float *d_logits, void *stream);
/* Mean soft cross-entropy against d_targets (n x n_out class probabilities):
 * gradients (flat, FP32) into d_grad; n must be a multiple of 32. */
int tqec_mlp_gradients(tqec_mlp *mlp, const uint8_t *d_syndromes, const float *d_targets, int64_t n,
                       float *d_grad, float *h_loss, void *stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss. */
int tqec_mlp_train_step(tqec_mlp *mlp, const uint8_t *d_syndromes, const float *d_targets, int64_t n, float lr,
                        float *h_loss, void *stream);
/* Classes (argmax), confidences (largest posterior) and the list of rows
 * whose confidence is below `threshold` (*d_reject_count is reset first). */
int tqec_mlp_gate(tqec_mlp *mlp, const uint8_t *d_syndromes, int64_t n, float threshold, int32_t *d_classes,
                  float *d_confidence, int32_t *d_reject, int32_t *d_reject_count, void *stream);
/* Kernel of tqec_mlp_forward / tqec_mlp_gate: 1 (default where it applies:
 * hidden 512, n_out <= 32, n_in <= 512) the fused kernel — all four layers
 * for 128 syndromes per 8-wave workgroup with the activations in LDS; 0 the
 * layer-by-layer kernels that training uses.  Logits agree to FP32
 * rounding (the fused kernel adds the bias before the products).
 * tqec_mlp_inference reports the kernel in use (0 or 1). */
int tqec_mlp_set_inference(tqec_mlp *mlp, int32_t kernel);
int tqec_mlp_inference(const tqec_mlp *mlp, int32_t *kernel);
void tqec_mlp_destroy(tqec_mlp *mlp);
/* Training targets from TN marginals: q = marginal / sum over classes (a row
 * of zeros gives the uniform distribution). */
int tqec_posteriors_from_marginals(const double *d_marginals, int64_t n, int32_t n_classes, float *d_q,
                                   void *stream);

/* ---- Mixed-code executor ------------------------------------------------
 * Batches of syndromes of different codes (different plans) decoded by one
 * persistent FP32 launch with in-launch FP64 re-checks, as
 * `decode(ct, syndrome)` (src/decoding/interfaces.jl:116-160) for a stream
 * of compiled decoders of uneven size; each batch's results are those of
 * tqec_decode_batch_device(plan, ..., TQEC_MODE_FUSED_F32, ...).  Plans whose
 * intermediates exceed the LDS run from a per-workgroup slab of device
 * memory.  Per-call tables are staged in pinned memory and, on a non-null
 * stream, the call (table copy, counter reset, both kernels) is recorded
 * once into a HIP graph and replayed.  Two calls may be in flight; a third
 * waits for the first.  No reference counterpart: the reference decodes one
 * syndrome of one code per call. */
typedef struct tqec_mixed tqec_mixed;

typedef struct {
    int32_t plan;              /* index into the plans given to tqec_mixed_create */
    int32_t reserved;
    int64_t n;                 /* syndromes in this batch (0 allowed) */
    const uint8_t *syndromes;  /* device [n][n_syndrome_bits of the plan] */
    double *marginals;         /* device [n][2^n_out] or NULL */
    int32_t *classes;          /* device [n] or NULL */
    uint8_t *ties;             /* device [n] or NULL */
} tqec_mixed_batch;

/* Binds the plans to the current device; at most `max_batches` batches and
 * `max_rows` syndromes in total per call.  The plans must outlive it.  Each
 * plan's tie and re-check tolerances are copied to the device here: a later
 * tqec_plan_set_tolerances on a plan does not reach an executor created
 * before it. */
int tqec_mixed_create(tqec_plan *const *plans, int32_t n_plans, int32_t max_batches, int64_t max_rows,
                      tqec_mixed **out);
/* Graph capture on (default) or off. */
int tqec_mixed_set_graph(tqec_mixed *m, int32_t enable);
/* Order in which the tiles of a call's batches are handed out: as given, or
 * cheapest batch first (default; batch size x the plan's work per syndrome),
 * which lets small batches finish before large ones.  Results and latencies
 * are reported in the caller's batch order either way. */
#define TQEC_MIXED_ORDER_GIVEN 0
#define TQEC_MIXED_ORDER_CHEAPEST_FIRST 1
int tqec_mixed_set_order(tqec_mixed *m, int32_t order);
/* Kernel of the executor: the generic interpreter of the step descriptors
 * (default), or one kernel generated from every plan's specialised code
 * (each plan's current SPEC_F32 tile, 128 threads, compiled with hiprtc
 * here; results bit-identical).  Waits for calls in flight. */
#define TQEC_MIXED_KERNEL_GENERIC 0
#define TQEC_MIXED_KERNEL_SPEC 1
int tqec_mixed_set_kernel(tqec_mixed *m, int32_t kind);
/* Source of that generated kernel for a set of plans (inspection and
 * offline compilation checks; no device needed).  Output as
 * tqec_plan_spec_source. */
int tqec_mixed_spec_source(tqec_plan *const *plans, int32_t n_plans, int32_t threads, char *buf, size_t cap,
                           size_t *len);
/* Tile (syndromes per workgroup) chosen for a plan and where its
 * intermediates live: bit 0 FP32 pass in a slab, bit 1 FP64 re-check in a slab. */
int tqec_mixed_plan_config(const tqec_mixed *m, int32_t plan, int32_t *tile, int32_t *slab);
/* Enqueue one call on `stream` (device pointers in the table). */
int tqec_mixed_decode_device(tqec_mixed *m, const tqec_mixed_batch *batches, int32_t n_batches, void *stream);
/* After the last call: per batch, microseconds from the first tile's start
 * to the batch's last result write on the device clock (0 for empty
 * batches), and the number of FP64 re-checks (each done by the workgroup
 * that decoded the syndrome, right after its tile).  Waits for the call. */
int tqec_mixed_last_call(tqec_mixed *m, double *batch_us, int32_t n_batches, int64_t *recheck_count);
void tqec_mixed_destroy(tqec_mixed *m);

/* Flags of a plan-specialised kernel configuration. */
#define TQEC_SPEC_TABLES_IN_LDS 1  /* constant and sliced tables staged in LDS        */
#define TQEC_SPEC_FP64 2           /* FP64 variant (tqec_plan_spec_source only)       */
#define TQEC_SPEC_SLAB 4           /* intermediates in a per-workgroup global slab,
                                      for plans too wide for LDS                      */
#define TQEC_SPEC_MFMA 8           /* GEMM-shaped steps on BF16 MFMA (tile 1); with
                                      tqec_plan_set_spec this pins the SPEC_BF16
                                      configuration                                   */
#define TQEC_SPEC_REG 16           /* one syndrome per quad of lanes with every
                                      tensor in registers (no LDS intermediates, no
                                      barriers); tile = threads / 4; FP32 (default
                                      SPEC_F32 kernel) or, with TQEC_SPEC_FP64 in
                                      tqec_plan_spec_source, FP64 (default SPEC_F64
                                      kernel).  FP32: plan-time static exponents
                                      (no per-step maxima or rescales; syndromes
                                      whose largest marginal falls below the
                                      plan's underflow bound are re-checked in
                                      FP64), within the FP32 tolerance of the
                                      other kernels; FP64: bit-identical to the
                                      other FP64 kernels                            */
#define TQEC_SPEC_DYNSCALE 32      /* with TQEC_SPEC_REG in FP32: per-step scaling
                                      from each operand's maximum instead, the
                                      generic FP32 kernel's arithmetic
                                      (bit-identical to TQEC_MODE_FUSED_F32)        */
#define TQEC_SPEC_MFMA32 64        /* SPEC_F32 tiles of one syndrome (LDS or slab):
                                      GEMM-shaped steps (M, N >= 16, K >= 4, or
                                      M, N >= 32, K >= 2 per syndrome) on FP32
                                      MFMA with FP32 operands;
                                      within the FP32 tolerance, not bit-identical
                                      (summation order)                              */

/* Source of the plan-specialised kernel for a tile configuration
 * (inspection and offline compilation checks); `flags` as above, with the
 * unroll of the per-lane output loop in bits 8-15 (0 = 1).  Writes at most `cap` bytes
 * (NUL-terminated when it fits) and the full length to *len.  No reference
 * counterpart. */
int tqec_plan_spec_source(const tqec_plan *plan, int32_t tile, int32_t threads, int32_t flags, char *buf,
                          size_t cap, size_t *len);

/* Pin / read the TQEC_MODE_SPEC_F32 configuration: tile, threads,
 * TQEC_SPEC_TABLES_IN_LDS | TQEC_SPEC_SLAB | TQEC_SPEC_REG | TQEC_SPEC_DYNSCALE |
 * TQEC_SPEC_MFMA32 flags, and `waves` = minimum
 * waves per SIMD requested from the compiler (bits 0-7, 0 = none) | unroll
 * of the per-lane output loop << 8 (1, 2, 4 or 8; 0 = 1).  No reference
 * counterpart. */
int tqec_plan_set_spec(tqec_plan *plan, int32_t tile, int32_t threads, int32_t flags, int32_t waves);
int tqec_plan_spec_config(tqec_plan *plan, int32_t *tile, int32_t *threads, int32_t *flags, int32_t *waves);

/* Max-product plans: run tqec_map_decode_batch[_device] on a kernel
 * generated for the plan (1, default; compiled with hiprtc at the first call
 * on a device, so that call pays the compile time and should not be the
 * first one inside a stream capture; bit-identical results) or on the
 * generic max-product kernel (0).  No reference counterpart. */
int tqec_plan_set_map_kernel(tqec_plan *plan, int32_t specialised);
/* Kernel the next max-product call will use (1 = plan-specialised).  When
 * the generated kernel could not be built (hiprtc or module-load failure)
 * the plan falls back to the generic kernel by itself and `note` (NULL to
 * skip; at most cap-1 bytes, NUL-terminated) receives the reason; it is
 * empty otherwise. */
int tqec_plan_map_kernel(tqec_plan *plan, int32_t *specialised, char *note, size_t cap);

/* Tunables of a plan (defaults: tie_tol 1e-12, recheck_tol 1e-3). */
int tqec_plan_set_tolerances(tqec_plan *plan, double tie_tol, double recheck_tol);

/* Syndromes re-contracted in FP64 by the last FP32 call (FUSED_F32,
 * WIDE_F32, SPEC_F32, SPEC_BF16) on `stream`; waits for that call. */
int tqec_last_recheck_count(tqec_plan *plan, void *stream, int64_t *count);

/* Project autotuner: time every feasible (syndromes per workgroup, threads
 * per workgroup) configuration of a fused mode on a trial batch already in
 * device memory, keep the fastest for later calls of that mode and report it
 * (tile / threads may be NULL).  Synchronises `stream`. */
int tqec_plan_autotune(tqec_plan *plan, int32_t mode, const uint8_t *d_syndromes, int64_t n, void *stream,
                       int32_t *tile, int32_t *threads);

/* Pin the tile configuration of a fused mode (tile: syndromes per workgroup,
 * a power of two <= 64; threads: 64, 128 or 256). */
int tqec_plan_set_tile(tqec_plan *plan, int32_t mode, int32_t tile, int32_t threads);

/* Kernel timing with HIP events recorded on the call's stream (off by
 * default).  While enabled, every decode call records events around its
 * main contraction (the fused kernel, or the whole step sequence of the
 * plain executor) and around the FP64 re-check kernel.
 * tqec_kernel_time_totals waits for the recorded events of all calls on
 * `stream` since the previous query and returns the summed milliseconds and
 * the number of calls, then clears them. */
int tqec_plan_set_timing(tqec_plan *plan, int enable);
int tqec_kernel_time_totals(tqec_plan *plan, void *stream, double *main_ms, double *recheck_ms,
                            int64_t *calls);

/* ---------------------------------------------------------------------------
 * On-device sampling.  Replaces the per-sample `random_error_pattern(em)` +
 * `syndrome_extraction` loop (src/decoding/error_model.jl:69-71, 97-117,
 * 131-146; the harness loop of test/decoding/tndecoder.jl:85-99).  Sample s,
 * variable j draws u from Philox4x32-10 with counter (j, s mod 2^32, s >> 32,
 * tag) and key (seed mod 2^32, seed >> 32): bit-identical to the host sampler.
 *   CSS (tag 0x51D0): Y if u < py, X if u < px+py, Z if u < px+py+pz;
 *     syndrome row [sx | sz] with sx = Hx ez, sz = Hz ex; class bit i =
 *     lx_i . ez, bit k+i = lz_i . ex; errors row [ex | ez] (2 nq bytes).
 *   flips (tag 0xF11B): mechanism j fires if u < p_j; syndrome = D e,
 *     class bit l = L_l . e; errors row e (nm bytes).
 * Matrices are dense row-major bytes.  Pointers passed to tqec_sample are
 * device pointers; d_classes and d_errors may be NULL.
 * ------------------------------------------------------------------------- */
typedef struct tqec_sampler tqec_sampler;
int tqec_sampler_create_css(int32_t nq, const double *px, const double *py, const double *pz, int32_t nsx,
                            const uint8_t *Hx, int32_t nsz, const uint8_t *Hz, int32_t k, const uint8_t *lx,
                            const uint8_t *lz, tqec_sampler **out);
int tqec_sampler_create_flips(int32_t nm, const double *p, int32_t nd, const uint8_t *D, int32_t nl,
                              const uint8_t *L, tqec_sampler **out);
int tqec_sample(tqec_sampler *sampler, uint64_t seed, int64_t first, int64_t n, uint8_t *d_syndromes,
                int32_t *d_classes, uint8_t *d_errors, void *stream);
void tqec_sampler_destroy(tqec_sampler *sampler);

/* ---------------------------------------------------------------------------
 * On-device error-pattern recovery.  Replaces `error_pattern`
 * (src/decoding/tndecoder.jl:167-174 CSS, :255-271 DEM) and its MIP pure
 * error (src/decoding/ipdecoder.jl:150-169): e = R s with R a GF(2) right
 * inverse of the check matrix, then the reference's logical fix driven by the
 * decided class.
 *   CSS: Rx (nq x nsz) maps sz to ex, Rz (nq x nsx) maps sx to ez; lx/lz are
 *     k x nq.  ex += lx_i when lz_i . ex differs from class bit k+i; ez +=
 *     lz_i when lx_i . ez differs from class bit i.  Output ex -> d_err_a,
 *     ez -> d_err_b (n x nq bytes each).
 *   DEM: R (nm x nd); l2q (nl x nm) marks the mechanisms touching logical l;
 *     when their parity differs from class bit l, every one of them is
 *     flipped (the reference's rule).  Output -> d_err_a (n x nm bytes).
 * ------------------------------------------------------------------------- */
typedef struct tqec_recovery tqec_recovery;
int tqec_recovery_create_css(int32_t nq, int32_t nsx, int32_t nsz, const uint8_t *Rx, const uint8_t *Rz,
                             int32_t k, const uint8_t *lx, const uint8_t *lz, tqec_recovery **out);
int tqec_recovery_create_dem(int32_t nm, int32_t nd, const uint8_t *R, int32_t nl, const uint8_t *l2q,
                             tqec_recovery **out);
int tqec_recover(tqec_recovery *rec, const uint8_t *d_syndromes, const int32_t *d_classes, int64_t n,
                 uint8_t *d_err_a, uint8_t *d_err_b, void *stream);
void tqec_recovery_destroy(tqec_recovery *rec);

/* ---------------------------------------------------------------------------
 * Logical-error-rate loop on the device.  Replaces the per-sample harness
 * loop `random_error_pattern` -> `syndrome_extraction` -> `decode` ->
 * `check_logical_error` (src/decoding/error_model.jl:97-181, driven by
 * src/decoding/threshold.jl:1-32 and the test loops of test/decoding/):
 * samples first .. first+n-1 of the sampler's stream are drawn, decoded
 * with `plan` in `mode`, their error patterns recovered, and the difference
 * between the sampled and the recovered pattern checked against every
 * logical operator, in chunks of `batch` syndromes; nothing crosses to the
 * host but four counts:
 *   counts[0]  samples with a logical error (`check_logical_error` true)
 *   counts[1]  of which in the X-error half (lz . (ex + ex') odd for a row;
 *              0 for detector error models)
 *   counts[2]  of which in the Z-error half (lx . (ez + ez') odd)
 *   counts[3]  recovered patterns that do not reproduce the sampled syndrome
 * The sampler, recovery and plan must describe the same code (checked) and
 * outlive the handle; device buffers for one chunk are allocated at
 * creation.  tqec_ler_run synchronises `stream`.
 * ------------------------------------------------------------------------- */
typedef struct tqec_ler tqec_ler;
int tqec_ler_create(tqec_plan *plan, tqec_sampler *sampler, tqec_recovery *recovery, int64_t batch,
                    tqec_ler **out);
int tqec_ler_run(tqec_ler *ler, int32_t mode, uint64_t seed, int64_t first, int64_t n, int64_t *counts,
                 void *stream);
void tqec_ler_destroy(tqec_ler *ler);

/* Thread-local description of the last error. */
const char *tqec_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TQEC_HIP_H */
