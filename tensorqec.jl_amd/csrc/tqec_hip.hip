// tqec_hip.hip — HIP kernels and C ABI of the TNMMAP contraction executor
// for MI355X (gfx950).
//
// Two executors share one compiled plan (plan_compile.cpp):
//
//  * plain  — one launch per syndrome-dependent plan step, FP64, every
//             intermediate in HBM laid out [entry][syndrome] so a wavefront
//             reads 64 consecutive syndromes of one entry (fully coalesced)
//             and all index arithmetic is wave-uniform.
//  * fused  — one persistent kernel runs the whole per-syndrome step program
//             for a tile of G syndromes per workgroup with every
//             intermediate resident in LDS ([entry][g], bank-conflict free),
//             one barrier per step.  FP32 carries a per-tensor, per-syndrome
//             power-of-two exponent (max tracked with 64-bit tagged LDS
//             atomics) so nothing underflows; syndromes whose top-2 class
//             margin is below the re-check threshold are queued and
//             re-contracted by the FP64 fused kernel.
//
// Reference behaviour replaced: src/decoding/tndecoder.jl:148-165 and
// :240-253 (evidence update, OMEinsum contraction, findmax).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "plan.hpp"
#include "plan_compile.hpp"
#include "spec_codegen.hpp"
#include "tqec_hip.h"

using namespace tqec;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define TQ_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_error(TQEC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

template <typename T>
struct Pools {
    const T *cpool;
    const T *spool;
};

// Transpose a chunk of syndromes [n][nbits] (row-major bytes) to [bit][ldw].
__global__ void transpose_syndromes_kernel(const uint8_t *__restrict__ syn, int64_t n, int nbits,
                                           uint8_t *__restrict__ synT, int64_t ldw) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const uint8_t *row = syn + b * nbits;
    for (int i = 0; i < nbits; ++i) synT[(int64_t)i * ldw + b] = row[i] & 1u;
}

// Element `idx` of an operand.  Workspace operands (kInter) live at
// base[idx * stride]; table and constant operands at base[idx].
template <typename T, bool WS>
__device__ __forceinline__ T op_load(const T *base, uint32_t idx, int64_t stride) {
    return WS ? base[(int64_t)idx * stride] : base[idx];
}

// Sum over the operand's single-side labels (only workspace operands carry
// any; table and constant operands are pre-summed at plan creation).
template <typename T, bool WS>
__device__ __forceinline__ T op_load_sum(const T *base, uint32_t idx, int64_t stride, const OperandDesc &op) {
    if (!WS || op.nsingle == 0) return op_load<T, WS>(base, idx, stride);
    T s = 0;
    const uint32_t ns = 1u << op.nsingle;
    for (uint32_t q = 0; q < ns; ++q) s += op_load<T, WS>(base, idx | (q << op.sshift), stride);
    return s;
}

// One output element: sum over the contracted labels.
template <typename T, bool AW, bool BW>
__device__ __forceinline__ T contract_one(const StepDesc &st, uint32_t c, const T *pa, const T *pb, int64_t stride) {
    const uint32_t ar = gather_bits(c, st.a);
    const uint32_t br = gather_bits(c, st.b);
    const uint32_t nkk = 1u << st.nk;
    T acc = 0;
    for (uint32_t k = 0; k < nkk; ++k) {
        const T va = op_load_sum<T, AW>(pa, ar | (k << st.a.kshift), stride, st.a);
        const T vb = op_load_sum<T, BW>(pb, br | (k << st.b.kshift), stride, st.b);
        acc = fma(va, vb, acc);
    }
    return acc;
}

// Base pointer of an operand for syndrome b of the plain executor.
template <typename T>
__device__ __forceinline__ const T *plain_base(const OperandDesc &op, int64_t b, const T *ws, int64_t ldw,
                                               Pools<T> pools, const uint8_t *__restrict__ synT) {
    if (op.kind == kInter) return ws + op.off * ldw + b;
    if (op.kind == kConst) return pools.cpool + op.off;
    uint32_t v = 0;
    for (int j = 0; j < op.nev; ++j) v |= (uint32_t)synT[(int64_t)op.ev_bit[j] * ldw + b] << j;
    return pools.spool + op.off + ((int64_t)v << op.rank);
}

// Per-slot, per-syndrome scale of the wide FP32 executor: the largest
// entry's bits (atomic max of non-negative floats) and the exponent, laid out
// [slot][ldw] like the workspace.
struct PlainScale {
    uint32_t *mx;
    int32_t *ex;
};

template <typename T, bool SCALE, bool AW, bool BW>
__device__ __forceinline__ T plain_loop(const StepDesc &st, const T *pa, const T *pb, T *pc, int64_t ldw,
                                        int shift) {
    const uint32_t nc = 1u << st.rc;
    const int cexp = st.a.cexp + st.b.cexp;
    T lmax = 0;
    for (uint32_t c = blockIdx.y; c < nc; c += gridDim.y) {
        T acc = contract_one<T, AW, BW>(st, c, pa, pb, ldw);
        if (SCALE) {
            acc = ldexp(acc, shift);
            lmax = acc > lmax ? acc : lmax;
        } else if (cexp != 0) {
            acc = ldexp(acc, cexp);
        }
        pc[(int64_t)c * ldw] = acc;
    }
    return lmax;
}

// One plan step for a chunk of syndromes: thread = (output entry c, syndrome
// b); c is uniform across the block, so all index arithmetic is scalar and a
// wavefront reads 64 consecutive syndromes of one entry.  With SCALE (FP32),
// every per-syndrome tensor carries a power-of-two exponent as in the fused
// executor; the producer's maximum is gathered with a global atomic max.
template <typename T, bool SCALE>
__global__ void __launch_bounds__(256) plain_step_kernel(StepDesc st, T *__restrict__ ws, int64_t ldw, int64_t nb,
                                                         Pools<T> pools, const uint8_t *__restrict__ synT,
                                                         PlainScale sc) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const T *pa = plain_base(st.a, b, ws, ldw, pools, synT);
    const T *pb = plain_base(st.b, b, ws, ldw, pools, synT);
    T *pc = ws + st.c_off * ldw + b;
    int shift = 0, ea = 0, eb = 0;
    if (SCALE) {
        int na = 0, nbx = 0;
        if (st.a.kind == kInter) {
            const float m = __uint_as_float(sc.mx[(int64_t)st.a.slot * ldw + b]);
            na = m > 0.f ? ilogbf(m) : 0;
            ea = sc.ex[(int64_t)st.a.slot * ldw + b] + na;
        } else {
            ea = st.a.cexp;
        }
        if (st.b.kind == kInter) {
            const float m = __uint_as_float(sc.mx[(int64_t)st.b.slot * ldw + b]);
            nbx = m > 0.f ? ilogbf(m) : 0;
            eb = sc.ex[(int64_t)st.b.slot * ldw + b] + nbx;
        } else {
            eb = st.b.cexp;
        }
        shift = -(na + nbx);
    }
    const bool aw = st.a.kind == kInter, bw = st.b.kind == kInter;
    T lmax;
    if (aw && bw) lmax = plain_loop<T, SCALE, true, true>(st, pa, pb, pc, ldw, shift);
    else if (aw) lmax = plain_loop<T, SCALE, true, false>(st, pa, pb, pc, ldw, shift);
    else if (bw) lmax = plain_loop<T, SCALE, false, true>(st, pa, pb, pc, ldw, shift);
    else lmax = plain_loop<T, SCALE, false, false>(st, pa, pb, pc, ldw, shift);
    if (SCALE) {
        atomicMax(&sc.mx[(int64_t)st.c_slot * ldw + b], __float_as_uint((float)lmax));
        if (blockIdx.y == 0) sc.ex[(int64_t)st.c_slot * ldw + b] = ea + eb;
    }
}

__global__ void zero_u32_kernel(uint32_t *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Wide FP32 result: marginals with the exponent applied, first arg-max, tie
// flag, and low-margin syndromes appended to the FP64 re-check list.
__global__ void plain_finalize32_kernel(const float *__restrict__ ws, int64_t ldw, int64_t nb, int64_t root_off,
                                        const int32_t *__restrict__ ex_root, int nout, double tie_tol,
                                        double recheck_tol, int64_t out_base, double *__restrict__ marginals,
                                        int32_t *__restrict__ classes, uint8_t *__restrict__ ties,
                                        int32_t *__restrict__ list, int32_t *__restrict__ count) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int e = ex_root[b];
    float best = -1, second = -1;
    int bi = 0;
    for (int j = 0; j < nout; ++j) {
        const float x = ws[(root_off + j) * ldw + b];
        if (marginals) marginals[(out_base + b) * nout + j] = ldexp((double)x, e);
        if (j == 0 || x > best) {
            if (j > 0) second = best;
            best = x;
            bi = j;
        } else if (x > second) {
            second = x;
        }
    }
    if (classes) classes[out_base + b] = bi;
    const bool low = nout > 1 && second >= 0 && (double)(best - second) <= tie_tol * (double)best;
    if (ties) ties[out_base + b] = low ? 1 : 0;
    const bool rc = !(best > 0) ||
                    (nout > 1 && second >= 0 && (double)(best - second) <= recheck_tol * (double)best);
    if (rc) {
        const int32_t slot = atomicAdd(count, 1);
        list[slot] = (int32_t)(out_base + b);
    }
}

__global__ void gather_rows_kernel(const uint8_t *__restrict__ syn, int nbits, const int32_t *__restrict__ list,
                                   int64_t n, uint8_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nbits) return;
    const int64_t r = i / nbits;
    out[i] = syn[(int64_t)list[r] * nbits + (i - r * nbits)];
}

__global__ void scatter_results_kernel(const int32_t *__restrict__ list, int64_t n, int nout,
                                       const double *__restrict__ mar, const int32_t *__restrict__ cls,
                                       const uint8_t *__restrict__ tie, double *__restrict__ d_mar,
                                       int32_t *__restrict__ d_cls, uint8_t *__restrict__ d_tie) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t b = list[r];
    if (d_mar)
        for (int j = 0; j < nout; ++j) d_mar[b * nout + j] = mar[r * nout + j];
    if (d_cls) d_cls[b] = cls[r];
    if (d_tie) d_tie[b] = tie[r];
}

// First maximum (Julia findmax order) and top-2 gap of one syndrome.
__global__ void plain_finalize_kernel(const double *__restrict__ ws, int64_t ldw, int64_t nb, int64_t root_off,
                                      int nout, double tie_tol, int64_t out_base, double *__restrict__ marginals,
                                      int32_t *__restrict__ classes, uint8_t *__restrict__ ties) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double best = -1, second = -1;
    int bi = 0;
    for (int j = 0; j < nout; ++j) {
        const double x = ws[(root_off + j) * ldw + b];
        if (marginals) marginals[(out_base + b) * nout + j] = x;
        if (j == 0 || x > best) {
            if (j > 0) second = best;
            best = x;
            bi = j;
        } else if (x > second) {
            second = x;
        }
    }
    if (classes) classes[out_base + b] = bi;
    if (ties) ties[out_base + b] = (nout > 1 && second >= 0 && (best - second) <= tie_tol * best) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Fused executor
// ---------------------------------------------------------------------------

struct FusedArgs {
    const StepDesc *steps;
    int nsteps;
    const void *cpool;
    const void *spool;
    const uint8_t *syn;       // [n][nbits]
    int64_t n;
    int nbits;
    int glog;                 // log2 G
    int nout;
    int64_t ws_entries;
    int nslots;
    int64_t root_off;
    int root_slot;
    const int32_t *list;      // optional: syndrome ids to process
    const int32_t *list_count;
    double *marginals;
    int32_t *classes;
    uint8_t *ties;
    double tie_tol;
    double recheck_tol;
    int32_t *recheck_list;    // FP32 pass: low-margin syndromes are appended here
    int32_t *recheck_count;
    int32_t *recheck_batch;   // mixed-code executor: batch of each listed row
    int64_t recheck_cap;      // capacity of the re-check list
    int32_t batch_id;
};

// Base pointer of an operand for tile slot g of the fused executor.
template <typename T>
__device__ __forceinline__ const T *fused_base(const OperandDesc &op, int g, int glog, const T *work,
                                               const T *__restrict__ cpool, const T *__restrict__ spool,
                                               const uint8_t *synl) {
    if (op.kind == kInter) return work + (op.off << glog) + g;
    if (op.kind == kConst) return cpool + op.off;
    uint32_t v = 0;
    for (int j = 0; j < op.nev; ++j) v |= (uint32_t)synl[(op.ev_bit[j] << glog) + g] << j;
    return spool + op.off + ((int64_t)v << op.rank);
}

template <typename T, bool SCALE, bool AW, bool BW>
__device__ __forceinline__ T fused_loop(const StepDesc &st, const T *pa, const T *pb, T *pc, int glog, int c0,
                                       int cstep, int shift) {
    const int nc = 1 << st.rc;
    const int G = 1 << glog;
    const int cexp = st.a.cexp + st.b.cexp;
    T lmax = 0;
    for (int c = c0; c < nc; c += cstep) {
        T acc = contract_one<T, AW, BW>(st, (uint32_t)c, pa, pb, (int64_t)G);
        if (SCALE) {
            acc = ldexp(acc, shift);
            lmax = acc > lmax ? acc : lmax;
        } else if (cexp != 0) {
            acc = ldexp(acc, cexp);
        }
        pc[c << glog] = acc;
    }
    return lmax;
}

__device__ __forceinline__ uint32_t max_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ uint32_t max_bits(double v) { return __float_as_uint((float)v); }

// LDS layout: work[ws_entries][G] T | mx[nslots][G] u64 | ex[nslots][G] i32 |
//             syn[nbits][G] u8 | lowf[G] i32
// Thread t serves tile slot g = t mod G.  With GW (slab mode) the work array
// lives instead in the workgroup's private slab of global memory `gwork`
// (same [entry][G] layout) and the LDS holds only the scales and syndromes;
// this serves plans whose intermediates exceed the LDS.
//
// One tile of G syndromes (rows b0.. of args.syn, or list entries b0.. of
// args.list) through the whole step program.  Starts and ends with a
// workgroup barrier, so consecutive tiles may reuse the LDS freely.
template <typename T, bool SCALE, bool GW = false>
__device__ void fused_tile(const FusedArgs &args, int64_t b0, int64_t count, uint32_t tag_base,
                           unsigned char *smem, T *gwork = nullptr) {
    const int G = 1 << args.glog;
    const int glog = args.glog;
    T *work = GW ? gwork : reinterpret_cast<T *>(smem);
    size_t off = GW ? 0 : (size_t)args.ws_entries * G * sizeof(T);
    off = (off + 15) & ~size_t(15);
    unsigned long long *mx = reinterpret_cast<unsigned long long *>(smem + off);
    int32_t *ex = reinterpret_cast<int32_t *>(smem + off + (SCALE ? (size_t)args.nslots * G * 8 : 0));
    uint8_t *synl = smem + off + (SCALE ? (size_t)args.nslots * G * 12 : 0);
    int32_t *lowf = reinterpret_cast<int32_t *>(synl + (((size_t)args.nbits * G + 3) & ~size_t(3)));
    const T *cpool = reinterpret_cast<const T *>(args.cpool);
    const T *spool = reinterpret_cast<const T *>(args.spool);
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int g = tid & (G - 1);
    const int c0 = tid >> glog;
    const int cstep = nthr >> glog;
    const int lane = tid & 63;
    // An operand whose largest entry is below 2^-100 means its step lost
    // more than 100 binary orders of magnitude: such syndromes are re-checked.
    constexpr int kLowExp = -100;
    // ---- syndrome bits of the tile -> LDS [bit][g]; clear the scales ----
    if (!args.list) {
        const int64_t rows = min((int64_t)G, count - b0);
        const uint8_t *src = args.syn + b0 * args.nbits;
        const int tot = G * args.nbits;
        for (int i = tid; i < tot; i += nthr) {
            const int r = i / args.nbits, bit = i - r * args.nbits;
            synl[(bit << glog) + r] = (r < rows) ? (src[i] & 1u) : 0u;
        }
    } else {
        const int tot = G * args.nbits;
        for (int i = tid; i < tot; i += nthr) {
            const int r = i / args.nbits, bit = i - r * args.nbits;
            const int64_t q = b0 + r;
            synl[(bit << glog) + r] = (q < count) ? (args.syn[(int64_t)args.list[q] * args.nbits + bit] & 1u) : 0u;
        }
    }
    if (SCALE) {
        for (int i = tid; i < args.nslots * G; i += nthr) mx[i] = 0ull;
        if (tid < G) lowf[tid] = 0;
    }
    __syncthreads();
    for (int s = 0; s < args.nsteps; ++s) {
        const StepDesc &st = args.steps[s];
        int ea = 0, eb = 0, shift = 0;
        if (SCALE) {
            int na = 0, nb = 0;
            if (st.a.kind == kInter) {
                const float m = __uint_as_float((uint32_t)(mx[(st.a.slot << glog) + g] & 0xffffffffull));
                na = (m > 0.f) ? ilogbf(m) : 0;
                ea = ex[(st.a.slot << glog) + g] + na;
            } else {
                ea = st.a.cexp;
            }
            if (st.b.kind == kInter) {
                const float m = __uint_as_float((uint32_t)(mx[(st.b.slot << glog) + g] & 0xffffffffull));
                nb = (m > 0.f) ? ilogbf(m) : 0;
                eb = ex[(st.b.slot << glog) + g] + nb;
            } else {
                eb = st.b.cexp;
            }
            shift = -(na + nb);
            if ((st.a.kind == kInter && na < kLowExp) || (st.b.kind == kInter && nb < kLowExp)) lowf[g] = 1;
        }
        const T *pa = fused_base(st.a, g, glog, work, cpool, spool, synl);
        const T *pb = fused_base(st.b, g, glog, work, cpool, spool, synl);
        T *pc = work + (st.c_off << glog) + g;
        const bool aw = st.a.kind == kInter, bw = st.b.kind == kInter;
        T lmax;
        if (aw && bw) lmax = fused_loop<T, SCALE, true, true>(st, pa, pb, pc, glog, c0, cstep, shift);
        else if (aw) lmax = fused_loop<T, SCALE, true, false>(st, pa, pb, pc, glog, c0, cstep, shift);
        else if (bw) lmax = fused_loop<T, SCALE, false, true>(st, pa, pb, pc, glog, c0, cstep, shift);
        else lmax = fused_loop<T, SCALE, false, false>(st, pa, pb, pc, glog, c0, cstep, shift);
        if (SCALE) {
            for (int o = G; o < 64; o <<= 1) {
                const float other = __shfl_xor((float)lmax, o);
                lmax = other > (float)lmax ? (T)other : lmax;
            }
            if (lane < G) {
                const unsigned long long tagged =
                    ((unsigned long long)(tag_base + (uint32_t)s) << 32) | (unsigned long long)max_bits(lmax);
                atomicMax(&mx[(st.c_slot << glog) + g], tagged);
            }
            if (tid < G) ex[(st.c_slot << glog) + g] = ea + eb;
        }
        __syncthreads();
    }
    // ---- per-syndrome result ----
    if (tid < G) {
        const int64_t q = b0 + tid;
        if (q < count) {
            const int64_t b = args.list ? (int64_t)args.list[q] : q;
            const int e = SCALE ? ex[(args.root_slot << glog) + g] : 0;
            T best = 0, second = -1;
            int bi = 0;
            for (int j = 0; j < args.nout; ++j) {
                const T x = work[((args.root_off + j) << glog) + g];
                if (args.marginals) args.marginals[b * args.nout + j] = ldexp((double)x, e);
                if (j == 0 || x > best) {
                    if (j > 0) second = best;
                    best = x;
                    bi = j;
                } else if (x > second) {
                    second = x;
                }
            }
            if (args.classes) args.classes[b] = bi;
            const bool low = args.nout > 1 && second >= 0 && (double)(best - second) <= args.tie_tol * (double)best;
            if (args.ties) args.ties[b] = low ? 1 : 0;
            if (SCALE && args.recheck_list) {
                const bool rc = !(best > 0) || lowf[tid] != 0 ||
                                (args.nout > 1 && second >= 0 &&
                                 (double)(best - second) <= args.recheck_tol * (double)best);
                if (rc) {
                    const int32_t slot = atomicAdd(args.recheck_count, 1);
                    if (slot < args.recheck_cap) {
                        args.recheck_list[slot] = (int32_t)b;
                        if (args.recheck_batch) args.recheck_batch[slot] = args.batch_id;
                    }
                }
            }
        }
    }
    __syncthreads();
}

// Persistent single-plan kernel: workgroups stride over the tiles.
template <typename T, bool SCALE>
__global__ void __launch_bounds__(256) fused_kernel(FusedArgs args) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t count = args.list ? (int64_t)(*args.list_count) : args.n;
    const int64_t ntiles = (count + (1 << args.glog) - 1) >> args.glog;
    uint32_t iter = 0;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x, ++iter)
        fused_tile<T, SCALE>(args, t << args.glog, count, iter * (uint32_t)args.nsteps + 1u, smem);
}

// ---------------------------------------------------------------------------
// Mixed-code executor: batches of different codes in one persistent launch
// ---------------------------------------------------------------------------

struct MixedPlan {
    const StepDesc *steps;
    const void *cpool32, *spool32, *cpool64, *spool64;
    int64_t ws_entries, root_off;
    int nsteps, nbits, nout, nslots, root_slot;
    int glog32;          // largest FP32 tile of the plan (a batch may use a smaller one)
    int gw32, gw64;      // 1: intermediates in the workgroup's global slab (FP32 pass / FP64 re-check)
    int pad;
    double tie_tol, recheck_tol;
};

struct MixedBatch {
    const uint8_t *syn;  // [n][nbits] on the device
    double *mar;
    int32_t *cls;
    uint8_t *tie;
    int64_t n;
    int64_t tile0;       // first global tile of this batch
    int plan;
    int glog;            // FP32 tile of this batch
};

// Device-resident call header: written by the host before each call (a
// captured graph replays the same kernel arguments, so everything that
// varies between calls lives here).
struct MixedHeader {
    int32_t nbatches;
    int32_t pad;
    int64_t ntiles;
};

// Per-call counters, cleared by mixed_reset_kernel at the start of a call.
struct MixedCounters {
    unsigned long long next_tile;       // FP32 work queue
    unsigned long long next_recheck;    // FP64 work queue
    int32_t rl_count;                   // re-check list length
    int32_t pad;
    unsigned long long t_start;         // device wall clock at the first tile
};

struct MixedArgs {
    const MixedPlan *plans;
    const MixedBatch *batches;
    const MixedHeader *hdr;
    MixedCounters *ctr;
    unsigned long long *t_done;          // per batch: device wall clock of its last write
    char *slab32, *slab64;               // per-workgroup global slabs
    int64_t slab32_bytes, slab64_bytes;  // bytes per workgroup
};

__device__ __forceinline__ FusedArgs mixed_fused_args(const MixedArgs &m, const MixedPlan &p, const MixedBatch &bt,
                                                      bool f32, int glog) {
    FusedArgs a;
    a.steps = p.steps;
    a.nsteps = p.nsteps;
    a.cpool = f32 ? p.cpool32 : p.cpool64;
    a.spool = f32 ? p.spool32 : p.spool64;
    a.syn = bt.syn;
    a.n = bt.n;
    a.nbits = p.nbits;
    a.glog = glog;
    a.nout = p.nout;
    a.ws_entries = p.ws_entries;
    a.nslots = p.nslots;
    a.root_off = p.root_off;
    a.root_slot = p.root_slot;
    a.list = nullptr;
    a.list_count = nullptr;
    a.marginals = bt.mar;
    a.classes = bt.cls;
    a.ties = bt.tie;
    a.tie_tol = p.tie_tol;
    a.recheck_tol = p.recheck_tol;
    a.recheck_list = nullptr;
    a.recheck_count = nullptr;
    a.recheck_batch = nullptr;
    a.recheck_cap = 0;
    a.batch_id = 0;
    return a;
}

// Next item of a work queue, broadcast to the workgroup through LDS.
__device__ __forceinline__ int64_t mixed_grab(unsigned long long *queue, int64_t *slot) {
    __syncthreads();
    if (threadIdx.x == 0) *slot = (int64_t)atomicAdd(queue, 1ull);
    __syncthreads();
    return *slot;
}

// Workgroups pull global tiles from a queue; tile t belongs to the batch
// whose [tile0, next tile0) range holds it.  All plans share one LDS
// allocation; slab plans use the workgroup's slab instead.  Syndromes of a
// tile whose FP32 margin is below the re-check threshold are re-contracted
// in FP64 by the same workgroup right after the tile (one syndrome at a
// time), so a batch is complete when its last tile is.
__global__ void __launch_bounds__(256) mixed_kernel(MixedArgs m) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int64_t grab;
    __shared__ int32_t loc_n;
    __shared__ int32_t loc_rows[64];
    const int nb = m.hdr->nbatches;
    const int64_t ntiles = m.hdr->ntiles;
    uint32_t tag = 1;
    bool first = true;
    if (threadIdx.x == 0) loc_n = 0;
    for (;;) {
        const int64_t t = mixed_grab(&m.ctr->next_tile, &grab);
        if (t >= ntiles) break;
        if (first && threadIdx.x == 0) atomicMin(&m.ctr->t_start, (unsigned long long)wall_clock64());
        first = false;
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (m.batches[mid].tile0 <= t) lo = mid;
            else hi = mid - 1;
        }
        const MixedBatch bt = m.batches[lo];
        const MixedPlan &p = m.plans[bt.plan];
        FusedArgs a = mixed_fused_args(m, p, bt, true, bt.glog);
        a.recheck_list = loc_rows;
        a.recheck_count = &loc_n;
        a.recheck_cap = 64;
        const int64_t b0 = (t - bt.tile0) << bt.glog;
        if (p.gw32)
            fused_tile<float, true, true>(a, b0, bt.n, tag, smem,
                                          reinterpret_cast<float *>(m.slab32 + (int64_t)blockIdx.x * m.slab32_bytes));
        else
            fused_tile<float, true, false>(a, b0, bt.n, tag, smem);
        tag += (uint32_t)p.nsteps;
        const int nre = loc_n;
        if (nre > 0) {
            FusedArgs r = mixed_fused_args(m, p, bt, false, 0);
            for (int e = 0; e < nre; ++e) {
                const int64_t row = loc_rows[e];
                if (p.gw64)
                    fused_tile<double, false, true>(
                        r, row, bt.n, 1u, smem, reinterpret_cast<double *>(m.slab64 + (int64_t)blockIdx.x * m.slab64_bytes));
                else
                    fused_tile<double, false, false>(r, row, bt.n, 1u, smem);
            }
            if (threadIdx.x == 0) {
                atomicAdd(&m.ctr->rl_count, nre);
                loc_n = 0;
            }
        }
        if (threadIdx.x == 0) atomicMax(&m.t_done[lo], (unsigned long long)wall_clock64());
    }
}

__global__ void mixed_reset_kernel(MixedCounters *ctr, unsigned long long *t_done, int cap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctr->next_tile = 0ull;
        ctr->next_recheck = 0ull;
        ctr->rl_count = 0;
        ctr->t_start = ~0ull;
    }
    if (i < cap) t_done[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Max-product executor (TNMAP)
// ---------------------------------------------------------------------------

struct MapArgs {
    const StepDesc *steps;
    const MapStep *msteps;
    const int32_t *labels;
    int nsteps;
    const double *cpool;
    const double *spool;
    const uint8_t *syn;
    int64_t n;
    int nbits;
    int glog;
    int64_t ws_entries;
    int64_t arg_entries;
    int asg_words;
    int64_t root_off;
    int n_config;
    uint8_t *config;   // [n][n_config]
    double *logp;      // [n] natural log of the maximum, or null
};

// Element of a max-product operand; single-side labels are enumerated by the
// caller (they are eliminated here and their arg-max is recorded).
template <bool AW>
__device__ __forceinline__ double map_load(const double *base, uint32_t idx, int glog) {
    return AW ? base[idx << glog] : base[idx];
}

// LDS: work[ws][G] f64 | arg[arg_entries][G] u16 | syn[nbits][G] u8 | asg[G][words] u32
// Forward: each output keeps the maximum over its contracted and single-side
// labels (first maximum in the order k, then sa, then sb) and records which.
// Backtrace (one thread per syndrome): from the scalar root down, each step's
// record fixes the labels it eliminated; the error bits are read at the end.
__global__ void __launch_bounds__(256) map_kernel(MapArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int glog = a.glog, G = 1 << glog;
    double *work = reinterpret_cast<double *>(smem);
    size_t off = (size_t)a.ws_entries * G * sizeof(double);
    uint16_t *arg = reinterpret_cast<uint16_t *>(smem + off);
    off += (((size_t)a.arg_entries * G * sizeof(uint16_t)) + 15) & ~size_t(15);
    uint8_t *synl = smem + off;
    off += (((size_t)a.nbits * G) + 15) & ~size_t(15);
    uint32_t *asg = reinterpret_cast<uint32_t *>(smem + off);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int g = tid & (G - 1), c0 = tid >> glog, cstep = nthr >> glog;
    const int64_t ntiles = (a.n + G - 1) >> glog;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t b0 = t << glog;
        const int64_t rows = min((int64_t)G, a.n - b0);
        const uint8_t *src = a.syn + b0 * a.nbits;
        for (int i = tid; i < G * a.nbits; i += nthr) {
            const int r = i / a.nbits, bit = i - r * a.nbits;
            synl[(bit << glog) + r] = (r < rows) ? (src[i] & 1u) : 0u;
        }
        for (int i = tid; i < G * a.asg_words; i += nthr) asg[i] = 0u;
        __syncthreads();
        for (int s = 0; s < a.nsteps; ++s) {
            const StepDesc &st = a.steps[s];
            const MapStep &ms = a.msteps[s];
            const double *pa = fused_base(st.a, g, glog, work, a.cpool, a.spool, synl);
            const double *pb = fused_base(st.b, g, glog, work, a.cpool, a.spool, synl);
            const bool aw = st.a.kind == kInter, bw = st.b.kind == kInter;
            const int cexp = st.a.cexp + st.b.cexp;
            const uint32_t nkk = 1u << st.nk, nsa = 1u << st.a.nsingle, nsb = 1u << st.b.nsingle;
            for (int c = c0; c < (1 << st.rc); c += cstep) {
                const uint32_t ar = gather_bits((uint32_t)c, st.a);
                const uint32_t br = gather_bits((uint32_t)c, st.b);
                double best = -1.0;
                uint32_t bi = 0;
                for (uint32_t k = 0; k < nkk; ++k) {
                    for (uint32_t qa = 0; qa < nsa; ++qa) {
                        const uint32_t ia = ar | (qa << st.a.sshift) | (k << st.a.kshift);
                        const double va = aw ? map_load<true>(pa, ia, glog) : map_load<false>(pa, ia, glog);
                        for (uint32_t qb = 0; qb < nsb; ++qb) {
                            const uint32_t ib = br | (qb << st.b.sshift) | (k << st.b.kshift);
                            const double vb = bw ? map_load<true>(pb, ib, glog) : map_load<false>(pb, ib, glog);
                            const double v = va * vb;
                            if (v > best) {
                                best = v;
                                bi = k | (qa << st.nk) | (qb << (st.nk + st.a.nsingle));
                            }
                        }
                    }
                }
                if (cexp != 0) best = ldexp(best, cexp);
                work[((st.c_off + c) << glog) + g] = best;
                arg[((ms.arg_off + c) << glog) + g] = (uint16_t)bi;
            }
            __syncthreads();
        }
        if (tid < G && b0 + tid < a.n) {
            uint32_t *as = asg + tid * a.asg_words;
            for (int s = a.nsteps - 1; s >= 0; --s) {
                const MapStep &ms = a.msteps[s];
                const int32_t *L = a.labels + ms.off;
                uint32_t c = 0;
                for (int j = 0; j < ms.nc; ++j) c |= ((as[L[j] >> 5] >> (L[j] & 31)) & 1u) << j;
                const uint32_t rec = arg[((ms.arg_off + c) << glog) + tid];
                const int ne = ms.nk + ms.nsa + ms.nsb;
                for (int j = 0; j < ne; ++j) {
                    const int l = L[ms.nc + j];
                    as[l >> 5] |= ((rec >> j) & 1u) << (l & 31);
                }
            }
            const int64_t b = b0 + tid;
            for (int v = 0; v < a.n_config; ++v) a.config[b * a.n_config + v] = (uint8_t)((as[v >> 5] >> (v & 31)) & 1u);
            if (a.logp) a.logp[b] = log(work[(a.root_off << glog) + tid]);
        }
        __syncthreads();
    }
}

__global__ void zero_i32_kernel(int32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

namespace tqec {
int set_error_shared(int code, const std::string &msg) { return set_error(code, msg); }
}  // namespace tqec

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

struct DeviceData {
    float *tab32 = nullptr;     // constant pool followed by the sliced pool (specialised kernels)
    double *tab64 = nullptr;
    MapStep *msteps = nullptr;
    int32_t *mlabels = nullptr;
    StepDesc *steps = nullptr;
    double *cpool64 = nullptr, *spool64 = nullptr;
    float *cpool32 = nullptr, *spool32 = nullptr;
};

struct Workspace {
    int device = -1;
    void *ws = nullptr;        // plain workspace
    size_t ws_bytes = 0;
    uint8_t *synT = nullptr;
    size_t synT_bytes = 0;
    int32_t *list = nullptr;   // recheck list + count
    int64_t list_cap = 0;
    int32_t *count = nullptr;
    int32_t *host_count = nullptr;
    int32_t *segc = nullptr;        // register-resident kernel: re-check rows per tile segment
    int64_t segc_cap = 0;
    int64_t count_nseg = 0;         // > 0: the last call's re-check count is the sum of segc[0 .. count_nseg)
    uint8_t *d_syn = nullptr; size_t d_syn_bytes = 0;   // staging for the host API
    double *d_mar = nullptr; size_t d_mar_bytes = 0;
    int32_t *d_cls = nullptr; size_t d_cls_bytes = 0;
    uint8_t *d_tie = nullptr; size_t d_tie_bytes = 0;
    int64_t last_recheck = -1;
    bool count_on_device = false;   // the last call's re-check count is in `count` (not yet read)
    void *smx = nullptr; size_t smx_bytes = 0;       // wide FP32 scales
    void *sex = nullptr; size_t sex_bytes = 0;
    uint8_t *g_syn = nullptr; size_t g_syn_bytes = 0; // wide FP32 re-check staging
    double *g_mar = nullptr; size_t g_mar_bytes = 0;
    int32_t *g_cls = nullptr; size_t g_cls_bytes = 0;
    uint8_t *g_tie = nullptr; size_t g_tie_bytes = 0;
    void *spec_slab = nullptr; size_t spec_slab_bytes = 0;   // slab configurations of the specialised kernels
    // Per-call HIP events (main start/stop, re-check start/stop), accumulated
    // while timing is enabled and consumed by tqec_kernel_time_totals.
    std::vector<std::array<hipEvent_t, 4>> ev;
    std::vector<std::array<char, 4>> ev_used;
    size_t n_ev = 0;
};

struct FusedCfg {
    int glog = -1;      // -1: unsupported
    int threads = 256;
    size_t lds = 0;
    int blocks = 0;
};

// A compiled plan-specialised kernel (TQEC_MODE_SPEC_F32).
struct SpecKernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    SpecCfg cfg;
    size_t lds = 0;
    int blocks = 0;
};

}  // namespace

struct tqec_plan {
    CompiledPlan cp;
    double tie_tol = 1e-12;
    double recheck_tol = 1e-3;
    int64_t plain_chunk_cap_bytes = int64_t(8) << 30;
    std::mutex mu;
    std::map<int, DeviceData> dev;
    std::map<std::pair<int, hipStream_t>, Workspace> ws;
    FusedCfg f64, f32;
    bool cfg_done = false;
    bool timing = false;
    size_t lds_limit = 0;
    int n_cu = 0;
    SpecCfg spec_cfg;
    bool spec_cfg_set = false;
    std::map<std::tuple<int, int, int, int, int, int, int>, SpecKernel> spec;   // (device, glog, threads, tab_lds, waves, fp64, flags)
    bool map_spec = true;               // max-product plans: plan-specialised kernel (else the generic one)
    std::string map_spec_note;          // why the specialised max-product kernel was given up, if it was
    SpecCfg spec64_cfg;                 // TQEC_MODE_SPEC_F64
    bool spec64_cfg_set = false;
    SpecCfg specb_cfg;                  // TQEC_MODE_SPEC_BF16
    bool specb_cfg_set = false;
};

namespace {

template <typename T>
int upload(const std::vector<double> &h, T **d) {
    std::vector<T> tmp(h.size() ? h.size() : 1, T(0));
    for (size_t i = 0; i < h.size(); ++i) tmp[i] = (T)h[i];
    TQ_HIP(hipMalloc((void **)d, tmp.size() * sizeof(T)));
    TQ_HIP(hipMemcpy(*d, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return TQEC_OK;
}

int ensure_device(tqec_plan *p, int device, DeviceData **out) {
    auto it = p->dev.find(device);
    if (it != p->dev.end()) {
        *out = &it->second;
        return TQEC_OK;
    }
    DeviceData dd;
    const auto &cp = p->cp;
    TQ_HIP(hipMalloc((void **)&dd.steps, std::max<size_t>(1, cp.steps.size()) * sizeof(StepDesc)));
    TQ_HIP(hipMemcpy(dd.steps, cp.steps.data(), cp.steps.size() * sizeof(StepDesc), hipMemcpyHostToDevice));
    int rc;
    if ((rc = upload<double>(cp.const_pool, &dd.cpool64))) return rc;
    if ((rc = upload<double>(cp.sliced_pool, &dd.spool64))) return rc;
    if ((rc = upload<float>(cp.const_pool, &dd.cpool32))) return rc;
    if ((rc = upload<float>(cp.sliced_pool, &dd.spool32))) return rc;
    {
        std::vector<double> both(cp.const_pool);
        both.insert(both.end(), cp.sliced_pool.begin(), cp.sliced_pool.end());
        if ((rc = upload<float>(both, &dd.tab32))) return rc;
        if ((rc = upload<double>(both, &dd.tab64))) return rc;
    }
    if (cp.semiring == TQEC_SEMIRING_MAX) {
        TQ_HIP(hipMalloc((void **)&dd.msteps, std::max<size_t>(1, cp.map_steps.size()) * sizeof(MapStep)));
        TQ_HIP(hipMemcpy(dd.msteps, cp.map_steps.data(), cp.map_steps.size() * sizeof(MapStep), hipMemcpyHostToDevice));
        TQ_HIP(hipMalloc((void **)&dd.mlabels, std::max<size_t>(1, cp.map_labels.size()) * sizeof(int32_t)));
        TQ_HIP(hipMemcpy(dd.mlabels, cp.map_labels.data(), cp.map_labels.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    p->dev[device] = dd;
    *out = &p->dev[device];
    return TQEC_OK;
}

// LDS of a fused tile besides the work array: scales, syndrome bits, flags.
size_t fused_lds_aux(const CompiledPlan &cp, int glog, bool scale) {
    const size_t G = size_t(1) << glog;
    size_t b = 0;
    if (scale) b += (size_t)cp.n_slots * G * 12;
    b += ((size_t)cp.n_syn * G + 3) & ~size_t(3);
    if (scale) b += G * sizeof(int32_t);
    return b;
}

size_t fused_lds_bytes(const CompiledPlan &cp, int glog, size_t elem, bool scale) {
    const size_t G = size_t(1) << glog;
    size_t b = (size_t)cp.ws_entries * G * elem;
    b = (b + 15) & ~size_t(15);
    return b + fused_lds_aux(cp, glog, scale);
}

// Workgroups per CU a configuration can keep resident (LDS and wave slots).
int resident_per_cu(size_t lds, int threads, size_t lds_limit) {
    const int by_lds = (int)(lds_limit / std::max<size_t>(lds, 1));
    const int by_waves = 32 / std::max(1, threads / 64);
    return std::max(1, std::min(by_lds, by_waves));
}

FusedCfg make_cfg(const CompiledPlan &cp, int glog, int threads, size_t elem, bool scale, size_t lds_limit,
                  int n_cu) {
    FusedCfg c;
    c.glog = glog;
    c.threads = threads;
    c.lds = fused_lds_bytes(cp, glog, elem, scale);
    c.blocks = n_cu * resident_per_cu(c.lds, threads, lds_limit);
    return c;
}

// Default before tuning: the largest tile whose LDS image lets at least four
// workgroups share a CU (else two, else one), 128 threads.
FusedCfg choose_fused(const CompiledPlan &cp, size_t elem, bool scale, size_t lds_limit, int n_cu) {
    for (int div : {4, 2, 1}) {
        for (int gl = 6; gl >= 0; --gl) {
            if (fused_lds_bytes(cp, gl, elem, scale) <= lds_limit / div)
                return make_cfg(cp, gl, std::max(64, std::min(128, 1 << gl)), elem, scale, lds_limit, n_cu);
        }
    }
    return FusedCfg();
}

int configure(tqec_plan *p) {
    if (p->cfg_done) return TQEC_OK;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    TQ_HIP(hipGetDeviceProperties(&prop, dev));
    size_t lds = prop.sharedMemPerBlock;
    if (prop.maxSharedMemoryPerMultiProcessor > 0) lds = std::min(lds, (size_t)prop.maxSharedMemoryPerMultiProcessor);
    p->lds_limit = lds;
    p->n_cu = prop.multiProcessorCount;
    p->f64 = choose_fused(p->cp, sizeof(double), false, lds, p->n_cu);
    p->f32 = choose_fused(p->cp, sizeof(float), true, lds, p->n_cu);
    // Opt both kernels in to the whole LDS once; a launch uses what it asks for.
    TQ_HIP(hipFuncSetAttribute((const void *)fused_kernel<double, false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    TQ_HIP(hipFuncSetAttribute((const void *)fused_kernel<float, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    p->cfg_done = true;
    return TQEC_OK;
}

template <typename T>
int grow(T **ptr, size_t *cap, size_t need) {
    if (*cap >= need) return TQEC_OK;
    if (*ptr) TQ_HIP(hipFree(*ptr));
    *ptr = nullptr;
    *cap = 0;
    TQ_HIP(hipMalloc((void **)ptr, need));
    *cap = need;
    return TQEC_OK;
}

// Record event i (0/1: main start/stop, 2/3: re-check start/stop) of the
// current call; event 0 opens a new call slot.
int record_or(hipError_t e) {
    if (e != hipSuccess) return set_error(TQEC_EDEVICE, std::string("HIP: ") + hipGetErrorString(e));
    return TQEC_OK;
}

int record(tqec_plan *p, Workspace &w, int i, hipStream_t stream) {
    if (!p->timing) return TQEC_OK;
    if (i == 0) {
        if (w.n_ev == w.ev.size()) {
            std::array<hipEvent_t, 4> e{};
            // timing only: no system-scope cache writeback when they are recorded
            for (auto &x : e) TQ_HIP(hipEventCreateWithFlags(&x, hipEventDisableSystemFence));
            w.ev.push_back(e);
            w.ev_used.push_back({0, 0, 0, 0});
        }
        w.ev_used[w.n_ev] = {0, 0, 0, 0};
        ++w.n_ev;
    }
    if (w.n_ev == 0) return TQEC_OK;
    TQ_HIP(hipEventRecord(w.ev[w.n_ev - 1][i], stream));
    w.ev_used[w.n_ev - 1][i] = 1;
    return TQEC_OK;
}

int ensure_count(Workspace &w, hipStream_t stream);
int read_recheck_count(Workspace &w, int64_t *out);

int run_plain(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, double *d_mar,
              int32_t *d_cls, uint8_t *d_tie, hipStream_t stream) {
    const CompiledPlan &cp = p->cp;
    const int64_t per = std::max<int64_t>(1, cp.ws_entries) * (int64_t)sizeof(double);
    int64_t chunk = std::max<int64_t>(256, p->plain_chunk_cap_bytes / per);
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(n, 1));
    chunk = (chunk + 255) / 256 * 256;
    int rc;
    if ((rc = grow((char **)&w.ws, &w.ws_bytes, (size_t)(per * chunk)))) return rc;
    if ((rc = grow(&w.synT, &w.synT_bytes, (size_t)std::max(1, cp.n_syn) * chunk))) return rc;
    Pools<double> pools{dd->cpool64, dd->spool64};
    const int nout = 1 << cp.n_out;
    if ((rc = record(p, w, 0, stream))) return rc;
    for (int64_t s0 = 0; s0 < n; s0 += chunk) {
        const int64_t nb = std::min(chunk, n - s0);
        const dim3 blk(256);
        const dim3 grd((unsigned)((nb + 255) / 256));
        hipLaunchKernelGGL(transpose_syndromes_kernel, grd, blk, 0, stream, d_syn + s0 * cp.n_syn, nb, cp.n_syn,
                           w.synT, chunk);
        for (const StepDesc &st : cp.steps) {
            const unsigned ny = (unsigned)std::min<int64_t>(int64_t(1) << st.rc, 65535);
            hipLaunchKernelGGL((plain_step_kernel<double, false>), dim3(grd.x, ny), blk, 0, stream, st,
                               (double *)w.ws, chunk, nb, pools, (const uint8_t *)w.synT, PlainScale{nullptr, nullptr});
        }
        hipLaunchKernelGGL(plain_finalize_kernel, grd, blk, 0, stream, (const double *)w.ws, chunk, nb, cp.root_off,
                           nout, p->tie_tol, s0, d_mar, d_cls, d_tie);
    }
    if ((rc = record(p, w, 1, stream))) return rc;
    TQ_HIP(hipGetLastError());
    return TQEC_OK;
}

// Wide FP32 executor: the plain executor's layout and launches in FP32 with
// per-tensor, per-syndrome power-of-two scaling; syndromes whose top-2 margin
// is below the re-check threshold are gathered and re-contracted by the plain
// FP64 executor (one host synchronisation for the list length).
int run_wide(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, double *d_mar,
             int32_t *d_cls, uint8_t *d_tie, hipStream_t stream) {
    const CompiledPlan &cp = p->cp;
    const int64_t per = std::max<int64_t>(1, cp.ws_entries) * (int64_t)sizeof(float);
    int64_t chunk = std::max<int64_t>(256, p->plain_chunk_cap_bytes / per);
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(n, 1));
    chunk = (chunk + 255) / 256 * 256;
    int rc;
    if ((rc = grow((char **)&w.ws, &w.ws_bytes, (size_t)(per * chunk)))) return rc;
    if ((rc = grow(&w.synT, &w.synT_bytes, (size_t)std::max(1, cp.n_syn) * chunk))) return rc;
    const size_t scale_bytes = (size_t)std::max(1, cp.n_slots) * chunk * 4;
    if ((rc = grow((char **)&w.smx, &w.smx_bytes, scale_bytes))) return rc;
    if ((rc = grow((char **)&w.sex, &w.sex_bytes, scale_bytes))) return rc;
    if (w.list_cap < n) {
        if (w.list) TQ_HIP(hipFree(w.list));
        w.list = nullptr;
        w.list_cap = 0;
        TQ_HIP(hipMalloc((void **)&w.list, (size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)));
        w.list_cap = n;
    }
    if ((rc = ensure_count(w, stream))) return rc;
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, stream, w.count, 1);
    w.count_nseg = 0;
    Pools<float> pools{dd->cpool32, dd->spool32};
    PlainScale sc{(uint32_t *)w.smx, (int32_t *)w.sex};
    const int nout = 1 << cp.n_out;
    if ((rc = record(p, w, 0, stream))) return rc;
    for (int64_t s0 = 0; s0 < n; s0 += chunk) {
        const int64_t nb = std::min(chunk, n - s0);
        const dim3 blk(256);
        const dim3 grd((unsigned)((nb + 255) / 256));
        hipLaunchKernelGGL(transpose_syndromes_kernel, grd, blk, 0, stream, d_syn + s0 * cp.n_syn, nb, cp.n_syn,
                           w.synT, chunk);
        for (const StepDesc &st : cp.steps) {
            hipLaunchKernelGGL(zero_u32_kernel, grd, blk, 0, stream, sc.mx + (int64_t)st.c_slot * chunk, nb);
            const unsigned ny = (unsigned)std::min<int64_t>(int64_t(1) << st.rc, 65535);
            hipLaunchKernelGGL((plain_step_kernel<float, true>), dim3(grd.x, ny), blk, 0, stream, st,
                               (float *)w.ws, chunk, nb, pools, (const uint8_t *)w.synT, sc);
        }
        hipLaunchKernelGGL(plain_finalize32_kernel, grd, blk, 0, stream, (const float *)w.ws, chunk, nb, cp.root_off,
                           (const int32_t *)sc.ex + (int64_t)cp.root_slot * chunk, nout, p->tie_tol, p->recheck_tol,
                           s0, d_mar, d_cls, d_tie, w.list, w.count);
    }
    if ((rc = record(p, w, 1, stream))) return rc;
    TQ_HIP(hipMemcpyAsync(w.host_count, w.count, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    TQ_HIP(hipStreamSynchronize(stream));
    const int64_t m = *w.host_count;
    w.last_recheck = m;
    w.count_on_device = false;
    if (m == 0) return TQEC_OK;
    // FP64 re-check of the listed rows
    if ((rc = grow(&w.g_syn, &w.g_syn_bytes, (size_t)m * std::max(1, cp.n_syn)))) return rc;
    if ((rc = grow(&w.g_mar, &w.g_mar_bytes, (size_t)m * nout * sizeof(double)))) return rc;
    if ((rc = grow(&w.g_cls, &w.g_cls_bytes, (size_t)m * sizeof(int32_t)))) return rc;
    if ((rc = grow(&w.g_tie, &w.g_tie_bytes, (size_t)m))) return rc;
    const int64_t tot = m * cp.n_syn;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, d_syn, cp.n_syn,
                       w.list, m, w.g_syn);
    const bool saved = p->timing;
    p->timing = false;
    rc = run_plain(p, dd, w, w.g_syn, m, w.g_mar, w.g_cls, w.g_tie, stream);
    p->timing = saved;
    if (rc) return rc;
    hipLaunchKernelGGL(scatter_results_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, w.list, m, nout,
                       w.g_mar, w.g_cls, w.g_tie, d_mar, d_cls, d_tie);
    TQ_HIP(hipGetLastError());
    return TQEC_OK;
}

FusedArgs base_args(tqec_plan *p, DeviceData *dd, const uint8_t *d_syn, int64_t n, int glog, bool f32) {
    const CompiledPlan &cp = p->cp;
    FusedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.steps = dd->steps;
    a.nsteps = (int)cp.steps.size();
    a.cpool = f32 ? (const void *)dd->cpool32 : (const void *)dd->cpool64;
    a.spool = f32 ? (const void *)dd->spool32 : (const void *)dd->spool64;
    a.syn = d_syn;
    a.n = n;
    a.nbits = cp.n_syn;
    a.glog = glog;
    a.nout = 1 << cp.n_out;
    a.ws_entries = cp.ws_entries;
    a.nslots = cp.n_slots;
    a.root_off = cp.root_off;
    a.root_slot = cp.root_slot;
    a.tie_tol = p->tie_tol;
    a.recheck_tol = p->recheck_tol;
    return a;
}

int prepare_recheck(Workspace &w, int64_t n, hipStream_t stream);
int enqueue_recheck(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, double *d_mar,
                    int32_t *d_cls, uint8_t *d_tie, hipStream_t stream);

// `in_list` / `in_count` (device): decode only the syndromes listed (at most
// n of them), writing their outputs at their own row.
int run_fused(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, int mode, double *d_mar,
              int32_t *d_cls, uint8_t *d_tie, hipStream_t stream, const int32_t *in_list = nullptr,
              const int32_t *in_count = nullptr) {
    const FusedCfg &c64 = p->f64;
    if (c64.glog < 0)
        return set_error(TQEC_EUNSUPPORTED, "plan needs more LDS per syndrome than one workgroup holds (FP64)");
    if (mode == TQEC_MODE_FUSED_F64) {
        FusedArgs a = base_args(p, dd, d_syn, n, c64.glog, false);
        a.list = in_list;
        a.list_count = in_count;
        a.marginals = d_mar;
        a.classes = d_cls;
        a.ties = d_tie;
        const int64_t ntiles = (n + (1 << c64.glog) - 1) >> c64.glog;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, c64.blocks));
        int rc0;
        if ((rc0 = record(p, w, 0, stream))) return rc0;
        hipLaunchKernelGGL((fused_kernel<double, false>), dim3(grid), dim3(c64.threads), c64.lds, stream, a);
        if ((rc0 = record(p, w, 1, stream))) return rc0;
        TQ_HIP(hipGetLastError());
        return TQEC_OK;
    }
    const FusedCfg &c32 = p->f32;
    if (c32.glog < 0)
        return set_error(TQEC_EUNSUPPORTED, "plan needs more LDS per syndrome than one workgroup holds (FP32)");
    int rc = TQEC_OK;
    if ((rc = prepare_recheck(w, n, stream))) return rc;
    FusedArgs a = base_args(p, dd, d_syn, n, c32.glog, true);
    a.list = in_list;
    a.list_count = in_count;
    a.marginals = d_mar;
    a.classes = d_cls;
    a.ties = d_tie;
    a.recheck_list = w.list;
    a.recheck_count = w.count;
    a.recheck_cap = w.list_cap;
    const int64_t ntiles = (n + (1 << c32.glog) - 1) >> c32.glog;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, c32.blocks));
    if ((rc = record(p, w, 0, stream))) return rc;
    hipLaunchKernelGGL((fused_kernel<float, true>), dim3(grid), dim3(c32.threads), c32.lds, stream, a);
    if ((rc = record(p, w, 1, stream))) return rc;
    return enqueue_recheck(p, dd, w, d_syn, n, d_mar, d_cls, d_tie, stream);
}

// Re-check list and counter of a (plan, stream) workspace, cleared on the stream.
int prepare_recheck(Workspace &w, int64_t n, hipStream_t stream) {
    if (w.list_cap < n) {
        if (w.list) TQ_HIP(hipFree(w.list));
        w.list = nullptr;
        w.list_cap = 0;
        TQ_HIP(hipMalloc((void **)&w.list, (size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)));
        w.list_cap = n;
    }
    int rc;
    if ((rc = ensure_count(w, stream))) return rc;
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, stream, w.count, 1);
    w.count_nseg = 0;
    return TQEC_OK;
}

// The single re-check counter: allocated on first use.
int ensure_count(Workspace &w, hipStream_t stream) {
    (void)stream;
    if (w.count) return TQEC_OK;
    TQ_HIP(hipMalloc((void **)&w.count, sizeof(int32_t)));
    TQ_HIP(hipHostMalloc((void **)&w.host_count, sizeof(int32_t)));
    return TQEC_OK;
}

// The last call's re-check count (after the stream's work has finished).
int read_recheck_count(Workspace &w, int64_t *out) {
    if (w.count_nseg > 0) {
        std::vector<int32_t> c((size_t)w.count_nseg);
        TQ_HIP(hipMemcpy(c.data(), w.segc, c.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        int64_t t = 0;
        for (int32_t x : c) t += x;
        *out = t;
        return TQEC_OK;
    }
    TQ_HIP(hipMemcpy(w.host_count, w.count, sizeof(int32_t), hipMemcpyDeviceToHost));
    *out = *w.host_count;
    return TQEC_OK;
}

// FP64 re-check of the syndromes listed by the FP32 pass (device-side list,
// no host synchronisation), then the list length is copied to the host.
int enqueue_recheck(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, double *d_mar,
                    int32_t *d_cls, uint8_t *d_tie, hipStream_t stream) {
    const FusedCfg &c64 = p->f64;
    int rc;
    FusedArgs r = base_args(p, dd, d_syn, n, c64.glog, false);
    r.list = w.list;
    r.list_count = w.count;
    r.marginals = d_mar;
    r.classes = d_cls;
    r.ties = d_tie;
    const unsigned rgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(
        (n + (1 << c64.glog) - 1) >> c64.glog, c64.blocks));
    if ((rc = record(p, w, 2, stream))) return rc;
    hipLaunchKernelGGL((fused_kernel<double, false>), dim3(rgrid), dim3(c64.threads), c64.lds, stream, r);
    if ((rc = record(p, w, 3, stream))) return rc;
    w.count_on_device = true;   // read on demand by tqec_last_recheck_count
    TQ_HIP(hipGetLastError());
    return TQEC_OK;
}

// ---- plan-specialised FP32 kernels ---------------------------------------

// Code objects by source text, shared by every plan of the process (equal
// plans and tile choices give equal sources).
std::mutex g_code_mu;
std::unordered_map<std::string, std::vector<char>> g_code_cache;

int spec_code(const std::string &src, std::vector<char> &code) {
    {
        std::lock_guard<std::mutex> lk(g_code_mu);
        auto it = g_code_cache.find(src);
        if (it != g_code_cache.end()) {
            code = it->second;
            return TQEC_OK;
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "tqec_spec.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return set_error(TQEC_EDEVICE, "hiprtcCreateProgram failed");
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
            log.resize(ls);
            hiprtcGetProgramLog(prog, &log[0]);
        }
        hiprtcDestroyProgram(&prog);
        return set_error(TQEC_EDEVICE, std::string("hiprtc: ") + hiprtcGetErrorString(r) + ": " + log.substr(0, 1500));
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.assign(cs, 0);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> lk(g_code_mu);
    g_code_cache[src] = code;
    return TQEC_OK;
}

// Live FP64 tensor registers per lane up to which the register-resident
// FP64 kernel is offered (512 VGPRs and AGPRs at one wave per SIMD).
// Slab memory of a single-plan specialised kernel's resident workgroups at
// most (bytes).
constexpr size_t kSpecSlabTotal = size_t(48) << 30;
constexpr int kRegBudgetF64 = 440;

bool spec_fits(const tqec_plan *p, const SpecCfg &c) {
    if (c.mfma32 && (c.mfma || c.reg || c.fp64 || c.glog != 0 || c.threads % 64 != 0)) return false;
    if (c.reg) {
        if (c.mfma || c.slab || c.threads % 64 != 0 || (1 << c.glog) * 4 != c.threads) return false;
        if (!spec_reg_supported(p->cp)) return false;
        // FP64 entries take two registers each
        if (c.fp64 && 2 * spec_reg_peak(p->cp) > kRegBudgetF64) return false;
    }
    return spec_lds_bytes(p->cp, c) <= p->lds_limit;
}

// Default before tuning: the register-resident FP32 kernel when the plan's
// live tensors fit a quad's registers (256 threads, tables in global memory,
// two waves per SIMD when the live set leaves room); else the largest tile whose LDS
// image (tables included when they fit) lets three workgroups share a CU
// (else two, else one), 16 threads per syndrome within [64, 128], blocks of
// 4 outputs per lane; else per-workgroup slabs, 256 threads, blocks of 4.
SpecCfg choose_spec(const tqec_plan *p, bool fp64 = false) {
    if (spec_reg_supported(p->cp)) {
        // every tensor of a syndrome fits a quad's registers: no LDS intermediates
        SpecCfg c;
        c.reg = true;
        c.fp64 = fp64;
        c.threads = 256;
        c.glog = 6;
        c.tab_lds = false;
        c.waves = (!fp64 && spec_reg_peak(p->cp) <= 200) ? 2 : 0;
        if (spec_fits(p, c)) return c;
    }
    for (int div : {3, 2, 1}) {
        for (int gl = 6; gl >= 0; --gl) {
            for (bool tab : {true, false}) {
                SpecCfg c;
                c.glog = gl;
                c.threads = std::max(64, std::min(128, 16 << gl));
                c.tab_lds = tab;
                c.fp64 = fp64;
                c.unroll = 4;
                if (spec_lds_bytes(p->cp, c) <= p->lds_limit / div) return c;
            }
        }
    }
    // intermediates too wide for LDS: per-workgroup slabs of at most 4 MiB;
    // the slab steps wait on dependent gathers from L2/HBM, so four outputs
    // per lane keep more loads in flight (1.4x over one on a 2^19-wide DEM)
    for (int gl = 4; gl >= 0; --gl) {
        SpecCfg c;
        c.glog = gl;
        c.threads = 256;
        c.tab_lds = false;
        c.fp64 = fp64;
        c.slab = true;
        c.unroll = 4;
        // GEMM-shaped steps of one-syndrome FP32 slabs on FP32 MFMA (toric L=5:
        // 2.2x; circuit-level d=5, 2 rounds: 1.08x after tuning)
        SpecCfg m = c;
        m.mfma32 = true;
        c.mfma32 = spec_plan_uses_mfma32(p->cp, m);   // false for FP64 and for tiles of several syndromes
        // two waves per SIMD (at most 256 registers) let one wave's MFMAs
        // cover the other's operand loads: toric L=5 1.4x over one wave
        if (c.mfma32) c.waves = 2;
        const size_t bytes = spec_slab_entries(p->cp, c) * (fp64 ? 8 : 4);
        if ((bytes <= (size_t(4) << 20) || gl == 0) && spec_lds_bytes(p->cp, c) <= p->lds_limit) return c;
    }
    SpecCfg c;
    c.glog = -1;
    return c;
}

std::tuple<int, int, int, int, int, int, int> spec_key(int dev, const SpecCfg &cfg) {
    return std::make_tuple(dev, cfg.glog, cfg.threads, cfg.tab_lds ? 1 : 0, cfg.waves, cfg.fp64 ? 1 : 0,
                           (cfg.slab ? 1 : 0) | (cfg.mfma ? 2 : 0) | (cfg.reg ? 4 : 0) | (cfg.unroll << 3) |
                               (cfg.dyn ? 1 << 8 : 0) | (cfg.mfma32 ? 1 << 9 : 0));
}

// The compiled kernel of a configuration on a device, or nullptr when it has
// not been built there yet (never compiles, never allocates).
SpecKernel *spec_find(tqec_plan *p, int dev, const SpecCfg &cfg) {
    auto it = p->spec.find(spec_key(dev, cfg));
    return it == p->spec.end() ? nullptr : &it->second;
}

// Largest generated source handed to hiprtc (beyond it a compile takes
// minutes; such plans stay on the generic executors).
constexpr size_t kMaxSpecSource = size_t(2) << 20;

int spec_get(tqec_plan *p, int dev, const SpecCfg &cfg, SpecKernel **out) {
    const auto key = spec_key(dev, cfg);
    auto it = p->spec.find(key);
    if (it != p->spec.end()) {
        *out = &it->second;
        return TQEC_OK;
    }
    if (cfg.glog < 0 || !spec_fits(p, cfg)) return set_error(TQEC_EUNSUPPORTED, "specialised tile does not fit in LDS");
    std::vector<char> code;
    int rc;
    const std::string src = spec_source(p->cp, cfg);
    if (src.empty()) return set_error(TQEC_EUNSUPPORTED, "plan does not suit this specialised kernel");
    if (src.size() > kMaxSpecSource)
        return set_error(TQEC_EUNSUPPORTED, "specialised kernel source of " + std::to_string(src.size() >> 10) +
                                                " KB is beyond the compile budget; use the generic executors");
    if ((rc = spec_code(src, code))) return rc;
    SpecKernel k;
    k.cfg = cfg;
    k.lds = spec_lds_bytes(p->cp, cfg);
    TQ_HIP(hipModuleLoadData(&k.mod, code.data()));
    const hipError_t e = hipModuleGetFunction(&k.fn, k.mod, spec_kernel_name(cfg));
    if (e != hipSuccess) {
        (void)hipModuleUnload(k.mod);
        return set_error(TQEC_EDEVICE, std::string("hipModuleGetFunction: ") + hipGetErrorString(e));
    }
    k.blocks = p->n_cu * resident_per_cu(k.lds, cfg.threads, p->lds_limit);
    if (cfg.slab) {
        // the kernel loops over tiles in as many workgroups as run: fewer of
        // them when their slabs would not fit the budget (plans of width 2^23:
        // 72 MB per workgroup)
        const size_t per = spec_slab_entries(p->cp, cfg) * (cfg.fp64 ? 8 : 4);
        if (per) k.blocks = (int)std::max<size_t>((size_t)p->n_cu, std::min<size_t>((size_t)k.blocks, kSpecSlabTotal / per));
    }
    if (cfg.reg) {
        // register-bound: the runtime's occupancy of the compiled kernel
        int per_cu = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k.fn, cfg.threads, 0) == hipSuccess &&
            per_cu > 0)
            k.blocks = p->n_cu * per_cu;
    }
    p->spec[key] = k;
    *out = &p->spec[key];
    return TQEC_OK;
}

// Slab bytes a specialised kernel needs for a batch of n.
size_t spec_slab_need(const tqec_plan *p, const SpecKernel *k, int64_t n) {
    if (!k->cfg.slab) return 0;
    const int64_t ntiles = (n + (1 << k->cfg.glog) - 1) >> k->cfg.glog;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ntiles, k->blocks));
    return spec_slab_entries(p->cp, k->cfg) * (k->cfg.fp64 ? 8 : 4) * (size_t)grid;
}

// Launch a specialised kernel over n syndromes (or the listed ones).
int spec_launch(tqec_plan *p, void *slab, size_t slab_bytes, SpecKernel *k, DeviceData *dd, const uint8_t *d_syn,
                int64_t n, const int32_t *in_list, const int32_t *in_count, double *d_mar, int32_t *d_cls,
                uint8_t *d_tie, int32_t *rl, int32_t *rcount, int64_t rcap, double tie_tol, double recheck_tol,
                hipStream_t stream, int64_t nseg = 0, int32_t segcap = 0) {
    const int64_t ntiles = nseg > 0 ? nseg : (n + (1 << k->cfg.glog) - 1) >> k->cfg.glog;
    // Register-resident kernels with tables in global memory have no per-
    // workgroup set-up, so they get one workgroup per tile and the dispatcher
    // balances the load (a tile that re-checks a row in FP64 takes longer);
    // the others loop over tiles in as many workgroups as stay resident (a
    // slab holds one workspace per resident workgroup).
    const bool per_tile = k->cfg.reg && !k->cfg.tab_lds && !k->cfg.slab;
    const unsigned grid = (unsigned)std::max<int64_t>(1, per_tile ? ntiles : std::min<int64_t>(ntiles, k->blocks));
    void *a_slab = nullptr;
    if (k->cfg.slab) {
        // the caller sized the slab before enqueueing anything that uses it
        if (spec_slab_need(p, k, n) > slab_bytes) return set_error(TQEC_EINVAL, "internal: slab too small");
        a_slab = slab;
    }
    const unsigned char *a_syn = d_syn;
    long long a_n = (long long)n;
    const int *a_list = in_list, *a_count = in_count;
    const void *a_tab = k->cfg.fp64 ? (const void *)dd->tab64 : (const void *)dd->tab32;
    double *a_mar = d_mar;
    int *a_cls = d_cls;
    unsigned char *a_tie = d_tie;
    int *a_rl = rl, *a_rc = rcount;
    long long a_cap = (long long)rcap;
    double a_tt = tie_tol, a_rt = recheck_tol;
    int a_nseg = (int)nseg, a_segcap = segcap;
    const double *a_tab64 = dd->tab64;
    void *args[] = {&a_syn, &a_n,  &a_list, &a_count, &a_tab, &a_mar,  &a_cls,    &a_tie,  &a_rl,
                    &a_rc,  &a_cap, &a_tt,  &a_rt,    &a_slab, &a_nseg, &a_segcap, &a_tab64};
    TQ_HIP(hipModuleLaunchKernel(k->fn, grid, 1, 1, (unsigned)k->cfg.threads, 1, 1, 0, stream, args, nullptr));
    return TQEC_OK;
}

// The same with the slab of a (plan, stream) workspace.
int spec_launch(tqec_plan *p, Workspace &w, SpecKernel *k, DeviceData *dd, const uint8_t *d_syn, int64_t n,
                const int32_t *in_list, const int32_t *in_count, double *d_mar, int32_t *d_cls, uint8_t *d_tie,
                int32_t *rl, int32_t *rcount, int64_t rcap, double tie_tol, double recheck_tol, hipStream_t stream,
                int64_t nseg = 0, int32_t segcap = 0) {
    return spec_launch(p, w.spec_slab, w.spec_slab_bytes, k, dd, d_syn, n, in_list, in_count, d_mar, d_cls, d_tie, rl,
                       rcount, rcap, tie_tol, recheck_tol, stream, nseg, segcap);
}

// FP64 re-check kernel: one syndrome per 256-thread workgroup (the list is
// short, so the latency of one syndrome's step chain matters: one output
// per thread on the widest steps), in a slab when too wide for LDS.
SpecCfg spec_recheck_cfg(tqec_plan *p) {
    SpecCfg c;
    c.glog = 0;
    c.threads = 256;
    c.tab_lds = false;
    c.fp64 = true;
    if (!spec_fits(p, c)) {
        // too wide for LDS: the re-check works in a global slab, 256 threads
        c.slab = true;
        c.threads = 256;
    }
    return c;
}

int spec_recheck_kernel(tqec_plan *p, int dev, SpecKernel **out) { return spec_get(p, dev, spec_recheck_cfg(p), out); }

int spec_recheck(tqec_plan *p, DeviceData *dd, Workspace &w, SpecKernel *k, const uint8_t *d_syn, int64_t n,
                 double *d_mar, int32_t *d_cls, uint8_t *d_tie, hipStream_t stream) {
    int rc;
    if ((rc = record(p, w, 2, stream))) return rc;
    if ((rc = spec_launch(p, w, k, dd, d_syn, n, w.list, w.count, d_mar, d_cls, d_tie, nullptr, nullptr, 0, p->tie_tol,
                          p->recheck_tol, stream)))
        return rc;
    if ((rc = record(p, w, 3, stream))) return rc;
    w.count_on_device = true;   // read on demand by tqec_last_recheck_count
    return TQEC_OK;
}

int run_spec64(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, double *d_mar,
               int32_t *d_cls, uint8_t *d_tie, hipStream_t stream, const int32_t *in_list = nullptr,
               const int32_t *in_count = nullptr) {
    if (!p->spec64_cfg_set) {
        p->spec64_cfg = choose_spec(p, true);
        p->spec64_cfg_set = true;
    }
    if (p->spec64_cfg.glog < 0)
        return set_error(TQEC_EUNSUPPORTED, "plan needs more LDS per syndrome than one workgroup holds (FP64)");
    SpecKernel *k;
    int rc;
    if ((rc = spec_get(p, w.device, p->spec64_cfg, &k))) return rc;
    if ((rc = grow((char **)&w.spec_slab, &w.spec_slab_bytes, spec_slab_need(p, k, n)))) return rc;
    if ((rc = record(p, w, 0, stream))) return rc;
    if ((rc = spec_launch(p, w, k, dd, d_syn, n, in_list, in_count, d_mar, d_cls, d_tie, nullptr, nullptr, 0,
                          p->tie_tol, p->recheck_tol, stream)))
        return rc;
    return record(p, w, 1, stream);
}

// Margin below which the BF16-MFMA mode re-checks a syndrome in FP64: BF16
// operands carry a relative error of 2^-9 each, so a step's products are off
// by up to 2^-8 and a chain of steps by a few percent at worst.
constexpr double kBf16RecheckTol = 0.05;

int run_spec(tqec_plan *p, DeviceData *dd, Workspace &w, const uint8_t *d_syn, int64_t n, double *d_mar,
             int32_t *d_cls, uint8_t *d_tie, hipStream_t stream, const int32_t *in_list = nullptr,
             const int32_t *in_count = nullptr, bool bf16 = false) {
    SpecCfg &cfg = bf16 ? p->specb_cfg : p->spec_cfg;
    bool &cfg_set = bf16 ? p->specb_cfg_set : p->spec_cfg_set;
    if (!cfg_set) {
        cfg = choose_spec(p);
        if (bf16 && cfg.glog >= 0) {
            // MFMA steps need tiles of one syndrome; BF16 steps replace the FP32 ones
            cfg.glog = 0;
            cfg.threads = 256;
            cfg.mfma = true;
            cfg.mfma32 = false;
        }
        cfg_set = true;
    }
    if (cfg.glog < 0)
        return set_error(TQEC_EUNSUPPORTED, "plan needs more LDS per syndrome than one workgroup holds (FP32)");
    SpecKernel *k, *k64 = nullptr;
    int rc;
    if ((rc = spec_get(p, w.device, cfg, &k))) return rc;
    // the register-resident kernel may re-check its flagged rows itself
    const bool inline_rc = spec_reg_inline_recheck(p->cp, cfg);
    if (!inline_rc && (rc = spec_recheck_kernel(p, w.device, &k64))) return rc;
    if ((rc = grow((char **)&w.spec_slab, &w.spec_slab_bytes,
                   std::max(spec_slab_need(p, k, n), k64 ? spec_slab_need(p, k64, n) : size_t(0)))))
        return rc;
    const double rtol = bf16 ? std::max(p->recheck_tol, kBf16RecheckTol) : p->recheck_tol;
    if (cfg.reg) {
        // Re-check list in per-tile segments, each written whole by the tile
        // that owns it: no clearing launch before the call.
        const int32_t G = 1 << cfg.glog;
        const int64_t nseg = (n + G - 1) / G;
        if (w.list_cap < nseg * G) {
            if (w.list) TQ_HIP(hipFree(w.list));
            w.list = nullptr;
            w.list_cap = 0;
            TQ_HIP(hipMalloc((void **)&w.list, (size_t)std::max<int64_t>(nseg * G, 1) * sizeof(int32_t)));
            w.list_cap = nseg * G;
        }
        if (w.segc_cap < nseg) {
            if (w.segc) TQ_HIP(hipFree(w.segc));
            w.segc = nullptr;
            w.segc_cap = 0;
            TQ_HIP(hipMalloc((void **)&w.segc, (size_t)std::max<int64_t>(nseg, 1) * sizeof(int32_t)));
            w.segc_cap = nseg;
        }
        if ((rc = ensure_count(w, stream))) return rc;
        if ((rc = record(p, w, 0, stream))) return rc;
        if ((rc = spec_launch(p, w, k, dd, d_syn, n, in_list, in_count, d_mar, d_cls, d_tie, w.list, w.segc, G,
                              p->tie_tol, rtol, stream)))
            return rc;
        if ((rc = record(p, w, 1, stream))) return rc;
        if (!inline_rc) {
            if ((rc = record(p, w, 2, stream))) return rc;
            if (nseg > 0 && (rc = spec_launch(p, w, k64, dd, d_syn, n, w.list, w.segc, d_mar, d_cls, d_tie, nullptr,
                                              nullptr, 0, p->tie_tol, p->recheck_tol, stream, nseg, G)))
                return rc;
            if ((rc = record(p, w, 3, stream))) return rc;
        }
        w.count_nseg = nseg;
        w.count_on_device = true;   // read on demand by tqec_last_recheck_count
        return TQEC_OK;
    }
    if ((rc = prepare_recheck(w, n, stream))) return rc;
    if ((rc = record(p, w, 0, stream))) return rc;
    if ((rc = spec_launch(p, w, k, dd, d_syn, n, in_list, in_count, d_mar, d_cls, d_tie, w.list, w.count, w.list_cap,
                          p->tie_tol, rtol, stream)))
        return rc;
    if ((rc = record(p, w, 1, stream))) return rc;
    return spec_recheck(p, dd, w, k64, d_syn, n, d_mar, d_cls, d_tie, stream);
}

// Compile a set of specialised configurations on host threads (hiprtc
// programs are independent); the code objects land in the process cache.
void spec_precompile(const tqec_plan *p, const std::vector<SpecCfg> &cands) {
    std::vector<std::future<int>> jobs;
    const size_t par = 8;
    for (size_t i = 0; i < cands.size(); ++i) {
        if (jobs.size() >= par) {
            jobs.front().get();
            jobs.erase(jobs.begin());
        }
        const std::string src = spec_source(p->cp, cands[i]);
        jobs.push_back(std::async(std::launch::async, [src]() {
            std::vector<char> code;
            return spec_code(src, code);
        }));
    }
    for (auto &j : jobs) j.get();
}

// Autotuner of the specialised kernel, in two stages: (tile 1..64, 64/128/256
// threads) and the register-resident kernel at 64/128/256 threads, all with
// tables in global memory; then, for an LDS tile, the best pair with tables
// in LDS, a 4-waves-per-SIMD register cap and 1/2/4/8 outputs per lane block,
// or, for the register-resident kernel, tables in LDS (when they fit) and a
// 2-waves-per-SIMD register cap.  Each configuration is timed
// twice on the trial batch and the fastest is kept.
int autotune_spec(tqec_plan *p, const uint8_t *d_syn, int64_t n, hipStream_t st, int32_t *tile, int32_t *threads,
                  bool bf16) {
    SpecCfg &tcfg = bf16 ? p->specb_cfg : p->spec_cfg;
    bool &tset = bf16 ? p->specb_cfg_set : p->spec_cfg_set;
    std::lock_guard<std::mutex> lk(p->mu);
    int rc;
    if ((rc = configure(p))) return rc;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    DeviceData *dd;
    if ((rc = ensure_device(p, dev, &dd))) return rc;
    Workspace &w = p->ws[std::make_pair(dev, st)];
    w.device = dev;
    const int nout = 1 << p->cp.n_out;
    if ((rc = grow(&w.d_mar, &w.d_mar_bytes, (size_t)n * nout * sizeof(double)))) return rc;
    if ((rc = grow(&w.d_cls, &w.d_cls_bytes, (size_t)n * sizeof(int32_t)))) return rc;
    if ((rc = grow(&w.d_tie, &w.d_tie_bytes, (size_t)n))) return rc;
    hipEvent_t e0, e1;
    TQ_HIP(hipEventCreate(&e0));
    TQ_HIP(hipEventCreate(&e1));
    const bool saved_timing = p->timing;
    p->timing = false;
    const SpecCfg saved = tcfg;
    const bool saved_set = tset;
    SpecCfg best;
    best.glog = -1;
    float best_ms = 1e30f;
    auto time_all = [&](const std::vector<SpecCfg> &cands) {
        spec_precompile(p, cands);
        for (const SpecCfg &c : cands) {
            tcfg = c;
            tset = true;
            float ms = 0;
            for (int rep = 0; rep < 2 && rc == TQEC_OK; ++rep) {
                if ((rc = record_or(hipEventRecord(e0, st)))) break;
                rc = run_spec(p, dd, w, d_syn, n, w.d_mar, w.d_cls, w.d_tie, st, nullptr, nullptr, bf16);
                if (rc) break;
                if ((rc = record_or(hipEventRecord(e1, st)))) break;
                if ((rc = record_or(hipEventSynchronize(e1)))) break;
                if ((rc = record_or(hipEventElapsedTime(&ms, e0, e1)))) break;
            }
            if (rc) return;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
    };
    std::vector<SpecCfg> stage1;
    for (int gl = 0; gl <= (bf16 ? 0 : 6); ++gl)
        for (int thr : {64, 128, 256}) {
            if (thr < (1 << gl)) continue;
            SpecCfg c;
            c.glog = gl;
            c.threads = thr;
            c.tab_lds = false;
            c.mfma = bf16;
            if (spec_fits(p, c)) stage1.push_back(c);
        }
    if (!bf16)
        for (int thr : {64, 128, 256}) {
            SpecCfg c;
            c.reg = true;
            c.threads = thr;
            c.glog = thr == 64 ? 4 : (thr == 128 ? 5 : 6);
            c.tab_lds = false;
            if (spec_fits(p, c)) stage1.push_back(c);
        }
    if (stage1.empty()) {
        // slab configurations (per-workgroup intermediates of at most 8 MiB)
        for (int gl = 0; gl <= (bf16 ? 0 : 4); ++gl)
            for (int thr : {128, 256}) {
                SpecCfg c;
                c.glog = gl;
                c.threads = thr;
                c.tab_lds = false;
                c.slab = true;
                c.mfma = bf16;
                if (spec_fits(p, c) && (gl == 0 || spec_slab_entries(p->cp, c) * 4 <= (size_t(8) << 20)))
                    stage1.push_back(c);
            }
    }
    if (stage1.empty()) rc = set_error(TQEC_EUNSUPPORTED, "no specialised tile fits");
    if (!rc) time_all(stage1);
    if (!rc && best.glog >= 0) {
        std::vector<SpecCfg> stage2;
        if (best.reg) {
            // register-resident: table placement and a two-waves-per-SIMD register cap
            for (bool tab : {false, true})
                for (int wv : {0, 2}) {
                    if (!tab && wv == 0) continue;
                    SpecCfg c = best;
                    c.tab_lds = tab;
                    c.waves = wv;
                    if (spec_fits(p, c)) stage2.push_back(c);
                }
        } else {
            for (bool tab : {false, true})
                for (int wv : {0, 4})
                    for (int un : {1, 2, 4, 8}) {
                        if (!tab && wv == 0 && un == 1) continue;
                        if (tab && best.slab) continue;
                        SpecCfg c = best;
                        c.tab_lds = tab;
                        c.waves = wv;
                        c.unroll = un;
                        if (spec_fits(p, c)) stage2.push_back(c);
                    }
        }
        time_all(stage2);
        // FP32 MFMA on the GEMM-shaped steps of the winner (tiles of one syndrome)
        if (!rc && !bf16 && !best.reg && best.glog == 0) {
            // FP32 MFMA on the GEMM-shaped steps of a one-syndrome winner; the
            // MFMA steps change the register budget, so the register cap and
            // the other steps' output blocks are retuned with them
            std::vector<SpecCfg> stage3;
            for (int wv : {0, 2, 4})
                for (int un : {2, 4}) {
                    SpecCfg c = best;
                    c.mfma32 = true;
                    c.waves = wv;
                    c.unroll = un;
                    if (spec_fits(p, c) && spec_plan_uses_mfma32(p->cp, c)) stage3.push_back(c);
                }
            time_all(stage3);
        }
    }
    p->timing = saved_timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        tcfg = saved;
        tset = saved_set;
        return rc;
    }
    tcfg = best;
    tset = true;
    if (tile) *tile = 1 << best.glog;
    if (threads) *threads = best.threads;
    return TQEC_OK;
}

int decode_device(tqec_plan *p, const uint8_t *d_syn, int64_t n, int mode, double *d_mar, int32_t *d_cls,
                  uint8_t *d_tie, hipStream_t stream, Workspace **wout) {
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    int rc;
    if ((rc = configure(p))) return rc;
    DeviceData *dd;
    if ((rc = ensure_device(p, dev, &dd))) return rc;
    Workspace &w = p->ws[std::make_pair(dev, stream)];
    w.device = dev;
    if (wout) *wout = &w;
    if (n == 0) return TQEC_OK;
    if (mode == TQEC_MODE_PLAIN_F64) return run_plain(p, dd, w, d_syn, n, d_mar, d_cls, d_tie, stream);
    if (mode == TQEC_MODE_WIDE_F32) return run_wide(p, dd, w, d_syn, n, d_mar, d_cls, d_tie, stream);
    if (mode == TQEC_MODE_FUSED_F64 || mode == TQEC_MODE_FUSED_F32)
        return run_fused(p, dd, w, d_syn, n, mode, d_mar, d_cls, d_tie, stream);
    if (mode == TQEC_MODE_SPEC_F32) return run_spec(p, dd, w, d_syn, n, d_mar, d_cls, d_tie, stream);
    if (mode == TQEC_MODE_SPEC_F64) return run_spec64(p, dd, w, d_syn, n, d_mar, d_cls, d_tie, stream);
    if (mode == TQEC_MODE_SPEC_BF16)
        return run_spec(p, dd, w, d_syn, n, d_mar, d_cls, d_tie, stream, nullptr, nullptr, true);
    return set_error(TQEC_EINVAL, "unknown mode " + std::to_string(mode));
}

// Mixed-code executor state.  Each of two call slots owns the per-call
// tables (header, batch table, counters, batch clocks) and a pinned host
// staging copy of header + batch table, so a call can be staged while the
// previous one still runs; the slabs are shared because calls execute in
// order.
struct MixedSlot {
    void *h_stage = nullptr;            // pinned: MixedHeader | MixedBatch[max_batches]
    void *d_table = nullptr;            // device copy of the same
    MixedCounters *ctr = nullptr;
    unsigned long long *t_done = nullptr;
    std::vector<int32_t> perm;          // staged position -> caller's batch index
    hipEvent_t staged = nullptr;        // the device has consumed h_stage
    hipStream_t graph_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t nbatches = 0;
    bool used = false;
};

struct MixedPlanCfg {
    int glog32 = 0;
    bool gw32 = false, gw64 = false;
    size_t lds32 = 0, lds64 = 0;        // dynamic LDS of the plan's tiles
    size_t slab32 = 0, slab64 = 0;      // slab bytes per workgroup (slab plans)
};

constexpr int kMixedThreads = 256;
// Slab memory of one mixed-code kernel's resident workgroups at most (bytes).
constexpr size_t kMixedSlabTotal = size_t(32) << 30;
constexpr size_t kMixedSlabBudget = size_t(4) << 20;   // FP32 slab per workgroup before shrinking the tile

}  // namespace

struct tqec_mixed {
    std::vector<tqec_plan *> plans;
    std::vector<MixedPlanCfg> cfg;
    int device = -1;
    int32_t max_batches = 0;
    int64_t max_rows = 0;
    size_t lds = 0;                     // dynamic LDS of the kernel (FP32 tiles and FP64 re-checks)
    int blocks = 0;
    int order = TQEC_MIXED_ORDER_CHEAPEST_FIRST;
    std::vector<double> cost;           // per plan: FP32 work per syndrome (multiply-adds + entries)
    MixedPlan *d_plans = nullptr;
    // specialised variant (TQEC_MIXED_KERNEL_SPEC): one generated kernel
    bool spec = false;
    hipModule_t spec_mod = nullptr;
    hipFunction_t spec_fn = nullptr;
    int spec_threads = 128, spec_blocks = 0;
    std::vector<int> spec_glog, spec_slab;   // per plan: FP32 tile, slab flags (bit 0 FP32, bit 1 FP64)
    const void **d_tabs32 = nullptr, **d_tabs64 = nullptr;
    float *spec_slab32 = nullptr;
    double *spec_slab64 = nullptr;
    char *slab32 = nullptr, *slab64 = nullptr;
    size_t slab32_bytes = 0, slab64_bytes = 0;
    std::array<MixedSlot, 2> slot;
    int cur = 0;                        // slot of the next call
    int last = -1;                      // slot of the last call
    bool use_graph = true;
    double clock_khz = 0;
};

namespace {

size_t mixed_table_bytes(int32_t max_batches) {
    return sizeof(MixedHeader) + (size_t)std::max(1, max_batches) * sizeof(MixedBatch);
}

MixedArgs mixed_args(tqec_mixed *m, MixedSlot &sl) {
    MixedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.plans = m->d_plans;
    a.hdr = reinterpret_cast<const MixedHeader *>(sl.d_table);
    a.batches = reinterpret_cast<const MixedBatch *>(reinterpret_cast<const char *>(sl.d_table) + sizeof(MixedHeader));
    a.ctr = sl.ctr;
    a.t_done = sl.t_done;
    a.slab32 = m->slab32;
    a.slab64 = m->slab64;
    a.slab32_bytes = (int64_t)m->slab32_bytes;
    a.slab64_bytes = (int64_t)m->slab64_bytes;
    return a;
}

// The whole call: stage the tables, clear the counters, decode.  Enqueued
// directly or recorded once into a graph per slot.
int mixed_enqueue(tqec_mixed *m, MixedSlot &sl, hipStream_t stream) {
    const size_t tb = mixed_table_bytes(m->max_batches);
    TQ_HIP(hipMemcpyAsync(sl.d_table, sl.h_stage, tb, hipMemcpyHostToDevice, stream));
    const int cap = std::max(1, m->max_batches);
    hipLaunchKernelGGL(mixed_reset_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, stream, sl.ctr, sl.t_done,
                       cap);
    if (m->spec) {
        const void *a_batches = reinterpret_cast<const char *>(sl.d_table) + sizeof(MixedHeader);
        const void *a_hdr = sl.d_table;
        void *a_ctr = sl.ctr, *a_done = sl.t_done;
        const void *a_t32 = m->d_tabs32, *a_t64 = m->d_tabs64;
        void *a_s32 = m->spec_slab32, *a_s64 = m->spec_slab64;
        void *args[] = {&a_batches, &a_hdr, &a_ctr, &a_done, &a_t32, &a_t64, &a_s32, &a_s64};
        TQ_HIP(hipModuleLaunchKernel(m->spec_fn, (unsigned)m->spec_blocks, 1, 1, (unsigned)m->spec_threads, 1, 1, 0,
                                     stream, args, nullptr));
    } else {
        MixedArgs a = mixed_args(m, sl);
        hipLaunchKernelGGL(mixed_kernel, dim3((unsigned)m->blocks), dim3(kMixedThreads), m->lds, stream, a);
    }
    TQ_HIP(hipGetLastError());
    return TQEC_OK;
}

void mixed_free_spec(tqec_mixed *m) {
    if (m->spec_mod) (void)hipModuleUnload(m->spec_mod);
    for (void *q : {(void *)m->d_tabs32, (void *)m->d_tabs64, (void *)m->spec_slab32, (void *)m->spec_slab64})
        if (q) (void)hipFree(q);
    m->spec_mod = nullptr;
    m->spec_fn = nullptr;
    m->d_tabs32 = m->d_tabs64 = nullptr;
    m->spec_slab32 = nullptr;
    m->spec_slab64 = nullptr;
    m->spec = false;
}

// Drop the captured graphs (they hold the old kernel) after the calls finish.
void mixed_drop_graphs(tqec_mixed *m) {
    for (MixedSlot &sl : m->slot) {
        if (sl.used && sl.staged) (void)hipEventSynchronize(sl.staged);
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        if (sl.graph) (void)hipGraphDestroy(sl.graph);
        sl.exec = nullptr;
        sl.graph = nullptr;
        sl.graph_stream = nullptr;
    }
}

// Per-plan tiles of the generated mixed kernel: each plan's current SPEC_F32
// LDS-tile configuration (or the default: a register-resident configuration
// does not apply here) with the kernel's thread count and no
// tables in LDS, its tile shrunk until it fits half the LDS (else a slab),
// and an FP64 tile of one syndrome for the re-checks.
int mixed_spec_layout(const std::vector<tqec_plan *> &plans, int nt, size_t lds_limit, std::vector<MixedSpecPlan> &sp,
                      std::vector<int> &glog, std::vector<int> &slabf, size_t &s32, size_t &s64) {
    sp.assign(plans.size(), MixedSpecPlan());
    glog.assign(plans.size(), 0);
    slabf.assign(plans.size(), 0);
    s32 = s64 = 0;
    const size_t budget = lds_limit / 2 - 1024;
    for (size_t i = 0; i < plans.size(); ++i) {
        tqec_plan *p = plans[i];
        if (p->cp.semiring != TQEC_SEMIRING_SUM)
            return set_error(TQEC_EINVAL, "plan " + std::to_string(i) + " is a max-product plan");
        SpecCfg c;
        const bool tuned = p->spec_cfg_set && p->spec_cfg.glog >= 0 && !p->spec_cfg.reg;
        if (tuned) {
            c = p->spec_cfg;
        } else {
            // the default tile of the largest size whose LDS image fits the budget
            c.glog = 6;
            c.threads = nt;
            c.tab_lds = false;
            while (c.glog > 0 && spec_lds_bytes(p->cp, c) > budget) --c.glog;
            if (spec_lds_bytes(p->cp, c) > budget) c.slab = true;
        }
        c.threads = nt;
        c.tab_lds = false;
        c.waves = 0;
        c.fp64 = false;
        if (c.glog > 6) c.glog = 6;
        if (!c.slab) {
            while (c.glog > 0 && spec_lds_bytes(p->cp, c) > budget) --c.glog;
            if (spec_lds_bytes(p->cp, c) > budget) c.slab = true;
        }
        if (c.slab) {
            c.glog = std::min(c.glog, 4);
            while (c.glog > 0 && spec_slab_entries(p->cp, c) * 4 > (size_t(4) << 20)) --c.glog;
        }
        if (!tuned && c.slab && c.glog == 0) {
            // the single-plan slab default: blocks of 4 outputs, FP32-MFMA GEMM steps
            c.unroll = 4;
            SpecCfg m = c;
            m.mfma32 = true;
            c.mfma32 = spec_plan_uses_mfma32(p->cp, m);
        }
        SpecCfg d;
        d.glog = 0;
        d.threads = nt;
        d.tab_lds = false;
        d.fp64 = true;
        d.unroll = c.unroll;
        if (spec_lds_bytes(p->cp, d) > budget) d.slab = true;
        sp[i].cp = &p->cp;
        sp[i].f32 = c;
        sp[i].f64 = d;
        sp[i].tie_tol = p->tie_tol;
        sp[i].recheck_tol = p->recheck_tol;
        glog[i] = c.glog;
        slabf[i] = (c.slab ? 1 : 0) | (d.slab ? 2 : 0);
        if (c.slab) s32 = std::max(s32, spec_slab_entries(p->cp, c));
        if (d.slab) s64 = std::max(s64, spec_slab_entries(p->cp, d));
    }
    return TQEC_OK;
}

void mixed_free(tqec_mixed *m) {
    if (m->device >= 0) (void)hipSetDevice(m->device);
    for (MixedSlot &sl : m->slot) {
        if (sl.staged) (void)hipEventSynchronize(sl.staged);
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        if (sl.graph) (void)hipGraphDestroy(sl.graph);
        if (sl.staged) (void)hipEventDestroy(sl.staged);
        if (sl.h_stage) (void)hipHostFree(sl.h_stage);
        for (void *q : {sl.d_table, (void *)sl.ctr, (void *)sl.t_done})
            if (q) (void)hipFree(q);
    }
    if (m->d_plans) (void)hipFree(m->d_plans);
    if (m->slab32) (void)hipFree(m->slab32);
    if (m->slab64) (void)hipFree(m->slab64);
    mixed_free_spec(m);
}

}  // namespace

namespace {

// Scratch of tqec_decode_batch_device_noalloc, one copy per device, held by
// the loaded code object until the process ends: the FP32 pass's re-check
// list (one int32 per row) followed by the re-check counts (register-resident
// kernel: one per tile of at least 16 rows; LDS-tile kernel: one list length).
constexpr int64_t kNoAllocRows = TQEC_NOALLOC_MAX_ROWS;
constexpr int64_t kNoAllocCounts = kNoAllocRows / 16 + 64;
static __device__ __align__(256) int32_t scratch[kNoAllocRows + kNoAllocCounts];

// Device address of `scratch` per device.  In host code the array's name only
// stands for the symbol; the address kernels may use comes from
// hipGetSymbolAddress, looked up once per device with that device current.
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kNoAllocMaxDevices] = {};

int scratch_address(int dev, int32_t **out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) return set_error(TQEC_EUNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(TQEC_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (int32_t *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return TQEC_OK;
}

}  // namespace

extern "C" {

int tqec_plan_create(const tqec_plan_desc *desc, tqec_plan **out) {
    if (!out) return set_error(TQEC_ENULL, "null output handle");
    *out = nullptr;
    std::unique_ptr<tqec_plan> p(new (std::nothrow) tqec_plan());
    if (!p) return set_error(TQEC_ENOMEM, "out of host memory");
    std::string err;
    int rc;
    try {
        rc = compile_plan(desc, p->cp, err);
        if (rc == TQEC_OK && p->cp.semiring == TQEC_SEMIRING_SUM) spec_reg_prepare(p->cp);
    } catch (const std::bad_alloc &) {
        return set_error(TQEC_ENOMEM, "out of host memory while compiling the plan");
    } catch (const std::exception &e) {
        return set_error(TQEC_EINVAL, e.what());
    }
    if (rc != TQEC_OK) return set_error(rc, err);
    *out = p.release();
    return TQEC_OK;
}

void tqec_plan_destroy(tqec_plan *p) {
    if (!p) return;
    for (auto &kv : p->ws) {
        Workspace &w = kv.second;
        if (hipSetDevice(w.device) != hipSuccess) continue;
        (void)hipStreamSynchronize(kv.first.second);
        if (w.ws) (void)hipFree(w.ws);
        if (w.synT) (void)hipFree(w.synT);
        if (w.list) (void)hipFree(w.list);
        if (w.count) (void)hipFree(w.count);
        if (w.segc) (void)hipFree(w.segc);
        if (w.host_count) (void)hipHostFree(w.host_count);
        if (w.d_syn) (void)hipFree(w.d_syn);
        if (w.d_mar) (void)hipFree(w.d_mar);
        if (w.d_cls) (void)hipFree(w.d_cls);
        if (w.d_tie) (void)hipFree(w.d_tie);
        for (void *q : {w.smx, w.sex, (void *)w.g_syn, (void *)w.g_mar, (void *)w.g_cls, (void *)w.g_tie, w.spec_slab})
            if (q) (void)hipFree(q);
        for (auto &a : w.ev)
            for (hipEvent_t e : a) (void)hipEventDestroy(e);
    }
    for (auto &kv : p->spec) {
        if (hipSetDevice(std::get<0>(kv.first)) != hipSuccess) continue;
        if (kv.second.mod) (void)hipModuleUnload(kv.second.mod);
    }
    for (auto &kv : p->dev) {
        if (hipSetDevice(kv.first) != hipSuccess) continue;
        DeviceData &d = kv.second;
        if (d.tab32) (void)hipFree(d.tab32);
        if (d.tab64) (void)hipFree(d.tab64);
        (void)hipFree(d.steps);
        if (d.msteps) (void)hipFree(d.msteps);
        if (d.mlabels) (void)hipFree(d.mlabels);
        (void)hipFree(d.cpool64);
        (void)hipFree(d.spool64);
        (void)hipFree(d.cpool32);
        (void)hipFree(d.spool32);
    }
    delete p;
}

int tqec_plan_stats(const tqec_plan *p, tqec_stats *s) {
    if (!p || !s) return set_error(TQEC_ENULL, "null plan or stats pointer");
    const CompiledPlan &cp = p->cp;
    std::memset(s, 0, sizeof(*s));
    s->tc = cp.tc;
    s->rw = cp.rw;
    s->sc = cp.sc;
    s->n_steps = cp.n_steps;
    s->n_dep_steps = (int32_t)cp.steps.size();
    s->dep_tc = cp.dep_tc;
    s->dep_rw = cp.dep_rw;
    s->plain_ws_entries = cp.ws_entries;
    s->fused_lds_entries = cp.ws_entries;
    s->n_out = cp.n_out;
    s->fused_tile_f64 = p->f64.glog >= 0 && p->cfg_done ? (1 << p->f64.glog) : 0;
    s->fused_tile_f32 = p->f32.glog >= 0 && p->cfg_done ? (1 << p->f32.glog) : 0;
    s->const_entries = (int64_t)cp.const_pool.size();
    s->table_entries = (int64_t)cp.sliced_pool.size();
    s->tabulate_bits = cp.fold_bits;
    return TQEC_OK;
}

int tqec_plan_set_tolerances(tqec_plan *p, double tie_tol, double recheck_tol) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (!(tie_tol >= 0) || !(recheck_tol >= 0) || !std::isfinite(tie_tol) || !std::isfinite(recheck_tol))
        return set_error(TQEC_EINVAL, "tolerances must be finite and non-negative");
    std::lock_guard<std::mutex> lk(p->mu);
    p->tie_tol = tie_tol;
    p->recheck_tol = recheck_tol;
    return TQEC_OK;
}

int tqec_decode_batch_device(tqec_plan *p, const uint8_t *d_syn, int64_t n, int32_t mode, double *d_mar,
                             int32_t *d_cls, uint8_t *d_tie, void *stream) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (n < 0) return set_error(TQEC_EINVAL, "negative batch size");
    if (n > 0 && !d_syn) return set_error(TQEC_ENULL, "null syndrome pointer");
    if (n > (int64_t(1) << 31) - 1) return set_error(TQEC_EINVAL, "batch too large (max 2^31-1)");
    if (p->cp.semiring != TQEC_SEMIRING_SUM)
        return set_error(TQEC_EINVAL, "a max-product plan is decoded by tqec_map_decode_batch");
    std::lock_guard<std::mutex> lk(p->mu);
    return decode_device(p, d_syn, n, mode, d_mar, d_cls, d_tie, (hipStream_t)stream, nullptr);
}

int tqec_decode_batch_device_noalloc(tqec_plan *p, const uint8_t *d_syn, int64_t n, int32_t mode, double *d_mar,
                                     int32_t *d_cls, uint8_t *d_tie, void *stream) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (n < 0) return set_error(TQEC_EINVAL, "negative batch size");
    if (n > 0 && !d_syn) return set_error(TQEC_ENULL, "null syndrome pointer");
    if (n > kNoAllocRows)
        return set_error(TQEC_EINVAL, "batch of " + std::to_string(n) + " rows exceeds the no-allocation limit of " +
                                          std::to_string(kNoAllocRows) + " rows");
    if (p->cp.semiring != TQEC_SEMIRING_SUM)
        return set_error(TQEC_EINVAL, "a max-product plan is decoded by tqec_map_decode_batch");
    if (mode != TQEC_MODE_SPEC_F32 && mode != TQEC_MODE_SPEC_F64)
        return set_error(TQEC_EUNSUPPORTED, "the no-allocation entry point runs TQEC_MODE_SPEC_F32 and "
                                            "TQEC_MODE_SPEC_F64 only");
    std::lock_guard<std::mutex> lk(p->mu);
    const hipStream_t st = (hipStream_t)stream;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    int32_t *scr;
    int rc;
    if ((rc = scratch_address(dev, &scr))) return rc;
    // Tables and kernels must already be on this device: building them
    // allocates, so it is left to a tqec_decode_batch[_device] call in the
    // same mode (or tqec_plan_autotune) made beforehand.
    const char *unprepared = "plan not prepared on this device: make one tqec_decode_batch_device call in this mode "
                             "first";
    auto dit = p->dev.find(dev);
    const bool f64 = mode == TQEC_MODE_SPEC_F64;
    if (dit == p->dev.end() || !(f64 ? p->spec64_cfg_set : p->spec_cfg_set)) return set_error(TQEC_EINVAL, unprepared);
    DeviceData *dd = &dit->second;
    const SpecCfg &cfg = f64 ? p->spec64_cfg : p->spec_cfg;
    if (cfg.glog < 0) return set_error(TQEC_EUNSUPPORTED, "plan does not fit the specialised kernel");
    SpecKernel *k = spec_find(p, dev, cfg);
    if (!k) return set_error(TQEC_EINVAL, unprepared);
    if (cfg.slab) return set_error(TQEC_EUNSUPPORTED, "plan runs from a global slab, which this entry point lacks");
    if (n == 0) return TQEC_OK;
    if (f64)
        return spec_launch(p, nullptr, 0, k, dd, d_syn, n, nullptr, nullptr, d_mar, d_cls, d_tie, nullptr, nullptr, 0,
                           p->tie_tol, p->recheck_tol, st);
    int32_t *list = scr, *counts = scr + kNoAllocRows;
    if (cfg.reg && spec_reg_inline_recheck(p->cp, cfg)) {
        // flagged rows are re-contracted in FP64 inside the kernel; only the
        // per-tile counts go to the scratch
        const int32_t G = 1 << cfg.glog;
        const int64_t nseg = (n + G - 1) / G;
        if (G < 16 || nseg > kNoAllocCounts) return set_error(TQEC_EINVAL, "internal: tile too small for the scratch");
        return spec_launch(p, nullptr, 0, k, dd, d_syn, n, nullptr, nullptr, d_mar, d_cls, d_tie, list, counts, G,
                           p->tie_tol, p->recheck_tol, st);
    }
    const SpecCfg rcfg = spec_recheck_cfg(p);
    SpecKernel *k64 = spec_find(p, dev, rcfg);
    if (!k64) return set_error(TQEC_EINVAL, unprepared);
    if (rcfg.slab) return set_error(TQEC_EUNSUPPORTED, "plan's FP64 re-check runs from a global slab");
    if (cfg.reg) {
        // per-tile segments written whole by each tile (no clearing pass);
        // a tile flags only its own valid rows, so list indices stay below n
        const int32_t G = 1 << cfg.glog;
        const int64_t nseg = (n + G - 1) / G;
        if (G < 16 || nseg > kNoAllocCounts) return set_error(TQEC_EINVAL, "internal: tile too small for the scratch");
        if ((rc = spec_launch(p, nullptr, 0, k, dd, d_syn, n, nullptr, nullptr, d_mar, d_cls, d_tie, list, counts, G,
                              p->tie_tol, p->recheck_tol, st)))
            return rc;
        return spec_launch(p, nullptr, 0, k64, dd, d_syn, n, list, counts, d_mar, d_cls, d_tie, nullptr, nullptr, 0,
                           p->tie_tol, p->recheck_tol, st, nseg, G);
    }
    // LDS tiles: one list with its length cleared by a fill kernel
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, st, counts, 1);
    TQ_HIP(hipGetLastError());
    if ((rc = spec_launch(p, nullptr, 0, k, dd, d_syn, n, nullptr, nullptr, d_mar, d_cls, d_tie, list, counts, n,
                          p->tie_tol, p->recheck_tol, st)))
        return rc;
    return spec_launch(p, nullptr, 0, k64, dd, d_syn, n, list, counts, d_mar, d_cls, d_tie, nullptr, nullptr, 0,
                       p->tie_tol, p->recheck_tol, st);
}

int tqec_decode_batch(tqec_plan *p, const uint8_t *syn, int64_t n, int32_t mode, double *mar, int32_t *cls,
                      uint8_t *tie, void *stream) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (n < 0) return set_error(TQEC_EINVAL, "negative batch size");
    if (n > 0 && !syn) return set_error(TQEC_ENULL, "null syndrome pointer");
    if (n > (int64_t(1) << 31) - 1) return set_error(TQEC_EINVAL, "batch too large (max 2^31-1)");
    if (mode < 0 || mode > TQEC_MODE_SPEC_BF16) return set_error(TQEC_EINVAL, "unknown mode " + std::to_string(mode));
    if (p->cp.semiring != TQEC_SEMIRING_SUM)
        return set_error(TQEC_EINVAL, "a max-product plan is decoded by tqec_map_decode_batch");
    std::lock_guard<std::mutex> lk(p->mu);
    hipStream_t st = (hipStream_t)stream;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    Workspace &w = p->ws[std::make_pair(dev, st)];
    w.device = dev;
    if (n == 0) return TQEC_OK;
    const int nout = 1 << p->cp.n_out;
    int rc;
    if ((rc = grow(&w.d_syn, &w.d_syn_bytes, (size_t)n * std::max(1, p->cp.n_syn)))) return rc;
    if (mar && (rc = grow(&w.d_mar, &w.d_mar_bytes, (size_t)n * nout * sizeof(double)))) return rc;
    if (cls && (rc = grow(&w.d_cls, &w.d_cls_bytes, (size_t)n * sizeof(int32_t)))) return rc;
    if (tie && (rc = grow(&w.d_tie, &w.d_tie_bytes, (size_t)n))) return rc;
    TQ_HIP(hipMemcpyAsync(w.d_syn, syn, (size_t)n * p->cp.n_syn, hipMemcpyHostToDevice, st));
    Workspace *wp = nullptr;
    if ((rc = decode_device(p, w.d_syn, n, mode, mar ? w.d_mar : nullptr, cls ? w.d_cls : nullptr,
                            tie ? w.d_tie : nullptr, st, &wp)))
        return rc;
    if (mar) TQ_HIP(hipMemcpyAsync(mar, w.d_mar, (size_t)n * nout * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cls) TQ_HIP(hipMemcpyAsync(cls, w.d_cls, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (tie) TQ_HIP(hipMemcpyAsync(tie, w.d_tie, (size_t)n, hipMemcpyDeviceToHost, st));
    TQ_HIP(hipStreamSynchronize(st));
    if (wp && wp->count_on_device && wp->host_count) {
        int64_t m;
        if ((rc = read_recheck_count(*wp, &m))) return rc;
        wp->last_recheck = m;
        wp->count_on_device = false;
    }
    return TQEC_OK;
}

int tqec_last_recheck_count(tqec_plan *p, void *stream, int64_t *count) {
    if (!p || !count) return set_error(TQEC_ENULL, "null argument");
    std::lock_guard<std::mutex> lk(p->mu);
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    auto it = p->ws.find(std::make_pair(dev, (hipStream_t)stream));
    if (it == p->ws.end() || !it->second.host_count) {
        *count = -1;
        return TQEC_OK;
    }
    Workspace &w = it->second;
    if (w.count_on_device) {
        TQ_HIP(hipStreamSynchronize((hipStream_t)stream));
        int64_t m;
        int rc;
        if ((rc = read_recheck_count(w, &m))) return rc;
        w.last_recheck = m;
        w.count_on_device = false;
    }
    *count = w.last_recheck;
    return TQEC_OK;
}

int tqec_plan_autotune(tqec_plan *p, int32_t mode, const uint8_t *d_syn, int64_t n, void *stream,
                       int32_t *tile, int32_t *threads) {
    if (!p || !d_syn) return set_error(TQEC_ENULL, "null argument");
    if (n <= 0 || n > (int64_t(1) << 31) - 1) return set_error(TQEC_EINVAL, "trial batch size out of range");
    if (mode == TQEC_MODE_SPEC_F32) return autotune_spec(p, d_syn, n, (hipStream_t)stream, tile, threads, false);
    if (mode == TQEC_MODE_SPEC_BF16) return autotune_spec(p, d_syn, n, (hipStream_t)stream, tile, threads, true);
    if (mode != TQEC_MODE_FUSED_F64 && mode != TQEC_MODE_FUSED_F32)
        return set_error(TQEC_EINVAL, "only the fused modes have tunable tiles");
    std::lock_guard<std::mutex> lk(p->mu);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = configure(p))) return rc;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    DeviceData *dd;
    if ((rc = ensure_device(p, dev, &dd))) return rc;
    Workspace &w = p->ws[std::make_pair(dev, st)];
    w.device = dev;
    const int nout = 1 << p->cp.n_out;
    const bool f32 = mode == TQEC_MODE_FUSED_F32;
    FusedCfg &cfg = f32 ? p->f32 : p->f64;
    if (cfg.glog < 0) return set_error(TQEC_EUNSUPPORTED, "plan does not fit the fused executor");
    if ((rc = grow(&w.d_mar, &w.d_mar_bytes, (size_t)n * nout * sizeof(double)))) return rc;
    if ((rc = grow(&w.d_cls, &w.d_cls_bytes, (size_t)n * sizeof(int32_t)))) return rc;
    if ((rc = grow(&w.d_tie, &w.d_tie_bytes, (size_t)n))) return rc;
    hipEvent_t e0, e1;
    TQ_HIP(hipEventCreate(&e0));
    TQ_HIP(hipEventCreate(&e1));
    const bool saved_timing = p->timing;
    p->timing = false;
    FusedCfg best = cfg;
    float best_ms = 1e30f;
    const size_t elem = f32 ? sizeof(float) : sizeof(double);
    for (int gl = 0; gl <= 6; ++gl) {
        if (fused_lds_bytes(p->cp, gl, elem, f32) > p->lds_limit) break;
        for (int thr : {64, 128, 256}) {
            if (thr < (1 << gl)) continue;
            cfg = make_cfg(p->cp, gl, thr, elem, f32, p->lds_limit, p->n_cu);
            float ms = 0;
            for (int rep = 0; rep < 2 && rc == TQEC_OK; ++rep) {
                if ((rc = record_or(hipEventRecord(e0, st)))) break;
                rc = run_fused(p, dd, w, d_syn, n, mode, w.d_mar, w.d_cls, w.d_tie, st);
                if (rc) break;
                if ((rc = record_or(hipEventRecord(e1, st)))) break;
                if ((rc = record_or(hipEventSynchronize(e1)))) break;
                if ((rc = record_or(hipEventElapsedTime(&ms, e0, e1)))) break;
            }
            if (rc) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = cfg;
            }
        }
        if (rc) break;
    }
    cfg = best;
    p->timing = saved_timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (tile) *tile = 1 << best.glog;
    if (threads) *threads = best.threads;
    return TQEC_OK;
}

int tqec_decode_list_device(tqec_plan *p, const uint8_t *d_syn, const int32_t *d_list, const int32_t *d_count,
                            int64_t n_max, int32_t mode, double *d_mar, int32_t *d_cls, uint8_t *d_tie,
                            void *stream) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (p->cp.semiring != TQEC_SEMIRING_SUM)
        return set_error(TQEC_EINVAL, "a max-product plan is decoded by tqec_map_decode_batch");
    if (mode != TQEC_MODE_FUSED_F64 && mode != TQEC_MODE_FUSED_F32 && mode != TQEC_MODE_SPEC_F32 &&
        mode != TQEC_MODE_SPEC_F64)
        return set_error(TQEC_EINVAL, "list decoding runs on the fused executors");
    if (n_max < 0 || n_max > (int64_t(1) << 31) - 1) return set_error(TQEC_EINVAL, "n_max out of range");
    if (n_max > 0 && (!d_syn || !d_list || !d_count)) return set_error(TQEC_ENULL, "null device pointer");
    std::lock_guard<std::mutex> lk(p->mu);
    int rc;
    if ((rc = configure(p))) return rc;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    DeviceData *dd;
    if ((rc = ensure_device(p, dev, &dd))) return rc;
    Workspace &w = p->ws[std::make_pair(dev, (hipStream_t)stream)];
    w.device = dev;
    if (n_max == 0) return TQEC_OK;
    if (mode == TQEC_MODE_SPEC_F32)
        return run_spec(p, dd, w, d_syn, n_max, d_mar, d_cls, d_tie, (hipStream_t)stream, d_list, d_count);
    if (mode == TQEC_MODE_SPEC_F64)
        return run_spec64(p, dd, w, d_syn, n_max, d_mar, d_cls, d_tie, (hipStream_t)stream, d_list, d_count);
    return run_fused(p, dd, w, d_syn, n_max, mode, d_mar, d_cls, d_tie, (hipStream_t)stream, d_list, d_count);
}

int tqec_map_decode_batch_device(tqec_plan *p, const uint8_t *d_syn, int64_t n, int32_t n_config,
                                 uint8_t *d_config, double *d_logp, void *stream) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (p->cp.semiring != TQEC_SEMIRING_MAX)
        return set_error(TQEC_EINVAL, "tqec_map_decode_batch needs a max-product plan");
    if (n < 0) return set_error(TQEC_EINVAL, "negative batch size");
    if (n > 0 && (!d_syn || (n_config > 0 && !d_config))) return set_error(TQEC_ENULL, "null pointer");
    if (n_config < 0 || n_config > p->cp.n_vars) return set_error(TQEC_EINVAL, "n_config out of range");
    std::lock_guard<std::mutex> lk(p->mu);
    int rc;
    if ((rc = configure(p))) return rc;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    DeviceData *dd;
    if ((rc = ensure_device(p, dev, &dd))) return rc;
    if (n == 0) return TQEC_OK;
    const CompiledPlan &cp = p->cp;
    const int words = (cp.n_vars + 31) / 32;
    auto lds_of = [&](int gl) {
        const size_t G = size_t(1) << gl;
        size_t b = (size_t)cp.ws_entries * G * 8;
        b += ((size_t)cp.arg_entries * G * 2 + 15) & ~size_t(15);
        b += ((size_t)cp.n_syn * G + 15) & ~size_t(15);
        b += G * words * 4;
        return b;
    };
    int gl = -1;
    for (int div : {4, 2, 1}) {
        for (int x = 6; x >= 0 && gl < 0; --x)
            if (lds_of(x) <= p->lds_limit / div) gl = x;
        if (gl >= 0) break;
    }
    if (gl < 0) return set_error(TQEC_EUNSUPPORTED, "max-product plan needs more LDS per syndrome than a workgroup holds");
    const size_t lds = lds_of(gl);
    const int threads = std::max(64, std::min(256, 1 << (gl + 4)));
    MapArgs a;
    std::memset(&a, 0, sizeof(a));
    a.steps = dd->steps;
    a.msteps = dd->msteps;
    a.labels = dd->mlabels;
    a.nsteps = (int)cp.steps.size();
    a.cpool = dd->cpool64;
    a.spool = dd->spool64;
    a.syn = d_syn;
    a.n = n;
    a.nbits = cp.n_syn;
    a.glog = gl;
    a.ws_entries = cp.ws_entries;
    a.arg_entries = cp.arg_entries;
    a.asg_words = words;
    a.root_off = cp.root_off;
    a.n_config = n_config;
    a.config = d_config;
    a.logp = d_logp;
    const int64_t ntiles = (n + (1 << gl) - 1) >> gl;
    const int per_cu = resident_per_cu(lds, threads, p->lds_limit);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)p->n_cu * per_cu));
    if (p->map_spec) {
        // plan-specialised max-product kernel (hiprtc, cached per device and tile)
        const auto key = std::make_tuple(dev, gl, threads, 0, 0, 1, 1 << 20);
        auto it = p->spec.find(key);
        if (it == p->spec.end()) {
            // a plan whose generated kernel cannot be built (hiprtc or module
            // load failure) falls back to the generic kernel for good; the
            // reason is kept in tqec_last_error's thread-local text only until
            // the next error, so it is also recorded on the plan
            std::vector<char> code;
            SpecKernel k;
            k.lds = lds;
            std::string why;
            if (spec_code(spec_map_source(cp, gl, threads), code)) {
                why = "hiprtc: " + g_last_error;
            } else {
                hipError_t e = hipModuleLoadData(&k.mod, code.data());
                if (e == hipSuccess) {
                    e = hipModuleGetFunction(&k.fn, k.mod, "tqec_map_spec");
                    if (e != hipSuccess) (void)hipModuleUnload(k.mod);
                }
                if (e != hipSuccess) why = std::string("module load: ") + hipGetErrorString(e);
            }
            if (why.empty()) {
                it = p->spec.emplace(key, k).first;
            } else {
                p->map_spec = false;
                p->map_spec_note = why;
            }
        }
    }
    if (p->map_spec) {
        auto it = p->spec.find(std::make_tuple(dev, gl, threads, 0, 0, 1, 1 << 20));
        const unsigned char *a_syn = d_syn;
        long long a_n = (long long)n;
        const double *a_cp = dd->cpool64, *a_sp = dd->spool64;
        unsigned char *a_cfg = d_config;
        int a_nc = n_config;
        double *a_lp = d_logp;
        void *args[] = {&a_syn, &a_n, &a_cp, &a_sp, &a_cfg, &a_nc, &a_lp};
        TQ_HIP(hipModuleLaunchKernel(it->second.fn, grid, 1, 1, (unsigned)threads, 1, 1, 0, (hipStream_t)stream, args,
                                     nullptr));
        return TQEC_OK;
    }
    TQ_HIP(hipFuncSetAttribute((const void *)map_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_limit));
    hipLaunchKernelGGL(map_kernel, dim3(grid), dim3(threads), lds, (hipStream_t)stream, a);
    TQ_HIP(hipGetLastError());
    return TQEC_OK;
}

int tqec_map_decode_batch(tqec_plan *p, const uint8_t *syn, int64_t n, int32_t n_config, uint8_t *config,
                          double *logp, void *stream) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (n < 0) return set_error(TQEC_EINVAL, "negative batch size");
    if (n > 0 && (!syn || (n_config > 0 && !config))) return set_error(TQEC_ENULL, "null pointer");
    if (n == 0) return TQEC_OK;
    hipStream_t st = (hipStream_t)stream;
    uint8_t *dsyn = nullptr, *dcfg = nullptr;
    double *dlp = nullptr;
    const size_t sb = (size_t)n * std::max(1, p->cp.n_syn), cb = (size_t)n * std::max(1, n_config);
    TQ_HIP(hipMalloc((void **)&dsyn, sb));
    int rc = TQEC_OK;
    hipError_t e = hipMalloc((void **)&dcfg, cb);
    if (e == hipSuccess && logp) e = hipMalloc((void **)&dlp, (size_t)n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(dsyn, syn, (size_t)n * p->cp.n_syn, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = set_error(TQEC_EDEVICE, hipGetErrorString(e));
    if (rc == TQEC_OK) rc = tqec_map_decode_batch_device(p, dsyn, n, n_config, dcfg, dlp, stream);
    if (rc == TQEC_OK && n_config > 0) {
        e = hipMemcpyAsync(config, dcfg, (size_t)n * n_config, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) rc = set_error(TQEC_EDEVICE, hipGetErrorString(e));
    }
    if (rc == TQEC_OK && logp) {
        e = hipMemcpyAsync(logp, dlp, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) rc = set_error(TQEC_EDEVICE, hipGetErrorString(e));
    }
    e = hipStreamSynchronize(st);
    if (rc == TQEC_OK && e != hipSuccess) rc = set_error(TQEC_EDEVICE, hipGetErrorString(e));
    (void)hipFree(dsyn);
    (void)hipFree(dcfg);
    if (dlp) (void)hipFree(dlp);
    return rc;
}

int tqec_plan_set_tile(tqec_plan *p, int32_t mode, int32_t tile, int32_t threads) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (mode != TQEC_MODE_FUSED_F64 && mode != TQEC_MODE_FUSED_F32 && mode != TQEC_MODE_SPEC_F32)
        return set_error(TQEC_EINVAL, "only the fused modes have tunable tiles");
    int gl = 0;
    while (gl <= 6 && (1 << gl) != tile) ++gl;
    if (gl > 6) return set_error(TQEC_EINVAL, "tile must be a power of two in [1, 64]");
    if (threads != 64 && threads != 128 && threads != 256) return set_error(TQEC_EINVAL, "threads must be 64, 128 or 256");
    if (threads < tile) return set_error(TQEC_EINVAL, "threads must be at least the tile");
    std::lock_guard<std::mutex> lk(p->mu);
    int rc;
    if ((rc = configure(p))) return rc;
    if (mode == TQEC_MODE_SPEC_F32) {
        SpecCfg c;
        c.glog = gl;
        c.threads = threads;
        c.tab_lds = true;
        if (!spec_fits(p, c)) c.tab_lds = false;
        if (!spec_fits(p, c)) return set_error(TQEC_EUNSUPPORTED, "tile does not fit in LDS");
        p->spec_cfg = c;
        p->spec_cfg_set = true;
        return TQEC_OK;
    }
    const bool f32 = mode == TQEC_MODE_FUSED_F32;
    const size_t elem = f32 ? sizeof(float) : sizeof(double);
    if (fused_lds_bytes(p->cp, gl, elem, f32) > p->lds_limit)
        return set_error(TQEC_EUNSUPPORTED, "tile does not fit in LDS");
    (f32 ? p->f32 : p->f64) = make_cfg(p->cp, gl, threads, elem, f32, p->lds_limit, p->n_cu);
    return TQEC_OK;
}

int tqec_plan_set_timing(tqec_plan *p, int enable) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    std::lock_guard<std::mutex> lk(p->mu);
    p->timing = enable != 0;
    return TQEC_OK;
}

int tqec_kernel_time_totals(tqec_plan *p, void *stream, double *main_ms, double *recheck_ms, int64_t *calls) {
    if (!p || !main_ms || !recheck_ms || !calls) return set_error(TQEC_ENULL, "null argument");
    std::lock_guard<std::mutex> lk(p->mu);
    *main_ms = 0;
    *recheck_ms = 0;
    *calls = 0;
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    auto it = p->ws.find(std::make_pair(dev, (hipStream_t)stream));
    if (it == p->ws.end()) return TQEC_OK;
    Workspace &w = it->second;
    for (size_t c = 0; c < w.n_ev; ++c) {
        float ms = 0;
        if (w.ev_used[c][0] && w.ev_used[c][1]) {
            TQ_HIP(hipEventSynchronize(w.ev[c][1]));
            TQ_HIP(hipEventElapsedTime(&ms, w.ev[c][0], w.ev[c][1]));
            *main_ms += ms;
        }
        if (w.ev_used[c][2] && w.ev_used[c][3]) {
            TQ_HIP(hipEventSynchronize(w.ev[c][3]));
            TQ_HIP(hipEventElapsedTime(&ms, w.ev[c][2], w.ev[c][3]));
            *recheck_ms += ms;
        }
    }
    *calls = (int64_t)w.n_ev;
    w.n_ev = 0;
    return TQEC_OK;
}

const char *tqec_last_error(void) { return g_last_error.c_str(); }


int tqec_mixed_create(tqec_plan *const *plans, int32_t n_plans, int32_t max_batches, int64_t max_rows,
                      tqec_mixed **out) {
    if (!out || !plans) return set_error(TQEC_ENULL, "null argument");
    *out = nullptr;
    if (n_plans < 1) return set_error(TQEC_EINVAL, "need at least one plan");
    if (max_batches < 1) return set_error(TQEC_EINVAL, "max_batches must be positive");
    if (max_rows < 1 || max_rows > (int64_t(1) << 31) - 1) return set_error(TQEC_EINVAL, "max_rows out of range");
    for (int i = 0; i < n_plans; ++i) {
        if (!plans[i]) return set_error(TQEC_ENULL, "null plan " + std::to_string(i));
        if (plans[i]->cp.semiring != TQEC_SEMIRING_SUM)
            return set_error(TQEC_EINVAL, "plan " + std::to_string(i) + " is a max-product plan");
    }
    std::unique_ptr<tqec_mixed> m(new (std::nothrow) tqec_mixed());
    if (!m) return set_error(TQEC_ENOMEM, "out of host memory");
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    m->device = dev;
    m->plans.assign(plans, plans + n_plans);
    m->max_batches = max_batches;
    m->max_rows = max_rows;
    int khz = 0;
    TQ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    m->clock_khz = khz > 0 ? (double)khz : 100000.0;
    size_t lds_limit = 0;
    int n_cu = 0;
    std::vector<MixedPlan> hp(n_plans);
    int rc;
    for (int i = 0; i < n_plans; ++i) {
        tqec_plan *p = plans[i];
        std::lock_guard<std::mutex> lk(p->mu);
        if ((rc = configure(p))) return rc;
        DeviceData *dd;
        if ((rc = ensure_device(p, dev, &dd))) return rc;
        lds_limit = p->lds_limit;
        n_cu = p->n_cu;
        const CompiledPlan &cp = p->cp;
        const size_t budget = lds_limit / 2;
        MixedPlanCfg c;
        int gl = p->f32.glog >= 0 ? std::min(p->f32.glog, 6) : 6;
        while (gl >= 0 && fused_lds_bytes(cp, gl, sizeof(float), true) > budget) --gl;
        if (gl >= 0) {
            c.glog32 = gl;
            c.lds32 = fused_lds_bytes(cp, gl, sizeof(float), true);
        } else {
            // slab mode: the LDS keeps the scales and syndrome bits only
            c.gw32 = true;
            const size_t ws = (size_t)std::max<int64_t>(1, cp.ws_entries);
            gl = 6;
            while (gl > 0 && ((ws * sizeof(float)) << gl > kMixedSlabBudget || fused_lds_aux(cp, gl, true) > budget))
                --gl;
            c.glog32 = gl;
            c.lds32 = fused_lds_aux(cp, gl, true);
            c.slab32 = (((ws << gl) * sizeof(float)) + 255) & ~size_t(255);
        }
        const size_t l64 = fused_lds_bytes(cp, 0, sizeof(double), false);
        if (l64 <= budget) {
            c.lds64 = l64;
        } else {
            c.gw64 = true;
            const size_t ws = (size_t)std::max<int64_t>(1, cp.ws_entries);
            c.lds64 = fused_lds_aux(cp, 0, false);
            c.slab64 = (ws * sizeof(double) + 255) & ~size_t(255);
        }
        m->cfg.push_back(c);
        MixedPlan &q = hp[i];
        std::memset(&q, 0, sizeof(q));
        q.steps = dd->steps;
        q.cpool32 = dd->cpool32;
        q.spool32 = dd->spool32;
        q.cpool64 = dd->cpool64;
        q.spool64 = dd->spool64;
        q.ws_entries = cp.ws_entries;
        q.root_off = cp.root_off;
        q.nsteps = (int)cp.steps.size();
        q.nbits = cp.n_syn;
        q.nout = 1 << cp.n_out;
        q.nslots = cp.n_slots;
        q.root_slot = cp.root_slot;
        q.glog32 = c.glog32;
        q.gw32 = c.gw32;
        q.gw64 = c.gw64;
        q.tie_tol = p->tie_tol;
        q.recheck_tol = p->recheck_tol;
        m->lds = std::max(m->lds, std::max(c.lds32, c.lds64));
        m->cost.push_back((double)cp.dep_tc + (double)cp.dep_rw);
        m->slab32_bytes = std::max(m->slab32_bytes, c.slab32);
        m->slab64_bytes = std::max(m->slab64_bytes, c.slab64);
    }
    // (the kernel's own static LDS: the work-queue slot and the re-check list)
    if (m->lds > lds_limit - 512) return set_error(TQEC_EUNSUPPORTED, "a plan's scales and syndrome bits exceed the LDS");
    TQ_HIP(hipFuncSetAttribute((const void *)mixed_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(lds_limit - 512)));
    m->blocks = n_cu * resident_per_cu(m->lds + 512, kMixedThreads, lds_limit);
    // the persistent kernel takes any grid: cap its workgroups so that the
    // slabs of a wide plan (toric L=5: 68 MB per workgroup) stay in budget
    if (const size_t per = m->slab32_bytes + m->slab64_bytes)
        m->blocks = (int)std::max<size_t>((size_t)n_cu, std::min<size_t>((size_t)m->blocks, kMixedSlabTotal / per));
    tqec_mixed *mp = m.get();
    auto fail = [&](int code) {
        mixed_free(mp);
        m.reset();
        return code;
    };
#define TQ_MX(call)                                                                               \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(set_error(TQEC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_))); \
    } while (0)
    TQ_MX(hipMalloc((void **)&mp->d_plans, hp.size() * sizeof(MixedPlan)));
    TQ_MX(hipMemcpy(mp->d_plans, hp.data(), hp.size() * sizeof(MixedPlan), hipMemcpyHostToDevice));
    if (mp->slab32_bytes) TQ_MX(hipMalloc((void **)&mp->slab32, mp->slab32_bytes * (size_t)mp->blocks));
    if (mp->slab64_bytes) TQ_MX(hipMalloc((void **)&mp->slab64, mp->slab64_bytes * (size_t)mp->blocks));
    const size_t tb = mixed_table_bytes(max_batches);
    for (MixedSlot &sl : mp->slot) {
        TQ_MX(hipHostMalloc(&sl.h_stage, tb, hipHostMallocDefault));
        std::memset(sl.h_stage, 0, tb);
        TQ_MX(hipMalloc(&sl.d_table, tb));
        TQ_MX(hipMalloc((void **)&sl.ctr, sizeof(MixedCounters)));
        TQ_MX(hipMalloc((void **)&sl.t_done, (size_t)max_batches * sizeof(unsigned long long)));
        TQ_MX(hipEventCreateWithFlags(&sl.staged, hipEventDisableTiming));
    }
#undef TQ_MX
    *out = m.release();
    return TQEC_OK;
}

int tqec_mixed_set_graph(tqec_mixed *m, int32_t enable) {
    if (!m) return set_error(TQEC_ENULL, "null executor");
    m->use_graph = enable != 0;
    return TQEC_OK;
}

int tqec_mixed_plan_config(const tqec_mixed *m, int32_t plan, int32_t *tile, int32_t *slab) {
    if (!m) return set_error(TQEC_ENULL, "null executor");
    if (plan < 0 || plan >= (int32_t)m->cfg.size()) return set_error(TQEC_EINVAL, "plan index out of range");
    if (m->spec) {
        if (tile) *tile = 1 << m->spec_glog[plan];
        if (slab) *slab = m->spec_slab[plan];
        return TQEC_OK;
    }
    if (tile) *tile = 1 << m->cfg[plan].glog32;
    if (slab) *slab = (m->cfg[plan].gw32 ? 1 : 0) | (m->cfg[plan].gw64 ? 2 : 0);
    return TQEC_OK;
}

int tqec_mixed_decode_device(tqec_mixed *m, const tqec_mixed_batch *batches, int32_t n_batches, void *stream) {
    if (!m) return set_error(TQEC_ENULL, "null executor");
    if (n_batches < 0 || n_batches > m->max_batches)
        return set_error(TQEC_EINVAL, "n_batches out of range [0, " + std::to_string(m->max_batches) + "]");
    if (n_batches > 0 && !batches) return set_error(TQEC_ENULL, "null batch table");
    int dev;
    TQ_HIP(hipGetDevice(&dev));
    if (dev != m->device) return set_error(TQEC_EINVAL, "executor was created on another device");
    int64_t rows = 0;
    for (int32_t i = 0; i < n_batches; ++i) {
        const tqec_mixed_batch &b = batches[i];
        if (b.plan < 0 || b.plan >= (int32_t)m->plans.size())
            return set_error(TQEC_EINVAL, "batch " + std::to_string(i) + ": plan index out of range");
        if (b.n < 0) return set_error(TQEC_EINVAL, "batch " + std::to_string(i) + ": negative size");
        if (b.n > 0 && !b.syndromes) return set_error(TQEC_ENULL, "batch " + std::to_string(i) + ": null syndromes");
        rows += b.n;
        if (rows > m->max_rows)
            return set_error(TQEC_EINVAL, "batches hold more than max_rows = " + std::to_string(m->max_rows) + " syndromes");
    }
    hipStream_t st = (hipStream_t)stream;
    MixedSlot &sl = m->slot[m->cur];
    if (sl.used) TQ_HIP(hipEventSynchronize(sl.staged));
    // Calls share the slabs: a call on another stream starts after the last one.
    const MixedSlot &prev = m->slot[m->cur ^ 1];
    if (prev.used) TQ_HIP(hipStreamWaitEvent(st, prev.staged, 0));
    char *h = reinterpret_cast<char *>(sl.h_stage);
    MixedHeader *hdr = reinterpret_cast<MixedHeader *>(h);
    MixedBatch *tb = reinterpret_cast<MixedBatch *>(h + sizeof(MixedHeader));
    sl.perm.resize(n_batches);
    for (int32_t i = 0; i < n_batches; ++i) sl.perm[i] = i;
    if (m->order == TQEC_MIXED_ORDER_CHEAPEST_FIRST)
        std::stable_sort(sl.perm.begin(), sl.perm.end(), [&](int32_t x, int32_t y) {
            return (double)batches[x].n * m->cost[batches[x].plan] < (double)batches[y].n * m->cost[batches[y].plan];
        });
    int64_t tiles = 0;
    for (int32_t i = 0; i < n_batches; ++i) {
        const tqec_mixed_batch &b = batches[sl.perm[i]];
        int gl = 0;
        if (m->spec)
            gl = m->spec_glog[b.plan];   // the generated tile has a fixed size
        else
            while (gl < m->cfg[b.plan].glog32 && (int64_t(1) << gl) < b.n) ++gl;
        MixedBatch &q = tb[i];
        q.syn = b.syndromes;
        q.mar = b.marginals;
        q.cls = b.classes;
        q.tie = b.ties;
        q.n = b.n;
        q.tile0 = tiles;
        q.plan = b.plan;
        q.glog = gl;
        tiles += (b.n + (int64_t(1) << gl) - 1) >> gl;
    }
    hdr->nbatches = n_batches;
    hdr->pad = 0;
    hdr->ntiles = tiles;
    int rc;
    if (m->use_graph && st != nullptr) {
        if (!sl.exec || sl.graph_stream != st) {
            if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
            if (sl.graph) (void)hipGraphDestroy(sl.graph);
            sl.exec = nullptr;
            sl.graph = nullptr;
            TQ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = mixed_enqueue(m, sl, st);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return set_error(TQEC_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            sl.graph = g;
            TQ_HIP(hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0));
            sl.graph_stream = st;
        }
        TQ_HIP(hipGraphLaunch(sl.exec, st));
    } else {
        if ((rc = mixed_enqueue(m, sl, st))) return rc;
    }
    TQ_HIP(hipEventRecord(sl.staged, st));
    sl.used = true;
    sl.nbatches = n_batches;
    m->last = m->cur;
    m->cur ^= 1;
    return TQEC_OK;
}

int tqec_mixed_last_call(tqec_mixed *m, double *batch_us, int32_t n_batches, int64_t *recheck_count) {
    if (!m) return set_error(TQEC_ENULL, "null executor");
    if (m->last < 0) return set_error(TQEC_EINVAL, "no call yet");
    MixedSlot &sl = m->slot[m->last];
    if (batch_us && n_batches != sl.nbatches)
        return set_error(TQEC_EINVAL, "the last call had " + std::to_string(sl.nbatches) + " batches");
    TQ_HIP(hipSetDevice(m->device));
    TQ_HIP(hipEventSynchronize(sl.staged));
    MixedCounters c;
    TQ_HIP(hipMemcpy(&c, sl.ctr, sizeof(c), hipMemcpyDeviceToHost));
    if (recheck_count) *recheck_count = c.rl_count;
    if (batch_us && n_batches > 0) {
        std::vector<unsigned long long> t(n_batches);
        TQ_HIP(hipMemcpy(t.data(), sl.t_done, (size_t)n_batches * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int32_t i = 0; i < n_batches; ++i)
            batch_us[sl.perm[i]] = (t[i] == 0 || c.t_start == ~0ull || t[i] < c.t_start)
                                       ? 0.0
                                       : (double)(t[i] - c.t_start) * 1000.0 / m->clock_khz;
    }
    return TQEC_OK;
}

void tqec_mixed_destroy(tqec_mixed *m) {
    if (!m) return;
    mixed_free(m);
    delete m;
}


int tqec_plan_spec_source(const tqec_plan *p, int32_t tile, int32_t threads, int32_t flags, char *buf, size_t cap,
                          size_t *len) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (p->cp.semiring != TQEC_SEMIRING_SUM) return set_error(TQEC_EINVAL, "max-product plans have no FP32 kernel");
    int gl = 0;
    while (gl <= 6 && (1 << gl) != tile) ++gl;
    if (gl > 6) return set_error(TQEC_EINVAL, "tile must be a power of two in [1, 64]");
    if ((threads != 64 && threads != 128 && threads != 256) || threads < tile)
        return set_error(TQEC_EINVAL, "threads must be 64, 128 or 256 and at least the tile");
    SpecCfg c;
    c.glog = gl;
    c.threads = threads;
    c.tab_lds = (flags & 1) != 0;
    c.fp64 = (flags & 2) != 0;
    c.slab = (flags & 4) != 0;
    c.mfma = (flags & 8) != 0;
    c.reg = (flags & 16) != 0;
    c.dyn = (flags & 32) != 0;
    c.mfma32 = (flags & 64) != 0;
    c.unroll = std::max(1, (flags >> 8) & 0xff);
    if (c.reg && (c.mfma || c.mfma32 || c.slab || tile * 4 != threads))
        return set_error(TQEC_EINVAL, "register-resident kernels run without slab or MFMA, tile = threads / 4");
    if (c.mfma && c.mfma32) return set_error(TQEC_EINVAL, "BF16 and FP32 MFMA steps are exclusive");
    std::string src;
    try {
        src = spec_source(p->cp, c);
    } catch (const std::exception &e) {
        return set_error(TQEC_ENOMEM, e.what());
    }
    std::string why;
    if (src.empty() && c.reg && !spec_reg_supported(p->cp, &why))
        return set_error(TQEC_EUNSUPPORTED, "plan does not suit the register-resident kernel: " + why);
    if (src.empty()) return set_error(TQEC_EUNSUPPORTED, "plan does not suit this specialised kernel");
    if (len) *len = src.size();
    if (buf && cap > 0) {
        const size_t m = std::min(cap - 1, src.size());
        std::memcpy(buf, src.data(), m);
        buf[m] = 0;
    }
    return TQEC_OK;
}


int tqec_plan_set_spec(tqec_plan *p, int32_t tile, int32_t threads, int32_t flags, int32_t waves) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    int gl = 0;
    while (gl <= 6 && (1 << gl) != tile) ++gl;
    if (gl > 6) return set_error(TQEC_EINVAL, "tile must be a power of two in [1, 64]");
    if ((threads != 64 && threads != 128 && threads != 256) || threads < tile)
        return set_error(TQEC_EINVAL, "threads must be 64, 128 or 256 and at least the tile");
    const int unroll = (waves >> 8) & 0xff;
    waves &= 0xff;
    if (waves > 8) return set_error(TQEC_EINVAL, "waves must be in [0, 8]");
    if (unroll != 0 && unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8)
        return set_error(TQEC_EINVAL, "unroll must be 1, 2, 4 or 8");
    std::lock_guard<std::mutex> lk(p->mu);
    int rc;
    if ((rc = configure(p))) return rc;
    SpecCfg c;
    c.glog = gl;
    c.threads = threads;
    c.tab_lds = (flags & 1) != 0;
    c.slab = (flags & 4) != 0;
    c.mfma = (flags & 8) != 0;
    c.reg = (flags & 16) != 0;
    c.dyn = (flags & 32) != 0;
    c.mfma32 = (flags & 64) != 0;
    c.waves = waves;
    c.unroll = unroll ? unroll : 1;
    if (c.reg && (c.mfma || c.mfma32 || c.slab || tile * 4 != threads))
        return set_error(TQEC_EINVAL, "register-resident kernels are FP32 without slab or MFMA, tile = threads / 4");
    if (c.mfma && c.mfma32) return set_error(TQEC_EINVAL, "BF16 and FP32 MFMA steps are exclusive");
    if (c.mfma32 && (gl != 0 || threads % 64 != 0))
        return set_error(TQEC_EINVAL, "FP32 MFMA steps need tiles of one syndrome");
    std::string why;
    if (c.reg && !spec_reg_supported(p->cp, &why))
        return set_error(TQEC_EUNSUPPORTED, "plan does not suit the register-resident kernel: " + why);
    if (!spec_fits(p, c)) return set_error(TQEC_EUNSUPPORTED, "tile does not fit in LDS");
    if (c.mfma) {
        if (c.glog != 0) return set_error(TQEC_EINVAL, "MFMA steps need tiles of one syndrome");
        p->specb_cfg = c;
        p->specb_cfg_set = true;
    } else {
        p->spec_cfg = c;
        p->spec_cfg_set = true;
    }
    return TQEC_OK;
}

int tqec_plan_set_map_kernel(tqec_plan *p, int32_t specialised) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    if (p->cp.semiring != TQEC_SEMIRING_MAX) return set_error(TQEC_EINVAL, "not a max-product plan");
    std::lock_guard<std::mutex> lk(p->mu);
    p->map_spec = specialised != 0;
    if (p->map_spec) p->map_spec_note.clear();
    return TQEC_OK;
}

int tqec_plan_map_kernel(tqec_plan *p, int32_t *specialised, char *note, size_t cap) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    std::lock_guard<std::mutex> lk(p->mu);
    if (specialised) *specialised = p->map_spec ? 1 : 0;
    if (note && cap > 0) {
        const size_t k = std::min(cap - 1, p->map_spec_note.size());
        std::memcpy(note, p->map_spec_note.data(), k);
        note[k] = 0;
    }
    return TQEC_OK;
}

int tqec_plan_spec_config(tqec_plan *p, int32_t *tile, int32_t *threads, int32_t *flags, int32_t *waves) {
    if (!p) return set_error(TQEC_ENULL, "null plan");
    std::lock_guard<std::mutex> lk(p->mu);
    int rc;
    if ((rc = configure(p))) return rc;
    if (!p->spec_cfg_set) {
        p->spec_cfg = choose_spec(p);
        p->spec_cfg_set = true;
    }
    if (p->spec_cfg.glog < 0) return set_error(TQEC_EUNSUPPORTED, "plan does not fit the specialised kernel");
    if (tile) *tile = 1 << p->spec_cfg.glog;
    if (threads) *threads = p->spec_cfg.threads;
    if (flags)
        *flags = (p->spec_cfg.tab_lds ? 1 : 0) | (p->spec_cfg.slab ? 4 : 0) | (p->spec_cfg.reg ? 16 : 0) |
                 (p->spec_cfg.dyn ? 32 : 0) | (p->spec_cfg.mfma32 ? 64 : 0);
    if (waves) *waves = p->spec_cfg.waves | (p->spec_cfg.unroll << 8);
    return TQEC_OK;
}

int tqec_mixed_set_order(tqec_mixed *m, int32_t order) {
    if (!m) return set_error(TQEC_ENULL, "null executor");
    if (order != TQEC_MIXED_ORDER_GIVEN && order != TQEC_MIXED_ORDER_CHEAPEST_FIRST)
        return set_error(TQEC_EINVAL, "unknown order");
    m->order = order;
    return TQEC_OK;
}

int tqec_mixed_set_kernel(tqec_mixed *m, int32_t kind) {
    if (!m) return set_error(TQEC_ENULL, "null executor");
    if (kind != TQEC_MIXED_KERNEL_GENERIC && kind != TQEC_MIXED_KERNEL_SPEC) return set_error(TQEC_EINVAL, "unknown kind");
    TQ_HIP(hipSetDevice(m->device));
    mixed_drop_graphs(m);
    mixed_free_spec(m);
    if (kind == TQEC_MIXED_KERNEL_GENERIC) return TQEC_OK;
    const int nt = m->spec_threads;
    size_t lds_limit = 0;
    int n_cu = 0;
    int rc;
    for (tqec_plan *p : m->plans) {
        std::lock_guard<std::mutex> lk(p->mu);
        if ((rc = configure(p))) return rc;
        lds_limit = p->lds_limit;
        n_cu = p->n_cu;
    }
    std::vector<MixedSpecPlan> sp;
    std::vector<int> glog, slabf;
    size_t s32 = 0, s64 = 0;
    if ((rc = mixed_spec_layout(m->plans, nt, lds_limit, sp, glog, slabf, s32, s64))) return rc;
    const size_t lds = spec_mixed_lds_bytes(sp) + 8 + 4 + 256;
    if (lds > lds_limit) return set_error(TQEC_EUNSUPPORTED, "the plans' tiles exceed the LDS");
    std::vector<char> code;
    if ((rc = spec_code(spec_mixed_source(sp, nt, std::max<size_t>(s32, 1), std::max<size_t>(s64, 1)), code))) return rc;
    TQ_HIP(hipModuleLoadData(&m->spec_mod, code.data()));
    hipError_t e = hipModuleGetFunction(&m->spec_fn, m->spec_mod, "tqec_mixed_spec");
    if (e != hipSuccess) {
        mixed_free_spec(m);
        return set_error(TQEC_EDEVICE, std::string("hipModuleGetFunction: ") + hipGetErrorString(e));
    }
    m->spec_blocks = n_cu * resident_per_cu(lds, nt, lds_limit);
    if (const size_t per = s32 * sizeof(float) + s64 * sizeof(double))
        m->spec_blocks = (int)std::max<size_t>((size_t)n_cu, std::min<size_t>((size_t)m->spec_blocks, kMixedSlabTotal / per));
    std::vector<const void *> t32(m->plans.size()), t64(m->plans.size());
    for (size_t i = 0; i < m->plans.size(); ++i) {
        DeviceData *dd;
        std::lock_guard<std::mutex> lk(m->plans[i]->mu);
        if ((rc = ensure_device(m->plans[i], m->device, &dd))) {
            mixed_free_spec(m);
            return rc;
        }
        t32[i] = dd->tab32;
        t64[i] = dd->tab64;
    }
#define TQ_SP(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            mixed_free_spec(m);                                                                  \
            return set_error(TQEC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
        }                                                                                        \
    } while (0)
    TQ_SP(hipMalloc((void **)&m->d_tabs32, t32.size() * sizeof(void *)));
    TQ_SP(hipMalloc((void **)&m->d_tabs64, t64.size() * sizeof(void *)));
    TQ_SP(hipMemcpy(m->d_tabs32, t32.data(), t32.size() * sizeof(void *), hipMemcpyHostToDevice));
    TQ_SP(hipMemcpy(m->d_tabs64, t64.data(), t64.size() * sizeof(void *), hipMemcpyHostToDevice));
    if (s32) TQ_SP(hipMalloc((void **)&m->spec_slab32, s32 * sizeof(float) * (size_t)m->spec_blocks));
    if (s64) TQ_SP(hipMalloc((void **)&m->spec_slab64, s64 * sizeof(double) * (size_t)m->spec_blocks));
#undef TQ_SP
    m->spec_glog = glog;
    m->spec_slab = slabf;
    m->spec = true;
    return TQEC_OK;
}

int tqec_mixed_spec_source(tqec_plan *const *plans, int32_t n_plans, int32_t threads, char *buf, size_t cap,
                           size_t *len) {
    if (!plans || n_plans < 1) return set_error(TQEC_EINVAL, "need at least one plan");
    if (threads != 64 && threads != 128 && threads != 256) return set_error(TQEC_EINVAL, "threads must be 64, 128 or 256");
    std::vector<tqec_plan *> v(plans, plans + n_plans);
    for (tqec_plan *p : v)
        if (!p) return set_error(TQEC_ENULL, "null plan");
    std::vector<MixedSpecPlan> sp;
    std::vector<int> glog, slabf;
    size_t s32 = 0, s64 = 0;
    int rc;
    // gfx950: 160 KiB of LDS per workgroup
    if ((rc = mixed_spec_layout(v, threads, size_t(160) << 10, sp, glog, slabf, s32, s64))) return rc;
    const std::string src = spec_mixed_source(sp, threads, std::max<size_t>(s32, 1), std::max<size_t>(s64, 1));
    if (len) *len = src.size();
    if (buf && cap > 0) {
        const size_t k = std::min(cap - 1, src.size());
        std::memcpy(buf, src.data(), k);
        buf[k] = 0;
    }
    return TQEC_OK;
}
}  // extern "C"
