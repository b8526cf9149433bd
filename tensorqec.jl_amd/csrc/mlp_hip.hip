// mlp_hip.hip — learned front-end: a syndrome -> logical-class MLP with
// hand-written HIP training and inference kernels for gfx950.
//
// Network: x (n_in bits, zero-padded to a multiple of 32) -> 3 hidden layers
// (ReLU) -> n_out logits.  Weights are kept in FP32 (master copy, Adam) and
// in BF16 (the copy every GEMM reads); activations are BF16; every product is
// accumulated in FP32 by v_mfma_f32_32x32x16_bf16.  Training minimises the
// soft cross-entropy against class posteriors produced by the tensor-network
// decoder.  A confidence gate accepts the network's class when its posterior
// reaches a threshold and queues the other syndromes for the TN path.
//
// This component has no counterpart in the reference (TensorQEC.jl has no
// learned decoder); it sits in front of the TNMMAP path
// (src/decoding/tndecoder.jl:160-165) and falls back to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tqec_hip.h"

namespace tqec {
int set_error_shared(int code, const std::string &msg);
}

namespace {

int fail(int code, const std::string &m) { return tqec::set_error_shared(code, m); }

#define TQM_HIP(call)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess) return fail(TQEC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef uint16_t bf16;
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (bf16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);   // round to nearest even
}

// ---------------------------------------------------------------------------
// GEMM: C[M][N] = op(A)[M][K] * op(B)[K][N], BF16 in, FP32 accumulate.
//   A: AT ? A[k*lda + m] : A[m*lda + k];   B: BT ? B[n*ldb + k] : B[k*ldb + n]
// Block tile 128x128, K step 32, 4 waves each owning 64x64 = 2x2 MFMA tiles
// of 32x32.  Tiles are staged in LDS as [row][k] (k contiguous) so each lane
// reads its 8 consecutive k values with one 16-byte LDS load.  Rows/columns
// beyond M/N are zero-filled on load and skipped on store; K must be a
// multiple of 32.
// ---------------------------------------------------------------------------

enum Epi : int {
    kEpiBiasReluBf16 = 0,  // out_bf16 = bf16(relu(acc + bias[col]))
    kEpiMaskBf16 = 1,      // out_bf16 = bf16(acc) where mask_bf16 > 0, else 0
    kEpiF32 = 2,           // out_f32 = acc
};

struct GemmArgs {
    const bf16 *A;
    const bf16 *B;
    int M, N, K, lda, ldb, ldc;
    const float *bias;
    const bf16 *mask;
    bf16 *outb;
    float *outf;
};

constexpr int BM = 128, BN = 128, BK = 32, LDSK = BK + 8;

template <bool T>
__device__ __forceinline__ void load_tile(bf16 *S, const bf16 *G, int rows0, int R, int k0, int K, int ld, int tid) {
    // S[r][k] for r in [0,128), k in [0,32)
    if (!T) {
        // G[r*ld + k]: 128 rows x 32 k = 512 chunks of 8
        for (int c = tid; c < 512; c += 256) {
            const int r = c >> 2, kc = (c & 3) * 8;
            const int gr = rows0 + r;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gr < R) v = *reinterpret_cast<const uint4 *>(G + (size_t)gr * ld + k0 + kc);
            *reinterpret_cast<uint4 *>(S + r * LDSK + kc) = v;
        }
    } else {
        // G[k*ld + r]: 32 k x 128 r = 512 chunks of 8 along r, transposed into S
        for (int c = tid; c < 512; c += 256) {
            const int k = c >> 4, rc = (c & 15) * 8;
            const int gr = rows0 + rc;
            bf16 v[8];
            if (gr + 7 < R) {
                const uint4 u = *reinterpret_cast<const uint4 *>(G + (size_t)(k0 + k) * ld + gr);
                memcpy(v, &u, 16);
            } else {
                for (int j = 0; j < 8; ++j) v[j] = (gr + j < R) ? G[(size_t)(k0 + k) * ld + gr + j] : (bf16)0;
            }
            for (int j = 0; j < 8; ++j) S[(rc + j) * LDSK + k] = v[j];
        }
    }
}

template <bool AT, bool BT, int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
    __shared__ __align__(16) bf16 As[BM * LDSK];
    __shared__ __align__(16) bf16 Bs[BN * LDSK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x16 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        load_tile<AT>(As, g.A, m0, g.M, k0, g.K, g.lda, tid);
        load_tile<BT ? false : true>(Bs, g.B, n0, g.N, k0, g.K, g.ldb, tid);
        __syncthreads();
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 a[2], b[2];
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(As + (wm + i * 32 + (lane & 31)) * LDSK + kk + 8 * (lane >> 5));
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(Bs + (wn + j * 32 + (lane & 31)) * LDSK + kk + 8 * (lane >> 5));
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout of the 32x32 tile: lane l holds column l%32, value v is row
    // 8*(v/4) + 4*(l/32) + v%4.
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + j * 32 + (lane & 31);
            if (col >= g.N) continue;
            for (int v = 0; v < 16; ++v) {
                const int row = m0 + wm + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row >= g.M) continue;
                const float x = acc[i][j][v];
                const size_t o = (size_t)row * g.ldc + col;
                if (EPI == kEpiBiasReluBf16) {
                    g.outb[o] = f2bf(fmaxf(x + g.bias[col], 0.f));
                } else if (EPI == kEpiMaskBf16) {
                    g.outb[o] = bf2f(g.mask[o]) > 0.f ? f2bf(x) : (bf16)0;
                } else {
                    g.outf[o] = x;
                }
            }
        }
}

// Forward GEMM of a hidden layer: H[M][N] = relu(A[M][K] * W + bias) with
// the weights given transposed, Wt[N][K], so both tiles are rows of 16-byte
// k-chunks.  Block tile 128 x (64 * NW), K step KS (64, or 32 when K is not
// a multiple of 64), 4 waves in a 2 x 2 grid each owning 64 x (32 * NW) =
// 2 x NW MFMA tiles;
// the next K step's tiles are loaded into registers while the current one
// is multiplied out of LDS.  Same MFMA sequence (k ascending in steps of 16)
// as gemm_bf16_kernel, so the results are identical.
template <int KS, int NW>
__global__ void __launch_bounds__(256) gemm_fwd_kernel(const bf16 *__restrict__ A, const bf16 *__restrict__ Wt,
                                                       int M, int N, int K, const float *__restrict__ bias,
                                                       bf16 *__restrict__ out) {
    constexpr int TN = 64 * NW;                // block tile columns
    constexpr int LK = KS + 8;                 // padded LDS row (bf16)
    constexpr int PA = 128 * KS / 8 / 256;     // 16-byte chunks per thread: A tile
    constexpr int PB = TN * KS / 8 / 256;      //                             B tile
    __shared__ __align__(16) bf16 As[128 * LK];
    __shared__ __align__(16) bf16 Bs[TN * LK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 128, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * (32 * NW);
    uint4 ra[PA], rb[PB];
    auto fetch = [&](int k0) {
        for (int j = 0; j < PA; ++j) {
            const int c = tid + j * 256;
            const int r = c / (KS / 8), kc = (c % (KS / 8)) * 8;
            const int gm = m0 + r;
            ra[j] = gm < M ? *reinterpret_cast<const uint4 *>(A + (size_t)gm * K + k0 + kc) : make_uint4(0, 0, 0, 0);
        }
        for (int j = 0; j < PB; ++j) {
            const int c = tid + j * 256;
            const int r = c / (KS / 8), kc = (c % (KS / 8)) * 8;
            const int gn = n0 + r;
            rb[j] = gn < N ? *reinterpret_cast<const uint4 *>(Wt + (size_t)gn * K + k0 + kc) : make_uint4(0, 0, 0, 0);
        }
    };
    auto stash = [&]() {
        for (int j = 0; j < PA; ++j) {
            const int c = tid + j * 256;
            *reinterpret_cast<uint4 *>(As + (c / (KS / 8)) * LK + (c % (KS / 8)) * 8) = ra[j];
        }
        for (int j = 0; j < PB; ++j) {
            const int c = tid + j * 256;
            *reinterpret_cast<uint4 *>(Bs + (c / (KS / 8)) * LK + (c % (KS / 8)) * 8) = rb[j];
        }
    };
    f32x16 acc[2][NW];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < NW; ++j)
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    fetch(0);
    stash();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += KS) {
        const bool more = k0 + KS < K;
        if (more) fetch(k0 + KS);
        for (int kk = 0; kk < KS; kk += 16) {
            bf16x8 a[2], b[NW];
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(As + (wm + i * 32 + (lane & 31)) * LK + kk + 8 * (lane >> 5));
            for (int j = 0; j < NW; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(Bs + (wn + j * 32 + (lane & 31)) * LK + kk + 8 * (lane >> 5));
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < NW; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            stash();
            __syncthreads();
        }
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < NW; ++j) {
            const int col = n0 + wn + j * 32 + (lane & 31);
            if (col >= N) continue;
            const float bc = bias[col];
            for (int v = 0; v < 16; ++v) {
                const int row = m0 + wm + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row >= M) continue;
                out[(size_t)row * N + col] = f2bf(fmaxf(acc[i][j][v] + bc, 0.f));
            }
        }
}

// ---------------------------------------------------------------------------
// Fused inference: the whole network for a tile of 128 syndromes in one
// workgroup, hidden width 512.  The tile's activations stay in LDS
// (128 x 512 BF16 = 128 KB, rows of 64 16-byte chunks, chunk c of row m
// stored at c ^ (m & 15) so that the 16 lanes of a ds_read_b128 pass hit
// distinct banks) and each layer overwrites them in place after a barrier.
// Every product runs on v_mfma_f32_32x32x16_bf16 with the operands swapped
// (D = W^T * act^T): wave w owns output columns [w * 512 / NW, ...) for all
// 128 rows, so no two waves read the same weights and each wave streams its
// weight fragments straight from global memory (L2-resident, 1.1 MB for the
// d=7 network) in the fragment order laid out once by pack_fused_kernel,
// two k-steps ahead; only the activations are shared, with no barrier inside
// a layer.  A lane of a D tile holds 4 consecutive output columns of one
// row, written back as one 8-byte LDS store per group.  Biases seed the
// accumulators.  The output layer (N <= 32, padded) runs the same way with
// one wave per 32 rows.  FP32 accumulation, BF16 activations, as the
// layer-by-layer kernels (same roundings; the bias enters before the
// products instead of after them).
// ---------------------------------------------------------------------------

constexpr int kFusedRows = 128, kFusedHid = 512;

__device__ __forceinline__ int act_off(int m, int chunk) { return m * (kFusedHid * 2) + ((chunk ^ (m & 15)) << 4); }

// Fragment-ordered weights.  Hidden layer l (K x 512): for each 16-row k-step
// s and 32-column tile t, 64 lanes x 8 BF16: lane L holds W[16 s + 8 (L >> 5)
// + e][32 t + (L & 31)], e = 0..7.  Output layer: one 32-column tile, columns
// >= nout zero.
__global__ void pack_fused_kernel(const bf16 *__restrict__ W, int K, int N, int nout_pad, bf16 *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int tiles = nout_pad / 32;
    if (i >= (int64_t)K * nout_pad) return;
    const int e = (int)(i & 7), L = (int)((i >> 3) & 63);
    const int64_t q = i >> 9;   // (s, t)
    const int t = (int)(q % tiles), s = (int)(q / tiles);
    const int k = 16 * s + 8 * (L >> 5) + e, n = 32 * t + (L & 31);
    out[i] = n < N ? W[(int64_t)k * N + n] : (bf16)0;
}

template <int NW>
__global__ void __launch_bounds__(64 * NW) mlp_fused_fwd_kernel(const uint8_t *__restrict__ syn, int64_t B, int n_in,
                                                                 int kin, const bf16 *__restrict__ wf1,
                                                                 const bf16 *__restrict__ wf2,
                                                                 const bf16 *__restrict__ wf3,
                                                                 const float *__restrict__ b1,
                                                                 const float *__restrict__ b2,
                                                                 const float *__restrict__ b3,
                                                                 const bf16 *__restrict__ wf4,
                                                                 const float *__restrict__ b4, int nout,
                                                                 float *__restrict__ z) {
    extern __shared__ __align__(16) unsigned char act[];
    constexpr int NT = 16 / NW;                  // 32-column tiles per wave
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m32 = lane & 31, h = lane >> 5;
    const int64_t row0 = (int64_t)blockIdx.x * kFusedRows;
    // syndrome bits -> BF16 0/1 features in columns [0, kin)
    const int cpr = kin >> 3;
    for (int q = tid; q < kFusedRows * cpr; q += 64 * NW) {
        const int m = q / cpr, c = q - m * cpr;
        const int64_t r = row0 + m;
        uint16_t v[8];
        for (int e = 0; e < 8; ++e) {
            const int f = 8 * c + e;
            v[e] = (r < B && f < n_in && (syn[r * n_in + f] & 1)) ? (uint16_t)0x3f80 : (uint16_t)0;
        }
        uint4 u;
        memcpy(&u, v, 16);
        *reinterpret_cast<uint4 *>(act + act_off(m, c)) = u;
    }
    __syncthreads();
    for (int l = 0; l < 3; ++l) {
        const int K = l == 0 ? kin : kFusedHid;
        const int steps = K >> 4;   // even: K is a multiple of 32
        const bf16 *W = l == 0 ? wf1 : (l == 1 ? wf2 : wf3);
        const float *bias = l == 0 ? b1 : (l == 1 ? b2 : b3);
        f32x16 acc[NT][4];
        for (int t = 0; t < NT; ++t) {
            // lane holds columns 32 T + 8 g + 4 h + r for v = 4 g + r
            const int n0 = 32 * (wave * NT + t) + 4 * h;
            for (int g = 0; g < 4; ++g) {
                const float4 bb = *reinterpret_cast<const float4 *>(bias + n0 + 8 * g);
                for (int mt = 0; mt < 4; ++mt) {
                    acc[t][mt][4 * g + 0] = bb.x;
                    acc[t][mt][4 * g + 1] = bb.y;
                    acc[t][mt][4 * g + 2] = bb.z;
                    acc[t][mt][4 * g + 3] = bb.w;
                }
            }
        }
        // weight fragments of k-steps s and s + 1 in flight (two register
        // sets, the loop unrolled by two so that every index is static)
        bf16x8 wa0[NT], wa1[NT];
        auto wload = [&](int s, bf16x8 *dst) {
            for (int t = 0; t < NT; ++t)
                dst[t] = *reinterpret_cast<const bf16x8 *>(W + ((((int64_t)s * 16 + wave * NT + t) << 6) + lane) * 8);
        };
        auto kstep = [&](int s, bf16x8 *wa) {
            bf16x8 xb[4], cur[NT];
            for (int mt = 0; mt < 4; ++mt)
                xb[mt] = *reinterpret_cast<const bf16x8 *>(act + act_off(32 * mt + m32, 2 * s + h));
            for (int t = 0; t < NT; ++t) cur[t] = wa[t];
            if (s + 2 < steps) wload(s + 2, wa);
            for (int t = 0; t < NT; ++t)
                for (int mt = 0; mt < 4; ++mt)
                    acc[t][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur[t], xb[mt], acc[t][mt], 0, 0, 0);
        };
        wload(0, wa0);
        wload(1, wa1);
        for (int s = 0; s < steps; s += 2) {
            kstep(s, wa0);
            kstep(s + 1, wa1);
        }
        __syncthreads();   // every wave is done reading this layer's input
        for (int t = 0; t < NT; ++t)
            for (int mt = 0; mt < 4; ++mt) {
                const int m = 32 * mt + m32;
                for (int g = 0; g < 4; ++g) {
                    const int n0 = 32 * (wave * NT + t) + 8 * g + 4 * h;
                    uint16_t v[4];
                    for (int r = 0; r < 4; ++r) v[r] = f2bf(fmaxf(acc[t][mt][4 * g + r], 0.f));
                    uint2 u;
                    memcpy(&u, v, 8);
                    *reinterpret_cast<uint2 *>(act + act_off(m, n0 >> 3) + ((n0 & 7) << 1)) = u;
                }
            }
        __syncthreads();
    }
    // output layer: wave w -> rows [32 w, 32 w + 32) of the tile, all (padded) 32 columns
    for (int mt = wave; mt < 4; mt += NW) {
        f32x16 acc;
        for (int v = 0; v < 16; ++v) {
            const int n = 8 * (v >> 2) + 4 * h + (v & 3);
            acc[v] = n < nout ? b4[n] : 0.f;
        }
        for (int s = 0; s < kFusedHid / 16; ++s) {
            const bf16x8 w = *reinterpret_cast<const bf16x8 *>(wf4 + (((int64_t)s << 6) + lane) * 8);
            const bf16x8 x = *reinterpret_cast<const bf16x8 *>(act + act_off(32 * mt + m32, 2 * s + h));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, x, acc, 0, 0, 0);
        }
        const int64_t r = row0 + 32 * mt + m32;
        if (r < B)
            for (int v = 0; v < 16; ++v) {
                const int n = 8 * (v >> 2) + 4 * h + (v & 3);
                if (n < nout) z[r * nout + n] = acc[v];
            }
    }
}

// Wt[n][k] = W[k][n] for the three hidden-layer weight matrices.
__global__ void transpose_bf16_kernel(const bf16 *__restrict__ W, int K, int N, bf16 *__restrict__ Wt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)K * N) return;
    const int n = (int)(i / K), k = (int)(i - (int64_t)n * K);
    Wt[i] = W[(int64_t)k * N + n];
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------

// Syndrome bytes [B][n_in] -> BF16 0/1 features [B][kin] (zero padded).
__global__ void embed_kernel(const uint8_t *__restrict__ syn, int64_t B, int n_in, int kin, bf16 *__restrict__ X) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * kin) return;
    const int64_t b = i / kin;
    const int f = (int)(i - b * kin);
    X[i] = (f < n_in && (syn[b * n_in + f] & 1)) ? (bf16)0x3f80 : (bf16)0;
}

// Output layer for hid = 512, nout <= 4: each lane keeps W[8*lane .. 8*lane+7][0..nout)
// in registers and reads its 8 hidden values of a row with one 16-byte load;
// waves stride over the rows.  Each lane sums its 8 products, then an
// xor-shuffle tree adds the lanes (FP32 throughout).
template <int NOUT>
__global__ void __launch_bounds__(256) out_fwd512_kernel(const bf16 *__restrict__ H, int64_t B,
                                                         const bf16 *__restrict__ W, const float *__restrict__ bias,
                                                         float *__restrict__ z) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    float w[8][NOUT];
    for (int j = 0; j < 8; ++j)
        for (int c = 0; c < NOUT; ++c) w[j][c] = bf2f(W[(8 * lane + j) * NOUT + c]);
    for (int64_t b = wave; b < B; b += nwaves) {
        const uint4 u = *reinterpret_cast<const uint4 *>(H + b * 512 + 8 * lane);
        bf16 h[8];
        memcpy(h, &u, 16);
        float part[NOUT];
        for (int c = 0; c < NOUT; ++c) part[c] = 0.f;
        for (int j = 0; j < 8; ++j) {
            const float x = bf2f(h[j]);
            for (int c = 0; c < NOUT; ++c) part[c] += x * w[j][c];
        }
        for (int c = 0; c < NOUT; ++c) {
            float s = part[c];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
            if (lane == 0) z[b * NOUT + c] = s + bias[c];
        }
    }
}

// Output layer: z[b][c] = sum_k H[b][k] W[k][c] + bias[c]; one wave per row.
__global__ void out_fwd_kernel(const bf16 *__restrict__ H, int64_t B, int hid, const bf16 *__restrict__ W,
                               const float *__restrict__ bias, int nout, float *__restrict__ z) {
    const int64_t b = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= B) return;
    float part[16];
    for (int c = 0; c < nout; ++c) part[c] = 0.f;
    for (int k = lane; k < hid; k += 64) {
        const float h = bf2f(H[b * hid + k]);
        for (int c = 0; c < nout; ++c) part[c] += h * bf2f(W[k * nout + c]);
    }
    for (int c = 0; c < nout; ++c) {
        float s = part[c];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) z[b * nout + c] = s + bias[c];
    }
}

// Softmax + soft cross-entropy: dz = (softmax(z) - q) / B (FP32), loss summed
// per block into loss_parts.
__global__ void xent_kernel(const float *__restrict__ z, const float *__restrict__ q, int64_t B, int nout,
                            float *__restrict__ dz, float *__restrict__ loss_parts) {
    __shared__ float red[256];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (b < B) {
        float m = -INFINITY;
        for (int c = 0; c < nout; ++c) m = fmaxf(m, z[b * nout + c]);
        float s = 0.f;
        for (int c = 0; c < nout; ++c) s += expf(z[b * nout + c] - m);
        const float ls = logf(s);
        for (int c = 0; c < nout; ++c) {
            const float lp = z[b * nout + c] - m - ls;
            const float qc = q[b * nout + c];
            l -= qc * lp;
            dz[b * nout + c] = (expf(lp) - qc) / (float)B;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0];
}

// dA[b][k] = (sum_c dz[b][c] W[k][c]) masked by H[b][k] > 0, as BF16.
__global__ void out_bwd_dx_kernel(const float *__restrict__ dz, const bf16 *__restrict__ W, const bf16 *__restrict__ H,
                                  int64_t B, int hid, int nout, bf16 *__restrict__ dA) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * hid) return;
    const int64_t b = i / hid;
    const int k = (int)(i - b * hid);
    float s = 0.f;
    for (int c = 0; c < nout; ++c) s += dz[b * nout + c] * bf2f(W[k * nout + c]);
    dA[i] = bf2f(H[i]) > 0.f ? f2bf(s) : (bf16)0;
}

// dW[k][c] = sum_b H[b][k] dz[b][c] (one thread per (k, c), loop over b).
__global__ void out_bwd_dw_kernel(const bf16 *__restrict__ H, const float *__restrict__ dz, int64_t B, int hid,
                                  int nout, float *__restrict__ dW) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hid * nout) return;
    const int k = i / nout, c = i - k * nout;
    float s = 0.f;
    for (int64_t b = 0; b < B; ++b) s += bf2f(H[b * hid + k]) * dz[b * nout + c];
    dW[i] = s;
}

// Column sums: db[c] = sum_b D[b][c] (BF16 or FP32 input).
template <typename T>
__global__ void colsum_kernel(const T *__restrict__ D, int64_t B, int n, float *__restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0.f;
    for (int64_t b = 0; b < B; ++b) {
        if constexpr (sizeof(T) == 2) s += bf2f(D[b * n + c]);
        else s += D[b * n + c];
    }
    db[c] = s;
}

// Adam on the FP32 master parameters; refreshes the BF16 copy.
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, bf16 *__restrict__ wb, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(const float *__restrict__ w, bf16 *__restrict__ wb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

// Confidence gate: the class is accepted when the largest posterior is at
// least `threshold`; otherwise the syndrome id is appended to `reject`.
__global__ void gate_kernel(const float *__restrict__ z, int64_t B, int nout, float threshold,
                            int32_t *__restrict__ classes, float *__restrict__ conf, int32_t *__restrict__ reject,
                            int32_t *__restrict__ reject_count) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float m = -INFINITY;
    int bi = 0;
    for (int c = 0; c < nout; ++c) {
        const float x = z[b * nout + c];
        if (x > m) {
            m = x;
            bi = c;
        }
    }
    float s = 0.f;
    for (int c = 0; c < nout; ++c) s += expf(z[b * nout + c] - m);
    const float pmax = 1.f / s;
    if (conf) conf[b] = pmax;
    if (classes) classes[b] = bi;
    if (pmax < threshold && reject) {
        const int32_t slot = atomicAdd(reject_count, 1);
        reject[slot] = (int32_t)b;
    }
}

__global__ void zero_count_kernel(int32_t *p) { *p = 0; }

// Class posteriors from TN marginals: q[b][c] = mar[b][c] / sum_c mar[b][c].
__global__ void normalize_kernel(const double *__restrict__ mar, int64_t B, int nout, float *__restrict__ q) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0;
    for (int c = 0; c < nout; ++c) s += mar[b * nout + c];
    for (int c = 0; c < nout; ++c) q[b * nout + c] = s > 0 ? (float)(mar[b * nout + c] / s) : 1.f / nout;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct tqec_mlp {
    int n_in = 0, kin = 0, hid = 0, nout = 0;
    int64_t nparams = 0;
    int64_t off[8] = {0};          // W1 b1 W2 b2 W3 b3 W4 b4 in the flat parameter vector
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wb = nullptr;
    bf16 *wt[3] = {nullptr, nullptr, nullptr};   // transposed BF16 weights of the hidden layers (inference GEMM)
    bool wt_stale = true;
    // fused inference (hidden width 512): fragment-ordered BF16 weights of the
    // four layers, rebuilt after every weight change
    bf16 *wf[4] = {nullptr, nullptr, nullptr, nullptr};
    bool wf_stale = true;
    bool fused = true;             // forward / gate on the fused kernel where it applies
    int64_t zcap = 0;              // rows the logits buffer z holds
    bool fused_attr = false;
    int64_t step = 0;
    // activations / scratch for up to `cap` rows
    int64_t cap = 0;
    bf16 *X = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dA = nullptr, *dB = nullptr;
    float *z = nullptr, *dz = nullptr, *loss_parts = nullptr;
    int64_t loss_cap = 0;
};

namespace {

int64_t param_count(int kin, int hid, int nout, int64_t off[8]) {
    int64_t o = 0;
    const int64_t sizes[8] = {(int64_t)kin * hid, hid, (int64_t)hid * hid, hid, (int64_t)hid * hid, hid,
                              (int64_t)hid * nout, nout};
    for (int i = 0; i < 8; ++i) {
        off[i] = o;
        o += sizes[i];
    }
    return o;
}

int ensure_cap(tqec_mlp *p, int64_t B) {
    if (p->cap >= B) return TQEC_OK;
    bf16 **bufs[6] = {&p->X, &p->H1, &p->H2, &p->H3, &p->dA, &p->dB};
    for (auto b : bufs)
        if (*b) (void)hipFree(*b);
    if (p->z) (void)hipFree(p->z);
    if (p->dz) (void)hipFree(p->dz);
    p->z = nullptr;
    p->zcap = 0;
    p->cap = 0;
    TQM_HIP(hipMalloc((void **)&p->X, (size_t)B * p->kin * sizeof(bf16)));
    for (int i = 1; i < 6; ++i) TQM_HIP(hipMalloc((void **)bufs[i], (size_t)B * p->hid * sizeof(bf16)));
    TQM_HIP(hipMalloc((void **)&p->z, (size_t)B * p->nout * sizeof(float)));
    p->zcap = B;
    TQM_HIP(hipMalloc((void **)&p->dz, (size_t)B * p->nout * sizeof(float)));
    const int64_t nb = (B + 255) / 256;
    if (p->loss_cap < nb) {
        if (p->loss_parts) (void)hipFree(p->loss_parts);
        TQM_HIP(hipMalloc((void **)&p->loss_parts, (size_t)nb * sizeof(float)));
        p->loss_cap = nb;
    }
    p->cap = B;
    return TQEC_OK;
}

template <bool AT, bool BT, int EPI>
void gemm(const GemmArgs &a, hipStream_t st) {
    dim3 grid((a.N + BN - 1) / BN, (a.M + BM - 1) / BM);
    hipLaunchKernelGGL((gemm_bf16_kernel<AT, BT, EPI>), grid, dim3(256), 0, st, a);
}

bool fused_ok(const tqec_mlp *p) { return p->hid == kFusedHid && p->nout <= 32 && p->kin <= kFusedHid; }

// Logits only (no activations kept): the fused kernel, one workgroup per 128 rows.
int forward_fused(tqec_mlp *p, const uint8_t *d_syn, int64_t B, hipStream_t st) {
    if (p->zcap < B) {
        // logits only; the activation buffers are allocated for training
        if (p->z) (void)hipFree(p->z);
        p->z = nullptr;
        p->zcap = 0;
        TQM_HIP(hipMalloc((void **)&p->z, (size_t)B * p->nout * sizeof(float)));
        p->zcap = B;
    }
    const size_t lds = (size_t)kFusedRows * kFusedHid * 2;
    if (!p->fused_attr) {
        TQM_HIP(hipFuncSetAttribute((const void *)mlp_fused_fwd_kernel<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        p->fused_attr = true;
    }
    if (p->wf_stale || !p->wf[0]) {
        const int kdim[4] = {p->kin, p->hid, p->hid, p->hid};
        for (int l = 0; l < 4; ++l) {
            const int npad = l < 3 ? p->hid : 32;
            if (!p->wf[l]) TQM_HIP(hipMalloc((void **)&p->wf[l], (size_t)kdim[l] * npad * sizeof(bf16)));
            const int64_t n = (int64_t)kdim[l] * npad;
            hipLaunchKernelGGL(pack_fused_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                               p->wb + p->off[2 * l], kdim[l], l < 3 ? p->hid : p->nout, npad, p->wf[l]);
        }
        p->wf_stale = false;
    }
    const unsigned grid = (unsigned)((B + kFusedRows - 1) / kFusedRows);
    const float *b[4] = {p->w + p->off[1], p->w + p->off[3], p->w + p->off[5], p->w + p->off[7]};
    hipLaunchKernelGGL(mlp_fused_fwd_kernel<8>, dim3(grid), dim3(512), lds, st, d_syn, B, p->n_in, p->kin, p->wf[0],
                       p->wf[1], p->wf[2], b[0], b[1], b[2], p->wf[3], b[3], p->nout, p->z);
    TQM_HIP(hipGetLastError());
    return TQEC_OK;
}

int forward(tqec_mlp *p, const uint8_t *d_syn, int64_t B, hipStream_t st, bool keep = true) {
    int rc;
    if (!keep && p->fused && fused_ok(p)) return forward_fused(p, d_syn, B, st);
    if ((rc = ensure_cap(p, B))) return rc;
    const int64_t nx = B * p->kin;
    hipLaunchKernelGGL(embed_kernel, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, d_syn, B, p->n_in, p->kin,
                       p->X);
    const bf16 *in[3] = {p->X, p->H1, p->H2};
    bf16 *out[3] = {p->H1, p->H2, p->H3};
    const int kdim[3] = {p->kin, p->hid, p->hid};
    if (p->wt_stale) {
        for (int l = 0; l < 3; ++l) {
            const int64_t n = (int64_t)kdim[l] * p->hid;
            hipLaunchKernelGGL(transpose_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                               p->wb + p->off[2 * l], kdim[l], p->hid, p->wt[l]);
        }
        p->wt_stale = false;
    }
    for (int l = 0; l < 3; ++l) {
        // 64 x 128 wave tiles (NW = 4) measured slower at d=7 sizes: 252 VGPRs leave one wave per SIMD
        const bool wide = false;
        const dim3 grid((unsigned)(p->hid / (wide ? 256 : 128)), (unsigned)((B + 127) / 128));
        const float *bias = p->w + p->off[2 * l + 1];
        if (kdim[l] % 64 == 0) {
            if (wide)
                hipLaunchKernelGGL((gemm_fwd_kernel<64, 4>), grid, dim3(256), 0, st, in[l], p->wt[l], (int)B, p->hid,
                                   kdim[l], bias, out[l]);
            else
                hipLaunchKernelGGL((gemm_fwd_kernel<64, 2>), grid, dim3(256), 0, st, in[l], p->wt[l], (int)B, p->hid,
                                   kdim[l], bias, out[l]);
        } else {
            if (wide)
                hipLaunchKernelGGL((gemm_fwd_kernel<32, 4>), grid, dim3(256), 0, st, in[l], p->wt[l], (int)B, p->hid,
                                   kdim[l], bias, out[l]);
            else
                hipLaunchKernelGGL((gemm_fwd_kernel<32, 2>), grid, dim3(256), 0, st, in[l], p->wt[l], (int)B, p->hid,
                                   kdim[l], bias, out[l]);
        }
    }
    const unsigned rows_per_block = 4;
    if (p->hid == 512 && p->nout <= 4) {
        const unsigned blocks = (unsigned)std::min<int64_t>((B + 3) / 4, 4096);
        if (p->nout == 4)
            hipLaunchKernelGGL(out_fwd512_kernel<4>, dim3(blocks), dim3(256), 0, st, p->H3, B, p->wb + p->off[6],
                               p->w + p->off[7], p->z);
        else if (p->nout == 2)
            hipLaunchKernelGGL(out_fwd512_kernel<2>, dim3(blocks), dim3(256), 0, st, p->H3, B, p->wb + p->off[6],
                               p->w + p->off[7], p->z);
        else
            hipLaunchKernelGGL(out_fwd_kernel, dim3((unsigned)((B + rows_per_block - 1) / rows_per_block)), dim3(256),
                               0, st, p->H3, B, p->hid, p->wb + p->off[6], p->w + p->off[7], p->nout, p->z);
    } else {
        hipLaunchKernelGGL(out_fwd_kernel, dim3((unsigned)((B + rows_per_block - 1) / rows_per_block)), dim3(256), 0,
                           st, p->H3, B, p->hid, p->wb + p->off[6], p->w + p->off[7], p->nout, p->z);
    }
    TQM_HIP(hipGetLastError());
    return TQEC_OK;
}

// Gradients of the mean soft cross-entropy into p->g (FP32, flat layout).
int backward(tqec_mlp *p, const float *d_q, int64_t B, hipStream_t st) {
    const int64_t nb = (B + 255) / 256;
    hipLaunchKernelGGL(xent_kernel, dim3((unsigned)nb), dim3(256), 0, st, p->z, d_q, B, p->nout, p->dz, p->loss_parts);
    // output layer
    hipLaunchKernelGGL(out_bwd_dw_kernel, dim3((unsigned)((p->hid * p->nout + 255) / 256)), dim3(256), 0, st, p->H3,
                       p->dz, B, p->hid, p->nout, p->g + p->off[6]);
    hipLaunchKernelGGL((colsum_kernel<float>), dim3(1), dim3(64), 0, st, p->dz, B, p->nout, p->g + p->off[7]);
    const int64_t nh = B * p->hid;
    hipLaunchKernelGGL(out_bwd_dx_kernel, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, p->dz,
                       p->wb + p->off[6], p->H3, B, p->hid, p->nout, p->dA);
    // hidden layers 3, 2, 1: dA holds dL/d(pre-activation) of layer l
    const bf16 *hin[3] = {p->X, p->H1, p->H2};
    const int kdim[3] = {p->kin, p->hid, p->hid};
    bf16 *cur = p->dA, *nxt = p->dB;
    for (int l = 2; l >= 0; --l) {
        // dW_l[K][N] = in_l^T [K][B] * dA [B][N]
        GemmArgs w{};
        w.A = hin[l];
        w.B = cur;
        w.M = kdim[l];
        w.N = p->hid;
        w.K = (int)B;
        w.lda = kdim[l];
        w.ldb = p->hid;
        w.ldc = p->hid;
        w.outf = p->g + p->off[2 * l];
        gemm<true, false, kEpiF32>(w, st);
        hipLaunchKernelGGL((colsum_kernel<bf16>), dim3((unsigned)((p->hid + 255) / 256)), dim3(256), 0, st, cur, B,
                           p->hid, p->g + p->off[2 * l + 1]);
        if (l > 0) {
            // d in_l [B][K] = dA [B][N] * W_l^T (W_l stored [K][N]), masked by in_l > 0
            GemmArgs x{};
            x.A = cur;
            x.B = p->wb + p->off[2 * l];
            x.M = (int)B;
            x.N = kdim[l];
            x.K = p->hid;
            x.lda = p->hid;
            x.ldb = p->hid;
            x.ldc = kdim[l];
            x.mask = hin[l];
            x.outb = nxt;
            gemm<false, true, kEpiMaskBf16>(x, st);
            std::swap(cur, nxt);
        }
    }
    TQM_HIP(hipGetLastError());
    return TQEC_OK;
}

// Training reduces over the batch inside the weight-gradient GEMM, whose K
// loop needs a multiple of 32 rows; inference takes any batch size.
int check_batch(tqec_mlp *p, int64_t B, bool training) {
    if (!p) return fail(TQEC_ENULL, "null network");
    if (B <= 0 || B > (int64_t(1) << 30)) return fail(TQEC_EINVAL, "batch size out of range");
    if (training && B % 32 != 0) return fail(TQEC_EINVAL, "training batch size must be a multiple of 32");
    return TQEC_OK;
}

}  // namespace

extern "C" {

int tqec_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, tqec_mlp **out) {
    if (!out) return fail(TQEC_ENULL, "null output handle");
    *out = nullptr;
    if (n_in <= 0 || n_in > 4096 || hidden <= 0 || hidden % 128 != 0 || n_out <= 0 || n_out > 16)
        return fail(TQEC_EINVAL, "n_in in [1, 4096], hidden a multiple of 128, n_out in [1, 16]");
    auto *p = new (std::nothrow) tqec_mlp();
    if (!p) return fail(TQEC_ENOMEM, "out of host memory");
    p->n_in = n_in;
    p->kin = (n_in + 31) / 32 * 32;
    p->hid = hidden;
    p->nout = n_out;
    p->nparams = param_count(p->kin, hidden, n_out, p->off);
    const size_t nb = (size_t)p->nparams;
    hipError_t e = hipMalloc((void **)&p->w, nb * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&p->g, nb * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&p->m, nb * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&p->v, nb * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&p->wb, nb * 2);
    for (int l = 0; l < 3 && e == hipSuccess; ++l)
        e = hipMalloc((void **)&p->wt[l], (size_t)(l == 0 ? p->kin : p->hid) * p->hid * 2);
    if (e == hipSuccess) e = hipMemset(p->w, 0, nb * 4);
    if (e == hipSuccess) e = hipMemset(p->m, 0, nb * 4);
    if (e == hipSuccess) e = hipMemset(p->v, 0, nb * 4);
    if (e == hipSuccess) e = hipMemset(p->wb, 0, nb * 2);
    if (e != hipSuccess) {
        tqec_mlp_destroy(p);
        return fail(TQEC_EDEVICE, hipGetErrorString(e));
    }
    *out = p;
    return TQEC_OK;
}

int64_t tqec_mlp_num_params(const tqec_mlp *p) { return p ? p->nparams : -1; }

int tqec_mlp_set_params(tqec_mlp *p, const float *host) {
    if (!p || !host) return fail(TQEC_ENULL, "null argument");
    // padded input rows of W1 (k >= n_in) must stay zero
    std::vector<float> w(host, host + p->nparams);
    for (int k = p->n_in; k < p->kin; ++k)
        for (int c = 0; c < p->hid; ++c) w[p->off[0] + (int64_t)k * p->hid + c] = 0.f;
    TQM_HIP(hipMemcpy(p->w, w.data(), (size_t)p->nparams * 4, hipMemcpyHostToDevice));
    TQM_HIP(hipMemset(p->m, 0, (size_t)p->nparams * 4));
    TQM_HIP(hipMemset(p->v, 0, (size_t)p->nparams * 4));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((p->nparams + 255) / 256)), dim3(256), 0, 0, p->w, p->wb,
                       p->nparams);
    TQM_HIP(hipDeviceSynchronize());
    p->step = 0;
    p->wt_stale = true;
    p->wf_stale = true;
    return TQEC_OK;
}

int tqec_mlp_get_params(tqec_mlp *p, float *host) {
    if (!p || !host) return fail(TQEC_ENULL, "null argument");
    TQM_HIP(hipDeviceSynchronize());
    TQM_HIP(hipMemcpy(host, p->w, (size_t)p->nparams * 4, hipMemcpyDeviceToHost));
    return TQEC_OK;
}

int tqec_mlp_forward(tqec_mlp *p, const uint8_t *d_syn, int64_t B, float *d_logits, void *stream) {
    int rc;
    if ((rc = check_batch(p, B, false))) return rc;
    if (!d_syn || !d_logits) return fail(TQEC_ENULL, "null device pointer");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = forward(p, d_syn, B, st, false))) return rc;
    TQM_HIP(hipMemcpyAsync(d_logits, p->z, (size_t)B * p->nout * 4, hipMemcpyDeviceToDevice, st));
    return TQEC_OK;
}

int tqec_mlp_gradients(tqec_mlp *p, const uint8_t *d_syn, const float *d_targets, int64_t B, float *d_grad,
                       float *h_loss, void *stream) {
    int rc;
    if ((rc = check_batch(p, B, true))) return rc;
    if (!d_syn || !d_targets) return fail(TQEC_ENULL, "null device pointer");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = forward(p, d_syn, B, st))) return rc;
    if ((rc = backward(p, d_targets, B, st))) return rc;
    if (d_grad) TQM_HIP(hipMemcpyAsync(d_grad, p->g, (size_t)p->nparams * 4, hipMemcpyDeviceToDevice, st));
    if (h_loss) {
        const int64_t nb = (B + 255) / 256;
        std::vector<float> parts(nb);
        TQM_HIP(hipMemcpyAsync(parts.data(), p->loss_parts, nb * 4, hipMemcpyDeviceToHost, st));
        TQM_HIP(hipStreamSynchronize(st));
        double s = 0;
        for (float x : parts) s += x;
        *h_loss = (float)(s / (double)B);
    }
    return TQEC_OK;
}

int tqec_mlp_train_step(tqec_mlp *p, const uint8_t *d_syn, const float *d_targets, int64_t B, float lr,
                        float *h_loss, void *stream) {
    int rc;
    if ((rc = tqec_mlp_gradients(p, d_syn, d_targets, B, nullptr, h_loss, stream))) return rc;
    hipStream_t st = (hipStream_t)stream;
    p->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((p->nparams + 255) / 256)), dim3(256), 0, st, p->w, p->g, p->m,
                       p->v, p->wb, p->nparams, lr, b1, b2, 1e-8f, c1, c2);
    p->wt_stale = true;
    p->wf_stale = true;
    // padded rows of W1 get zero gradient (zero inputs) and so stay zero
    TQM_HIP(hipGetLastError());
    return TQEC_OK;
}

int tqec_mlp_gate(tqec_mlp *p, const uint8_t *d_syn, int64_t B, float threshold, int32_t *d_classes,
                  float *d_confidence, int32_t *d_reject, int32_t *d_reject_count, void *stream) {
    int rc;
    if ((rc = check_batch(p, B, false))) return rc;
    if (!d_syn || (d_reject && !d_reject_count)) return fail(TQEC_ENULL, "null device pointer");
    hipStream_t st = (hipStream_t)stream;
    if ((rc = forward(p, d_syn, B, st, false))) return rc;
    if (d_reject_count) hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st, d_reject_count);
    hipLaunchKernelGGL(gate_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, p->z, B, p->nout, threshold,
                       d_classes, d_confidence, d_reject, d_reject_count);
    TQM_HIP(hipGetLastError());
    return TQEC_OK;
}

int tqec_posteriors_from_marginals(const double *d_marginals, int64_t n, int32_t n_classes, float *d_q,
                                   void *stream) {
    if (n < 0 || n_classes <= 0) return fail(TQEC_EINVAL, "bad sizes");
    if (n > 0 && (!d_marginals || !d_q)) return fail(TQEC_ENULL, "null device pointer");
    if (n == 0) return TQEC_OK;
    hipLaunchKernelGGL(normalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_marginals, n, n_classes, d_q);
    TQM_HIP(hipGetLastError());
    return TQEC_OK;
}

int tqec_mlp_set_inference(tqec_mlp *p, int32_t kernel) {
    if (!p) return fail(TQEC_ENULL, "null network");
    if (kernel != 0 && kernel != 1) return fail(TQEC_EINVAL, "inference kernel must be 0 (layer by layer) or 1 (fused)");
    if (kernel == 1 && !fused_ok(p))
        return fail(TQEC_EUNSUPPORTED, "the fused kernel needs hidden width 512, n_out <= 32 and n_in <= 512");
    p->fused = kernel == 1;
    return TQEC_OK;
}

int tqec_mlp_inference(const tqec_mlp *p, int32_t *kernel) {
    if (!p || !kernel) return fail(TQEC_ENULL, "null argument");
    *kernel = (p->fused && fused_ok(p)) ? 1 : 0;
    return TQEC_OK;
}

void tqec_mlp_destroy(tqec_mlp *p) {
    if (!p) return;
    void *bufs[] = {p->w,  p->g,  p->m,  p->v,  p->wb, p->X,  p->H1,         p->H2,    p->H3,
                    p->dA, p->dB, p->z,  p->dz, p->loss_parts, p->wt[0], p->wt[1], p->wt[2],
                    p->wf[0], p->wf[1], p->wf[2], p->wf[3]};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

}  // extern "C"
