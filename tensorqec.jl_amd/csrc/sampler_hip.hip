// sampler_hip.hip — on-device syndrome sampling and error-pattern recovery.
//
// Sampler: replaces the reference's per-sample loop of
// `random_error_pattern(em)` (src/decoding/error_model.jl:69-71, 97-117)
// followed by `syndrome_extraction` (:131-146).  Draws come from
// Philox4x32-10 with counter (column j, sample mod 2^32, sample >> 32,
// stream tag) and key (seed mod 2^32, seed >> 32), converted to a double with
// 53 random bits, so every sample is bit-identical to the host sampler.
// Each thread produces one sample: it XORs the precomputed syndrome/logical
// column of every flipped variable into a register bitset.
//
// Recovery: replaces `error_pattern` (src/decoding/tndecoder.jl:167-174) and
// its MIP pure-error solve (src/decoding/ipdecoder.jl:150-169) by a GF(2)
// right inverse: e = R s, then the reference's logical fix per class bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tqec_hip.h"

namespace tqec {
int set_error_shared(int code, const std::string &msg);  // tqec_hip.hip: backs tqec_last_error()
}

namespace {

int fail(int code, const std::string &m) { return tqec::set_error_shared(code, m); }

constexpr int kMaxSynWords = 8;   // up to 512 syndrome bits
constexpr int kMaxErrWords = 16;  // up to 1024 error variables per pattern half

__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double uniform53(uint32_t x0, uint32_t x1) {
    const double a = (double)(x0 >> 5), b = (double)(x1 >> 6);
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

struct SamplerArgs {
    int kind;            // 0: CSS depolarizing, 1: independent flips
    int nvar;            // qubits (kind 0) or mechanisms (kind 1)
    int nsyn;            // syndrome bits
    int nlog;            // logical class bits
    int synw;            // 64-bit words per syndrome column
    const double *t1, *t2, *t3;   // kind 0: py, px+py, px+py+pz; kind 1: t1 = p
    // kind 0: per qubit, the syndrome column of an X error and of a Z error
    // and the logical-class column of each; kind 1: column of the mechanism
    const unsigned long long *colx, *colz;   // [nvar][synw]
    const uint32_t *logx, *logz;             // [nvar]
    uint32_t seed_lo, seed_hi, stream_tag;
    int64_t first, n;
    uint8_t *syn;        // [n][nsyn]
    int32_t *cls;        // [n] or null
    uint8_t *err;        // kind 0: [n][2 nvar] (x bits then z bits); kind 1: [n][nvar]; or null
};

__global__ void sample_kernel(SamplerArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t s = a.first + i;
    unsigned long long acc[kMaxSynWords];
    for (int w = 0; w < kMaxSynWords; ++w) acc[w] = 0ull;
    uint32_t lg = 0;
    for (int j = 0; j < a.nvar; ++j) {
        uint32_t c[4] = {(uint32_t)j, (uint32_t)(s & 0xffffffffll), (uint32_t)((unsigned long long)s >> 32),
                         a.stream_tag};
        philox10(c, a.seed_lo, a.seed_hi);
        const double u = uniform53(c[0], c[1]);
        bool fx, fz;
        if (a.kind == 0) {
            const bool y = u < a.t1[j];
            const bool x = !y && u < a.t2[j];
            const bool z = !y && !x && u < a.t3[j];
            fx = y || x;
            fz = y || z;
        } else {
            fx = u < a.t1[j];
            fz = false;
        }
        if (fx) {
            for (int w = 0; w < a.synw; ++w) acc[w] ^= a.colx[(int64_t)j * a.synw + w];
            lg ^= a.logx[j];
        }
        if (fz) {
            for (int w = 0; w < a.synw; ++w) acc[w] ^= a.colz[(int64_t)j * a.synw + w];
            lg ^= a.logz[j];
        }
        if (a.err) {
            if (a.kind == 0) {
                a.err[i * 2 * a.nvar + j] = fx;
                a.err[i * 2 * a.nvar + a.nvar + j] = fz;
            } else {
                a.err[i * a.nvar + j] = fx;
            }
        }
    }
    uint8_t *row = a.syn + i * a.nsyn;
    for (int b = 0; b < a.nsyn; ++b) row[b] = (uint8_t)((acc[b >> 6] >> (b & 63)) & 1ull);
    if (a.cls) a.cls[i] = (int32_t)lg;
}

struct RecoveryArgs {
    int nq;              // error variables per half
    int nsa, nsb;        // syndrome bits feeding half a (ex <- sz) and half b (ez <- sx)
    int ew;              // 64-bit words per error pattern half
    int k;               // logical pairs (CSS) or logicals (DEM)
    int kind;            // 0: CSS; 1: DEM with the reference's logical fix
    int nsyn;            // bytes per syndrome row
    int offa, offb;      // first syndrome bit of each half's syndrome
    const unsigned long long *Ra, *Rb;   // [nsa][ew], [nsb][ew]: columns of the right inverses
    const unsigned long long *la, *lb;   // [k][ew]: CSS: lz (test ex), lx (test ez); DEM: l2q masks
    const unsigned long long *fa, *fb;   // [k][ew]: CSS: lx (added to ex), lz (added to ez)
    const uint8_t *syn;
    const int32_t *cls;
    int64_t n;
    uint8_t *ea, *eb;    // [n][nq] each (eb unused for DEM)
};

__global__ void recovery_kernel(RecoveryArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint8_t *row = a.syn + i * a.nsyn;
    unsigned long long ea[kMaxErrWords], eb[kMaxErrWords];
    for (int w = 0; w < kMaxErrWords; ++w) ea[w] = eb[w] = 0ull;
    for (int j = 0; j < a.nsa; ++j)
        if (row[a.offa + j] & 1)
            for (int w = 0; w < a.ew; ++w) ea[w] ^= a.Ra[(int64_t)j * a.ew + w];
    for (int j = 0; j < a.nsb; ++j)
        if (row[a.offb + j] & 1)
            for (int w = 0; w < a.ew; ++w) eb[w] ^= a.Rb[(int64_t)j * a.ew + w];
    const int32_t c = a.cls[i];
    for (int l = 0; l < a.k; ++l) {
        if (a.kind == 0) {
            // tndecoder.jl:170: lz_l . ex must equal class bit k+l, else ex += lx_l
            int par = 0;
            for (int w = 0; w < a.ew; ++w) par ^= __popcll(ea[w] & a.la[l * a.ew + w]) & 1;
            if (par != ((c >> (a.k + l)) & 1))
                for (int w = 0; w < a.ew; ++w) ea[w] ^= a.fa[l * a.ew + w];
            // tndecoder.jl:171: lx_l . ez must equal class bit l, else ez += lz_l
            par = 0;
            for (int w = 0; w < a.ew; ++w) par ^= __popcll(eb[w] & a.lb[l * a.ew + w]) & 1;
            if (par != ((c >> l) & 1))
                for (int w = 0; w < a.ew; ++w) eb[w] ^= a.fb[l * a.ew + w];
        } else {
            // tndecoder.jl:263-269: when the parity of the mechanisms touching
            // logical l disagrees with the class, every one of them is flipped
            int par = 0;
            for (int w = 0; w < a.ew; ++w) par ^= __popcll(ea[w] & a.la[l * a.ew + w]) & 1;
            if (par != ((c >> l) & 1))
                for (int w = 0; w < a.ew; ++w) ea[w] ^= a.la[l * a.ew + w];
        }
    }
    for (int q = 0; q < a.nq; ++q) a.ea[i * a.nq + q] = (uint8_t)((ea[q >> 6] >> (q & 63)) & 1ull);
    if (a.kind == 0 && a.eb)
        for (int q = 0; q < a.nq; ++q) a.eb[i * a.nq + q] = (uint8_t)((eb[q >> 6] >> (q & 63)) & 1ull);
}

// Logical-error check of a batch: per sample, the difference d between the
// sampled and the recovered error pattern; its logical class bits (the XOR of
// the class columns of the variables where d is set) are `check_logical_error`
// (src/decoding/error_model.jl:161-181) of every logical row at once, and its
// syndrome (XOR of the syndrome columns) must be zero when the recovered
// pattern reproduces the sampled syndrome.  counts: [0] samples with a
// logical error, [1] with one in the X-error half (lz . dx odd for some row;
// DEM: 0), [2] in the Z-error half (lx . dz odd), [3] recovered patterns that
// do not reproduce the syndrome.
struct CheckArgs {
    int kind, nvar, synw, k;
    const unsigned long long *colx, *colz;
    const uint32_t *logx, *logz;
    const uint8_t *err;        // sampled: kind 0 [n][2 nvar] (x, z); kind 1 [n][nvar]
    const uint8_t *ea, *eb;    // recovered: ex / ez (kind 0) or e (kind 1), [n][nvar]
    int64_t n;
    unsigned long long *counts;
};

__global__ void ler_check_kernel(CheckArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool lerr = false, xerr = false, zerr = false, bad = false;
    if (i < a.n) {
        unsigned long long syn[kMaxSynWords];
        for (int w = 0; w < kMaxSynWords; ++w) syn[w] = 0ull;
        uint32_t lg = 0;
        const int stride = a.kind == 0 ? 2 * a.nvar : a.nvar;
        const uint8_t *e = a.err + i * stride;
        for (int j = 0; j < a.nvar; ++j) {
            if ((e[j] ^ a.ea[i * a.nvar + j]) & 1) {
                for (int w = 0; w < a.synw; ++w) syn[w] ^= a.colx[(int64_t)j * a.synw + w];
                lg ^= a.logx[j];
            }
            if (a.kind == 0 && ((e[a.nvar + j] ^ a.eb[i * a.nvar + j]) & 1)) {
                for (int w = 0; w < a.synw; ++w) syn[w] ^= a.colz[(int64_t)j * a.synw + w];
                lg ^= a.logz[j];
            }
        }
        unsigned long long any = 0ull;
        for (int w = 0; w < a.synw; ++w) any |= syn[w];
        bad = any != 0ull;
        lerr = lg != 0u;
        if (a.kind == 0) {
            zerr = (lg & ((1u << a.k) - 1u)) != 0u;
            xerr = (lg >> a.k) != 0u;
        }
    }
    // one atomic per wave and counter
    const unsigned long long bl = __ballot(lerr), bx = __ballot(xerr), bz = __ballot(zerr), bb = __ballot(bad);
    if ((threadIdx.x & 63) == 0) {
        if (bl) atomicAdd(&a.counts[0], (unsigned long long)__popcll(bl));
        if (bx) atomicAdd(&a.counts[1], (unsigned long long)__popcll(bx));
        if (bz) atomicAdd(&a.counts[2], (unsigned long long)__popcll(bz));
        if (bb) atomicAdd(&a.counts[3], (unsigned long long)__popcll(bb));
    }
}

__global__ void ler_zero_kernel(unsigned long long *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0ull;
}

#define TQS_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return fail(TQEC_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int upload(const std::vector<T> &h, T **d) {
    TQS_HIP(hipMalloc((void **)d, std::max<size_t>(1, h.size()) * sizeof(T)));
    if (!h.empty()) TQS_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return TQEC_OK;
}

std::vector<unsigned long long> pack_rows(const uint8_t *M, int rows, int cols, int words) {
    // M is rows x cols (row-major); returns [rows][words] bit-packed rows
    std::vector<unsigned long long> out((size_t)rows * words, 0ull);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c)
            if (M[(size_t)r * cols + c] & 1) out[(size_t)r * words + (c >> 6)] |= 1ull << (c & 63);
    return out;
}

}  // namespace

struct tqec_sampler {
    int kind = 0, nvar = 0, nsyn = 0, nlog = 0, synw = 0;
    uint32_t stream_tag = 0;
    double *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
    unsigned long long *colx = nullptr, *colz = nullptr;
    uint32_t *logx = nullptr, *logz = nullptr;
};

struct tqec_recovery {
    RecoveryArgs a{};
    unsigned long long *Ra = nullptr, *Rb = nullptr, *la = nullptr, *lb = nullptr, *fa = nullptr, *fb = nullptr;
};

struct tqec_ler {
    tqec_plan *plan = nullptr;
    tqec_sampler *smp = nullptr;
    tqec_recovery *rec = nullptr;
    int64_t batch = 0;
    int nout = 0;
    uint8_t *syn = nullptr, *err = nullptr, *ea = nullptr, *eb = nullptr, *tie = nullptr;
    int32_t *cls = nullptr;
    double *mar = nullptr;
    unsigned long long *counts = nullptr;
};

extern "C" {

int tqec_sampler_create_css(int32_t nq, const double *px, const double *py, const double *pz, int32_t nsx,
                            const uint8_t *Hx, int32_t nsz, const uint8_t *Hz, int32_t k, const uint8_t *lx,
                            const uint8_t *lz, tqec_sampler **out) {
    if (!out || !px || !py || !pz || (nsx && !Hx) || (nsz && !Hz) || (k && (!lx || !lz)))
        return fail(TQEC_ENULL, "null argument");
    *out = nullptr;
    const int ns = nsx + nsz;
    if (nq <= 0 || ns <= 0 || ns > 64 * kMaxSynWords || k < 0 || 2 * k > 31)
        return fail(TQEC_EUNSUPPORTED, "sampler sizes out of range");
    for (int j = 0; j < nq; ++j)
        if (!(px[j] >= 0 && py[j] >= 0 && pz[j] >= 0 && px[j] + py[j] + pz[j] <= 1))
            return fail(TQEC_EINVAL, "probabilities must be non-negative with px+py+pz <= 1");
    auto *s = new (std::nothrow) tqec_sampler();
    if (!s) return fail(TQEC_ENOMEM, "out of host memory");
    s->kind = 0;
    s->nvar = nq;
    s->nsyn = ns;
    s->nlog = 2 * k;
    s->synw = (ns + 63) / 64;
    s->stream_tag = 0x51D0;
    std::vector<double> t1(nq), t2(nq), t3(nq);
    for (int j = 0; j < nq; ++j) {
        t1[j] = py[j];
        t2[j] = px[j] + py[j];
        t3[j] = px[j] + py[j] + pz[j];
    }
    // An X error on qubit j flips the Z checks holding j (sz = Hz ex) and the
    // class bit k+i when lz_i holds j; a Z error flips X checks and bit i.
    std::vector<unsigned long long> cx((size_t)nq * s->synw, 0), cz((size_t)nq * s->synw, 0);
    std::vector<uint32_t> lgx(nq, 0), lgz(nq, 0);
    for (int j = 0; j < nq; ++j) {
        for (int r = 0; r < nsx; ++r)
            if (Hx[(size_t)r * nq + j] & 1) cz[(size_t)j * s->synw + (r >> 6)] |= 1ull << (r & 63);
        for (int r = 0; r < nsz; ++r) {
            const int b = nsx + r;
            if (Hz[(size_t)r * nq + j] & 1) cx[(size_t)j * s->synw + (b >> 6)] |= 1ull << (b & 63);
        }
        for (int i = 0; i < k; ++i) {
            if (lz[(size_t)i * nq + j] & 1) lgx[j] |= 1u << (k + i);
            if (lx[(size_t)i * nq + j] & 1) lgz[j] |= 1u << i;
        }
    }
    int rc;
    if ((rc = upload(t1, &s->t1)) || (rc = upload(t2, &s->t2)) || (rc = upload(t3, &s->t3)) ||
        (rc = upload(cx, &s->colx)) || (rc = upload(cz, &s->colz)) || (rc = upload(lgx, &s->logx)) ||
        (rc = upload(lgz, &s->logz))) {
        tqec_sampler_destroy(s);
        return rc;
    }
    *out = s;
    return TQEC_OK;
}

int tqec_sampler_create_flips(int32_t nm, const double *p, int32_t nd, const uint8_t *D, int32_t nl,
                              const uint8_t *Lm, tqec_sampler **out) {
    if (!out || !p || (nd && !D) || (nl && !Lm)) return fail(TQEC_ENULL, "null argument");
    *out = nullptr;
    if (nm <= 0 || nd <= 0 || nd > 64 * kMaxSynWords || nl < 0 || nl > 31)
        return fail(TQEC_EUNSUPPORTED, "sampler sizes out of range");
    for (int j = 0; j < nm; ++j)
        if (!(p[j] >= 0 && p[j] <= 1)) return fail(TQEC_EINVAL, "probabilities must be in [0, 1]");
    auto *s = new (std::nothrow) tqec_sampler();
    if (!s) return fail(TQEC_ENOMEM, "out of host memory");
    s->kind = 1;
    s->nvar = nm;
    s->nsyn = nd;
    s->nlog = nl;
    s->synw = (nd + 63) / 64;
    s->stream_tag = 0xF11B;
    std::vector<double> t1(p, p + nm);
    std::vector<unsigned long long> cx((size_t)nm * s->synw, 0);
    std::vector<uint32_t> lg(nm, 0);
    for (int j = 0; j < nm; ++j) {
        for (int r = 0; r < nd; ++r)
            if (D[(size_t)r * nm + j] & 1) cx[(size_t)j * s->synw + (r >> 6)] |= 1ull << (r & 63);
        for (int l = 0; l < nl; ++l)
            if (Lm[(size_t)l * nm + j] & 1) lg[j] |= 1u << l;
    }
    int rc;
    if ((rc = upload(t1, &s->t1)) || (rc = upload(cx, &s->colx)) || (rc = upload(lg, &s->logx))) {
        tqec_sampler_destroy(s);
        return rc;
    }
    *out = s;
    return TQEC_OK;
}

int tqec_sample(tqec_sampler *s, uint64_t seed, int64_t first, int64_t n, uint8_t *d_syndromes,
                int32_t *d_classes, uint8_t *d_errors, void *stream) {
    if (!s || (n > 0 && !d_syndromes)) return fail(TQEC_ENULL, "null argument");
    if (n < 0 || first < 0) return fail(TQEC_EINVAL, "negative count or offset");
    if (n == 0) return TQEC_OK;
    SamplerArgs a{};
    a.kind = s->kind;
    a.nvar = s->nvar;
    a.nsyn = s->nsyn;
    a.nlog = s->nlog;
    a.synw = s->synw;
    a.t1 = s->t1;
    a.t2 = s->t2;
    a.t3 = s->t3;
    a.colx = s->colx;
    a.colz = s->colz;
    a.logx = s->logx;
    a.logz = s->logz;
    a.seed_lo = (uint32_t)(seed & 0xffffffffull);
    a.seed_hi = (uint32_t)(seed >> 32);
    a.stream_tag = s->stream_tag;
    a.first = first;
    a.n = n;
    a.syn = d_syndromes;
    a.cls = d_classes;
    a.err = d_errors;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(sample_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
    TQS_HIP(hipGetLastError());
    return TQEC_OK;
}

void tqec_sampler_destroy(tqec_sampler *s) {
    if (!s) return;
    (void)hipFree(s->t1);
    (void)hipFree(s->t2);
    (void)hipFree(s->t3);
    (void)hipFree(s->colx);
    (void)hipFree(s->colz);
    (void)hipFree(s->logx);
    (void)hipFree(s->logz);
    delete s;
}

int tqec_recovery_create_css(int32_t nq, int32_t nsx, int32_t nsz, const uint8_t *Rx, const uint8_t *Rz,
                             int32_t k, const uint8_t *lx, const uint8_t *lz, tqec_recovery **out) {
    if (!out || !Rx || !Rz || (k && (!lx || !lz))) return fail(TQEC_ENULL, "null argument");
    *out = nullptr;
    if (nq <= 0 || nq > 64 * kMaxErrWords || nsx < 0 || nsz < 0 || k < 0 || 2 * k > 31)
        return fail(TQEC_EUNSUPPORTED, "recovery sizes out of range");
    auto *r = new (std::nothrow) tqec_recovery();
    if (!r) return fail(TQEC_ENOMEM, "out of host memory");
    const int ew = (nq + 63) / 64;
    // Rx: nq x nsz maps sz -> ex; columns packed as [nsz][ew]
    std::vector<uint8_t> tx((size_t)nsz * nq), tz((size_t)nsx * nq);
    for (int q = 0; q < nq; ++q) {
        for (int j = 0; j < nsz; ++j) tx[(size_t)j * nq + q] = Rx[(size_t)q * nsz + j];
        for (int j = 0; j < nsx; ++j) tz[(size_t)j * nq + q] = Rz[(size_t)q * nsx + j];
    }
    int rc;
    if ((rc = upload(pack_rows(tx.data(), nsz, nq, ew), &r->Ra)) ||
        (rc = upload(pack_rows(tz.data(), nsx, nq, ew), &r->Rb)) ||
        (rc = upload(pack_rows(lz, k, nq, ew), &r->la)) || (rc = upload(pack_rows(lx, k, nq, ew), &r->lb)) ||
        (rc = upload(pack_rows(lx, k, nq, ew), &r->fa)) || (rc = upload(pack_rows(lz, k, nq, ew), &r->fb))) {
        tqec_recovery_destroy(r);
        return rc;
    }
    RecoveryArgs &a = r->a;
    a.nq = nq;
    a.nsa = nsz;
    a.nsb = nsx;
    a.ew = ew;
    a.k = k;
    a.kind = 0;
    a.nsyn = nsx + nsz;
    a.offa = nsx;
    a.offb = 0;
    a.Ra = r->Ra;
    a.Rb = r->Rb;
    a.la = r->la;
    a.lb = r->lb;
    a.fa = r->fa;
    a.fb = r->fb;
    *out = r;
    return TQEC_OK;
}

int tqec_recovery_create_dem(int32_t nm, int32_t nd, const uint8_t *R, int32_t nl, const uint8_t *l2q,
                             tqec_recovery **out) {
    if (!out || !R || (nl && !l2q)) return fail(TQEC_ENULL, "null argument");
    *out = nullptr;
    if (nm <= 0 || nm > 64 * kMaxErrWords || nd <= 0 || nl < 0 || nl > 31)
        return fail(TQEC_EUNSUPPORTED, "recovery sizes out of range");
    auto *r = new (std::nothrow) tqec_recovery();
    if (!r) return fail(TQEC_ENOMEM, "out of host memory");
    const int ew = (nm + 63) / 64;
    std::vector<uint8_t> t((size_t)nd * nm);
    for (int q = 0; q < nm; ++q)
        for (int j = 0; j < nd; ++j) t[(size_t)j * nm + q] = R[(size_t)q * nd + j];
    int rc;
    if ((rc = upload(pack_rows(t.data(), nd, nm, ew), &r->Ra)) || (rc = upload(pack_rows(l2q, nl, nm, ew), &r->la))) {
        tqec_recovery_destroy(r);
        return rc;
    }
    RecoveryArgs &a = r->a;
    a.nq = nm;
    a.nsa = nd;
    a.nsb = 0;
    a.ew = ew;
    a.k = nl;
    a.kind = 1;
    a.nsyn = nd;
    a.offa = 0;
    a.Ra = r->Ra;
    a.la = r->la;
    *out = r;
    return TQEC_OK;
}

int tqec_recover(tqec_recovery *r, const uint8_t *d_syndromes, const int32_t *d_classes, int64_t n,
                 uint8_t *d_err_a, uint8_t *d_err_b, void *stream) {
    if (!r || (n > 0 && (!d_syndromes || !d_classes || !d_err_a))) return fail(TQEC_ENULL, "null argument");
    if (n < 0) return fail(TQEC_EINVAL, "negative count");
    if (n == 0) return TQEC_OK;
    RecoveryArgs a = r->a;
    a.syn = d_syndromes;
    a.cls = d_classes;
    a.n = n;
    a.ea = d_err_a;
    a.eb = d_err_b;
    hipLaunchKernelGGL(recovery_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, (hipStream_t)stream, a);
    TQS_HIP(hipGetLastError());
    return TQEC_OK;
}

int tqec_ler_create(tqec_plan *plan, tqec_sampler *smp, tqec_recovery *rec, int64_t batch, tqec_ler **out) {
    if (!out || !plan || !smp || !rec) return fail(TQEC_ENULL, "null argument");
    *out = nullptr;
    if (batch <= 0 || batch > (int64_t(1) << 24)) return fail(TQEC_EINVAL, "batch must be in [1, 2^24]");
    tqec_stats st{};
    int rc;
    if ((rc = tqec_plan_stats(plan, &st))) return rc;
    const RecoveryArgs &ra = rec->a;
    if (ra.kind != smp->kind || ra.nq != smp->nvar || ra.nsyn != smp->nsyn)
        return fail(TQEC_EINVAL, "sampler and recovery describe different codes");
    if (ra.kind == 0 ? (2 * ra.k != smp->nlog) : (ra.k != smp->nlog))
        return fail(TQEC_EINVAL, "sampler and recovery have different logical counts");
    if (st.n_out != smp->nlog) return fail(TQEC_EINVAL, "plan and sampler have different logical counts");
    auto *h = new (std::nothrow) tqec_ler();
    if (!h) return fail(TQEC_ENOMEM, "out of host memory");
    h->plan = plan;
    h->smp = smp;
    h->rec = rec;
    h->batch = batch;
    h->nout = 1 << st.n_out;
    const size_t B = (size_t)batch, nv = (size_t)smp->nvar;
    hipError_t e = hipMalloc((void **)&h->syn, B * smp->nsyn);
    if (e == hipSuccess) e = hipMalloc((void **)&h->err, B * nv * (smp->kind == 0 ? 2 : 1));
    if (e == hipSuccess) e = hipMalloc((void **)&h->ea, B * nv);
    if (e == hipSuccess && smp->kind == 0) e = hipMalloc((void **)&h->eb, B * nv);
    if (e == hipSuccess) e = hipMalloc((void **)&h->tie, B);
    if (e == hipSuccess) e = hipMalloc((void **)&h->cls, B * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&h->mar, B * h->nout * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&h->counts, 4 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        tqec_ler_destroy(h);
        return fail(TQEC_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    *out = h;
    return TQEC_OK;
}

int tqec_ler_run(tqec_ler *h, int32_t mode, uint64_t seed, int64_t first, int64_t n, int64_t *counts, void *stream) {
    if (!h || !counts) return fail(TQEC_ENULL, "null argument");
    if (n < 0 || first < 0) return fail(TQEC_EINVAL, "negative count or offset");
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(ler_zero_kernel, dim3(1), dim3(64), 0, st, h->counts);
    TQS_HIP(hipGetLastError());
    int rc;
    for (int64_t off = 0; off < n; off += h->batch) {
        const int64_t nb = std::min(h->batch, n - off);
        if ((rc = tqec_sample(h->smp, seed, first + off, nb, h->syn, nullptr, h->err, stream))) return rc;
        if ((rc = tqec_decode_batch_device(h->plan, h->syn, nb, mode, h->mar, h->cls, h->tie, stream))) return rc;
        if ((rc = tqec_recover(h->rec, h->syn, h->cls, nb, h->ea, h->eb, stream))) return rc;
        CheckArgs a{};
        a.kind = h->smp->kind;
        a.nvar = h->smp->nvar;
        a.synw = h->smp->synw;
        a.k = h->rec->a.k;
        a.colx = h->smp->colx;
        a.colz = h->smp->colz;
        a.logx = h->smp->logx;
        a.logz = h->smp->logz;
        a.err = h->err;
        a.ea = h->ea;
        a.eb = h->eb;
        a.n = nb;
        a.counts = h->counts;
        hipLaunchKernelGGL(ler_check_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, a);
        TQS_HIP(hipGetLastError());
    }
    unsigned long long c[4];
    TQS_HIP(hipMemcpyAsync(c, h->counts, sizeof(c), hipMemcpyDeviceToHost, st));
    TQS_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < 4; ++i) counts[i] = (int64_t)c[i];
    return TQEC_OK;
}

void tqec_ler_destroy(tqec_ler *h) {
    if (!h) return;
    (void)hipFree(h->syn);
    (void)hipFree(h->err);
    (void)hipFree(h->ea);
    (void)hipFree(h->eb);
    (void)hipFree(h->tie);
    (void)hipFree(h->cls);
    (void)hipFree(h->mar);
    (void)hipFree(h->counts);
    delete h;
}

void tqec_recovery_destroy(tqec_recovery *r) {
    if (!r) return;
    (void)hipFree(r->Ra);
    (void)hipFree(r->Rb);
    (void)hipFree(r->la);
    (void)hipFree(r->lb);
    (void)hipFree(r->fa);
    (void)hipFree(r->fb);
    delete r;
}

}  // extern "C"
