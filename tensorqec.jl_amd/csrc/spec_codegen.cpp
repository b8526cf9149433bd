// spec_codegen.cpp — see spec_codegen.hpp.
//
// The generated kernel follows the generic FP32 fused kernel of
// tqec_hip.hip step for step (syndrome staging, per-tensor power-of-two
// scaling with tagged LDS maxima, low-exponent flag, finalisation and FP64
// re-check list), with every descriptor field substituted by its value.
#include "spec_codegen.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cctype>
#include <cmath>
#include <sstream>
#include <vector>

namespace tqec {

namespace {

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
    size_t work = 0, tab = 0, mx = 0, ex = 0, syn = 0, low = 0, total = 0;
    size_t tab_entries = 0;
};

Layout layout(const CompiledPlan &cp, const SpecCfg &cfg) {
    const size_t G = size_t(1) << cfg.glog;
    const size_t elem = cfg.fp64 ? 8 : 4;
    Layout L;
    L.tab_entries = cp.const_pool.size() + cp.sliced_pool.size();
    size_t off = 0;
    L.work = off;
    if (!cfg.slab) off += round_up((size_t)std::max<int64_t>(1, cp.ws_entries) * G * elem, 16);
    L.tab = off;
    if (cfg.tab_lds) off += round_up(std::max<size_t>(1, L.tab_entries) * elem, 16);
    L.mx = off;
    if (!cfg.fp64) off += (size_t)std::max(1, cp.n_slots) * G * 8;
    L.ex = off;
    if (!cfg.fp64) off += (size_t)std::max(1, cp.n_slots) * G * 4;
    L.syn = off;
    off += round_up((size_t)std::max(1, cp.n_syn) * G, 4);
    L.low = off;
    if (!cfg.fp64) off += G * 4;
    L.total = round_up(off, 16);
    return L;
}

// gather_bits(c, op) as an expression of `c`.
std::string gather_expr(const OperandDesc &op) {
    std::ostringstream s;
    bool any = false;
    for (int i = 0; i < op.nseg; ++i) {
        const Seg &g = op.seg[i];
        if (g.width <= 0) continue;
        const unsigned mask = (g.width >= 32) ? 0xffffffffu : ((1u << g.width) - 1u);
        if (any) s << " | ";
        s << "(((c >> " << g.src << ") & " << mask << "u) << " << g.dst << ")";
        any = true;
    }
    if (!any) return "0u";
    return s.str();
}

// Base pointer of an operand (fused_base).
void emit_base(std::ostringstream &o, const char *name, const OperandDesc &op, size_t const_n) {
    if (op.kind == kInter) {
        o << "    const T *" << name << " = work + " << ((long long)op.off << 0) << " * G + g;\n";
    } else if (op.kind == kConst) {
        o << "    const T *" << name << " = tab + " << (long long)op.off << ";\n";
    } else {
        o << "    unsigned v_" << name << " = 0u;\n";
        for (int j = 0; j < op.nev; ++j)
            o << "    v_" << name << " |= (unsigned)synl[" << op.ev_bit[j] << " * G + g] << " << j << ";\n";
        o << "    const T *" << name << " = tab + " << (long long)(const_n + (size_t)op.off) << " + (v_" << name
          << " << " << op.rank << ");\n";
    }
}

// Bits an operand index can take from the output index.
uint32_t gathered_mask(const OperandDesc &op) {
    uint32_t m = 0;
    for (int i = 0; i < op.nseg; ++i) {
        const Seg &g = op.seg[i];
        if (g.width <= 0) continue;
        const uint32_t w = (g.width >= 32) ? 0xffffffffu : ((1u << g.width) - 1u);
        m |= w << g.dst;
    }
    return m;
}

// op_load_sum for element index (idx_var | kconst).  The index is written
// as a sum when the constant's bits cannot meet the gathered bits, which
// lets the compiler fold it into the LDS instruction's offset.
std::string load_expr(const OperandDesc &op, const char *base, const char *idx, uint32_t kconst,
                      std::ostringstream &pre, const std::string &tmp) {
    std::ostringstream s;
    const uint32_t gm = gathered_mask(op);
    uint32_t allk = kconst;
    if (op.kind == kInter)
        for (uint32_t q = 0; q < (1u << op.nsingle); ++q) allk |= kconst | (q << op.sshift);
    const char *join = (gm & allk) ? " | " : " + ";
    if (op.kind != kInter) {
        s << base << "[" << idx << join << kconst << "u]";
        return s.str();
    }
    if (op.nsingle == 0) {
        s << base << "[(" << idx << join << kconst << "u) << GLOG]";
        return s.str();
    }
    // T s = 0; for q: s += base[(idx | q << sshift) * G]
    pre << "        T " << tmp << " = 0;\n";
    const uint32_t ns = 1u << op.nsingle;
    for (uint32_t q = 0; q < ns; ++q)
        pre << "        " << tmp << " += " << base << "[(" << idx << join << (kconst | (q << op.sshift))
            << "u) << GLOG];\n";
    return tmp;
}

// Output bits of a step that feed an operand's index (the segments' sources).
uint32_t source_mask(const OperandDesc &op) {
    uint32_t m = 0;
    for (int i = 0; i < op.nseg; ++i) {
        const Seg &g = op.seg[i];
        if (g.width <= 0) continue;
        const uint32_t w = (g.width >= 32) ? 0xffffffffu : ((1u << g.width) - 1u);
        m |= w << g.src;
    }
    return m;
}

// `expr` with every whole-word `c` replaced by `by`.
std::string subst_c(const std::string &expr, const std::string &by) {
    std::string r;
    for (size_t i = 0; i < expr.size(); ++i) {
        if (expr[i] == 'c' && (i == 0 || !isalnum((unsigned char)expr[i - 1])) &&
            (i + 1 == expr.size() || !isalnum((unsigned char)expr[i + 1])))
            r += "(" + by + ")";
        else
            r += expr[i];
    }
    return r;
}

// The output loop of a step.  Lane c0 (< CSTEP) owns the outputs whose low
// bits are c0.  Up to `block` of them are computed together: the block
// varies output bits that feed only A or only B, so an A row (or B row)
// loaded once serves several outputs (register blocking of the small
// per-syndrome contraction); all loads come before the fma chains and the
// stores after them.  Every output keeps the generic kernel's order of
// operations (k ascending, single-side sums first), so results are
// bit-identical.
// Contraction terms per output unrolled in full; longer contractions run a
// loop over chunks of this many terms (same order of the multiply-adds).
constexpr uint32_t kMaxUnrollK = 16;

void emit_outputs(std::ostringstream &o, const StepDesc &st, int cstep, int block, bool fp64) {
    const OperandDesc &A = st.a, &B = st.b;
    const int rc = st.rc;
    const long long nc = 1ll << rc;
    const std::string ga = gather_expr(A), gb = gather_expr(B);
    const uint32_t nkk = 1u << st.nk;
    const int cexp = A.cexp + B.cexp;
    int lc = 0;
    while ((1 << lc) < cstep) ++lc;
    // block bits S among the high bits [lc, rc), the loop runs over the rest R
    std::vector<int> S, R;
    if (nc > (long long)cstep) {
        const uint32_t am = source_mask(A), bm = source_mask(B);
        std::vector<int> aonly, bonly, other;
        for (int bit = lc; bit < rc; ++bit) {
            const bool ia = (am >> bit) & 1u, ib = (bm >> bit) & 1u;
            if (ia && !ib) aonly.push_back(bit);
            else if (ib && !ia) bonly.push_back(bit);
            else other.push_back(bit);
        }
        int want = 0;
        while ((2 << want) <= std::max(1, block)) ++want;
        size_t ia = 0, ib = 0, io = 0;
        while ((int)S.size() < want && (ia < aonly.size() || ib < bonly.size() || io < other.size())) {
            if (ia < aonly.size() && (ia <= ib || ib >= bonly.size())) S.push_back(aonly[ia++]);
            else if (ib < bonly.size()) S.push_back(bonly[ib++]);
            else S.push_back(other[io++]);
        }
        for (int bit = lc; bit < rc; ++bit)
            if (std::find(S.begin(), S.end(), bit) == S.end()) R.push_back(bit);
    }
    const uint32_t am = source_mask(A), bm = source_mask(B);
    const int ns = (int)S.size();
    // subsets of S that matter to A / to B
    std::vector<uint32_t> acombo, bcombo;   // bit patterns (in output-index bits)
    auto subsets = [&](uint32_t mask) {
        std::vector<uint32_t> v;
        std::vector<int> bits;
        for (int x : S)
            if ((mask >> x) & 1u) bits.push_back(x);
        for (uint32_t t = 0; t < (1u << bits.size()); ++t) {
            uint32_t m = 0;
            for (size_t j = 0; j < bits.size(); ++j)
                if ((t >> j) & 1u) m |= 1u << bits[j];
            v.push_back(m);
        }
        return v;
    };
    acombo = subsets(am);
    bcombo = subsets(bm);
    std::ostringstream b;
    for (size_t x = 0; x < acombo.size(); ++x)
        b << "      const unsigned ar" << x << " = " << subst_c(ga, "c | " + std::to_string(acombo[x]) + "u") << ";\n";
    for (size_t y = 0; y < bcombo.size(); ++y)
        b << "      const unsigned br" << y << " = " << subst_c(gb, "c | " + std::to_string(bcombo[y]) + "u") << ";\n";
    // contraction terms in chunks of kch (all of them at once when short);
    // operand values loaded once per (row, k) of the chunk
    const uint32_t kch = std::min(nkk, kMaxUnrollK);
    const bool kloop = kch < nkk;
    const std::string ia = kloop ? "(ar%1 | (kb << " + std::to_string(A.kshift) + "))" : "ar%1";
    const std::string ib = kloop ? "(br%1 | (kb << " + std::to_string(B.kshift) + "))" : "br%1";
    auto idx_of = [](const std::string &pat, size_t v) {
        std::string s = pat;
        const size_t at = s.find("%1");
        return s.replace(at, 2, std::to_string(v));
    };
    std::vector<size_t> out_x, out_y;
    for (uint32_t t = 0; t < (1u << ns); ++t) {
        uint32_t m = 0;
        for (int j = 0; j < ns; ++j)
            if ((t >> j) & 1u) m |= 1u << S[j];
        out_x.push_back(std::find(acombo.begin(), acombo.end(), m & am) - acombo.begin());
        out_y.push_back(std::find(bcombo.begin(), bcombo.end(), m & bm) - bcombo.begin());
        b << "      T acc" << t << " = 0;\n";
    }
    if (kloop) b << "      #pragma unroll 1\n      for (unsigned kb = 0; kb < " << nkk << "u; kb += " << kch << "u) {\n";
    for (uint32_t k = 0; k < kch; ++k) {
        for (size_t x = 0; x < acombo.size(); ++x) {
            std::ostringstream pre;
            const std::string nm = idx_of(ia, x);
            const std::string tmp = "sa" + std::to_string(k) + "_" + std::to_string(x);
            const std::string v = load_expr(A, "pa", nm.c_str(), k << A.kshift, pre, tmp);
            b << pre.str();
            if (v != tmp) b << "      const T " << tmp << " = " << v << ";\n";
        }
        for (size_t y = 0; y < bcombo.size(); ++y) {
            std::ostringstream pre;
            const std::string nm = idx_of(ib, y);
            const std::string tmp = "sb" + std::to_string(k) + "_" + std::to_string(y);
            const std::string v = load_expr(B, "pb", nm.c_str(), k << B.kshift, pre, tmp);
            b << pre.str();
            if (v != tmp) b << "      const T " << tmp << " = " << v << ";\n";
        }
    }
    for (uint32_t t = 0; t < (1u << ns); ++t)
        for (uint32_t k = 0; k < kch; ++k)
            b << "      acc" << t << " = " << (fp64 ? "fma(" : "fmaf(") << "sa" << k << "_" << out_x[t] << ", sb" << k << "_"
              << out_y[t] << ", acc" << t << ");\n";
    if (kloop) b << "      }\n";
    // outputs of the block
    for (uint32_t t = 0; t < (1u << ns); ++t) {
        uint32_t m = 0;
        for (int j = 0; j < ns; ++j)
            if ((t >> j) & 1u) m |= 1u << S[j];
        const std::string acc = "acc" + std::to_string(t);
        if (fp64) {
            if (cexp != 0) b << "      " << acc << " = ldexp(" << acc << ", " << cexp << ");\n";
        } else {
            b << "      " << acc << " = ldexpf(" << acc << ", shift);\n";
            b << "      lmax = " << acc << " > lmax ? " << acc << " : lmax;\n";
        }
        b << "      pc[(c | " << m << "u) << GLOG] = " << acc << ";\n";
    }
    if (nc <= (long long)cstep) {
        o << "    if (c0 < " << nc << ") {\n      const unsigned c = (unsigned)c0;\n" << b.str() << "    }\n";
        return;
    }
    // c = c0 | deposit(i, R)
    std::string dep = "(unsigned)c0";
    for (size_t j = 0; j < R.size(); ++j)
        dep += " | (((i >> " + std::to_string(j) + ") & 1u) << " + std::to_string(R[j]) + ")";
    o << "    #pragma unroll 1\n";
    o << "    for (unsigned i = 0; i < " << (1u << R.size()) << "u; ++i) {\n";
    o << "      const unsigned c = " << dep << ";\n";
    o << b.str();
    o << "    }\n";
}

// Bits of `mask` (ascending).
std::vector<int> bits_of(uint32_t mask) {
    std::vector<int> v;
    for (int b = 0; b < 32; ++b)
        if ((mask >> b) & 1u) v.push_back(b);
    return v;
}

// `(x)` spread onto the given bit positions: bit t of x goes to pos[t].
std::string deposit(const std::string &x, const std::vector<int> &pos) {
    if (pos.empty()) return "0u";
    std::string r;
    for (size_t t = 0; t < pos.size(); ++t) {
        if (t) r += " | ";
        r += "(((" + x + ") >> " + std::to_string(t) + ") & 1u) << " + std::to_string(pos[t]);
    }
    return "(" + r + ")";
}

// GEMM-shaped step on BF16 MFMA: per syndrome, C[i, j] = sum_k A[i, k] B[k, j]
// with i over the output bits that feed only A, j over those that feed only
// B, and the remaining output bits (shared, or beyond the 32 x 32 tile)
// enumerating tiles.  Operands are rounded to BF16 (nearest even) as they are
// read, the products accumulate in FP32 inside the MFMA, the output is scaled
// like every FP32 step.  Layouts (v_mfma_f32_32x32x16_bf16, wave64): lane l
// holds A[l % 32][8 (l / 32) .. +7], B[8 (l / 32) .. +7][l % 32], and
// D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32] for v = 0..15.
void emit_outputs_mfma(std::ostringstream &o, const StepDesc &st, int threads) {
    const OperandDesc &A = st.a, &B = st.b;
    const uint32_t am = source_mask(A), bm = source_mask(B);
    const uint32_t outm = (st.rc >= 32) ? 0xffffffffu : ((1u << st.rc) - 1u);
    const std::vector<int> aon = bits_of(am & ~bm & outm), bon = bits_of(bm & ~am & outm);
    const std::vector<int> ai(aon.begin(), aon.begin() + 5), bj(bon.begin(), bon.begin() + 5);
    std::vector<int> rest;
    for (int b = 0; b < st.rc; ++b)
        if (std::find(ai.begin(), ai.end(), b) == ai.end() && std::find(bj.begin(), bj.end(), b) == bj.end())
            rest.push_back(b);
    const unsigned ntiles = 1u << rest.size();
    const int nwaves = threads / 64;
    const int chunks = 1 << (st.nk - 4);
    const std::string ga = gather_expr(A), gb = gather_expr(B);
    auto ld = [&](const OperandDesc &op, const char *base, const char *idx, const std::string &k) {
        const std::string e = std::string(idx) + " | ((" + k + ") << " + std::to_string(op.kshift) + ")";
        if (op.kind == kInter) return std::string(base) + "[(" + e + ") << GLOG]";
        return std::string(base) + "[" + e + "]";
    };
    o << "    {\n";
    o << "      const int l = tid & 63, w = tid >> 6;\n";
    o << "      for (unsigned t = (unsigned)w; t < " << ntiles << "u; t += " << nwaves << "u) {\n";
    o << "        const unsigned tb = " << deposit("t", rest) << ";\n";
    o << "        const unsigned ci = tb | " << deposit("(unsigned)(l & 31)", ai) << ";\n";
    o << "        const unsigned cj = tb | " << deposit("(unsigned)(l & 31)", bj) << ";\n";
    o << "        const unsigned ar = " << subst_c(ga, "ci") << ";\n";
    o << "        const unsigned br = " << subst_c(gb, "cj") << ";\n";
    o << "        f32x16 acc = {};\n";
    for (int ch = 0; ch < chunks; ++ch) {
        o << "        {\n          const unsigned kb = " << ch * 16 << "u + 8u * (unsigned)(l >> 5);\n";
        o << "          bf16x8 av, bv;\n";
        for (int e = 0; e < 8; ++e) {
            const std::string k = "kb + " + std::to_string(e) + "u";
            o << "          av[" << e << "] = (short)to_bf16(" << ld(A, "pa", "ar", k) << ");\n";
            o << "          bv[" << e << "] = (short)to_bf16(" << ld(B, "pb", "br", k) << ");\n";
        }
        o << "          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);\n        }\n";
    }
    o << "        #pragma unroll\n";
    o << "        for (int v = 0; v < 16; ++v) {\n";
    o << "          const unsigned i = 8u * (unsigned)(v >> 2) + 4u * (unsigned)(l >> 5) + (unsigned)(v & 3);\n";
    o << "          const unsigned c = tb | " << deposit("i", ai) << " | " << deposit("(unsigned)(l & 31)", bj) << ";\n";
    o << "          const float x = ldexpf(acc[v], shift);\n";
    o << "          lmax = x > lmax ? x : lmax;\n";
    o << "          pc[c << GLOG] = x;\n";
    o << "        }\n      }\n    }\n";
}

// Exclusive output bits of a step: those feeding only A and only B.
void mfma_side_bits(const StepDesc &st, std::vector<int> &aon, std::vector<int> &bon) {
    const uint32_t am = source_mask(st.a), bm = source_mask(st.b);
    const uint32_t outm = (st.rc >= 32) ? 0xffffffffu : ((1u << st.rc) - 1u);
    aon = bits_of(am & ~bm & outm);
    bon = bits_of(bm & ~am & outm);
}

// GEMM-shaped step on FP32 MFMA, full FP32 operands: per syndrome
// C[i, j] = sum_k A[i, k] B[k, j] as above, in tiles of 32 x 32
// (v_mfma_f32_32x32x2f32, K = 2 per instruction) when both sides have five
// exclusive output bits, else 16 x 16 (v_mfma_f32_16x16x4f32, K = 4).  Each
// wave owns a block of ra x rb tiles (up to two more A-only output bits and
// one more B-only bit: 4 x 2 tiles, the best of 2 x 2, 4 x 2 and 2 x 4 on
// toric L=5; 4 x 4 spills), so every operand value a lane loads feeds rb (or
// ra) MFMAs.  The k loop takes ku instructions' worth of contracted values
// per iteration and loads the next iteration's operands before issuing its
// own MFMAs (at one or two waves per SIMD few other waves hide the loads).
// Layouts (wave64), with T the tile edge (64 / T contracted values per
// instruction):
// lane l holds A[l % T][l / T] and B[l / T][l % T]; T = 32: D[v] =
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32], v = 0..15; T = 16: D[v] =
// C[4 (l / 16) + v][l % 16], v = 0..3.  Every index stays inside the
// operand and output ranges by construction (i, j < T over that many
// output bits, k < 2^nk, and every lane runs the same number of k
// iterations), so a layout mistake shows as wrong values, never as an
// out-of-range access.
void emit_outputs_mfma32(std::ostringstream &o, const StepDesc &st, int threads) {
    const OperandDesc &A = st.a, &B = st.b;
    std::vector<int> aon, bon;
    mfma_side_bits(st, aon, bon);
    const bool big = aon.size() >= 5 && bon.size() >= 5;
    const int tbits = big ? 5 : 4;                 // tile edge 2^tbits
    const unsigned tedge = 1u << tbits, kq = 64u / tedge;   // k values per instruction
    const std::vector<int> ai(aon.begin(), aon.begin() + tbits), bj(bon.begin(), bon.begin() + tbits);
    const int xa = std::min<int>(2, (int)aon.size() - tbits), xb = std::min<int>(1, (int)bon.size() - tbits);
    const std::vector<int> a2(aon.begin() + tbits, aon.begin() + tbits + xa);
    const std::vector<int> b2(bon.begin() + tbits, bon.begin() + tbits + xb);
    std::vector<int> rest;
    for (int b = 0; b < st.rc; ++b) {
        const bool used = std::find(ai.begin(), ai.end(), b) != ai.end() ||
                          std::find(bj.begin(), bj.end(), b) != bj.end() ||
                          std::find(a2.begin(), a2.end(), b) != a2.end() ||
                          std::find(b2.begin(), b2.end(), b) != b2.end();
        if (!used) rest.push_back(b);
    }
    const int ra = 1 << xa, rb = 1 << xb;
    const unsigned ntiles = 1u << rest.size();
    const int nwaves = threads / 64;
    const unsigned nkk = 1u << st.nk;
    const unsigned ku = std::min(4u, nkk / kq);   // instructions' k groups per loop iteration
    const unsigned kstep = kq * ku;
    const std::string tm = std::to_string(tedge - 1) + "u", lsh = std::to_string(tbits);
    const char *acc_t = big ? "f32x16" : "f32x4";
    const char *mfma = big ? "__builtin_amdgcn_mfma_f32_32x32x2f32" : "__builtin_amdgcn_mfma_f32_16x16x4f32";
    const std::string ga = gather_expr(A), gb = gather_expr(B);
    auto ld = [&](const OperandDesc &op, const char *base, const std::string &idx, const std::string &k) {
        const std::string e = idx + " | ((" + k + ") << " + std::to_string(op.kshift) + ")";
        if (op.kind == kInter) return std::string(base) + "[(" + e + ") << GLOG]";
        return std::string(base) + "[" + e + "]";
    };
    o << "    {\n";
    o << "      const int l = tid & 63, w = tid >> 6;\n";
    o << "      #pragma unroll 1\n";
    o << "      for (unsigned t = (unsigned)w; t < " << ntiles << "u; t += " << nwaves << "u) {\n";
    o << "        const unsigned tb = " << deposit("t", rest) << ";\n";
    for (int x = 0; x < ra; ++x)
        o << "        const unsigned ar" << x << " = "
          << subst_c(ga, "(tb | " + deposit("(unsigned)(l & " + tm + ")", ai) + " | " +
                             deposit(std::to_string(x) + "u", a2) + ")")
          << ";\n";
    for (int y = 0; y < rb; ++y)
        o << "        const unsigned br" << y << " = "
          << subst_c(gb, "(tb | " + deposit("(unsigned)(l & " + tm + ")", bj) + " | " +
                             deposit(std::to_string(y) + "u", b2) + ")")
          << ";\n";
    for (int x = 0; x < ra; ++x)
        for (int y = 0; y < rb; ++y) o << "        " << acc_t << " acc" << x << "_" << y << " = {};\n";
    auto loads = [&](const std::string &kbase, const char *pre, bool decl) {
        for (unsigned u = 0; u < ku; ++u) {
            const std::string k = kbase + " + " + std::to_string(kq * u) + "u";
            for (int x = 0; x < ra; ++x)
                o << "          " << (decl ? "float " : "") << pre << "a" << u << "_" << x << " = "
                  << ld(A, "pa", "ar" + std::to_string(x), k) << ";\n";
            for (int y = 0; y < rb; ++y)
                o << "          " << (decl ? "float " : "") << pre << "b" << u << "_" << y << " = "
                  << ld(B, "pb", "br" + std::to_string(y), k) << ";\n";
        }
    };
    o << "        {\n";
    loads("(unsigned)(l >> " + lsh + ")", "n", true);
    o << "        #pragma unroll 1\n";
    o << "        for (unsigned kb = (unsigned)(l >> " << lsh << "); kb < " << nkk << "u; kb += " << kstep << "u) {\n";
    for (unsigned u = 0; u < ku; ++u) {
        for (int x = 0; x < ra; ++x) o << "          const float a" << u << "_" << x << " = na" << u << "_" << x << ";\n";
        for (int y = 0; y < rb; ++y) o << "          const float b" << u << "_" << y << " = nb" << u << "_" << y << ";\n";
    }
    // the same for every lane (kb differs across the wave by l >> tbits < kq <= kstep): uniform
    o << "          if (kb + " << kstep << "u < " << nkk << "u) {\n";
    loads("kb + " + std::to_string(kstep) + "u", "n", false);
    o << "          }\n";
    for (unsigned u = 0; u < ku; ++u)
        for (int x = 0; x < ra; ++x)
            for (int y = 0; y < rb; ++y)
                o << "          acc" << x << "_" << y << " = " << mfma << "(a" << u << "_" << x << ", b" << u << "_" << y
                  << ", acc" << x << "_" << y << ", 0, 0, 0);\n";
    o << "        }\n";
    o << "        }\n";
    for (int x = 0; x < ra; ++x)
        for (int y = 0; y < rb; ++y) {
            o << "        #pragma unroll\n";
            o << "        for (int v = 0; v < " << (big ? 16 : 4) << "; ++v) {\n";
            if (big)
                o << "          const unsigned i = 8u * (unsigned)(v >> 2) + 4u * (unsigned)(l >> 5) + (unsigned)(v & 3);\n";
            else
                o << "          const unsigned i = 4u * (unsigned)(l >> 4) + (unsigned)v;\n";
            o << "          const unsigned c = tb | " << deposit("i", ai) << " | " << deposit(std::to_string(x) + "u", a2)
              << " | " << deposit("(unsigned)(l & " + tm + ")", bj) << " | " << deposit(std::to_string(y) + "u", b2)
              << ";\n";
            o << "          const float x = ldexpf(acc" << x << "_" << y << "[v], shift);\n";
            o << "          lmax = x > lmax ? x : lmax;\n";
            o << "          pc[c << GLOG] = x;\n";
            o << "        }\n";
        }
    o << "      }\n    }\n";
}

void emit_step(std::ostringstream &o, const StepDesc &st, int s, size_t const_n, int cstep, bool fp64, int unroll,
               const SpecCfg *mfma_cfg, bool opaque_tid = false) {
    const OperandDesc &A = st.a, &B = st.b;
    o << "  { // step " << s << ": rank " << st.rc << ", contracted " << st.nk << "\n";
    if (opaque_tid)
        // the step's own copy of the thread index, opaque to the optimiser:
        // nothing derived from it is hoisted out of the kernel's loop over
        // tiles or shared with other steps, so per-thread index values live
        // for one step only
        o << "    int tid = (int)threadIdx.x;\n    asm volatile(\"\" : \"+v\"(tid));\n"
             "    const int g = tid & (G - 1), c0 = tid >> GLOG, lane = tid & 63;\n"
             "    (void)g; (void)c0; (void)lane;\n";
    if (fp64) {
        // generic FP64 step: no scaling, the folded constant exponents restored per entry
        emit_base(o, "pa", A, const_n);
        emit_base(o, "pb", B, const_n);
        o << "    T *pc = work + " << (long long)st.c_off << " * G + g;\n";
        emit_outputs(o, st, cstep, unroll, true);
        o << "  }\n";
        o << "  __syncthreads();\n";
        return;
    }
    o << "    int ea, eb, na = 0, nb = 0;\n";
    for (int side = 0; side < 2; ++side) {
        const OperandDesc &op = side ? B : A;
        const char *n = side ? "nb" : "na";
        const char *e = side ? "eb" : "ea";
        if (op.kind == kInter) {
            o << "    { const float m = __uint_as_float((unsigned)(mx[" << op.slot
              << " * G + g] & 0xffffffffull)); " << n << " = (m > 0.f) ? ilogbf(m) : 0; " << e << " = ex["
              << op.slot << " * G + g] + " << n << "; }\n";
        } else {
            o << "    " << e << " = " << op.cexp << ";\n";
        }
    }
    o << "    const int shift = -(na + nb);\n";
    {
        std::string cond;
        if (A.kind == kInter) cond += "na < -100";
        if (B.kind == kInter) cond += std::string(cond.empty() ? "" : " || ") + "nb < -100";
        if (!cond.empty()) o << "    if (" << cond << ") lowf[g] = 1;\n";
    }
    emit_base(o, "pa", A, const_n);
    emit_base(o, "pb", B, const_n);
    o << "    T *pc = work + " << (long long)st.c_off << " * G + g;\n";
    o << "    T lmax = 0;\n";
    if (mfma_cfg && spec_step_uses_mfma(st, *mfma_cfg))
        emit_outputs_mfma(o, st, mfma_cfg->threads);
    else if (mfma_cfg && spec_step_uses_mfma32(st, *mfma_cfg))
        emit_outputs_mfma32(o, st, mfma_cfg->threads);
    else
        emit_outputs(o, st, cstep, unroll, false);
    o << "    for (int w = G; w < 64; w <<= 1) { const float other = __shfl_xor(lmax, w); lmax = other > lmax ? other : lmax; }\n";
    o << "    if (lane < G) atomicMax(&mx[" << st.c_slot << " * G + g], ((unsigned long long)(tag_base + " << s
      << "u) << 32) | (unsigned long long)__float_as_uint(lmax));\n";
    o << "    if (tid < G) ex[" << st.c_slot << " * G + g] = ea + eb;\n";
    o << "  }\n";
    o << "  __syncthreads();\n";
}

// ---- register-resident variant ------------------------------------------
//
// A quad of lanes owns one syndrome.  Tensor X of rank r keeps entry i in
// lane sum_t bit(i, pos[t]) << t of the quad and in register pext(i, other
// bits) of that lane; a missing pos[t] replicates X along lane bit t.  The
// consumer of X chooses X's lane bits: an output lane bit that feeds X is
// placed in the same lane slot of X, so those index bits never cross lanes
// and the register an operand entry lives in never depends on the lane.  A
// fetch is then one DPP quad permutation (or a plain register read when the
// permutation is the identity).  Layouts are decided from the root down.

struct RLay {
    int pos[2] = {-1, -1};   // tensor index bit held by lane bit t (-1: replicated along t)
};

int seg_image(const OperandDesc &op, int u) {
    for (int i = 0; i < op.nseg; ++i) {
        const Seg &g = op.seg[i];
        if (g.width > 0 && u >= g.src && u < g.src + g.width) return g.dst + (u - g.src);
    }
    return -1;
}

uint32_t gather_host(uint32_t c, const OperandDesc &op) {
    uint32_t r = 0;
    for (int i = 0; i < op.nseg; ++i) {
        const Seg &g = op.seg[i];
        if (g.width <= 0) continue;
        const uint32_t w = (g.width >= 32) ? 0xffffffffu : ((1u << g.width) - 1u);
        r |= ((c >> g.src) & w) << g.dst;
    }
    return r;
}

int lane_bits(const RLay &l) { return (l.pos[0] >= 0) + (l.pos[1] >= 0); }

std::vector<int> reg_bits(const RLay &l, int rank) {
    std::vector<int> v;
    for (int b = 0; b < rank; ++b)
        if (b != l.pos[0] && b != l.pos[1]) v.push_back(b);
    return v;
}

uint32_t pext(uint32_t x, const std::vector<int> &bits) {
    uint32_t r = 0;
    for (size_t j = 0; j < bits.size(); ++j) r |= ((x >> bits[j]) & 1u) << j;
    return r;
}

uint32_t pdep(uint32_t x, const std::vector<int> &bits) {
    uint32_t r = 0;
    for (size_t j = 0; j < bits.size(); ++j) r |= ((x >> j) & 1u) << bits[j];
    return r;
}

// Registers per lane above which the variant is not offered (512 VGPRs and
// AGPRs per lane at one wave per SIMD, less the step temporaries).
constexpr int kRegBudget = 400;

struct RegPlan {
    std::vector<int> prod[2];   // producing step of each step's intermediate A / B, or -1
    std::vector<RLay> lay;      // per step output
    int root = -1;
    int peak = 0;               // live tensor registers per lane
    std::string err;
};

// Live tensor registers per lane in execution order; false above the budget.
bool reg_peak_ok(const CompiledPlan &cp, RegPlan &rp) {
    int live = 0;
    rp.peak = 0;
    for (int s = 0; s < (int)cp.steps.size(); ++s) {
        const StepDesc &st = cp.steps[s];
        live += 1 << (st.rc - lane_bits(rp.lay[s]));
        rp.peak = std::max(rp.peak, live);
        for (int side = 0; side < 2; ++side) {
            const int p = rp.prod[side][s];
            if (p >= 0) live -= 1 << (cp.steps[p].rc - lane_bits(rp.lay[p]));
        }
    }
    if (rp.peak > kRegBudget) return rp.err = "more live tensor entries per lane than registers", false;
    return true;
}

// quad_perm control: destination lane d of each quad reads lane src[d]
int quad_ctrl(const int src[4]) { return src[0] | (src[1] << 2) | (src[2] << 4) | (src[3] << 6); }

constexpr int kIdentityCtrl = 0xE4;   // quad_perm [0, 1, 2, 3]

// DPP control with which a consumer writing layout LC reads entry `idx`
// (lane-independent bits) of operand `op` stored in layout LX; -1 when an
// output lane bit that feeds the operand does not sit in the same lane slot
// of LX (the register would then depend on the lane).
int fetch_ctrl(const OperandDesc &op, const RLay &LC, const RLay &LX, uint32_t idx) {
    int img[2];
    for (int t = 0; t < 2; ++t) {
        img[t] = LC.pos[t] >= 0 ? seg_image(op, LC.pos[t]) : -1;
        if (img[t] >= 0 && img[t] != LX.pos[t]) return -1;
    }
    int src[4];
    for (int d = 0; d < 4; ++d) {
        src[d] = 0;
        for (int t = 0; t < 2; ++t) {
            const int bit = (LX.pos[t] < 0 || img[t] >= 0) ? (d >> t) & 1 : (int)((idx >> LX.pos[t]) & 1u);
            src[d] |= bit << t;
        }
    }
    return quad_ctrl(src);
}

// Per-lane instruction estimate of one operand's reads in a step writing
// layout LC: distinct DPP moves (intermediate stored in *LX) or distinct
// table loads (LX null), summed over the contraction index as the emitter
// caches them.  -1 when LX is not a legal layout for this consumer.
long fetch_cost(const StepDesc &st, int side, const RLay &LC, const RLay *LX) {
    const OperandDesc &op = side ? st.b : st.a;
    if (LX)
        for (int t = 0; t < 2; ++t)
            if (LC.pos[t] >= 0) {
                const int img = seg_image(op, LC.pos[t]);
                if (img >= 0 && img != LX->pos[t]) return -1;
            }
    const std::vector<int> RB = reg_bits(LC, st.rc);
    const std::vector<int> RX = LX ? reg_bits(*LX, op.rank) : std::vector<int>();
    const uint32_t nreg = 1u << RB.size(), nkk = 1u << st.nk;
    const uint32_t nq = (LX && op.nsingle) ? (1u << op.nsingle) : 1u;
    long cost = 0;
    std::vector<uint64_t> keys;
    for (uint32_t k = 0; k < nkk; ++k) {
        keys.clear();
        for (uint32_t r = 0; r < nreg; ++r) {
            const uint32_t base = gather_host(pdep(r, RB), op) | (k << op.kshift);
            if (!LX) {
                keys.push_back(base);
                continue;
            }
            for (uint32_t q = 0; q < nq; ++q) {
                const uint32_t idx = base | (q << op.sshift);
                const int ctrl = fetch_ctrl(op, LC, *LX, idx);
                if (ctrl != kIdentityCtrl) keys.push_back(((uint64_t)pext(idx, RX) << 8) | (uint64_t)ctrl);
            }
        }
        std::sort(keys.begin(), keys.end());
        cost += (long)(std::unique(keys.begin(), keys.end()) - keys.begin());
    }
    return cost;
}

std::vector<RLay> layout_candidates(int rank) {
    std::vector<RLay> v;
    for (int a = -1; a < rank; ++a)
        for (int b = -1; b < rank; ++b) {
            if (a >= 0 && a == b) continue;
            RLay L;
            L.pos[0] = a;
            L.pos[1] = b;
            v.push_back(L);
        }
    return v;
}

// Lane layouts minimising the per-lane instruction estimate of the whole
// step program: for every step and output layout, its multiply-adds, scale
// and maximum per output register and table loads, plus for each
// intermediate operand the cheapest legal layout of it (its DPP moves here
// plus the best cost of its own subtree in that layout).  Exact over the
// tree (each tensor has one consumer); replication (a missing lane bit) is
// one of the options.  Only for tensors of rank <= 9.
bool reg_layout_dp(const CompiledPlan &cp, RegPlan &rp) {
    const int ns = (int)cp.steps.size();
    for (const StepDesc &st : cp.steps)
        if (st.rc > 9) return false;
    std::vector<std::vector<RLay>> cand(ns);
    std::vector<std::vector<long>> best(ns);
    std::vector<std::vector<std::array<int, 2>>> arg(ns);
    for (int s = 0; s < ns; ++s) {
        const StepDesc &st = cp.steps[s];
        cand[s] = layout_candidates(st.rc);
        best[s].assign(cand[s].size(), -1);
        arg[s].assign(cand[s].size(), {-1, -1});
        const long nkk = 1l << st.nk;
        for (size_t i = 0; i < cand[s].size(); ++i) {
            const RLay &L = cand[s][i];
            long total = (1l << (st.rc - lane_bits(L))) * (nkk + 2);
            bool ok = true;
            for (int side = 0; side < 2 && ok; ++side) {
                const int p = rp.prod[side][s];
                if (p < 0) {
                    total += fetch_cost(st, side, L, nullptr);
                    continue;
                }
                long bc = -1;
                int bj = -1;
                for (size_t j = 0; j < cand[p].size(); ++j) {
                    if (best[p][j] < 0) continue;
                    const long c = fetch_cost(st, side, L, &cand[p][j]);
                    if (c < 0) continue;
                    if (bc < 0 || c + best[p][j] < bc) {
                        bc = c + best[p][j];
                        bj = (int)j;
                    }
                }
                if (bj < 0) ok = false;
                total += bc;
                arg[s][i][side] = bj;
            }
            if (ok) best[s][i] = total;
        }
    }
    int bi = -1;
    for (size_t i = 0; i < cand[rp.root].size(); ++i)
        if (best[rp.root][i] >= 0 && (bi < 0 || best[rp.root][i] < best[rp.root][bi])) bi = (int)i;
    if (bi < 0) return false;
    std::vector<int> choice(ns, -1);
    choice[rp.root] = bi;
    for (int s = ns - 1; s >= 0; --s) {
        if (choice[s] < 0) return false;
        rp.lay[s] = cand[s][choice[s]];
        for (int side = 0; side < 2; ++side)
            if (rp.prod[side][s] >= 0) choice[rp.prod[side][s]] = arg[s][choice[s]][side];
    }
    return true;
}

bool reg_plan(const CompiledPlan &cp, RegPlan &rp) {
    const int ns = (int)cp.steps.size();
    if (ns == 0) return rp.err = "no per-syndrome steps", false;
    if (cp.n_out > 4) return rp.err = "more than 16 classes", false;
    rp.prod[0].assign(ns, -1);
    rp.prod[1].assign(ns, -1);
    rp.lay.assign(ns, RLay());
    std::vector<int> uses(ns, 0);
    for (int s = 0; s < ns; ++s) {
        const StepDesc &st = cp.steps[s];
        if (st.rc > 16) return rp.err = "tensor too wide for registers", false;
        for (int side = 0; side < 2; ++side) {
            const OperandDesc &op = side ? st.b : st.a;
            if (op.kind != kInter) continue;
            int p = -1;
            for (int j = s - 1; j >= 0; --j)
                if (cp.steps[j].c_off == op.off) {
                    p = j;
                    break;
                }
            if (p < 0 || cp.steps[p].rc != op.rank) return rp.err = "operand without a producing step", false;
            rp.prod[side][s] = p;
            ++uses[p];
        }
        if (st.is_root) rp.root = s;
    }
    if (rp.root < 0 || cp.steps[rp.root].c_off != cp.root_off || cp.steps[rp.root].rc != cp.n_out)
        return rp.err = "root step not found", false;
    for (int s = 0; s < ns; ++s)
        if (uses[s] != (s == rp.root ? 0 : 1)) return rp.err = "step output not used exactly once", false;
    if (cp.reg_checked && cp.reg_lay.empty()) return rp.err = cp.reg_err, false;
    if (cp.reg_lay.size() == 2 * (size_t)ns) {
        // prepared with the plan (spec_reg_prepare)
        for (int s = 0; s < ns; ++s) {
            rp.lay[s].pos[0] = cp.reg_lay[2 * s];
            rp.lay[s].pos[1] = cp.reg_lay[2 * s + 1];
        }
        return reg_peak_ok(cp, rp);
    }
    if (reg_layout_dp(cp, rp) && reg_peak_ok(cp, rp)) return true;
    rp.err.clear();
    rp.peak = 0;
    // greedy layouts, consumers first
    RLay &lr = rp.lay[rp.root];
    for (int t = 0; t < 2; ++t)
        if (cp.steps[rp.root].rc > t) lr.pos[t] = t;
    for (int s = ns - 1; s >= 0; --s) {
        const StepDesc &st = cp.steps[s];
        const RLay LC = rp.lay[s];
        for (int side = 0; side < 2; ++side) {
            const int p = rp.prod[side][s];
            if (p < 0) continue;
            const OperandDesc &op = side ? st.b : st.a;
            RLay L;
            for (int t = 0; t < 2; ++t)
                if (LC.pos[t] >= 0) L.pos[t] = seg_image(op, LC.pos[t]);
            // free slots: the bit that feeds most of the producer's intermediate operands
            const StepDesc &ps = cp.steps[p];
            for (int t = 0; t < 2; ++t) {
                if (L.pos[t] >= 0) continue;
                int best = -1, best_score = -1;
                for (int v = 0; v < op.rank; ++v) {
                    if (v == L.pos[0] || v == L.pos[1]) continue;
                    int score = 0;
                    if (ps.a.kind == kInter && seg_image(ps.a, v) >= 0) ++score;
                    if (ps.b.kind == kInter && seg_image(ps.b, v) >= 0) ++score;
                    if (score > best_score) {
                        best = v;
                        best_score = score;
                    }
                }
                L.pos[t] = best;
            }
            rp.lay[p] = L;
        }
    }
    return reg_peak_ok(cp, rp);
}

// Expression for entry `idx` (lane-independent bits) of intermediate X, made
// by step p with layout LX, read by a consumer with output layout LC.
bool reg_fetch(const OperandDesc &op, const RLay &LC, const RLay &LX, int p, uint32_t idx, std::string &out,
               std::string &err) {
    const int ctrl = fetch_ctrl(op, LC, LX, idx);
    if (ctrl < 0) return err = "lane bit of an operand moved", false;
    const std::string name = "x" + std::to_string(p) + "_" + std::to_string(pext(idx, reg_bits(LX, op.rank)));
    out = (ctrl == kIdentityCtrl) ? name : "QP(" + name + ", " + std::to_string(ctrl) + ")";
    return true;
}

// Plan-time scaling of the register-resident FP32 kernel (static mode).
//
// Every entry of every table is non-negative, so contracting the plan in
// FP64 with each sliced table replaced by the sum of its variants (the
// evidence of all syndromes at once) bounds every intermediate entry of every
// syndrome from above (ub).  The kernel stores intermediate X as
// X_true * 2^-E_X, E_X being the sum of the folded exponents of X's tables
// plus any static shifts below it.  A step shifts its output (one multiply
// per output register) only where its pre-shift bound
// log2 max ub - (E_a + E_b) would leave [kShiftLo, kShiftHi], so every value
// and every partial sum stays below 2^127 whatever the syndrome, and the
// plan is refused (the dynamic mode is used instead) when a bound cannot be
// kept below kOverflow.
//
// Underflow: an operation whose result drops below FP32's normal range errs
// by at most 2^-126 of its unit 2^E; carried to the root by the environment
// bound (env: the adjoint of the ub contraction, again over all syndromes),
// all of them together err by at most errsum in absolute terms (true
// units), also counting the FP32 rounding of the tables.  A syndrome whose
// largest root entry is below 2^theta, theta = log2 errsum + kErrBits, is
// re-checked in FP64, so accepted marginals err by at most 2^-kErrBits of
// the row maximum from underflow: one comparison at the root replaces the
// per-step maxima, exponents and rescales of the dynamic mode.
struct RegScale {
    std::vector<int> shift;   // per step: output scaled by 2^-shift
    std::vector<int> E;       // per step: exponent of the stored values
    std::vector<double> pre;  // per step: pre-shift bound, log2
    int e_root = 0;
    double theta = 0;         // log2 of the smallest root maximum accepted without re-check
};

constexpr double kShiftHi = 60, kShiftLo = -30, kOverflow = 120, kErrBits = 21;

// True values (in memory order) of a table operand, summed over its
// variants for a sliced one.
std::vector<double> table_bound(const CompiledPlan &cp, const OperandDesc &op) {
    const size_t n = size_t(1) << op.rank;
    std::vector<double> v(n, 0.0);
    const double sc = std::ldexp(1.0, op.cexp);
    if (op.kind == kConst) {
        for (size_t i = 0; i < n; ++i) v[i] = cp.const_pool[(size_t)op.off + i] * sc;
    } else {
        for (uint32_t var = 0; var < (1u << op.nev); ++var)
            for (size_t i = 0; i < n; ++i) v[i] += cp.sliced_pool[(size_t)op.off + ((size_t)var << op.rank) + i] * sc;
    }
    return v;
}

bool reg_static_scales(const CompiledPlan &cp, const RegPlan &rp, RegScale &rs, std::string &err) {
    const int ns = (int)cp.steps.size();
    std::vector<std::vector<double>> ub(ns), env(ns);
    std::vector<std::array<std::vector<double>, 2>> opv(ns);
    std::vector<std::array<int, 2>> opE(ns);
    rs.shift.assign(ns, 0);
    rs.E.assign(ns, 0);
    rs.pre.assign(ns, 0);
    auto lg = [](const std::vector<double> &v) {
        double m = 0;
        for (double x : v) m = std::max(m, x);
        return m > 0 ? std::log2(m) : -HUGE_VAL;
    };
    // forward: upper bounds and exponents
    for (int s = 0; s < ns; ++s) {
        const StepDesc &st = cp.steps[s];
        for (int side = 0; side < 2; ++side) {
            const OperandDesc &op = side ? st.b : st.a;
            const int p = rp.prod[side][s];
            opv[s][side] = p >= 0 ? ub[p] : table_bound(cp, op);
            opE[s][side] = p >= 0 ? rs.E[p] : op.cexp;
            const double single = lg(opv[s][side]) - opE[s][side] + op.nsingle;
            if (single > kOverflow) return err = "operand bound too large for static scaling", false;
        }
        const OperandDesc &A = st.a, &B = st.b;
        const std::vector<double> &va = opv[s][0], &vb = opv[s][1];
        std::vector<double> &c = ub[s];
        c.assign(size_t(1) << st.rc, 0.0);
        for (uint32_t ci = 0; ci < (1u << st.rc); ++ci) {
            const uint32_t ia = gather_host(ci, A), ib = gather_host(ci, B);
            double acc = 0;
            for (uint32_t k = 0; k < (1u << st.nk); ++k) {
                double sa = 0, sb = 0;
                for (uint32_t q = 0; q < (1u << A.nsingle); ++q) sa += va[ia | (q << A.sshift) | (k << A.kshift)];
                for (uint32_t q = 0; q < (1u << B.nsingle); ++q) sb += vb[ib | (q << B.sshift) | (k << B.kshift)];
                acc += sa * sb;
            }
            c[ci] = acc;
        }
        const int eab = opE[s][0] + opE[s][1];
        const double pre = lg(c) - eab;
        if (!(pre < HUGE_VAL)) return err = "non-finite bound", false;
        rs.pre[s] = pre;
        int sh = 0;
        if (pre > kShiftHi || (pre < kShiftLo && pre > -HUGE_VAL)) sh = (int)std::floor(pre);
        if (pre > kOverflow) return err = "step bound too large for static scaling", false;
        rs.shift[s] = sh;
        rs.E[s] = eab + sh;
    }
    // adjoint: environment bounds, and the underflow error they carry
    env[rp.root].assign(size_t(1) << cp.steps[rp.root].rc, 1.0);
    double errsum = 0;
    for (int s = ns - 1; s >= 0; --s) {
        const StepDesc &st = cp.steps[s];
        const OperandDesc &A = st.a, &B = st.b;
        const std::vector<double> &va = opv[s][0], &vb = opv[s][1], &ec = env[s];
        std::vector<double> ea(va.size(), 0.0), eb(vb.size(), 0.0);
        double envsum = 0;
        for (uint32_t ci = 0; ci < (1u << st.rc); ++ci) {
            envsum += ec[ci];
            const uint32_t ia = gather_host(ci, A), ib = gather_host(ci, B);
            for (uint32_t k = 0; k < (1u << st.nk); ++k) {
                double sa = 0, sb = 0;
                for (uint32_t q = 0; q < (1u << A.nsingle); ++q) sa += va[ia | (q << A.sshift) | (k << A.kshift)];
                for (uint32_t q = 0; q < (1u << B.nsingle); ++q) sb += vb[ib | (q << B.sshift) | (k << B.kshift)];
                for (uint32_t q = 0; q < (1u << A.nsingle); ++q) ea[ia | (q << A.sshift) | (k << A.kshift)] += ec[ci] * sb;
                for (uint32_t q = 0; q < (1u << B.nsingle); ++q) eb[ib | (q << B.sshift) | (k << B.kshift)] += ec[ci] * sa;
            }
        }
        // this step's operations: per output entry 2^nk products and the
        // single-side sums, each within 2^-126 of its unit when it underflows
        const double ops = std::ldexp(1.0, st.nk) * (1.0 + std::ldexp(1.0, A.nsingle) + std::ldexp(1.0, B.nsingle));
        const int unit = opE[s][0] + opE[s][1] + std::max(rs.shift[s], 0);
        errsum += ops * envsum * std::ldexp(1.0, unit - 126);
        for (int side = 0; side < 2; ++side) {
            const std::vector<double> &e = side ? eb : ea;
            const int p = rp.prod[side][s];
            if (p >= 0) {
                env[p] = e;
            } else {
                // FP32 rounding of the table's entries below the normal range
                double es = 0;
                for (double x : e) es += x;
                errsum += es * std::ldexp(1.0, opE[s][side] - 126);
            }
        }
    }
    if (!std::isfinite(errsum)) return err = "non-finite error bound", false;
    rs.e_root = rs.E[rp.root];
    rs.theta = errsum > 0 ? std::log2(errsum) + kErrBits : -HUGE_VAL;
    return true;
}

// One step of the register-resident kernel.  FP32 with `rs`: the static
// exponents and shifts of reg_static_scales; FP32 without: per-tensor
// power-of-two scaling from each operand's maximum (the generic FP32
// kernel's); FP64 (`fp64`): no scaling, the constant exponents folded out of
// table data restored per entry (the generic FP64 kernel's).
bool emit_reg_step(std::ostringstream &o, const CompiledPlan &cp, const RegPlan &rp, int s, size_t const_n,
                   bool fp64, const RegScale *rs, std::string &err) {
    const StepDesc &st = cp.steps[s];
    const RLay &LC = rp.lay[s];
    const std::vector<int> RB = reg_bits(LC, st.rc);
    const uint32_t nreg = 1u << RB.size();
    const uint32_t nkk = 1u << st.nk;
    const char *T = fp64 ? "double" : "float";
    const char *zero = fp64 ? "0.0" : "0.f";
    const bool dyn = !fp64 && !rs;
    if (dyn) o << "    float m" << s << ";\n    int e" << s << ";\n";
    o << "    " << T;
    for (uint32_t r = 0; r < nreg; ++r) o << (r ? ", " : " ") << "x" << s << "_" << r;
    o << ";\n";
    o << "    { // step " << s << ": rank " << st.rc << ", contracted " << st.nk << ", " << nreg
      << " registers per lane";
    if (rs) o << "; bound 2^" << (int)std::ceil(rs->pre[s]) << " before a shift of " << rs->shift[s] << ", exponent "
              << rs->E[s];
    o << "\n";
    const char *ab[2] = {"a", "b"};
    for (int side = 0; side < 2 && dyn; ++side) {
        const OperandDesc &op = side ? st.b : st.a;
        const int p = rp.prod[side][s];
        if (op.kind == kInter)
            o << "      const int n" << ab[side] << " = (m" << p << " > 0.f) ? ilogbf(m" << p << ") : 0, e" << ab[side]
              << " = e" << p << " + n" << ab[side] << ";\n";
        else
            o << "      const int n" << ab[side] << " = 0, e" << ab[side] << " = " << op.cexp << ";\n";
    }
    if (dyn) o << "      const int shift = -(na + nb);\n";
    if (dyn) {
        std::string cond;
        if (st.a.kind == kInter) cond += "na < -100";
        if (st.b.kind == kInter) cond += std::string(cond.empty() ? "" : " || ") + "nb < -100";
        if (!cond.empty()) o << "      if (" << cond << ") lowf = 1;\n";
    }
    bool has_lane_off[2] = {false, false};
    for (int side = 0; side < 2; ++side) {
        const OperandDesc &op = side ? st.b : st.a;
        if (op.kind == kInter) continue;
        if (op.kind == kConst) {
            o << "      const " << T << " *p" << ab[side] << " = tab + " << (long long)op.off << ";\n";
        } else {
            o << "      const unsigned v" << ab[side] << " = 0u";
            for (int j = 0; j < op.nev; ++j) o << " | (s" << op.ev_bit[j] << " << " << j << ")";
            o << ";\n      const " << T << " *p" << ab[side] << " = tab + " << (long long)(const_n + (size_t)op.off) << " + (v"
              << ab[side] << " << " << op.rank << ");\n";
        }
        std::string lo;
        for (int t = 0; t < 2; ++t) {
            if (LC.pos[t] < 0) continue;
            const uint32_t gbit = gather_host(1u << LC.pos[t], op);
            if (!gbit) continue;
            int sh = 0;
            while ((1u << sh) != gbit) ++sh;
            lo += std::string(lo.empty() ? "" : " | ") + "((((unsigned)q >> " + std::to_string(t) + ") & 1u) << " +
                  std::to_string(sh) + ")";
        }
        if (!lo.empty()) {
            o << "      const unsigned l" << ab[side] << " = " << lo << ";\n";
            has_lane_off[side] = true;
        }
    }
    o << "      " << T;
    for (uint32_t r = 0; r < nreg; ++r) o << (r ? ", " : " ") << "acc" << r << " = " << zero;
    o << ";\n";
    int ntmp = 0;
    for (uint32_t k = 0; k < nkk; ++k) {
        o << "      {\n";
        std::vector<std::pair<std::string, std::string>> cache;   // expression -> temporary
        auto value = [&](int side, uint32_t cst, std::string &out) -> bool {
            const OperandDesc &op = side ? st.b : st.a;
            const uint32_t idx = gather_host(cst, op) | (k << op.kshift);
            std::string expr;
            if (op.kind != kInter) {
                expr = std::string("p") + ab[side] + "[" + (has_lane_off[side] ? std::string("l") + ab[side] + " + " : "") +
                       std::to_string(idx) + "u]";
            } else {
                const int p = rp.prod[side][s];
                if (op.nsingle == 0) {
                    if (!reg_fetch(op, LC, rp.lay[p], p, idx, expr, err)) return false;
                    if (expr.compare(0, 3, "QP(") != 0) {
                        out = expr;
                        return true;
                    }
                } else {
                    // single-side labels summed first, ascending (as the generic kernel)
                    expr = "S";
                    std::vector<std::string> terms;
                    for (uint32_t qq = 0; qq < (1u << op.nsingle); ++qq) {
                        std::string t1;
                        if (!reg_fetch(op, LC, rp.lay[p], p, idx | (qq << op.sshift), t1, err)) return false;
                        terms.push_back(t1);
                        expr += "|" + t1;
                    }
                    for (auto &c : cache)
                        if (c.first == expr) {
                            out = c.second;
                            return true;
                        }
                    const std::string nm = "f" + std::to_string(ntmp++);
                    o << "        " << T << " " << nm << " = " << zero << ";\n";
                    for (auto &t1 : terms) o << "        " << nm << " += " << t1 << ";\n";
                    cache.emplace_back(expr, nm);
                    out = nm;
                    return true;
                }
            }
            for (auto &c : cache)
                if (c.first == expr) {
                    out = c.second;
                    return true;
                }
            const std::string nm = "f" + std::to_string(ntmp++);
            o << "        const " << T << " " << nm << " = " << expr << ";\n";
            cache.emplace_back(expr, nm);
            out = nm;
            return true;
        };
        for (uint32_t r = 0; r < nreg; ++r) {
            const uint32_t cst = pdep(r, RB);
            std::string va, vb;
            if (!value(0, cst, va) || !value(1, cst, vb)) return false;
            o << "        acc" << r << " = " << (fp64 ? "fma(" : "fmaf(") << va << ", " << vb << ", acc" << r << ");\n";
        }
        o << "      }\n";
    }
    if (fp64) {
        const int cexp = st.a.cexp + st.b.cexp;
        for (uint32_t r = 0; r < nreg; ++r) {
            o << "      x" << s << "_" << r << " = ";
            if (cexp != 0)
                o << "ldexp(acc" << r << ", " << cexp << ");\n";
            else
                o << "acc" << r << ";\n";
        }
        o << "    }\n";
        return true;
    }
    if (rs) {
        const int sh = rs->shift[s];
        for (uint32_t r = 0; r < nreg; ++r) {
            o << "      x" << s << "_" << r << " = ";
            if (sh != 0)
                o << "ldexpf(acc" << r << ", " << -sh << ");\n";
            else
                o << "acc" << r << ";\n";
        }
        o << "    }\n";
        return true;
    }
    // Entries are >= 0 or NaN, and NaN-ignoring fmaxf over them equals the
    // generic kernel's running `x > m ? x : m` from 0 (max is exact and order
    // free); a tree of three-input maxima halves the instructions.
    std::vector<std::string> mx;
    for (uint32_t r = 0; r < nreg; ++r) {
        o << "      x" << s << "_" << r << " = ldexpf(acc" << r << ", shift);\n";
        mx.push_back("x" + std::to_string(s) + "_" + std::to_string(r));
    }
    mx.push_back("0.f");
    int nm = 0;
    while (mx.size() > 1) {
        std::vector<std::string> nx;
        for (size_t i = 0; i < mx.size(); i += 3) {
            if (i + 1 >= mx.size()) {
                nx.push_back(mx[i]);
                continue;
            }
            const std::string t = "mx" + std::to_string(nm++);
            o << "      const float " << t << " = ";
            if (i + 2 < mx.size())
                o << "fmaxf(fmaxf(" << mx[i] << ", " << mx[i + 1] << "), " << mx[i + 2] << ");\n";
            else
                o << "fmaxf(" << mx[i] << ", " << mx[i + 1] << ");\n";
            nx.push_back(t);
        }
        mx.swap(nx);
    }
    o << "      float lmax = " << mx[0] << ";\n";
    o << "      lmax = fmaxf(QP(lmax, 177), lmax);\n      lmax = fmaxf(QP(lmax, 78), lmax);\n";
    o << "      m" << s << " = lmax;\n      e" << s << " = ea + eb;\n    }\n";
    return true;
}

void emit_tile_fn(std::ostringstream &o, const CompiledPlan &cp, const SpecCfg &cfg, const std::string &name,
                  bool noinline = false);

// FP64 configuration of the register-resident kernel's inline re-check: one
// syndrome at a time over the whole workgroup, intermediates in LDS.
SpecCfg inline_recheck_cfg(const SpecCfg &cfg) {
    SpecCfg c;
    c.glog = 0;
    c.threads = cfg.threads;
    c.tab_lds = false;
    c.fp64 = true;
    return c;
}

// LDS a workgroup of the register-resident kernel may spend on its inline
// FP64 re-check (beyond it, the separate re-check kernel runs instead).
constexpr size_t kInlineRecheckLds = 48 * 1024;

std::string reg_source(const CompiledPlan &cp, const SpecCfg &cfg, std::string &err) {
    RegPlan rp;
    if (!reg_plan(cp, rp)) {
        err = rp.err;
        return "";
    }
    const size_t const_n = cp.const_pool.size();
    const size_t tab_n = cp.const_pool.size() + cp.sliced_pool.size();
    const int G = cfg.threads / 4;
    const int nout = 1 << cp.n_out;
    const bool fp64 = cfg.fp64;
    const char *T = fp64 ? "double" : "float";
    std::ostringstream o;
    o << "// plan-specialised " << (fp64 ? "FP64" : "FP32")
      << " TNMMAP kernel, one syndrome per quad of lanes, tensors in registers (generated)\n";
    RegScale rsc;
    bool stat = false;
    if (!fp64 && !cfg.dyn) {
        std::string why;
        stat = reg_static_scales(cp, rp, rsc, why);
        if (stat)
            o << "// static exponents: root exponent " << rsc.e_root << ", FP64 re-check below 2^"
              << (int)std::ceil(rsc.theta) << " (true units)\n";
        else
            o << "// static scaling not applicable (" << why << "): per-step scaling\n";
    }
    const bool dyn = !fp64 && !stat;
    if (fp64)
        o << "#define QP(x, c) __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(x), (c), 0xf, 0xf, false), "
             "__builtin_amdgcn_mov_dpp(__double2loint(x), (c), 0xf, 0xf, false))\n";
    else
        o << "#define QP(x, c) __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), (c), 0xf, 0xf, false))\n";
    const bool inl = spec_reg_inline_recheck(cp, cfg);
    const SpecCfg c64 = inline_recheck_cfg(cfg);
    // a call of its own, so that its registers do not crowd the FP32 program's
    if (inl) emit_tile_fn(o, cp, c64, "recheck64", true);
    o << "extern \"C\" __global__ void __launch_bounds__(" << cfg.threads;
    if (cfg.waves > 0) o << ", " << cfg.waves;
    o << ") " << spec_kernel_name(cfg) << "(\n"
         "    const unsigned char *__restrict__ syn, long long n, const int *__restrict__ list,\n"
         "    const int *__restrict__ list_count, const " << T << " *__restrict__ gtab, double *__restrict__ mar,\n"
         "    int *__restrict__ cls, unsigned char *__restrict__ ties, int *__restrict__ rl,\n"
         "    int *__restrict__ rcount, long long rcap, double tie_tol, double recheck_tol, " << T
      << " *__restrict__ gslab,\n"
         "    int nseg, int segcap, const double *__restrict__ gtab64) {\n";
    o << "  constexpr int G = " << G << ", NBITS = " << std::max(1, cp.n_syn) << ", NOUT = " << nout << ";\n";
    o << "  (void)gslab; (void)nseg; (void)segcap; (void)gtab64;\n";
    o << "  __shared__ int seg_n;\n";
    if (inl) {
        // rows this tile flags for the FP64 re-check, done by the whole
        // workgroup at the end of the tile
        o << "  __shared__ int flagged[G];\n";
        o << "  __shared__ __attribute__((aligned(16))) unsigned char smem64[" << layout(cp, c64).total << "];\n";
    }
    if (cfg.tab_lds) {
        o << "  __shared__ " << T << " tab[" << std::max<size_t>(1, tab_n) << "];\n";
        o << "  for (int i = threadIdx.x; i < " << tab_n << "; i += " << cfg.threads << ") tab[i] = gtab[i];\n";
        o << "  __syncthreads();\n";
    } else {
        o << "  const " << T << " *tab = gtab;\n";
    }
    // Re-check list in per-tile segments: tile t's flagged rows go to
    // rl[t * rcap ..] and their number to rcount[t], written by every tile
    // the launch may own (those beyond a listed count write 0), so the list
    // needs no clearing between calls.
    o << "  const int q = threadIdx.x & 3;\n"
         "  const long long count = list ? (long long)(*list_count) : n;\n"
         "  const long long ntiles = (count + G - 1) / G;\n"
         "  const long long nseg_all = (n + G - 1) / G;\n"
         "  for (long long t = blockIdx.x; t < nseg_all; t += gridDim.x) {\n"
         "   if (threadIdx.x == 0) seg_n = 0;\n"
         "   __syncthreads();\n"
         "   if (t < ntiles) {\n"
         "    const long long qi = t * G + (threadIdx.x >> 2);\n"
         "    const bool valid = qi < count;\n"
         "    const long long b = valid ? (list ? (long long)list[qi] : qi) : 0ll;\n"
         "    const unsigned char *row = syn + b * NBITS;\n";
    std::vector<char> used(std::max(1, cp.n_syn), 0);
    for (const StepDesc &st : cp.steps)
        for (const OperandDesc *op : {&st.a, &st.b})
            if (op->kind == kSliced)
                for (int j = 0; j < op->nev; ++j) used[op->ev_bit[j]] = 1;
    for (int j = 0; j < cp.n_syn; ++j)
        if (used[j]) o << "    const unsigned s" << j << " = valid ? (unsigned)(row[" << j << "] & 1u) : 0u;\n";
    if (dyn) o << "    int lowf = 0;\n";
    for (int s = 0; s < (int)cp.steps.size(); ++s)
        if (!emit_reg_step(o, cp, rp, s, const_n, fp64, stat ? &rsc : nullptr, err)) return "";
    // every lane gathers the root's entries; the quad's lane 0 finishes the row
    const RLay &LR = rp.lay[rp.root];
    const std::vector<int> RR = reg_bits(LR, cp.steps[rp.root].rc);
    o << "    " << T << " rv[NOUT];\n";
    for (int j = 0; j < nout; ++j) {
        int src[4];
        for (int d = 0; d < 4; ++d) {
            src[d] = 0;
            for (int t = 0; t < 2; ++t) {
                const int bit = LR.pos[t] >= 0 ? ((j >> LR.pos[t]) & 1) : ((d >> t) & 1);
                src[d] |= bit << t;
            }
        }
        const int ctrl = quad_ctrl(src);
        const std::string nm = "x" + std::to_string(rp.root) + "_" + std::to_string(pext((uint32_t)j, RR));
        o << "    rv[" << j << "] = " << (ctrl == kIdentityCtrl ? nm : "QP(" + nm + ", " + std::to_string(ctrl) + ")") << ";\n";
    }
    o << "    if (q == 0 && valid) {\n"
         "      const int e = "
      << (fp64 ? std::string("0") : stat ? std::to_string(rsc.e_root) : "e" + std::to_string(rp.root)) << ";\n"
         "      " << T << " best = 0, second = -1;\n"
         "      int bi = 0;\n"
         "      for (int j = 0; j < NOUT; ++j) {\n"
         "        const " << T << " x = rv[j];\n"
         "        if (mar) mar[b * NOUT + j] = ldexp((double)x, e);\n"
         "        if (j == 0 || x > best) {\n"
         "          if (j > 0) second = best;\n"
         "          best = x;\n"
         "          bi = j;\n"
         "        } else if (x > second) {\n"
         "          second = x;\n"
         "        }\n"
         "      }\n"
         "      if (cls) cls[b] = bi;\n"
         "      const bool low = NOUT > 1 && second >= 0 && (double)(best - second) <= tie_tol * (double)best;\n"
         "      if (ties) ties[b] = low ? 1 : 0;\n";
    std::string low = "lowf != 0";
    if (stat) {
        // the root maximum in stored units against 2^(theta - e_root)
        const double lim = std::ceil(rsc.theta - rsc.e_root);
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.17g", lim < -1074 ? 0.0 : std::ldexp(1.0, (int)std::min(lim, 1000.0)));
        low = std::string("(double)best < ") + buf;
    }
    if (fp64)
        o << "      (void)recheck_tol;\n";
    else if (inl)
        o << "      {\n"
             "        const bool rc = !(best > 0) || " << low << " ||\n"
             "                        (NOUT > 1 && second >= 0 && (double)(best - second) <= recheck_tol * (double)best);\n"
             "        if (rc) flagged[atomicAdd(&seg_n, 1)] = (int)b;\n"
             "      }\n";
    else
        o << "      if (rl) {\n"
             "        const bool rc = !(best > 0) || " << low << " ||\n"
             "                        (NOUT > 1 && second >= 0 && (double)(best - second) <= recheck_tol * (double)best);\n"
             "        if (rc) {\n"
             "          const int slot = atomicAdd(&seg_n, 1);\n"
             "          if (slot < rcap) rl[t * rcap + slot] = (int)b;\n"
             "        }\n"
             "      }\n";
    o << "    }\n"
         "   }\n"
         "   __syncthreads();\n";
    if (inl)
        // at most G flagged rows (one per valid row of the tile), re-contracted
        // in FP64 one after the other by the whole workgroup
        o << "   const int nf = seg_n;\n"
             "   for (int i = 0; i < nf; ++i)\n"
             "     recheck64(syn, nf, flagged, i, gtab64, mar, cls, ties, nullptr, nullptr, 0, tie_tol, recheck_tol, "
             "smem64, nullptr, 0u);\n"
             "   if (threadIdx.x == 0 && rcount) rcount[t] = nf;\n"
             "   __syncthreads();\n";
    else
        o << "   if (threadIdx.x == 0 && rcount) rcount[t] = min(seg_n, (int)rcap);\n";
    o << "  }\n"
         "}\n";
    return o.str();
}

}  // namespace

bool spec_reg_inline_recheck(const CompiledPlan &cp, const SpecCfg &cfg) {
    return cfg.reg && !cfg.fp64 && cp.ws_entries > 0 && layout(cp, inline_recheck_cfg(cfg)).total <= kInlineRecheckLds;
}

size_t spec_lds_bytes(const CompiledPlan &cp, const SpecCfg &cfg) {
    if (cfg.reg) return cfg.tab_lds ? round_up(std::max<size_t>(1, cp.const_pool.size() + cp.sliced_pool.size()) * 4, 16) : 0;
    return layout(cp, cfg).total;
}

bool spec_reg_supported(const CompiledPlan &cp, std::string *why) {
    RegPlan rp;
    const bool ok = reg_plan(cp, rp);
    if (!ok && why) *why = rp.err;
    return ok;
}

void spec_reg_prepare(CompiledPlan &cp) {
    cp.reg_lay.clear();
    cp.reg_checked = false;
    RegPlan rp;
    const bool ok = reg_plan(cp, rp);
    if (ok)
        for (const RLay &L : rp.lay) {
            cp.reg_lay.push_back(L.pos[0]);
            cp.reg_lay.push_back(L.pos[1]);
        }
    cp.reg_err = ok ? std::string() : rp.err;
    cp.reg_checked = true;
}

int spec_reg_peak(const CompiledPlan &cp) {
    RegPlan rp;
    reg_plan(cp, rp);
    return rp.peak;
}

size_t spec_slab_entries(const CompiledPlan &cp, const SpecCfg &cfg) {
    return cfg.slab ? (size_t)std::max<int64_t>(1, cp.ws_entries) << cfg.glog : 0;
}

namespace {

// One tile of G syndromes (rows b0.., or list entries b0..) through the
// whole step program, as a device function: the body of the generic fused
// kernel's tile with every descriptor value substituted.  Starts and ends
// with a workgroup barrier.
void emit_tile_fn(std::ostringstream &o, const CompiledPlan &cp, const SpecCfg &cfg, const std::string &name,
                  bool noinline) {
    const Layout L = layout(cp, cfg);
    const size_t const_n = cp.const_pool.size();
    const int nsteps = (int)cp.steps.size();
    const char *tn = cfg.fp64 ? "double" : "float";
    o << "__device__ " << (noinline ? "__attribute__((noinline))" : "__forceinline__") << " void " << name << "(\n"
      << "    const unsigned char *__restrict__ syn, long long count, const int *__restrict__ list, long long b0,\n"
      << "    const " << tn << " *__restrict__ tab, double *__restrict__ mar, int *__restrict__ cls,\n"
      << "    unsigned char *__restrict__ ties, int *rl, int *rcount, long long rcap, double tie_tol,\n"
      << "    double recheck_tol, unsigned char *smem, " << tn << " *__restrict__ slabw, unsigned tag_base) {\n";
    o << "  using T = " << tn << ";\n";
    o << "  constexpr int G = " << (1 << cfg.glog) << ", GLOG = " << cfg.glog << ", NT = " << cfg.threads
      << ", CSTEP = " << (cfg.threads >> cfg.glog) << ", NBITS = " << std::max(1, cp.n_syn) << ", NOUT = "
      << (1 << cp.n_out) << ", NSLOTS = " << std::max(1, cp.n_slots) << ";\n";
    if (cfg.slab)
        o << "  T *work = slabw;\n";
    else
        o << "  T *work = reinterpret_cast<T *>(smem + " << L.work << ");\n";
    if (!cfg.fp64) {
        o << "  unsigned long long *mx = reinterpret_cast<unsigned long long *>(smem + " << L.mx << ");\n";
        o << "  int *ex = reinterpret_cast<int *>(smem + " << L.ex << ");\n";
        o << "  int *lowf = reinterpret_cast<int *>(smem + " << L.low << ");\n";
    }
    o << "  unsigned char *synl = smem + " << L.syn << ";\n";
    o << "  const int tid = threadIdx.x;\n  const int g = tid & (G - 1);\n  const int c0 = tid >> GLOG;\n"
         "  const int lane = tid & 63;\n  (void)lane; (void)c0; (void)tag_base; (void)rl; (void)rcount; (void)rcap;\n"
         "  (void)recheck_tol; (void)NSLOTS;\n";
    o << "  if (!list) {\n"
         "    const long long rows = min((long long)G, count - b0);\n"
         "    const unsigned char *src = syn + b0 * NBITS;\n"
         "    for (int i = tid; i < G * NBITS; i += NT) {\n"
         "      const int r = i / NBITS, bit = i - r * NBITS;\n"
         "      synl[(bit << GLOG) + r] = (r < rows) ? (src[i] & 1u) : 0u;\n"
         "    }\n"
         "  } else {\n"
         "    for (int i = tid; i < G * NBITS; i += NT) {\n"
         "      const int r = i / NBITS, bit = i - r * NBITS;\n"
         "      const long long q = b0 + r;\n"
         "      synl[(bit << GLOG) + r] = (q < count) ? (syn[(long long)list[q] * NBITS + bit] & 1u) : 0u;\n"
         "    }\n"
         "  }\n";
    if (!cfg.fp64)
        o << "  for (int i = tid; i < NSLOTS * G; i += NT) mx[i] = 0ull;\n"
             "  if (tid < G) lowf[tid] = 0;\n";
    o << "  __syncthreads();\n";
    for (int s = 0; s < nsteps; ++s)
        emit_step(o, cp.steps[s], s, const_n, cfg.threads >> cfg.glog, cfg.fp64, std::max(1, cfg.unroll),
                  ((cfg.mfma || cfg.mfma32) && !cfg.fp64) ? &cfg : nullptr, cfg.slab);
    o << "  if (tid < G) {\n"
         "    const long long q = b0 + tid;\n"
         "    if (q < count) {\n"
         "      const long long b = list ? (long long)list[q] : q;\n";
    if (cfg.fp64)
        o << "      const int e = 0;\n";
    else
        o << "      const int e = ex[" << cp.root_slot << " * G + g];\n";
    o << "      T best = 0, second = -1;\n"
         "      int bi = 0;\n"
         "      for (int j = 0; j < NOUT; ++j) {\n";
    o << "        const T x = work[(" << (long long)cp.root_off << " + j) * G + g];\n";
    o << "        if (mar) mar[b * NOUT + j] = ldexp((double)x, e);\n"
         "        if (j == 0 || x > best) {\n"
         "          if (j > 0) second = best;\n"
         "          best = x;\n"
         "          bi = j;\n"
         "        } else if (x > second) {\n"
         "          second = x;\n"
         "        }\n"
         "      }\n"
         "      if (cls) cls[b] = bi;\n"
         "      const bool low = NOUT > 1 && second >= 0 && (double)(best - second) <= tie_tol * (double)best;\n"
         "      if (ties) ties[b] = low ? 1 : 0;\n";
    if (!cfg.fp64)
        o << "      if (rl) {\n"
         "        const bool rc = !(best > 0) || lowf[tid] != 0 ||\n"
         "                        (NOUT > 1 && second >= 0 && (double)(best - second) <= recheck_tol * (double)best);\n"
         "        if (rc) {\n"
         "          const int slot = atomicAdd(rcount, 1);\n"
         "          if (slot < rcap) rl[slot] = (int)b;\n"
         "        }\n"
         "      }\n";
    o << "    }\n"
         "  }\n"
         "  __syncthreads();\n"
         "}\n\n";
}

std::string dbl(double x) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", x);
    return buf;
}

// Types and helpers of the generated MFMA steps.
const char *kMfmaPrelude =
    "typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;\n"
    "typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;\n"
    "typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;\n"
    "__device__ __forceinline__ unsigned short to_bf16(float f) {\n"
    "  const unsigned u = __float_as_uint(f);\n"
    "  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);\n"
    "}\n\n";

}  // namespace

bool spec_step_uses_mfma32(const StepDesc &st, const SpecCfg &cfg) {
    if (!cfg.mfma32 || cfg.mfma || cfg.fp64 || cfg.reg || cfg.glog != 0 || cfg.threads % 64 != 0) return false;
    if ((st.a.kind == kInter && st.a.nsingle) || (st.b.kind == kInter && st.b.nsingle)) return false;
    std::vector<int> aon, bon;
    mfma_side_bits(st, aon, bon);
    if (aon.size() >= 5 && bon.size() >= 5) return st.nk >= 1;   // 32 x 32 x 2
    return aon.size() >= 4 && bon.size() >= 4 && st.nk >= 2;      // 16 x 16 x 4
}

bool spec_plan_uses_mfma32(const CompiledPlan &cp, const SpecCfg &cfg) {
    for (const StepDesc &st : cp.steps)
        if (spec_step_uses_mfma32(st, cfg)) return true;
    return false;
}

bool spec_step_uses_mfma(const StepDesc &st, const SpecCfg &cfg) {
    if (!cfg.mfma || cfg.fp64 || cfg.glog != 0 || cfg.threads % 64 != 0) return false;
    if (st.nk < 4) return false;
    if ((st.a.kind == kInter && st.a.nsingle) || (st.b.kind == kInter && st.b.nsingle)) return false;
    const uint32_t am = source_mask(st.a), bm = source_mask(st.b);
    const uint32_t outm = (st.rc >= 32) ? 0xffffffffu : ((1u << st.rc) - 1u);
    return __builtin_popcount(am & ~bm & outm) >= 5 && __builtin_popcount(bm & ~am & outm) >= 5;
}

std::string spec_source(const CompiledPlan &cp, const SpecCfg &cfg) {
    if (cfg.reg) {
        std::string err;
        return reg_source(cp, cfg, err);
    }
    const Layout L = layout(cp, cfg);
    const int nsteps = (int)cp.steps.size();
    std::ostringstream o;
    o << "// plan-specialised " << (cfg.fp64 ? "FP64" : "FP32") << " TNMMAP kernel (generated)\n";
    if (cfg.mfma || cfg.mfma32) o << kMfmaPrelude;
    emit_tile_fn(o, cp, cfg, "tile");
    const char *tn = cfg.fp64 ? "double" : "float";
    const int G = 1 << cfg.glog;
    o << "extern \"C\" __global__ void __launch_bounds__(" << cfg.threads;
    if (cfg.waves > 0) o << ", " << cfg.waves;
    o << ") " << spec_kernel_name(cfg) << "(\n"
      << "    const unsigned char *__restrict__ syn, long long n, const int *__restrict__ list,\n"
      << "    const int *__restrict__ list_count, const " << tn << " *__restrict__ gtab, double *__restrict__ mar,\n"
      << "    int *__restrict__ cls, unsigned char *__restrict__ ties, int *__restrict__ rl,\n"
      << "    int *__restrict__ rcount, long long rcap, double tie_tol, double recheck_tol, " << tn
      << " *__restrict__ gslab, int nseg, int segcap, const double *__restrict__ gtab64) {\n";
    o << "  (void)gtab64;\n";
    o << "  __shared__ __attribute__((aligned(16))) unsigned char smem[" << L.total << "];\n";
    if (cfg.tab_lds) {
        o << "  " << tn << " *tab = reinterpret_cast<" << tn << " *>(smem + " << L.tab << ");\n";
        o << "  for (int i = threadIdx.x; i < " << L.tab_entries << "; i += " << cfg.threads << ") tab[i] = gtab[i];\n";
    } else {
        o << "  const " << tn << " *tab = gtab;\n";
    }
    if (cfg.slab)
        o << "  " << tn << " *slabw = gslab + (unsigned long long)blockIdx.x * " << spec_slab_entries(cp, cfg) << "ull;\n";
    else
        o << "  " << tn << " *slabw = nullptr;\n";
    o << "  unsigned iter = 0;\n";
    // nseg > 0: the rows listed in per-tile segments (the register-resident
    // kernel's re-check list): segment w holds min(list_count[w], segcap)
    // rows at list[w * segcap ..]; each is decoded as row b0 of the syndrome
    // array (any row: a listed call's rows may lie beyond its count)
    o << "  if (nseg > 0) {\n"
         "    for (long long w = blockIdx.x; w < nseg; w += gridDim.x) {\n"
         "      const int c = min(list_count[w], segcap);\n"
         "      for (int e = 0; e < c; ++e, ++iter) {\n"
         "        const long long row = (long long)list[w * segcap + e];\n"
         "        tile(syn, row + 1, nullptr, row, tab, mar, cls, ties, rl, rcount, rcap, tie_tol, recheck_tol, smem, slabw,"
         " iter * " << nsteps << "u + 1u);\n"
         "      }\n"
         "    }\n"
         "    return;\n"
         "  }\n";
    o << "  const long long count = list ? (long long)(*list_count) : n;\n";
    o << "  const long long ntiles = (count + " << G - 1 << ") >> " << cfg.glog << ";\n";
    o << "  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x, ++iter)\n";
    o << "    tile(syn, count, list, t << " << cfg.glog << ", tab, mar, cls, ties, rl, rcount, rcap, tie_tol, recheck_tol,"
         " smem, slabw, iter * " << nsteps << "u + 1u);\n";
    o << "}\n";
    return o.str();
}

std::string spec_map_source(const CompiledPlan &cp, int glog, int threads) {
    const int G = 1 << glog;
    const int ns = (int)cp.steps.size();
    const int words = (cp.n_vars + 31) / 32;
    const size_t work_b = (size_t)std::max<int64_t>(1, cp.ws_entries) * G * 8;
    const size_t arg_off = round_up(work_b, 16);
    const size_t syn_off = arg_off + round_up((size_t)std::max<int64_t>(1, cp.arg_entries) * G * 2, 16);
    const size_t total = round_up(syn_off + (size_t)std::max(1, cp.n_syn) * G, 16);
    std::ostringstream o;
    o << "// plan-specialised max-product (TNMAP) kernel (generated)\n";
    o << "extern \"C\" __global__ void __launch_bounds__(" << threads << ") tqec_map_spec(\n"
         "    const unsigned char *__restrict__ syn, long long n, const double *__restrict__ cpool,\n"
         "    const double *__restrict__ spool, unsigned char *__restrict__ config, int n_config,\n"
         "    double *__restrict__ logp) {\n";
    o << "  constexpr int G = " << G << ", GLOG = " << glog << ", NT = " << threads << ", CSTEP = " << (threads >> glog)
      << ", NBITS = " << std::max(1, cp.n_syn) << ", NV = " << cp.n_vars << ";\n";
    o << "  __shared__ __attribute__((aligned(16))) unsigned char smem[" << total << "];\n";
    o << "  double *work = reinterpret_cast<double *>(smem);\n";
    o << "  unsigned short *arg = reinterpret_cast<unsigned short *>(smem + " << arg_off << ");\n";
    o << "  unsigned char *synl = smem + " << syn_off << ";\n";
    o << "  const int tid = threadIdx.x, g = tid & (G - 1), c0 = tid >> GLOG;\n";
    o << "  const long long ntiles = (n + G - 1) >> GLOG;\n";
    o << "  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {\n"
         "    const long long b0 = t << GLOG;\n"
         "    const long long rows = min((long long)G, n - b0);\n"
         "    const unsigned char *src = syn + b0 * NBITS;\n"
         "    for (int i = tid; i < G * NBITS; i += NT) {\n"
         "      const int r = i / NBITS, bit = i - r * NBITS;\n"
         "      synl[(bit << GLOG) + r] = (r < rows) ? (src[i] & 1u) : 0u;\n"
         "    }\n"
         "    __syncthreads();\n";
    // forward: every output keeps the first maximum over (k, sa, sb), as the generic kernel
    for (int s = 0; s < ns; ++s) {
        const StepDesc &st = cp.steps[s];
        const MapStep &ms = cp.map_steps[s];
        o << "    { // step " << s << ": rank " << st.rc << ", eliminated " << (st.nk + st.a.nsingle + st.b.nsingle)
          << "\n";
        const char *nm[2] = {"pa", "pb"};
        for (int side = 0; side < 2; ++side) {
            const OperandDesc &op = side ? st.b : st.a;
            if (op.kind == kInter) {
                o << "      const double *" << nm[side] << " = work + (" << (long long)op.off << "ll << GLOG) + g;\n";
            } else if (op.kind == kConst) {
                o << "      const double *" << nm[side] << " = cpool + " << (long long)op.off << ";\n";
            } else {
                o << "      unsigned v" << nm[side] << " = 0u;\n";
                for (int j = 0; j < op.nev; ++j)
                    o << "      v" << nm[side] << " |= (unsigned)synl[(" << op.ev_bit[j] << " << GLOG) + g] << " << j
                      << ";\n";
                o << "      const double *" << nm[side] << " = spool + " << (long long)op.off << " + ((long long)v"
                  << nm[side] << " << " << op.rank << ");\n";
            }
        }
        auto ld = [&](const OperandDesc &op, const char *base, const std::string &idx) {
            return op.kind == kInter ? std::string(base) + "[(" + idx + ") << GLOG]" : std::string(base) + "[" + idx + "]";
        };
        const int cexp = st.a.cexp + st.b.cexp;
        o << "      for (int c = c0; c < " << (1 << st.rc) << "; c += CSTEP) {\n";
        o << "        const unsigned ar = " << subst_c(gather_expr(st.a), "(unsigned)c") << ";\n";
        o << "        const unsigned br = " << subst_c(gather_expr(st.b), "(unsigned)c") << ";\n";
        o << "        double best = -1.0;\n        unsigned bi = 0u;\n";
        o << "        for (unsigned k = 0; k < " << (1u << st.nk) << "u; ++k)\n";
        o << "          for (unsigned qa = 0; qa < " << (1u << st.a.nsingle) << "u; ++qa) {\n";
        o << "            const double va = "
          << ld(st.a, "pa", "ar | (qa << " + std::to_string(st.a.sshift) + ") | (k << " + std::to_string(st.a.kshift) + ")")
          << ";\n";
        o << "            for (unsigned qb = 0; qb < " << (1u << st.b.nsingle) << "u; ++qb) {\n";
        o << "              const double vb = "
          << ld(st.b, "pb", "br | (qb << " + std::to_string(st.b.sshift) + ") | (k << " + std::to_string(st.b.kshift) + ")")
          << ";\n";
        o << "              const double v = va * vb;\n";
        o << "              if (v > best) {\n                best = v;\n                bi = k | (qa << " << st.nk
          << ") | (qb << " << (st.nk + st.a.nsingle) << ");\n              }\n";
        o << "            }\n          }\n";
        if (cexp != 0) o << "        best = ldexp(best, " << cexp << ");\n";
        o << "        work[((" << (long long)st.c_off << "ll + c) << GLOG) + g] = best;\n";
        o << "        arg[((" << (long long)ms.arg_off << "ll + c) << GLOG) + g] = (unsigned short)bi;\n";
        o << "      }\n    }\n    __syncthreads();\n";
    }
    // backtrace: one thread per syndrome, from the root down, labels as constants
    o << "    if (tid < G && b0 + tid < n) {\n";
    o << "      unsigned as[" << std::max(1, words) << "];\n";
    o << "      #pragma unroll\n      for (int w = 0; w < " << std::max(1, words) << "; ++w) as[w] = 0u;\n";
    for (int s = ns - 1; s >= 0; --s) {
        const MapStep &ms = cp.map_steps[s];
        const int32_t *L = cp.map_labels.data() + ms.off;
        o << "      {\n        unsigned c = 0u;\n";
        for (int j = 0; j < ms.nc; ++j)
            o << "        c |= ((as[" << (L[j] >> 5) << "] >> " << (L[j] & 31) << ") & 1u) << " << j << ";\n";
        o << "        const unsigned rec = arg[((" << (long long)ms.arg_off << "ll + c) << GLOG) + tid];\n";
        const int ne = ms.nk + ms.nsa + ms.nsb;
        for (int j = 0; j < ne; ++j) {
            const int l = L[ms.nc + j];
            o << "        as[" << (l >> 5) << "] |= ((rec >> " << j << ") & 1u) << " << (l & 31) << ";\n";
        }
        o << "      }\n";
    }
    o << "      const long long b = b0 + tid;\n";
    o << "      #pragma unroll\n      for (int v = 0; v < NV; ++v)\n"
         "        if (v < n_config) config[b * n_config + v] = (unsigned char)((as[v >> 5] >> (v & 31)) & 1u);\n";
    o << "      if (logp) logp[b] = log(work[(" << (long long)cp.root_off << "ll << GLOG) + tid]);\n";
    o << "    }\n    __syncthreads();\n  }\n}\n";
    return o.str();
}

size_t spec_mixed_lds_bytes(const std::vector<MixedSpecPlan> &plans) {
    size_t m = 16;
    for (const MixedSpecPlan &q : plans) {
        m = std::max(m, layout(*q.cp, q.f32).total);
        m = std::max(m, layout(*q.cp, q.f64).total);
    }
    return m;
}

std::string spec_mixed_source(const std::vector<MixedSpecPlan> &plans, int threads, size_t slab32_entries,
                              size_t slab64_entries) {
    std::ostringstream o;
    o << "// mixed-code kernel: plan-specialised tiles of " << plans.size() << " plans (generated)\n";
    bool any_mfma = false;
    for (const MixedSpecPlan &q : plans) any_mfma = any_mfma || q.f32.mfma || q.f32.mfma32;
    if (any_mfma) o << kMfmaPrelude;
    // the layouts of the host's batch table, header and counters
    o << "struct MBatch { const unsigned char *syn; double *mar; int *cls; unsigned char *tie; long long n;\n"
         "                long long tile0; int plan; int glog; };\n"
         "struct MHeader { int nbatches; int pad; long long ntiles; };\n"
         "struct MCounters { unsigned long long next_tile; unsigned long long next_recheck; int rl_count; int pad;\n"
         "                   unsigned long long t_start; };\n\n";
    for (size_t i = 0; i < plans.size(); ++i) {
        emit_tile_fn(o, *plans[i].cp, plans[i].f32, "p" + std::to_string(i) + "_f32");
        emit_tile_fn(o, *plans[i].cp, plans[i].f64, "p" + std::to_string(i) + "_f64");
    }
    o << "extern \"C\" __global__ void __launch_bounds__(" << threads << ") tqec_mixed_spec(\n"
         "    const MBatch *__restrict__ batches, const MHeader *__restrict__ hdr, MCounters *ctr,\n"
         "    unsigned long long *t_done, const float *const *tabs32, const double *const *tabs64, float *slab32,\n"
         "    double *slab64) {\n";
    o << "  __shared__ __attribute__((aligned(16))) unsigned char smem[" << spec_mixed_lds_bytes(plans) << "];\n";
    o << "  __shared__ long long grab;\n  __shared__ int loc_n;\n  __shared__ int loc_rows[64];\n";
    o << "  float *s32 = slab32 ? slab32 + (unsigned long long)blockIdx.x * " << slab32_entries << "ull : nullptr;\n";
    o << "  double *s64 = slab64 ? slab64 + (unsigned long long)blockIdx.x * " << slab64_entries << "ull : nullptr;\n";
    o << "  const int nb = hdr->nbatches;\n  const long long ntiles = hdr->ntiles;\n"
         "  unsigned tag = 1;\n  bool first = true;\n"
         "  if (threadIdx.x == 0) loc_n = 0;\n"
         "  for (;;) {\n"
         "    __syncthreads();\n"
         "    if (threadIdx.x == 0) grab = (long long)atomicAdd(&ctr->next_tile, 1ull);\n"
         "    __syncthreads();\n"
         "    const long long t = grab;\n"
         "    if (t >= ntiles) break;\n"
         "    if (first && threadIdx.x == 0) atomicMin(&ctr->t_start, (unsigned long long)wall_clock64());\n"
         "    first = false;\n"
         "    int lo = 0, hi = nb - 1;\n"
         "    while (lo < hi) {\n"
         "      const int mid = (lo + hi + 1) >> 1;\n"
         "      if (batches[mid].tile0 <= t) lo = mid; else hi = mid - 1;\n"
         "    }\n"
         "    const MBatch bt = batches[lo];\n"
         "    switch (bt.plan) {\n";
    for (size_t i = 0; i < plans.size(); ++i) {
        const MixedSpecPlan &q = plans[i];
        o << "    case " << i << ":\n      p" << i << "_f32(bt.syn, bt.n, nullptr, (t - bt.tile0) << " << q.f32.glog
          << ", tabs32[" << i << "], bt.mar, bt.cls, bt.tie, loc_rows, &loc_n, 64, " << dbl(q.tie_tol) << ", "
          << dbl(q.recheck_tol) << ", smem, s32, tag);\n      tag += " << q.cp->steps.size() << "u;\n      break;\n";
    }
    o << "    default: break;\n    }\n";
    o << "    const int nre = loc_n;\n"
         "    if (nre > 0) {\n"
         "      for (int e = 0; e < nre; ++e) {\n"
         "        const long long row = loc_rows[e];\n"
         "        switch (bt.plan) {\n";
    for (size_t i = 0; i < plans.size(); ++i) {
        const MixedSpecPlan &q = plans[i];
        o << "        case " << i << ":\n          p" << i << "_f64(bt.syn, bt.n, nullptr, row, tabs64[" << i
          << "], bt.mar, bt.cls, bt.tie, nullptr, nullptr, 0, " << dbl(q.tie_tol) << ", " << dbl(q.recheck_tol)
          << ", smem, s64, 1u);\n          break;\n";
    }
    o << "        default: break;\n        }\n"
         "      }\n"
         "      if (threadIdx.x == 0) {\n"
         "        atomicAdd(&ctr->rl_count, nre);\n"
         "        loc_n = 0;\n"
         "      }\n"
         "    }\n"
         "    if (threadIdx.x == 0) atomicMax(&t_done[lo], (unsigned long long)wall_clock64());\n"
         "  }\n"
         "}\n";
    return o.str();
}

}  // namespace tqec
