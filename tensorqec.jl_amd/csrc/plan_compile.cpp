// plan_compile.cpp — validation, constant folding, layout choice and memory
// planning for a flattened TNMMAP contraction plan.
//
// The reference evaluates every pairwise einsum of the NestedEinsum per
// syndrome (OMEinsum, called at src/decoding/tndecoder.jl:162).  Here the
// syndrome-independent subtrees are contracted once, in FP64 on the host,
// and only the steps that depend on the evidence are kept for the device.
#include "plan_compile.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>

namespace tqec {
namespace {

std::string fmt(const char *f, long a = 0, long b = 0, long c = 0) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), f, a, b, c);
    return buf;
}

// Operand index of a joint index x, byte by byte: t[k][v] holds the operand
// index bits that byte k of x, equal to v, sets (operand bit i is bit pos[i]
// of x).  A few table reads per element instead of a shift per label.
struct ByteGather {
    std::vector<uint64_t> t;   // nbytes x 256
    int nbytes = 0;
    ByteGather(const std::vector<int> &pos, int nu) {
        nbytes = (nu + 7) / 8;
        t.assign(size_t(nbytes) * 256, 0);
        for (int k = 0; k < nbytes; ++k)
            for (int v = 0; v < 256; ++v) {
                uint64_t r = 0;
                for (size_t i = 0; i < pos.size(); ++i) {
                    const int q = pos[i] - 8 * k;
                    if (q >= 0 && q < 8) r |= uint64_t((v >> q) & 1) << i;
                }
                t[size_t(k) * 256 + v] = r;
            }
    }
    uint64_t operator()(uint64_t x) const {
        uint64_t r = 0;
        for (int k = 0; k < nbytes; ++k) r |= t[size_t(k) * 256 + ((x >> (8 * k)) & 255)];
        return r;
    }
};

// Contract two host tensors (each in its own label order) into `cl` order:
// C[c] = sum over all other labels of A * B, accumulated in increasing joint
// index.  Used for the constant and tabulated subtrees.
std::vector<double> host_contract(const std::vector<double> &A, const std::vector<int> &al,
                                  const std::vector<double> &B, const std::vector<int> &bl,
                                  const std::vector<int> &cl) {
    std::vector<int> u = cl;
    for (int l : al)
        if (std::find(u.begin(), u.end(), l) == u.end()) u.push_back(l);
    for (int l : bl)
        if (std::find(u.begin(), u.end(), l) == u.end()) u.push_back(l);
    const int nu = (int)u.size();
    std::vector<int> pa(al.size()), pb(bl.size());
    for (size_t i = 0; i < al.size(); ++i) pa[i] = (int)(std::find(u.begin(), u.end(), al[i]) - u.begin());
    for (size_t i = 0; i < bl.size(); ++i) pb[i] = (int)(std::find(u.begin(), u.end(), bl[i]) - u.begin());
    const ByteGather ga(pa, nu), gb(pb, nu);
    std::vector<double> C(size_t(1) << cl.size(), 0.0);
    const uint64_t cmask = (uint64_t(1) << cl.size()) - 1;
    for (uint64_t x = 0; x < (uint64_t(1) << nu); ++x) C[x & cmask] += A[ga(x)] * B[gb(x)];
    return C;
}

// Sum out labels not in `keep` and permute to `keep` order.
std::vector<double> host_reduce(const std::vector<double> &A, const std::vector<int> &al,
                                const std::vector<int> &keep) {
    std::vector<int> pos(keep.size());
    for (size_t i = 0; i < keep.size(); ++i) pos[i] = (int)(std::find(al.begin(), al.end(), keep[i]) - al.begin());
    std::vector<double> C(size_t(1) << keep.size(), 0.0);
    for (uint64_t x = 0; x < (uint64_t(1) << al.size()); ++x) {
        uint64_t ic = 0;
        for (size_t i = 0; i < pos.size(); ++i) ic |= ((x >> pos[i]) & 1) << i;
        C[ic] += A[x];
    }
    return C;
}

// Power-of-two normalisation of constant data: returns e with max(A*2^-e) in [1, 2).
int normalise(std::vector<double> &A) {
    double m = 0;
    for (double v : A) m = std::max(m, v);
    if (!(m > 0)) return 0;
    int e = std::ilogb(m);
    for (double &v : A) v = std::ldexp(v, -e);
    return e;
}

bool contains(const std::vector<int> &v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

// Interval allocator over per-syndrome entry offsets (first fit, coalescing).
struct Allocator {
    std::map<int64_t, int64_t> free_;  // off -> size
    int64_t top = 0;
    int64_t alloc(int64_t n) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                int64_t off = it->first, sz = it->second;
                free_.erase(it);
                if (sz > n) free_[off + n] = sz - n;
                return off;
            }
        }
        int64_t off = top;
        top += n;
        return off;
    }
    void release(int64_t off, int64_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
        // shrink the top when the last block is free
        if (!free_.empty()) {
            auto last = std::prev(free_.end());
            if (last->first + last->second == top) {
                top = last->first;
                free_.erase(last);
            }
        }
    }
};

struct SlotAllocator {
    std::vector<int> freel;
    int n = 0;
    int alloc() {
        if (!freel.empty()) {
            int s = freel.back();
            freel.pop_back();
            return s;
        }
        return n++;
    }
    void release(int s) { freel.push_back(s); }
};

}  // namespace

int compile_plan(const tqec_plan_desc *d, CompiledPlan &p, std::string &err) {
    if (!d) { err = "null descriptor"; return TQEC_ENULL; }
    p = CompiledPlan();
    const int NV = d->n_vars, NL = d->n_leaves, NS = d->n_steps;
    if (NV <= 0 || NV > (1 << 24)) { err = fmt("n_vars out of range (%ld)", NV); return TQEC_EINVAL; }
    if (NL < 2) { err = fmt("a plan needs at least 2 leaves (got %ld)", NL); return TQEC_EINVAL; }
    if (NS != NL - 1) { err = fmt("n_steps must be n_leaves-1 (%ld vs %ld)", NS, NL); return TQEC_EINVAL; }
    if (!d->leaf_rank || !d->leaf_labels || !d->step_lhs || !d->step_rhs || !d->step_out_rank ||
        !d->out_labels) {
        err = "null array in descriptor";
        return TQEC_ENULL;
    }
    if (d->n_out < 0 || d->n_out > kMaxOut) { err = fmt("n_out out of range (%ld)", d->n_out); return TQEC_EUNSUPPORTED; }
    if (d->n_syndrome_bits < 0 || d->n_syndrome_bits > (1 << 20)) { err = "n_syndrome_bits out of range"; return TQEC_EINVAL; }
    if (d->n_sliced < 0 || (d->n_sliced > 0 && (!d->sliced_label || !d->sliced_bit))) { err = "bad sliced arrays"; return TQEC_EINVAL; }
    if (d->tabulate_bits < 0 || d->tabulate_bits > kMaxEv) { err = fmt("tabulate_bits must be in [0, %ld]", kMaxEv); return TQEC_EINVAL; }
    p.n_vars = NV; p.n_leaves = NL; p.n_steps = NS; p.n_out = d->n_out; p.n_syn = d->n_syndrome_bits;
    p.fold_limit = d->tabulate_bits;
    if (d->semiring != TQEC_SEMIRING_SUM && d->semiring != TQEC_SEMIRING_MAX) { err = "unknown semiring"; return TQEC_EINVAL; }
    p.semiring = d->semiring;
    if (p.semiring == TQEC_SEMIRING_MAX) {
        p.fold_limit = 0;
        if (d->n_out != 0) { err = "a max-product plan must contract to a scalar (n_out = 0)"; return TQEC_EINVAL; }
    }
    p.slice_bit_of_label.assign(NV, -1);
    for (int i = 0; i < d->n_sliced; ++i) {
        int l = d->sliced_label[i], b = d->sliced_bit[i];
        if (l < 0 || l >= NV) { err = fmt("sliced label %ld out of range", l); return TQEC_EINVAL; }
        if (b < 0 || b >= p.n_syn) { err = fmt("sliced bit %ld out of range", b); return TQEC_EINVAL; }
        if (p.slice_bit_of_label[l] >= 0) { err = fmt("label %ld sliced twice", l); return TQEC_EINVAL; }
        p.slice_bit_of_label[l] = b;
    }
    // ---- leaves ----------------------------------------------------------
    size_t lab_off = 0, dat_off = 0;
    p.raw_labels.resize(NL);
    p.raw_data.resize(NL);
    p.leaf_ev.assign(NL, -1);
    p.labels.resize(NL + NS);
    for (int t = 0; t < NL; ++t) {
        int r = d->leaf_rank[t];
        if (r < 0 || r > 30) { err = fmt("leaf %ld: rank %ld out of range", t, r); return TQEC_EINVAL; }
        std::vector<int> L(d->leaf_labels + lab_off, d->leaf_labels + lab_off + r);
        lab_off += r;
        for (int l : L)
            if (l < 0 || l >= NV) { err = fmt("leaf %ld: label %ld out of range", t, l); return TQEC_EINVAL; }
        { std::set<int> s(L.begin(), L.end()); if ((int)s.size() != r) { err = fmt("leaf %ld: repeated label", t); return TQEC_EINVAL; } }
        p.raw_labels[t] = L;
        int ev = d->leaf_evidence ? d->leaf_evidence[t] : -1;
        if (ev >= 0) {
            if (r != 1) { err = fmt("evidence leaf %ld must have rank 1 (has %ld)", t, r); return TQEC_EINVAL; }
            if (ev >= p.n_syn) { err = fmt("evidence leaf %ld: syndrome bit %ld out of range", t, ev); return TQEC_EINVAL; }
            if (p.slice_bit_of_label[L[0]] >= 0) { err = fmt("evidence leaf %ld carries a sliced label", t); return TQEC_EINVAL; }
            p.leaf_ev[t] = ev;
        } else if (ev < -1) {
            err = fmt("leaf %ld: bad evidence bit %ld", t, ev); return TQEC_EINVAL;
        }
        size_t n = size_t(1) << r;
        if (ev < 0) {
            if (!d->leaf_data) { err = "leaf_data is null"; return TQEC_ENULL; }
            p.raw_data[t].assign(d->leaf_data + dat_off, d->leaf_data + dat_off + n);
            for (double v : p.raw_data[t])
                if (!std::isfinite(v) || v < 0) { err = fmt("leaf %ld: entries must be finite and non-negative", t); return TQEC_EINVAL; }
        }
        dat_off += n;
        std::vector<int> eff;
        for (int l : L)
            if (p.slice_bit_of_label[l] < 0) eff.push_back(l);
        p.labels[t] = eff;
    }
    // ---- steps -----------------------------------------------------------
    std::vector<int> used(NL + NS, 0);
    size_t out_off = 0;
    p.lhs.resize(NS); p.rhs.resize(NS);
    for (int i = 0; i < NS; ++i) {
        int a = d->step_lhs[i], b = d->step_rhs[i], node = NL + i;
        if (a < 0 || a >= node || b < 0 || b >= node || a == b) { err = fmt("step %ld: bad operands (%ld, %ld)", i, a, b); return TQEC_EINVAL; }
        if (used[a]++ || used[b]++) { err = fmt("step %ld: node used twice", i); return TQEC_EINVAL; }
        p.lhs[i] = a; p.rhs[i] = b;
        int r = d->step_out_rank[i];
        if (r < 0 || r > 30) { err = fmt("step %ld: rank %ld out of range", i, r); return TQEC_EINVAL; }
        if (!d->step_out_labels && r > 0) { err = "step_out_labels is null"; return TQEC_ENULL; }
        std::vector<int> L(d->step_out_labels + out_off, d->step_out_labels + out_off + r);
        out_off += r;
        std::set<int> s(L.begin(), L.end());
        if ((int)s.size() != r) { err = fmt("step %ld: repeated output label", i); return TQEC_EINVAL; }
        for (int l : L) {
            if (l < 0 || l >= NV) { err = fmt("step %ld: label %ld out of range", i, l); return TQEC_EINVAL; }
            if (!contains(p.labels[a], l) && !contains(p.labels[b], l)) { err = fmt("step %ld: output label %ld not in its operands", i, l); return TQEC_EINVAL; }
        }
        p.labels[node] = L;
    }
    // label completeness: a label may only disappear once every tensor that
    // holds it has been contracted, and output labels never disappear.
    {
        std::vector<int> total(NV, 0);
        for (int t = 0; t < NL; ++t) for (int l : p.labels[t]) total[l]++;
        std::vector<char> is_out(NV, 0);
        for (int i = 0; i < d->n_out; ++i) {
            int l = d->out_labels[i];
            if (l < 0 || l >= NV) { err = fmt("output label %ld out of range", l); return TQEC_EINVAL; }
            if (is_out[l]) { err = "repeated output label"; return TQEC_EINVAL; }
            if (p.slice_bit_of_label[l] >= 0) { err = "an output label is sliced"; return TQEC_EINVAL; }
            is_out[l] = 1;
            p.out_labels.push_back(l);
        }
        std::vector<std::vector<int>> cnt(NL + NS);
        for (int t = 0; t < NL; ++t) {
            cnt[t].assign(NV, 0);
            for (int l : p.labels[t]) cnt[t][l] = 1;
        }
        for (int i = 0; i < NS; ++i) {
            int node = NL + i, a = p.lhs[i], b = p.rhs[i];
            cnt[node].resize(NV);
            for (int l = 0; l < NV; ++l) cnt[node][l] = cnt[a][l] + cnt[b][l];
            for (int l : p.labels[a]) {
                if (!contains(p.labels[node], l) && (is_out[l] || cnt[node][l] < total[l])) {
                    err = fmt("step %ld drops label %ld that is still needed", i, l); return TQEC_EINVAL;
                }
            }
            for (int l : p.labels[b]) {
                if (!contains(p.labels[node], l) && (is_out[l] || cnt[node][l] < total[l])) {
                    err = fmt("step %ld drops label %ld that is still needed", i, l); return TQEC_EINVAL;
                }
            }
            std::vector<int>().swap(cnt[a]);
            std::vector<int>().swap(cnt[b]);
        }
        const std::vector<int> &rl = p.labels[NL + NS - 1];
        std::set<int> rs(rl.begin(), rl.end()), os(p.out_labels.begin(), p.out_labels.end());
        if (rs != os) { err = "root labels differ from out_labels"; return TQEC_EINVAL; }
    }
    // ---- syndrome dependence ----------------------------------------------
    // bits[x]: syndrome bits node x depends on.  Nodes that depend on no bit
    // are contracted once here (constant folding); nodes that depend on at
    // most `fold_bits` bits are tabulated here for every assignment of those
    // bits (the device reads the right variant); only the rest run per
    // syndrome on the device.
    p.bits.assign(NL + NS, {});
    for (int t = 0; t < NL; ++t) {
        std::vector<int> &B = p.bits[t];
        if (p.leaf_ev[t] >= 0) B.push_back(p.leaf_ev[t]);
        for (int l : p.raw_labels[t])
            if (p.slice_bit_of_label[l] >= 0) B.push_back(p.slice_bit_of_label[l]);
        std::sort(B.begin(), B.end());
        B.erase(std::unique(B.begin(), B.end()), B.end());
        if ((int)B.size() > kMaxEv) {
            err = fmt("leaf %ld depends on %ld syndrome bits (limit %ld)", t, (long)B.size(), kMaxEv);
            return TQEC_EUNSUPPORTED;
        }
    }
    for (int i = 0; i < NS; ++i) {
        std::vector<int> u = p.bits[p.lhs[i]];
        u.insert(u.end(), p.bits[p.rhs[i]].begin(), p.bits[p.rhs[i]].end());
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        p.bits[NL + i] = u;
    }
    const int root = NL + NS - 1;
    if (p.bits[root].empty()) { err = "the result does not depend on any syndrome bit"; return TQEC_EUNSUPPORTED; }
    p.fold_bits = std::min<int>(p.fold_limit, (int)p.bits[root].size() - 1);
    p.dep.assign(NL + NS, 0);
    for (int x = NL; x < NL + NS; ++x) p.dep[x] = (int)p.bits[x].size() > p.fold_bits;
    if (p.semiring == TQEC_SEMIRING_MAX) {
        // The backtrace needs the arg-max of every eliminated label, so a step
        // that eliminates one (or has such a step below it) runs on the device
        // even when it does not depend on the syndrome.
        std::vector<char> elim(NL + NS, 0);
        for (int i = 0; i < NS; ++i) {
            const int node = NL + i, a = p.lhs[i], b = p.rhs[i];
            bool e = elim[a] || elim[b];
            for (int l : p.labels[a]) e = e || !contains(p.labels[node], l);
            for (int l : p.labels[b]) e = e || !contains(p.labels[node], l);
            elim[node] = e;
            p.dep[node] = p.dep[node] || e;
        }
    }

    // ---- statistics, constant folding and tabulation ----------------------
    p.value.resize(NL + NS);
    p.var.resize(NL + NS);
    auto restrict_bits = [&](int v, const std::vector<int> &from, const std::vector<int> &to) {
        int r = 0;
        for (size_t j = 0; j < to.size(); ++j) {
            const int pos = (int)(std::find(from.begin(), from.end(), to[j]) - from.begin());
            r |= ((v >> pos) & 1) << j;
        }
        return r;
    };
    for (int t = 0; t < NL; ++t) {
        p.sc = std::max(p.sc, (int)p.labels[t].size());
        if (p.bits[t].empty()) {
            p.value[t] = p.raw_data[t];
            continue;
        }
        const int nv = 1 << p.bits[t].size();
        p.var[t].resize(nv);
        if (p.leaf_ev[t] >= 0) {
            p.var[t][0] = {1.0, 0.0};
            p.var[t][1] = {0.0, 1.0};
            continue;
        }
        const std::vector<int> &raw = p.raw_labels[t];
        const size_t n_eff = size_t(1) << p.labels[t].size();
        for (int v = 0; v < nv; ++v) p.var[t][v].assign(n_eff, 0.0);
        for (uint64_t xi = 0; xi < (uint64_t(1) << raw.size()); ++xi) {
            int v = 0;
            uint64_t ki = 0;
            int kpos = 0;
            for (size_t q = 0; q < raw.size(); ++q) {
                const int bit = (int)((xi >> q) & 1);
                const int sb = p.slice_bit_of_label[raw[q]];
                if (sb >= 0) {
                    const int j = (int)(std::find(p.bits[t].begin(), p.bits[t].end(), sb) - p.bits[t].begin());
                    v |= bit << j;
                } else {
                    ki |= uint64_t(bit) << kpos++;
                }
            }
            p.var[t][v][ki] += p.raw_data[t][xi];
        }
    }
    for (int i = 0; i < NS; ++i) {
        int node = NL + i, a = p.lhs[i], b = p.rhs[i];
        std::set<int> u(p.labels[a].begin(), p.labels[a].end());
        u.insert(p.labels[b].begin(), p.labels[b].end());
        double w = std::ldexp(1.0, (int)u.size());
        double r = std::ldexp(1.0, (int)p.labels[a].size()) + std::ldexp(1.0, (int)p.labels[b].size()) +
                   std::ldexp(1.0, (int)p.labels[node].size());
        p.tc += w; p.rw += r;
        p.sc = std::max(p.sc, (int)p.labels[node].size());
        if (p.dep[node]) {
            p.dep_tc += w;
            p.dep_rw += (p.dep[a] ? std::ldexp(1.0, (int)p.labels[a].size()) : 0.0) +
                        (p.dep[b] ? std::ldexp(1.0, (int)p.labels[b].size()) : 0.0) +
                        std::ldexp(1.0, (int)p.labels[node].size());
            if ((int)p.labels[node].size() > kMaxRank) {
                err = fmt("step %ld: per-syndrome tensor of rank %ld exceeds the limit %ld", i, (long)p.labels[node].size(), kMaxRank);
                return TQEC_EUNSUPPORTED;
            }
            continue;
        }
        if (u.size() > 26) { err = fmt("host-evaluated step %ld too wide (%ld labels)", i, (long)u.size()); return TQEC_EUNSUPPORTED; }
        if (p.bits[node].empty()) {
            p.value[node] = host_contract(p.value[a], p.labels[a], p.value[b], p.labels[b], p.labels[node]);
        } else {
            const int nv = 1 << p.bits[node].size();
            p.var[node].resize(nv);
            for (int v = 0; v < nv; ++v) {
                const std::vector<double> &A = p.bits[a].empty() ? p.value[a] : p.var[a][restrict_bits(v, p.bits[node], p.bits[a])];
                const std::vector<double> &B = p.bits[b].empty() ? p.value[b] : p.var[b][restrict_bits(v, p.bits[node], p.bits[b])];
                p.var[node][v] = host_contract(A, p.labels[a], B, p.labels[b], p.labels[node]);
            }
        }
        for (int x : {a, b}) {
            std::vector<double>().swap(p.value[x]);
            std::vector<std::vector<double>>().swap(p.var[x]);
        }
    }

    // ---- layouts, top-down -----------------------------------------------
    p.mem.assign(NL + NS, {});
    p.mem[root] = p.out_labels;
    std::vector<StepDesc> rev;  // built in reverse order
    std::vector<MapStep> rev_map;
    for (int i = NS - 1; i >= 0; --i) {
        const int node = NL + i;
        if (!p.dep[node]) continue;
        const int a = p.lhs[i], b = p.rhs[i];
        const std::vector<int> &C = p.mem[node];
        const std::vector<int> &LA = p.labels[a], &LB = p.labels[b];
        std::vector<int> K, SA, SB, AR, BR;
        for (int l : LA) if (contains(LB, l) && !contains(C, l)) K.push_back(l);
        std::sort(K.begin(), K.end());
        for (int l : LA) if (!contains(LB, l) && !contains(C, l)) SA.push_back(l);
        for (int l : LB) if (!contains(LA, l) && !contains(C, l)) SB.push_back(l);
        std::sort(SA.begin(), SA.end()); std::sort(SB.begin(), SB.end());
        for (int l : C) { if (contains(LA, l)) AR.push_back(l); if (contains(LB, l)) BR.push_back(l); }
        StepDesc st;
        std::memset(&st, 0, sizeof(st));
        st.rc = (int)C.size();
        st.nk = (int)K.size();
        st.is_root = (node == root);
        auto build = [&](int x, const std::vector<int> &S, const std::vector<int> &R, OperandDesc &op) -> int {
            std::memset(&op, 0, sizeof(op));
            op.slot = -1;
            const bool inter = p.dep[x];
            // memory order: kept labels in C's order (low bits), then labels
            // summed on this side only, then the contracted labels (high bits).
            // Sum-product sums the single-side labels of host data here;
            // max-product keeps them (their arg-max is needed on the device).
            const bool keep_single = inter || p.semiring == TQEC_SEMIRING_MAX;
            std::vector<int> order = R;
            if (keep_single) order.insert(order.end(), S.begin(), S.end());
            order.insert(order.end(), K.begin(), K.end());
            p.mem[x] = order;
            op.rank = (int)order.size();
            op.nsingle = keep_single ? (int)S.size() : 0;
            op.sshift = (int)R.size();
            op.kshift = (int)R.size() + op.nsingle;
            int dst = 0;
            int nseg = 0;
            for (size_t ci = 0; ci < C.size();) {
                if (!contains(R, C[ci])) { ++ci; continue; }
                size_t cj = ci;
                while (cj < C.size() && contains(R, C[cj])) ++cj;
                if (nseg >= kMaxSeg) return -1;
                op.seg[nseg].src = (int16_t)ci;
                op.seg[nseg].width = (int16_t)(cj - ci);
                op.seg[nseg].dst = (int16_t)dst;
                dst += (int)(cj - ci);
                ++nseg;
                ci = cj;
            }
            op.nseg = nseg;
            if (inter) {
                op.kind = kInter;
            } else if (p.bits[x].empty()) {
                op.kind = kConst;
                std::vector<double> v = host_reduce(p.value[x], p.labels[x], order);
                op.cexp = normalise(v);
                op.off = (int64_t)p.const_pool.size();
                p.const_pool.insert(p.const_pool.end(), v.begin(), v.end());
            } else {
                op.kind = kSliced;
                op.nev = (int)p.bits[x].size();
                for (int j = 0; j < op.nev; ++j) op.ev_bit[j] = p.bits[x][j];
                std::vector<std::vector<double>> vars;
                double m = 0;
                for (const auto &vv : p.var[x]) {
                    vars.push_back(host_reduce(vv, p.labels[x], order));
                    for (double q : vars.back()) m = std::max(m, q);
                }
                op.cexp = (m > 0) ? std::ilogb(m) : 0;
                op.off = (int64_t)p.sliced_pool.size();
                for (auto &r : vars)
                    for (double q : r) p.sliced_pool.push_back(std::ldexp(q, -op.cexp));
            }
            return 0;
        };
        int ea = build(a, SA, AR, st.a);
        int eb = build(b, SB, BR, st.b);
        if (ea == -1 || eb == -1) { err = fmt("step %ld: too many index segments", i); return TQEC_EUNSUPPORTED; }
        rev.push_back(st);
        // labels of the backtrace: C in memory order, then K, SA, SB
        MapStep ms;
        ms.off = (int32_t)p.map_labels.size();
        ms.nc = (int32_t)C.size();
        ms.nk = (int32_t)K.size();
        ms.nsa = st.a.nsingle;
        ms.nsb = st.b.nsingle;
        for (int l : C) p.map_labels.push_back(l);
        for (int l : K) p.map_labels.push_back(l);
        if (st.a.nsingle) for (int l : SA) p.map_labels.push_back(l);
        if (st.b.nsingle) for (int l : SB) p.map_labels.push_back(l);
        rev_map.push_back(ms);
    }
    // ---- memory plan (execution order) ------------------------------------
    std::vector<int> dep_steps;
    for (int i = 0; i < NS; ++i) if (p.dep[NL + i]) dep_steps.push_back(i);
    std::reverse(rev.begin(), rev.end());
    p.steps = rev;
    std::reverse(rev_map.begin(), rev_map.end());
    p.map_steps = rev_map;
    if (p.semiring == TQEC_SEMIRING_MAX) {
        int64_t ao = 0;
        for (size_t j = 0; j < p.map_steps.size(); ++j) {
            MapStep &ms = p.map_steps[j];
            if (ms.nk + ms.nsa + ms.nsb > 16) {
                err = "a max-product step eliminates more than 16 labels at once";
                return TQEC_EUNSUPPORTED;
            }
            ms.arg_off = ao;
            ao += int64_t(1) << ms.nc;
        }
        p.arg_entries = ao;
    }
    Allocator al;
    SlotAllocator sa;
    std::vector<int64_t> off(NL + NS, -1);
    std::vector<int> slot(NL + NS, -1);
    int64_t peak = 0;
    for (size_t j = 0; j < dep_steps.size(); ++j) {
        int i = dep_steps[j], node = NL + i;
        StepDesc &st = p.steps[j];
        int64_t n = int64_t(1) << st.rc;
        off[node] = al.alloc(n);
        slot[node] = sa.alloc();
        peak = std::max(peak, al.top);
        st.c_off = off[node];
        st.c_slot = slot[node];
        int a = p.lhs[i], b = p.rhs[i];
        if (st.a.kind == kInter) { st.a.off = off[a]; st.a.slot = slot[a]; }
        if (st.b.kind == kInter) { st.b.off = off[b]; st.b.slot = slot[b]; }
        if (st.a.kind == kInter) { al.release(off[a], int64_t(1) << st.a.rank); sa.release(slot[a]); }
        if (st.b.kind == kInter) { al.release(off[b], int64_t(1) << st.b.rank); sa.release(slot[b]); }
    }
    p.ws_entries = peak;
    p.n_slots = sa.n;
    p.root_off = off[root];
    p.root_slot = slot[root];
    return TQEC_OK;
}

}  // namespace tqec
