"""Plans built exactly as the Julia binding builds them
(tensorqec.jl_amd/julia/TensorQECHIP.jl `flatten` / `HipPlan`), through the C ABI.

The Julia side hands over the reference's own networks and trees: every
tensor of `ct.tensors` as a leaf with its labels and column-major data, the
evidence vectors as one-hot leaves (the last ns tensors, data [1, 0] that
the library ignores), steps numbered in the post-order of a recursive visit
of the NestedEinsum from its root, each intermediate in the label order
OMEinsum chose (here: descending, unlike this package's planner), the root
in the network's `iy`, and `n_sliced = 0`, tabulation depth 4 — except for
TNMAP, whose evidence variables are sliced labels of the UAI model's leaves
(tndecoder.jl:33-57) in the max-product semiring, and for a contraction
re-optimized with a slicer (an OMEinsum `SlicedEinsum`), whose sliced
labels become syndrome bits after the evidence vectors' own.  Covered networks: the CSS
TNMMAP network without factorization (tndecoder.jl:97-146), the DEM network
with the reference's default chain factorization (:186-238), and the TNMAP
UAI model.

CPU: such plans are accepted, and their statistics equal those of this
package's plan over the same tree.  GPU: decoding them reproduces the
oracle — d=3 golden marginals for all 256 syndromes, d=5 against the
oracle's FP64 executor, `dem.dem` against brute force over its 2^21
mechanism subsets, TNMAP maxima against brute force over all 4^9 errors.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode
from tensorqec.dem import parse_dem_file
from tensorqec.errors import CSSErrorPattern, css_syndrome_rows, iid_error, sample_depolarizing, syndrome_extraction
from tensorqec.gf2 import logical_operator
from tensorqec.network import tnmmap_css_network, tnmmap_dem_network
from tensorqec.planner import flatten
from tensorqec.tnmap import reduce2general, uai_network

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def julia_flatten(graph, tree, evidence=None, sliced=None):
    """TensorQECHIP.jl `flatten(code, tensors; evidence, sliced, ixs)`:
    recursive post-order visit from the root; intermediate labels in
    descending order (an order other than the planner's), root labels = iy.
    `evidence`: tensor -> syndrome bit; `sliced`: label -> syndrome bit."""
    evidence = evidence or {}
    sliced = sliced or {}
    Ln = tree.n_leaves
    lhs, rhs, outs = [], [], []

    def labels_of(node):
        return [l for l in range(tree.masks[node].bit_length()) if (tree.masks[node] >> l) & 1 and l not in sliced]

    def visit(node):
        if node < Ln:
            return node
        a, b = tree.children[node - Ln]
        ia, ib = visit(a), visit(b)
        lhs.append(ia)
        rhs.append(ib)
        outs.append(list(graph.iy) if node == tree.root else sorted(labels_of(node), reverse=True))
        return Ln + len(lhs) - 1

    visit(tree.root)
    ranks = np.array([len(ix) for ix in graph.ixs], np.int32)
    labels = np.array([l for ix in graph.ixs for l in ix], np.int32)
    data = np.concatenate([np.asarray(t, np.float64).reshape(-1) for t in graph.tensors])
    ev = np.full(Ln, -1, np.int32)
    for t, bit in evidence.items():
        ev[t] = bit
    sl = sorted(sliced.items())
    return dict(ranks=ranks, labels=labels, data=data, ev=ev, lhs=np.array(lhs, np.int32),
                rhs=np.array(rhs, np.int32), orank=np.array([len(o) for o in outs], np.int32),
                olab=np.array([l for o in outs for l in o] or [0], np.int32),
                iy=np.array(list(graph.iy) or [0], np.int32), n_out=len(graph.iy),
                sl_l=np.array([a for a, _ in sl] or [0], np.int32), sl_b=np.array([b for _, b in sl] or [0], np.int32),
                n_sliced=len(sl), nvars=int(max([labels.max()] + [max(o) for o in outs if o] + [a for a, _ in sl])) + 1)


class JuliaShapedPlan(L.HipPlan):
    """A tqec_plan created from a julia_flatten descriptor (same methods as
    the package's HipPlan)."""

    def __init__(self, f, n_syndrome, tabulate_bits=4, semiring=0):
        p = lambda a, t: a.ctypes.data_as(t)  # noqa: E731
        i32 = ctypes.POINTER(ctypes.c_int32)
        desc = L.PlanDesc(f["nvars"], len(f["ranks"]), p(f["ranks"], i32), p(f["labels"], i32),
                          p(f["data"], ctypes.POINTER(ctypes.c_double)), p(f["ev"], i32), len(f["lhs"]),
                          p(f["lhs"], i32), p(f["rhs"], i32), p(f["orank"], i32), p(f["olab"], i32), f["n_out"],
                          p(f["iy"], i32), n_syndrome, f["n_sliced"], p(f["sl_l"], i32), p(f["sl_b"], i32),
                          tabulate_bits, semiring)
        h = ctypes.c_void_p()
        L.check(L.load().tqec_plan_create(ctypes.byref(desc), ctypes.byref(h)))
        self._keep = f
        self._h = h
        self.n_out = f["n_out"]
        self.n_syn = n_syndrome


def css_case(d, p):
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p, p, p, t.nq)
    lx, lz = logical_operator(t)
    g = tnmmap_css_network(t, em, factorize=False, lxlz=(lx, lz))
    ns = t.ns
    nt = len(g.ixs)
    # the reference network keeps its evidence vectors last (tndecoder.jl:133-139)
    assert list(g.evidence_leaves) == list(range(nt - ns, nt))
    tree = D.optimize_tree(g.ixs, g.iy, D.TreeSA(seed=5))
    f = julia_flatten(g, tree, evidence={nt - ns + i: i for i in range(ns)})
    return t, em, g, tree, f


def dem_case():
    dem = parse_dem_file(os.path.join(GOLD, "dem.dem"))
    g = tnmmap_dem_network(dem, factorize=True)
    ns = len(g.evidence_labels)
    nt = len(g.ixs)
    assert list(g.evidence_leaves) == list(range(nt - ns, nt))
    tree = D.optimize_tree(g.ixs, g.iy, D.TreeSA(seed=3))
    return dem, g, tree, julia_flatten(g, tree, evidence={nt - ns + i: i for i in range(ns)})


def tnmap_case():
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    p = 0.05 / 3
    em = iid_error(p, p, p, t.nq)
    g = uai_network(reduce2general(t, em))
    sliced = {l: i for i, l in enumerate(g.evidence_labels)}
    ixs = [[l for l in ix if l not in sliced] for ix in g.ixs]
    tree = D.optimize_tree(ixs, g.iy, D.TreeSA(seed=2))
    return t, em, g, tree, julia_flatten(g, tree, sliced=sliced)


@pytest.mark.parametrize("d", [3, 5])
def test_css_julia_plan_stats_match_package_plan(d):
    t, em, g, tree, f = css_case(d, 0.05 / 3)
    jp = JuliaShapedPlan(f, t.ns)
    pp = L.HipPlan(flatten(g, tree, False), tabulate_bits=4)
    a, b = jp.stats(), pp.stats()
    for k in ("tc", "rw", "sc", "n_steps", "n_dep_steps", "dep_tc", "n_out", "tabulate_bits"):
        assert a[k] == b[k], (k, a[k], b[k])


def test_dem_and_tnmap_julia_plans_accepted():
    dem, g, tree, f = dem_case()
    jp = JuliaShapedPlan(f, len(g.evidence_labels))
    pp = L.HipPlan(flatten(g, tree, False), tabulate_bits=4)
    assert jp.stats()["tc"] == pp.stats()["tc"] and jp.stats()["n_out"] == 1
    t, em, g, tree, f = tnmap_case()
    mp = JuliaShapedPlan(f, t.ns, tabulate_bits=0, semiring=1)
    assert mp.stats()["n_out"] == 0


def test_julia_sliced_einsum_plan_matches_sliced_package_plan():
    """TensorQECHIP.jl `HipPlan(::SlicedEinsum, ...)`: the inner tree
    flattened with the sliced labels dropped from every intermediate and
    mapped to syndrome bits ns + j after the evidence vectors' own bits, one
    row per slice (`sliced_rows`: syndrome b, slice s -> row b * 2^k + s,
    bit j of s as syndrome bit ns + j).  The library accepts it, and it is
    the package's sliced plan of the same tree (tensorqec.slicing), whose
    decode the GPU tests check against the oracle."""
    from tensorqec.slicing import choose_slices, expand_rows, slice_plan
    t, em, g, tree, _ = css_case(5, 0.05 / 3)
    ns, nt = t.ns, len(g.ixs)
    plan = flatten(g, tree, False)
    labels = choose_slices(plan, plan.stats()["sc_dep"] - 2)
    k = len(labels)
    assert k >= 1
    f = julia_flatten(g, tree, evidence={nt - ns + i: i for i in range(ns)},
                      sliced={l: ns + j for j, l in enumerate(labels)})
    jp = JuliaShapedPlan(f, ns + k)
    pp = L.HipPlan(slice_plan(plan, labels), tabulate_bits=4)
    a, b = jp.stats(), pp.stats()
    for key in ("tc", "rw", "sc", "n_steps", "n_dep_steps", "dep_tc", "n_out", "tabulate_bits"):
        assert a[key] == b[key], (key, a[key], b[key])
    # the binding's row expansion, column by column as `sliced_rows` writes it
    S = css_syndrome_rows(sample_depolarizing(em, seed=4, count=3), t)
    m = 1 << k
    cols = []
    for b_ in range(len(S)):
        for s_ in range(m):
            cols.append(np.concatenate([S[b_], [(s_ >> j) & 1 for j in range(k)]]).astype(np.uint8))
    assert np.array_equal(np.array(cols), expand_rows(S, k))


@pytest.mark.gpu
def test_css_julia_plan_decodes_like_the_oracle():
    golden = np.load(os.path.join(GOLD, "golden.npz"))
    t, em, g, tree, f = css_case(3, 0.05 / 3)
    jp = JuliaShapedPlan(f, t.ns)
    S, ref = golden["d3_p005_syn"], golden["d3_p005_mar"]
    for mode in ("plain_f64", "spec_f64", "spec_f32"):
        say(f"case julia-shaped d=3 {mode}")
        mar, cls, tie = jp.decode(S, mode=mode)
        err = float(np.max(np.abs(mar - ref) / ref.max(axis=1, keepdims=True)))
        assert err < (2e-5 if mode.endswith("f32") else 1e-12), (mode, err)
        ok = O.tie_flags(ref, 1e-12) == 0
        assert np.array_equal(cls[ok], O.findmax(ref)[ok]), mode
    t, em, g, tree, f = css_case(5, 0.05 / 3)
    jp = JuliaShapedPlan(f, t.ns)
    S = css_syndrome_rows(sample_depolarizing(em, seed=55, count=2048), t)
    ref, _, _ = OraclePlan(flatten(g, tree, False)).decode(S)
    for mode in ("plain_f64", "spec_f64", "spec_f32"):
        say(f"case julia-shaped d=5 {mode}")
        mar, cls, tie = jp.decode(S, mode=mode)
        err = float(np.max(np.abs(mar - ref) / ref.max(axis=1, keepdims=True)))
        assert err < (2e-5 if mode.endswith("f32") else 1e-12), (mode, err)
        ok = O.tie_flags(ref, 1e-12) == 0
        assert np.array_equal(cls[ok], O.findmax(ref)[ok]), mode


@pytest.mark.gpu
def test_dem_julia_plan_matches_bruteforce():
    dem, g, tree, f = dem_case()
    ns = len(g.evidence_labels)
    jp = JuliaShapedPlan(f, ns)
    bf = O.brute_force_dem(*O.parse_dem_text(open(os.path.join(GOLD, "dem.dem")).read()))
    S = np.array(sorted(bf), dtype=np.uint8)
    ref = np.array([bf[tuple(s)] for s in S])
    for mode in ("plain_f64", "spec_f64", "spec_f32"):
        say(f"case julia-shaped dem.dem {mode}: {len(S)} syndromes")
        mar, cls, _ = jp.decode(S, mode=mode)
        err = float(np.max(np.abs(mar - ref) / ref.max(axis=1, keepdims=True)))
        assert err < (2e-5 if mode.endswith("f32") else 1e-12), (mode, err)
        ok = O.tie_flags(ref, 1e-12) == 0
        assert np.array_equal(cls[ok], O.findmax(ref)[ok]), mode


@pytest.mark.gpu
def test_tnmap_julia_plan_attains_bruteforce_max():
    t, em, g, tree, f = tnmap_case()
    mp = JuliaShapedPlan(f, t.ns, tabulate_bits=0, semiring=1)
    bf = O.brute_force_map_css(t.stgx.H, t.stgz.H, em.px, em.py, em.pz)
    S = np.array(sorted(k for k, v in bf.items() if v > 0), dtype=np.uint8)
    say(f"case julia-shaped TNMAP d=3: {len(S)} syndromes")
    cfg, logp = mp.map_decode(S, 2 * t.nq)
    ex, ez = cfg[:, :t.nq], cfg[:, t.nq:]
    sx, sz = syndrome_extraction(CSSErrorPattern(ex, ez), t)
    assert np.array_equal(np.concatenate([sx, sz], axis=1), S)
    ref = np.array([bf[tuple(s)] for s in S])
    got = O.css_config_probability(ex, ez, em.px, em.py, em.pz)
    assert np.max(np.abs(got - ref) / ref) < 1e-12
    assert np.max(np.abs(np.exp(logp) - ref) / ref) < 1e-12
