"""Nested-dissection networks and trees (network.dissected_parity_network,
planner.dissect, decoders.Dissection) on the CPU.

The dissected network rewrites each parity check of the reference network
(src/decoding/tndecoder.jl:97-146 for CSS codes, :186-238 for detector error
models) as a tree of rank-3 parity tensors along a bisection of the error
units, so its marginals must be the reference network's: checked against
brute force (the `dem.dem` fixture of test/decoding/test_circuits, 2^21
mechanism subsets; the d=3 surface code, 4^9 Pauli errors) and against the
oracle's FP64 executor on the reference network's plan (circuit-level d=3
memory experiment, toric code L=3).  The tree must be a valid binary
contraction of every tensor, and cheaper than the time-sweep tree of the
reference network on the circuit-level DEM.
"""
import os

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.dem import parse_dem_file
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing
from tensorqec.gf2 import logical_operator
from tensorqec.network import dissected_css_network, dissected_dem_network, tnmmap_css_network, tnmmap_dem_network
from tensorqec.planner import check_tree, dissect, flatten, tree_cost

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_dissect_is_a_binary_tree_over_all_items():
    rng = np.random.default_rng(0)
    n = 57
    edges = [sorted(set(rng.integers(0, n, size=rng.integers(1, 5)).tolist())) for _ in range(80)]
    ch = dissect(n, edges, [False] * len(edges), seed=3)
    assert len(ch) == n - 1
    seen = []
    for a, b in ch:
        seen += [a, b]
    assert sorted(seen) == list(range(2 * n - 2))      # every node but the root used once
    for i, (a, b) in enumerate(ch):
        assert a < n + i and b < n + i                  # post-order


def test_dem_fixture_matches_bruteforce():
    dem = parse_dem_file(os.path.join(GOLD, "dem.dem"))
    g, tree = dissected_dem_network(dem, seed=1)
    ev = set(g.evidence_labels)
    check_tree(tree, [[l for l in ix if l not in ev] for ix in g.ixs[:tree.n_leaves]], g.iy)
    plan = flatten(g.without_evidence_leaves(), tree, True)
    bf = O.brute_force_dem(*O.parse_dem_text(open(os.path.join(GOLD, "dem.dem")).read()))
    S = np.array(sorted(bf), dtype=np.uint8)
    ref = np.array([bf[tuple(s)] for s in S])
    got, _, _ = OraclePlan(plan).decode(S)
    assert np.max(np.abs(got - ref) / ref.max(axis=1, keepdims=True)) < 1e-12
    assert np.array_equal(got == 0, ref == 0)


def test_circuit_level_dem_matches_reference_network_and_is_narrower():
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    g, tree = dissected_dem_network(dem, seed=0)
    plan_d = flatten(g.without_evidence_leaves(), tree, True)
    gr = tnmmap_dem_network(dem, factorize=True)
    plan_r = D.build_plan(gr, D.TNMMAP(optimizer=D.TreeSA(primary=tuple(range(len(dem.error_rates))))))
    rng = np.random.default_rng(5)
    E = (rng.random((128, len(dem.error_rates))) < 20 * np.asarray(dem.error_rates)[None, :]).astype(np.int64)
    S = ((E @ dem.detector_matrix().T) & 1).astype(np.uint8)
    a, _, _ = OraclePlan(plan_d).decode(S)
    b, _, _ = OraclePlan(plan_r).decode(S)
    assert np.array_equal(a == 0, b == 0)
    nz = b > 0
    assert np.max(np.abs(a - b)[nz] / b[nz]) < 1e-12
    # annealed as the compile does: fewer multiply-adds than the reference network's tree
    _, plan_a, _ = D.dissection_plan(lambda sd: dissected_dem_network(dem, seed=sd), D.Dissection())
    assert plan_a.stats()["tc"] < plan_r.stats()["tc"]


@pytest.mark.parametrize("code,p", [(SurfaceCode(3, 3), 0.05), (ToricCode(3, 3), 0.07)])
def test_css_dissection_matches_reference_network(code, p):
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    lx, lz = logical_operator(t)
    g, tree = dissected_css_network(t, em, lxlz=(lx, lz), seed=2)
    plan_d = flatten(g.without_evidence_leaves(), tree, True)
    S = css_syndrome_rows(sample_depolarizing(em, seed=9, count=200), t)
    a, _, _ = OraclePlan(plan_d).decode(S)
    gr = tnmmap_css_network(t, em, factorize=False, lxlz=(lx, lz))
    b, _, _ = OraclePlan(D.build_plan(gr, D.TNMMAP(factorize=False))).decode(S)
    assert np.array_equal(a == 0, b == 0)
    nz = b > 0
    assert np.max(np.abs(a - b)[nz] / b[nz]) < 1e-12
    if code.__class__.__name__ == "SurfaceCode":
        bf = O.brute_force_css(t.stgx.H, t.stgz.H, lx, lz, em.px, em.py, em.pz)
        for row, m in zip(S, a):
            want = bf[tuple(int(x) for x in row)]
            assert np.allclose(m, want, rtol=1e-12, atol=0)


def test_compile_keeps_the_cheaper_network():
    """`compile(TNMMAP(), dem)` with the default optimizer compares the
    reference network's TreeSA plan (with the time-sweep candidate) with the
    dissected one and keeps the one with fewer multiply-adds; an explicit
    Dissection optimizer always uses the dissected network."""
    dem = detector_error_model_from_stim(rotated_memory_z(3, 2, 0.001))
    ct_d = D.compile(D.TNMMAP(optimizer=D.Dissection()), dem)
    assert ct_d.graph.meta.get("dissected")
    gr = tnmmap_dem_network(dem, factorize=True)
    plan_r = D.build_plan(gr, D.TNMMAP(optimizer=D.TreeSA(primary=tuple(range(len(dem.error_rates))))))
    auto = D.compile(D.TNMMAP(), dem)
    assert auto.plan.stats()["tc"] == min(ct_d.plan.stats()["tc"], plan_r.stats()["tc"])


def test_toric5_fixture_inputs():
    """tests/golden/toric5.npz (the oracle's exact marginals of toric L=5 that
    the GPU tests check against): its syndromes are the seeded sampler's, and
    its marginals are probabilities of those syndromes (positive, summing to
    at most one over the 16 classes)."""
    g = np.load(os.path.join(GOLD, "toric5.npz"))
    t = CSSTannerGraph.from_code(ToricCode(5, 5))
    p = float(g["p"])
    S = css_syndrome_rows(sample_depolarizing(iid_error(p, p, p, t.nq), seed=5, count=16), t)
    assert np.array_equal(S, g["syn"])
    assert g["mar"].shape == (16, 16)
    assert np.all(g["mar"] > 0) and np.all(g["mar"].sum(axis=1) <= 1.0)


def test_toric6_fixture_inputs():
    """tests/golden/toric6.npz (the oracle's exact marginals of toric L=6,
    contracted in slices): its syndromes are the seeded sampler's, and its
    marginals are probabilities of those syndromes."""
    g = np.load(os.path.join(GOLD, "toric6.npz"))
    t = CSSTannerGraph.from_code(ToricCode(6, 6))
    p = float(g["p"])
    S = css_syndrome_rows(sample_depolarizing(iid_error(p, p, p, t.nq), seed=6, count=8), t)
    assert np.array_equal(S, g["syn"])
    assert g["mar"].shape == (8, 16)
    assert np.all(g["mar"] > 0) and np.all(g["mar"].sum(axis=1) <= 1.0)



def test_dem_d5r3_fixture_inputs():
    """tests/golden/dem_d5r3.npz (the oracle's exact marginals of the
    circuit-level d=5, 3-round model, contracted in slices): its syndromes
    are the generator's (mechanism flips at 10x their rates, seed 31), and
    its marginals are probabilities of those syndromes."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_dem_d5r3", os.path.join(GOLD, "make_dem_d5r3.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    g = np.load(os.path.join(GOLD, "dem_d5r3.npz"))
    dem = detector_error_model_from_stim(rotated_memory_z(5, 3, float(g["p"])))
    assert np.array_equal(mk.syndromes(dem), g["syn"])
    assert g["syn"].shape == (2, 72) and g["mar"].shape == (2, 2)
    assert np.all(g["mar"] > 0) and np.all(g["mar"].sum(axis=1) <= 1.0)
