"""Label slicing (tensorqec/slicing.py) on the CPU: a plan decoded in 2^k
slices (k inner labels fixed by virtual syndrome bits, rows expanded, FP64
sums) must give the unsliced plan's marginals, checked with the C oracle on
both plans (toric code L=3 over the nested-dissection network, d=3 surface
code against brute force), and a planted mistake (one slice left out of the
sum) must fail the comparison.  The reference's slicer is the `TreeSASlicer`
hook of src/nonclifford/tensornetwork.jl:174-192."""
import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing
from tensorqec.gf2 import logical_operator
from tensorqec.network import dissected_css_network
from tensorqec.planner import flatten
from tensorqec.slicing import (_sliced_cost, _step_masks, choose_slices, combine, expand_rows, slice_plan,
                                sliced_dissection_plan)


def _plan(code, p, seed=0):
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p, p, p, t.nq)
    lx, lz = logical_operator(t)
    g, tree = dissected_css_network(t, em, lxlz=(lx, lz), seed=seed)
    return t, em, lx, lz, flatten(g.without_evidence_leaves(), tree, True)


def test_choose_slices_reaches_the_target_width():
    _, _, _, _, plan = _plan(ToricCode(3, 3), 0.03)
    sc = plan.stats()["sc_dep"]
    labels = choose_slices(plan, sc - 3)
    assert 1 <= len(labels) <= 12 and len(set(labels)) == len(labels)
    sp = slice_plan(plan, labels)
    assert sp.stats()["sc_dep"] <= sc - 3
    assert sp.n_syndrome_bits == plan.n_syndrome_bits + len(labels)
    assert not set(labels) & set(plan.out_labels_final)
    # the slicer's cost model is the sliced plan's own count: every row-dependent
    # step (constant subtrees that carry a sliced label included), 2^k rows
    st = sp.stats()
    assert _sliced_cost(plan, _step_masks(plan), sum(1 << l for l in labels)) == \
        (st["tc_dep"] * 2.0 ** len(labels), st["sc_dep"])
    with pytest.raises(ValueError):
        slice_plan(plan, [plan.out_labels_final[0]])


def test_expand_rows_layout():
    S = np.array([[1, 0, 1], [0, 1, 1]], dtype=np.uint8)
    E = expand_rows(S, 2)
    assert E.shape == (8, 5)
    assert np.array_equal(E[:4, :3], np.repeat(S[:1], 4, axis=0))
    assert [int(r[3]) + 2 * int(r[4]) for r in E[4:]] == [0, 1, 2, 3]


@pytest.mark.parametrize("code,p,cut", [(ToricCode(3, 3), 0.03, 3), (SurfaceCode(3, 3), 0.05 / 3, 2)])
def test_sliced_marginals_equal_unsliced(code, p, cut):
    t, em, lx, lz, plan = _plan(code, p, seed=1)
    labels = choose_slices(plan, plan.stats()["sc_dep"] - cut)
    k = len(labels)
    S = css_syndrome_rows(sample_depolarizing(em, seed=4, count=64), t)
    ref, _, _ = OraclePlan(plan).decode(S)
    sp = slice_plan(plan, labels)
    part, _, _ = OraclePlan(sp).decode(expand_rows(S, k))
    mar, cls, tie = combine(part, len(S), k)
    nz = ref > 0
    assert np.array_equal(mar == 0, ref == 0)
    assert np.max(np.abs(mar[nz] - ref[nz]) / ref[nz]) < 1e-12
    ok = O.tie_flags(ref, 1e-12) == 0
    assert np.array_equal(cls[ok], O.findmax(ref)[ok])
    assert np.array_equal(tie, O.tie_flags(ref, 1e-12))
    if isinstance(code, SurfaceCode):
        bf = O.brute_force_css(t.stgx.H, t.stgz.H, lx, lz, em.px, em.py, em.pz)
        for row, m in zip(S, mar):
            assert np.allclose(m, bf[tuple(int(x) for x in row)], rtol=1e-12, atol=0)
    # planted mistake: one slice left out of the sum
    bad = part.reshape(len(S), 1 << k, -1).copy()
    bad[:, -1, :] = 0
    wrong = bad.sum(axis=1)
    assert np.max(np.abs(wrong[nz] - ref[nz]) / ref[nz]) > 1e-6


def test_sliced_dissection_plan_on_a_dem_matches_unsliced():
    """Tree and slices annealed together (sliced_dissection_plan) on the
    circuit-level d=3, 3-round detector error model: the per-row width meets
    the target, and the sliced plan's summed marginals equal the unsliced
    dissection plan's on 48 syndromes (C oracle on both)."""
    from tensorqec import decoders as D
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    from tensorqec.network import dissected_dem_network
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    build = lambda s: dissected_dem_network(dem, seed=s)  # noqa: E731
    _, ref_plan, _ = D.dissection_plan(build, D.Dissection(seeds=1, sweeps=50))
    target = ref_plan.stats()["sc_dep"] - 3
    plan, labels, tc, sc = sliced_dissection_plan(build, target, seeds=1, sweeps=50, iters=3, slice_sweeps=50)
    sp = slice_plan(plan, labels)
    assert sc <= target and sp.stats()["sc_dep"] == sc and len(labels) >= 1
    assert tc == sp.stats()["tc_dep"] * 2.0 ** len(labels)
    rng = np.random.default_rng(3)
    e = (rng.random((48, len(dem.error_rates))) < 20 * np.asarray(dem.error_rates)[None, :]).astype(np.int64)
    S = ((e @ dem.tanner().H.T.astype(np.int64)) & 1).astype(np.uint8)
    ref, _, _ = OraclePlan(ref_plan).decode(S)
    part, _, _ = OraclePlan(sp).decode(expand_rows(S, len(labels)))
    mar, _, _ = combine(part, len(S), len(labels))
    nz = ref > 0
    assert np.array_equal(mar == 0, ~nz)
    assert np.max(np.abs(mar[nz] - ref[nz]) / ref[nz]) < 1e-12


def test_slicing_a_plan_with_one_hot_evidence_leaves():
    """The drop-in network (the reference's, evidence as one-hot leaves
    selected by syndrome bits, tndecoder.jl:148-158) sliced on two labels:
    the virtual syndrome bits follow the leaves' own, and the summed slices
    equal the unsliced plan (C oracle on both)."""
    from tensorqec import decoders as D
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    em = iid_error(0.05, 0.05, 0.05, t.nq)
    plan = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="plain_f64", tabulate_bits=0), t, em).plan
    assert any(s >= 0 for s in plan.leaf_slot) and not plan.evidence_label_bits
    labels = choose_slices(plan, plan.stats()["sc_dep"] - 2)
    assert labels
    S = css_syndrome_rows(sample_depolarizing(em, seed=3, count=32), t)
    ref, _, _ = OraclePlan(plan).decode(S)
    part, _, _ = OraclePlan(slice_plan(plan, labels)).decode(expand_rows(S, len(labels)))
    mar, _, _ = combine(part, len(S), len(labels))
    nz = ref > 0
    assert np.array_equal(mar == 0, ~nz)
    assert np.max(np.abs(mar[nz] - ref[nz]) / ref[nz]) < 1e-12


def test_sliced_decoder_decides_close_sums_again_in_fp64():
    """SlicedDecoder's precision rule, with the library's decode replaced by
    the oracle on the sliced plan: in "spec_f32" the marginal of each
    syndrome's leading class is planted 3e-4 too low in every slice;
    syndromes whose summed top-2 gap is within 1e-3 of the maximum (toric
    L=3, p=0.03: two of 128 exactly, the closest at 1.0e-4, and whichever
    the planted error pushes below it) are decoded again in "spec_f64"
    (exact) and only those, so every class matches the oracle.  With the
    re-check switched off the planted error flips the closest syndrome's
    class."""
    from tensorqec.slicing import SlicedDecoder, needs_fp64
    t, em, lx, lz, plan = _plan(ToricCode(3, 3), 0.03, seed=1)
    labels = choose_slices(plan, plan.stats()["sc_dep"] - 2)
    S = css_syndrome_rows(sample_depolarizing(em, seed=9, count=128), t)
    ref, _, _ = OraclePlan(plan).decode(S)
    flagged = needs_fp64(ref, 1e-3)
    assert flagged.sum() == 2
    ok = O.tie_flags(ref, 1e-12) == 0
    m = 1 << len(labels)

    def decoder(recheck_tol):
        dec = SlicedDecoder(plan, labels, recheck_tol=recheck_tol)
        orc = OraclePlan(dec.plan)
        calls = []

        class OracleLibrary:
            def decode(self, rows, mode, want_marginals):
                calls.append((mode, len(rows)))
                mar, cls, tie = orc.decode(rows)
                if mode == "spec_f32":
                    g = mar.reshape(-1, m, mar.shape[1])
                    top = g.sum(axis=1).argmax(axis=1)
                    g[np.arange(len(g))[:, None], np.arange(m)[None, :], top[:, None]] *= 1.0 - 3e-4
                    mar = g.reshape(mar.shape)
                return mar, cls, tie

        dec.hip = OracleLibrary()
        return dec, calls

    dec, calls = decoder(1e-3)
    tot, cls, tie = dec.decode(S, "spec_f32")
    # the flags come from the FP32 sums: every close syndrome, plus any the
    # planted error pushed below the margin, and nothing else
    n_re = dec.last_rechecked
    assert 2 <= n_re <= 4
    assert calls == [("spec_f32", len(S) * m), ("spec_f64", n_re * m)]
    redone = np.all(np.abs(tot - ref) <= 1e-12 * ref.max(axis=1, keepdims=True), axis=1)
    assert redone.sum() == n_re and np.all(redone[flagged])
    assert np.array_equal(cls[ok], O.findmax(ref)[ok])
    assert np.array_equal(tie, O.tie_flags(ref, 1e-12))
    # planted mistake: no re-check of close sums
    dec0, _ = decoder(0.0)
    _, cls0, _ = dec0.decode(S, "spec_f32")
    assert dec0.last_rechecked == 0
    assert not np.array_equal(cls0[ok], O.findmax(ref)[ok])
