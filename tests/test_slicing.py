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
from tensorqec.slicing import choose_slices, combine, expand_rows, slice_plan


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
