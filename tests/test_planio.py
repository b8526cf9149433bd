"""Plan files (tensorqec/planio.py and the committed d=7 plans under
tensorqec.jl_amd/plans/): a plan survives the round trip bit for bit; the
committed files are what the seeded planner still produces (tree, label
orders, leaf tensors) and carry the library's own per-syndrome work, so
the benchmark's W_flops can be recomputed from them; the tree is the same
at every p; and the file's plan computes the reference network's marginals
(the C oracle on the reference network along another tree, 256 syndromes
per p)."""
import importlib.util
import json
import os

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec.codes import CSSTannerGraph, SurfaceCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing
from tensorqec.gf2 import logical_operator
from tensorqec.planio import load_plan, plan_from_dict, plan_to_dict, plan_work

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLANS = os.path.join(REPO, "tensorqec.jl_amd", "plans")
PS = (0.02, 0.05, 0.08)


def _make_plans():
    spec = importlib.util.spec_from_file_location("make_plans", os.path.join(PLANS, "make_plans.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _file(p):
    with open(_make_plans().path_for(p)) as f:
        return json.load(f)


def test_plan_round_trip_is_exact():
    plan = _make_plans().d7_plan(0.05)
    back = plan_from_dict(json.loads(json.dumps(plan_to_dict(plan))))
    assert plan_to_dict(back) == plan_to_dict(plan)
    for a, b in zip(plan.leaf_data, back.leaf_data):
        assert (a is None and b is None) or np.array_equal(np.asarray(a, np.float64).reshape(-1), b)


@pytest.mark.parametrize("p", PS)
def test_committed_plan_is_what_the_planner_produces(p):
    mp = _make_plans()
    d = _file(p)
    fresh = mp.d7_plan(p)
    want = plan_to_dict(fresh)
    assert {k: v for k, v in d.items() if k not in ("meta", "work")} == want
    assert d["work"] == json.loads(json.dumps(plan_work(fresh, mp.TABS)))
    # the headline kernel's algorithmic flops per syndrome at the tuner's depths
    assert d["work"]["8"]["w_flops_per_syndrome"] == 2 * d["work"]["8"]["dep_tc"]


def test_tree_is_the_same_at_every_p():
    keys = ("lhs", "rhs", "out_labels", "leaf_labels", "leaf_slot", "evidence_label_bits", "out_labels_final")
    ds = [_file(p) for p in PS]
    for d in ds[1:]:
        assert all(d[k] == ds[0][k] for k in keys)
        assert d["work"] == ds[0]["work"]


@pytest.mark.parametrize("p", PS)
def test_committed_plan_computes_the_reference_marginals(p):
    """Against the reference network itself (tndecoder.jl:97-146, evidence
    as one-hot leaves) contracted along another tree, both by the C oracle."""
    from tensorqec import decoders as D
    from tensorqec.network import tnmmap_css_network
    from tensorqec.planner import flatten
    plan, meta, work = load_plan(_make_plans().path_for(p))
    t = CSSTannerGraph.from_code(SurfaceCode(7, 7))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    lx, lz = logical_operator(t)
    g = tnmmap_css_network(t, em, factorize=False, lxlz=(lx, lz))
    other = flatten(g, D.optimize_tree(g.ixs, g.iy, D.TreeSA(seed=5)), False)
    S = css_syndrome_rows(sample_depolarizing(em, seed=17, count=256), t)
    ref, _, _ = OraclePlan(other).decode(S)
    mar, _, _ = OraclePlan(plan).decode(S)
    assert np.max(np.abs(mar - ref) / ref.max(axis=1, keepdims=True)) < 1e-12
    ok = O.tie_flags(ref, 1e-12) == 0
    assert np.array_equal(O.findmax(mar)[ok], O.findmax(ref)[ok])


def test_dem_and_sliced_plans_round_trip():
    """A detector-error-model plan (evidence as label slices, nested
    dissection) and its sliced form (virtual syndrome bits) survive the file
    round trip, and the library compiles the copy to the same plan."""
    import tempfile
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    from tensorqec.planio import save_plan
    from tensorqec.slicing import choose_slices, slice_plan
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    plan = D.compile(D.TNMMAP(), dem).plan
    sliced = slice_plan(plan, choose_slices(plan, plan.stats()["sc_dep"] - 2))
    for p_ in (plan, sliced):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "plan.json")
            save_plan(p_, path, meta={"k": 1}, tabulate_bits=(0, 4))
            back, meta, work = load_plan(path)
        assert plan_to_dict(back) == plan_to_dict(p_) and meta == {"k": 1}
        assert back.evidence_label_bits == p_.evidence_label_bits
        a, b = L.HipPlan(back, tabulate_bits=4).stats(), L.HipPlan(p_, tabulate_bits=4).stats()
        assert a == b and work["4"]["dep_tc"] == a["dep_tc"]
