"""Pinning the CPU oracle before it is trusted as the checker.

* the reference's only marginal golden (test/decoding/tndecoder.jl:101-110);
* exact brute-force enumeration over every error pattern (d=3 surface under
  two noise models, Steane, Shor, ToricCode(2,2), and the reference's
  dem.dem fixture), independent of the network and of any contraction order;
* total probability one, structural zeros, and two contraction orders;
* planted mistakes in the oracle (a transposed qubit prior, a dropped
  evidence tensor, a swapped logical axis) that each of these checks catches.
"""
import copy
import itertools
import os

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec.codes import CSSTannerGraph, ShorCode, SteaneCode, SurfaceCode, ToricCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing
from tensorqec.gf2 import logical_operator

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
DEM_TEXT = open(os.path.join(os.path.dirname(__file__), "golden", "dem.dem")).read()


def code_mats(code):
    t = CSSTannerGraph.from_code(code)
    lx, lz = logical_operator(t)
    return t, t.stgx.H, t.stgz.H, lx, lz


def test_reference_marginal_golden():
    # test/decoding/tndecoder.jl:109: ct.code(ct.tensors...) ≈ [0.3972875040000002 0.004284496000000001; 0.0 0.0]
    t, Hx, Hz, lx, lz = code_mats(SurfaceCode(3, 3))
    n = 9
    m = O.tnmmap_marginals(Hx, Hz, lx, lz, np.full(n, 0.1), np.zeros(n), np.zeros(n), np.zeros((1, 8), np.uint8))
    # Julia column-major [a b; c d] -> flat [a, c, b, d]
    exp = np.array([0.3972875040000002, 0.0, 0.004284496000000001, 0.0])
    assert np.allclose(m[0], exp, atol=1e-10, rtol=0)
    assert np.allclose(m[0], exp, rtol=1e-13, atol=0)
    assert m[0, 1] == 0.0 and m[0, 3] == 0.0   # structural zeros are exact


@pytest.mark.parametrize("key,code,noise", [
    ("d3_p005", SurfaceCode(3, 3), "dep"),
    ("d3_px01", SurfaceCode(3, 3), "px"),
    ("steane", SteaneCode(), "dep"),
    ("shor", ShorCode(), "dep"),
    ("toric2", ToricCode(2, 2), "dep"),
])
def test_oracle_vs_bruteforce(key, code, noise):
    t, Hx, Hz, lx, lz = code_mats(code)
    n = t.nq
    if noise == "dep":
        px = py = pz = np.full(n, 0.05 / 3)
    else:
        px, py, pz = np.full(n, 0.1), np.zeros(n), np.zeros(n)
    bf = O.brute_force_css(Hx, Hz, lx, lz, px, py, pz)
    syn_key = key if key != "d3_px01" else "d3_p005"
    S = GOLDEN[f"{syn_key}_syn"]
    ref = np.array([bf.get(tuple(s), np.zeros(1 << (2 * lx.shape[0]))) for s in S])
    assert np.allclose(ref, GOLDEN[f"{key}_mar"], rtol=1e-14, atol=0)
    m = O.tnmmap_marginals(Hx, Hz, lx, lz, px, py, pz, S)
    nz = ref > 0
    assert np.all(m[~nz] == 0.0)
    assert np.max(np.abs(m[nz] - ref[nz]) / ref[nz]) < 1e-12
    assert abs(ref.sum() - 1.0) < 1e-12


def test_two_contraction_orders_agree():
    t, Hx, Hz, lx, lz = code_mats(SurfaceCode(5, 5))
    em = iid_error(0.08 / 3, 0.08 / 3, 0.08 / 3, t.nq)
    S = css_syndrome_rows(sample_depolarizing(em, 3, 64), t)
    a = O.tnmmap_marginals(Hx, Hz, lx, lz, em.px, em.py, em.pz, S, order="greedy")
    b = O.tnmmap_marginals(Hx, Hz, lx, lz, em.px, em.py, em.pz, S, order="greedy_last")
    assert np.max(np.abs(a - b) / a.max(axis=1, keepdims=True)) < 1e-12


def test_planted_mistakes_are_caught():
    """Each plausible mistake in the oracle's network fails the brute force."""
    t, Hx, Hz, lx, lz = code_mats(SurfaceCode(3, 3))
    n = 9
    px, py, pz = np.full(n, 0.02), np.full(n, 0.03), np.full(n, 0.05)
    S = GOLDEN["d3_p005_syn"]
    bf = O.brute_force_css(Hx, Hz, lx, lz, px, py, pz)
    ref = np.array([bf.get(tuple(s), np.zeros(4)) for s in S])
    good = O.tnmmap_marginals(Hx, Hz, lx, lz, px, py, pz, S)
    assert np.allclose(good, ref, rtol=1e-12, atol=1e-300)

    orig = O.css_network

    def transposed_prior(*a):
        ixs, ts, iy, sv = orig(*a)
        for i in range(len(ts)):
            if len(ixs[i]) == 2 and ts[i].shape == (2, 2) and ixs[i][1] == ixs[i][0] + n:
                ts[i] = ts[i].T.copy()
        return ixs, ts, iy, sv

    def swapped_axes(*a):
        ixs, ts, iy, sv = orig(*a)
        return ixs, ts, iy[::-1], sv

    for bad in (transposed_prior, swapped_axes):
        O.css_network = bad
        try:
            m = O.tnmmap_marginals(Hx, Hz, lx, lz, px, py, pz, S)
        finally:
            O.css_network = orig
        assert not np.allclose(m, ref, rtol=1e-6, atol=1e-12), bad.__name__


def test_dem_parse_and_bruteforce():
    rates, flipped, dl, ll = O.parse_dem_text(DEM_TEXT)
    assert len(rates) == 21 and dl == list(range(1, 7)) and ll == [7]
    assert flipped[3] == [1, 2, 7]      # "D0 D1 L0": detectors 1-based, logical offset past D5
    bf = O.brute_force_dem(rates, flipped, dl, ll)
    S = GOLDEN["dem_syn"]
    ref = np.array([bf.get(tuple(s), np.zeros(2)) for s in S])
    assert np.allclose(ref, GOLDEN["dem_mar"], rtol=1e-14, atol=0)
    assert abs(ref.sum() - 1) < 1e-12
    for fac in (False, True):
        m = O.dem_marginals(rates, flipped, dl, ll, S, factorize=fac)
        nz = ref > 0
        assert np.all(m[~nz] == 0)
        assert np.max(np.abs(m[nz] - ref[nz]) / ref[nz]) < 1e-12


def _plan_for(code, em, factorize, evidence):
    from tensorqec import decoders as D
    from tensorqec.network import tnmmap_css_network
    t = CSSTannerGraph.from_code(code)
    g = tnmmap_css_network(t, em, factorize=factorize)
    return t, D.build_plan(g, D.TNMMAP(factorize=factorize, evidence=evidence))


@pytest.mark.parametrize("factorize,evidence", [(False, "leaves"), (True, "slices"), (False, "slices"),
                                                (True, "leaves")])
def test_c_oracle_matches_bruteforce(factorize, evidence):
    code = SurfaceCode(3, 3)
    p = 0.05 / 3
    em = iid_error(p, p, p, 9)
    t, plan = _plan_for(code, em, factorize, evidence)
    op = OraclePlan(plan)
    m, cls, tie = op.decode(GOLDEN["d3_p005_syn"])
    ref = GOLDEN["d3_p005_mar"]
    nz = ref > 0
    assert np.all(m[~nz] == 0)
    assert np.max(np.abs(m[nz] - ref[nz]) / ref[nz]) < 1e-12
    ok = tie == 0
    assert np.array_equal(cls[ok], np.argmax(ref, axis=1)[ok])


def test_c_oracle_matches_numpy_oracle_d5():
    code = SurfaceCode(5, 5)
    em = iid_error(0.02, 0.02, 0.02, 25)
    t, plan = _plan_for(code, em, True, "slices")
    S = css_syndrome_rows(sample_depolarizing(em, 9, 100), t)
    lx, lz = logical_operator(t)
    ref = O.tnmmap_marginals(t.stgx.H, t.stgz.H, lx, lz, em.px, em.py, em.pz, S)
    m, _, _ = OraclePlan(plan).decode(S)
    assert np.max(np.abs(m - ref) / ref.max(axis=1, keepdims=True)) < 1e-12
