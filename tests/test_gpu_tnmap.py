"""TNMAP (max-product contraction + arg-max backtrace) through the C ABI.

Pins: exact brute-force maxima over all 4^9 Pauli errors of the d=3 surface
code for every syndrome; the reference's unique-MAP known answers
(test/decoding/ipdecoder.jl:17-23); the TNMAP-vs-TNMMAP class decision of
test/decoding/tndecoder.jl:75-83; agreement of two networks (reference and
chain-factorized).  Which of several equally probable configurations is
returned is not pinned (the reference's tie-breaking lives in
TensorInference): only that the returned one attains the maximum.
"""
import numpy as np
import pytest

from oracle import tn_oracle as O
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.errors import (CSSErrorPattern, IndependentFlipError, check_logical_error, css_syndrome_rows,
                              iid_error, sample_depolarizing, syndrome_extraction, syndrome_extraction_H)
from tensorqec.gf2 import logical_operator
from tensorqec.tnmap import TNMAP, compile_tnmap, decode_tnmap, decode_tnmap_batch
from tensorqec import decoders as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("factorize", [False, True])
@pytest.mark.parametrize("px,py,pz", [(0.05 / 3,) * 3, (0.1, 0.0, 0.0), (0.02, 0.03, 0.05)])
def test_d3_all_syndromes_attain_bruteforce_max(factorize, px, py, pz):
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    em = iid_error(px, py, pz, 9)
    ct = compile_tnmap(TNMAP(factorize=factorize), t, em)
    bf = O.brute_force_map_css(t.stgx.H, t.stgz.H, em.px, em.py, em.pz)
    # syndromes of probability zero (e.g. any sx under X-only noise) have no MAP
    S = np.array(sorted(k for k, v in bf.items() if v > 0), dtype=np.uint8)
    cfg, logp = decode_tnmap_batch(ct, S)
    ex, ez = cfg[:, :9], cfg[:, 9:]
    sx, sz = syndrome_extraction(CSSErrorPattern(ex, ez), t)
    assert np.array_equal(np.concatenate([sx, sz], axis=1), S)          # consistent
    ref = np.array([bf[tuple(s)] for s in S])
    got = O.css_config_probability(ex, ez, em.px, em.py, em.pz)
    assert np.max(np.abs(got - ref) / ref) < 1e-12                       # attains the maximum
    assert np.max(np.abs(np.exp(logp) - ref) / ref) < 1e-12


def test_reference_unique_map_known_answers():
    # test/decoding/ipdecoder.jl:17-23 (classical problem on the Z checks)
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    e = np.array([0, 0, 0, 1, 0, 0, 0, 0, 0], np.uint8)
    syd = syndrome_extraction_H(e, t.stgz.H)
    ct = compile_tnmap(TNMAP(), t.stgz, IndependentFlipError(0.02 * np.arange(1, 10)))
    assert decode_tnmap(ct, D.SimpleSyndrome(syd)).error_pattern.tolist() == [0, 0, 0, 0, 0, 0, 1, 0, 0]
    ct = compile_tnmap(TNMAP(), t.stgz, IndependentFlipError(0.02 * np.arange(9, 0, -1)))
    assert decode_tnmap(ct, D.SimpleSyndrome(syd)).error_pattern.tolist() == [0, 0, 0, 1, 0, 0, 0, 0, 0]


def test_reference_tnmap_vs_tnmmap_decision():
    # test/decoding/tndecoder.jl:60-83: with p2 = p3 = 0.26 TNMAP picks the
    # logically different class, with a consistent syndrome
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    lx, lz = logical_operator(t)
    ep = CSSErrorPattern(np.array([1, 0, 0, 0, 0, 0, 0, 0, 0], np.uint8), np.zeros(9, np.uint8))
    sx, sz = syndrome_extraction(ep, t)
    pv = np.full(9, 0.1)
    pv[1] = pv[2] = 0.26
    ct = compile_tnmap(TNMAP(), t, iid_error(0.0, 0.0, 0.0, 9).__class__(pv, np.zeros(9), np.zeros(9)))
    res = decode_tnmap(ct, D.CSSSyndrome(np.zeros(4, np.uint8), sz))
    assert check_logical_error(res.error_pattern, ep, lx, lz)
    sx2, sz2 = syndrome_extraction(res.error_pattern, t)
    assert not sx2.any() and np.array_equal(sz2, sz)


@pytest.mark.parametrize("code,p", [(SurfaceCode(5, 5), 0.05), (ToricCode(3, 3), 0.08), (SurfaceCode(7, 7), 0.05)])
def test_networks_agree_and_configs_are_valid(code, p):
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    S = css_syndrome_rows(sample_depolarizing(em, 31, 512), t)
    outs = []
    for fac in ((False, True) if t.nq <= 25 else (True,)):
        ct = compile_tnmap(TNMAP(factorize=fac), t, em)
        cfg, logp = decode_tnmap_batch(ct, S)
        ex, ez = cfg[:, :t.nq], cfg[:, t.nq:]
        sx, sz = syndrome_extraction(CSSErrorPattern(ex, ez), t)
        assert np.array_equal(np.concatenate([sx, sz], axis=1), S)
        pr = O.css_config_probability(ex, ez, em.px, em.py, em.pz)
        assert np.max(np.abs(np.log(pr) - logp)) < 1e-9
        outs.append(logp)
    if len(outs) == 2:
        assert np.max(np.abs(outs[0] - outs[1])) < 1e-9


@pytest.mark.parametrize("code,p,fac", [(SurfaceCode(3, 3), 0.1, True), (SurfaceCode(5, 5), 0.05, True),
                                        (SurfaceCode(5, 5), 0.05, False), (ToricCode(3, 3), 0.08, True),
                                        (SurfaceCode(7, 7), 0.05, True), (SurfaceCode(9, 9), 0.05, True)])
def test_specialised_map_kernel_bit_identical_to_generic(code, p, fac):
    """The max-product kernel generated for the plan performs the generic
    kernel's comparisons in the same order and the same backtrace: identical
    configurations (tie-breaking included) and log-maxima."""
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    S = css_syndrome_rows(sample_depolarizing(em, 77 + t.nq, 3000 if t.nq < 81 else 1000), t)
    ct = compile_tnmap(TNMAP(factorize=fac), t, em)
    ct.hip.set_map_kernel(False)
    assert ct.hip.map_kernel() == (False, "")
    cfg0, lp0 = decode_tnmap_batch(ct, S)
    ct.hip.set_map_kernel(True)
    cfg1, lp1 = decode_tnmap_batch(ct, S)
    # the generated kernel was built and used (no fall-back to the generic one)
    assert ct.hip.map_kernel() == (True, "")
    assert np.array_equal(cfg0, cfg1)
    assert np.array_equal(lp0, lp1)
