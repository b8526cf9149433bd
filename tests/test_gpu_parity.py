"""Parity of the HIP executors with the CPU oracle, through the C ABI.

Tolerances (stated per the decoder's contract):
  * FP64 modes (plain_f64, fused_f64, spec_f64): every marginal within 1e-12 relative of
    the oracle, structural zeros exactly zero, classes bit-exact on every
    syndrome that is not a tie (top-2 relative gap <= 1e-12), tie flags equal.
  * fused_f32, spec_f32 (plan-specialised kernel), wide_f32: marginals within
    2e-5 relative of the row maximum (FP32 with power-of-two scaling); classes
    bit-exact on every non-tie syndrome, since low-margin syndromes are
    re-contracted in FP64.
"""
import os
import sys

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, ShorCode, SteaneCode, SurfaceCode, ToricCode
from tensorqec.errors import (CSSErrorPattern, check_logical_error, css_syndrome_rows, iid_error,
                              sample_depolarizing, syndrome_extraction)
from tensorqec.gf2 import logical_operator

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
MODES = ["plain_f64", "fused_f64", "fused_f32", "wide_f32", "spec_f32", "spec_f64"]
NETS = [dict(factorize=False, evidence="leaves"), dict(factorize=True, evidence="slices")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_marginals(m, ref, mode):
    ref = np.asarray(ref)
    if mode.endswith("f32"):
        valid = ref.max(axis=1) > 0          # impossible syndromes: every class exactly 0
        assert np.all(m[~valid] == 0.0)
        err = np.max(np.abs(m[valid] - ref[valid]) / ref[valid].max(axis=1, keepdims=True))
        assert err < 2e-5, err
        return
    nz = ref > 0
    assert np.all(m[~nz] == 0.0)
    assert np.max(np.abs(m[nz] - ref[nz]) / ref[nz]) < 1e-12


def check_classes(cls, tie, ref_mar, tie_tol=1e-12):
    ref_tie = O.tie_flags(ref_mar, tie_tol)
    ok = ref_tie == 0
    assert np.array_equal(cls[ok], O.findmax(ref_mar)[ok])
    return int((~ok).sum())


@pytest.mark.parametrize("net", NETS, ids=["reference_net", "factorized_net"])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("key,code,noise", [("d3_p005", SurfaceCode(3, 3), "dep"), ("d3_px01", SurfaceCode(3, 3), "px"),
                                            ("steane", SteaneCode(), "dep"), ("shor", ShorCode(), "dep"),
                                            ("toric2", ToricCode(2, 2), "dep")])
def test_golden_all_syndromes(key, code, noise, mode, net):
    say(f"case golden {key} {mode} {net}")
    t = CSSTannerGraph.from_code(code)
    n = t.nq
    p = 0.05 / 3
    em = iid_error(p, p, p, n) if noise == "dep" else iid_error(0.1, 0.0, 0.0, n)
    ct = D.compile(D.TNMMAP(mode=mode, **net), t, em)
    S = GOLDEN[("d3_p005" if key == "d3_px01" else key) + "_syn"]
    ref = GOLDEN[key + "_mar"]
    r = D.decode_batch(ct, S)
    check_marginals(r.marginals, ref, mode)
    # impossible syndromes (all-zero rows) have no defined class; compare the rest
    valid = ref.max(axis=1) > 0
    check_classes(r.classes[valid], r.ties[valid], ref[valid])
    if not mode.endswith("f32"):
        assert np.array_equal(r.ties[valid], O.tie_flags(ref[valid], 1e-12))


def test_reference_marginal_golden_on_gpu():
    # test/decoding/tndecoder.jl:101-110 through the library
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    ct = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="plain_f64"), t, iid_error(0.1, 0.0, 0.0, 9))
    r = D.decode_batch(ct, np.zeros((1, 8), np.uint8))
    assert np.allclose(r.marginals[0], [0.3972875040000002, 0.0, 0.004284496000000001, 0.0], atol=1e-10)


def test_reference_tnmmap_decision():
    # test/decoding/tndecoder.jl:60-73: X1 error, px-only 0.1 -> no logical error, consistent syndrome
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    lx, lz = logical_operator(t)
    ex = np.array([1, 0, 0, 0, 0, 0, 0, 0, 0], np.uint8)
    ep = CSSErrorPattern(ex, np.zeros(9, np.uint8))
    for mode in MODES:
        ct = D.compile(D.TNMMAP(mode=mode), t, iid_error(0.1, 0.0, 0.0, 9))
        sx, sz = syndrome_extraction(ep, t)
        res = D.decode(ct, D.CSSSyndrome(np.zeros(4, np.uint8), sz))
        assert res.success_tag
        assert not check_logical_error(res.error_pattern, ep, lx, lz)
        sx2, sz2 = syndrome_extraction(res.error_pattern, t)
        assert not sx2.any() and np.array_equal(sz2, sz)


@pytest.mark.parametrize("code,p", [(SurfaceCode(5, 5), 0.05), (SurfaceCode(7, 7), 0.02), (SurfaceCode(7, 7), 0.05),
                                    (SurfaceCode(7, 7), 0.08), (ToricCode(3, 3), 0.05), (ToricCode(4, 4), 0.1)])
@pytest.mark.parametrize("net", NETS, ids=["reference_net", "factorized_net"])
def test_seeded_batches_vs_c_oracle(code, p, net):
    say(f"case seeded {code} p={p} {net}")
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ep = sample_depolarizing(em, seed=20260 + t.nq, count=1024)
    S = css_syndrome_rows(ep, t)
    ct = D.compile(D.TNMMAP(**net), t, em)
    ref, _, _ = OraclePlan(ct.plan).decode(S)
    lds_f64 = ct.hip.stats()["fused_lds_entries"] * 8
    for mode in MODES:
        if mode.startswith("fused") and lds_f64 > 160 * 1024:
            # a tensor this wide cannot stay in one workgroup's LDS: the generic
            # fused executors refuse the plan instead of spilling (the
            # specialised ones run it from per-workgroup slabs)
            with pytest.raises(L.TqecError) as e:
                D.decode_batch(ct, S, mode=mode)
            assert e.value.code == L.TQEC_EUNSUPPORTED
            continue
        r = D.decode_batch(ct, S, mode=mode)
        check_marginals(r.marginals, ref, mode)
        check_classes(r.classes, r.ties, ref)
        # DecodingResult contract: consistent syndrome, error pattern in the decided class
        sx, sz = syndrome_extraction(r.error_patterns, t)
        assert np.array_equal(np.concatenate([sx, sz], axis=1), S)
        lx, lz = ct.lx, ct.lz
        k = lx.shape[0]
        got = np.zeros(len(S), np.int64)
        for i in range(k):
            got |= ((r.error_patterns.zerror.astype(np.int64) @ lx[i]) & 1) << i
            got |= ((r.error_patterns.xerror.astype(np.int64) @ lz[i]) & 1) << (k + i)
        assert np.array_equal(got, r.classes)


def test_d7_full_batch_fp32_matches_fp64():
    """65,536 syndromes at d=7, p=0.05: the FP32 path (with its FP64 re-check)
    makes the same decision as plain FP64 on every non-tie syndrome."""
    t = CSSTannerGraph.from_code(SurfaceCode(7, 7))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    S = css_syndrome_rows(sample_depolarizing(em, seed=7, count=65536), t)
    plain = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="plain_f64"), t, em)
    fast = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
    a = D.decode_batch(plain, S, error_patterns=False)
    b = D.decode_batch(fast, S, error_patterns=False)
    ok = a.ties == 0
    assert np.array_equal(a.classes[ok], b.classes[ok])
    rel = np.abs(a.marginals - b.marginals) / a.marginals.max(axis=1, keepdims=True)
    say(f"d7 fp32 max rel marginal err {rel.max():.3e}, ties {int((~ok).sum())}, "
        f"rechecked {fast.hip.last_recheck_count()}")
    assert rel.max() < 2e-5


def test_device_pointer_api_matches_host_api():
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    em = iid_error(0.02, 0.02, 0.02, t.nq)
    S = css_syndrome_rows(sample_depolarizing(em, seed=3, count=5000), t)
    ct = D.compile(D.TNMMAP(), t, em)
    host = D.decode_batch(ct, S, error_patterns=False)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dm = torch.empty((len(S), 4), dtype=torch.float64, device=dev)
    dc = torch.empty(len(S), dtype=torch.int32, device=dev)
    dt = torch.empty(len(S), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    ct.hip.decode_device(dS.data_ptr(), len(S), "fused_f32", dm.data_ptr(), dc.data_ptr(), dt.data_ptr(), stream)
    torch.cuda.synchronize()
    assert np.array_equal(dc.cpu().numpy(), host.classes)
    assert np.allclose(dm.cpu().numpy(), host.marginals, rtol=1e-12, atol=0)


def test_empty_and_ragged_batches():
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    ct = D.compile(D.TNMMAP(), t, iid_error(0.05 / 3, 0.05 / 3, 0.05 / 3, 9))
    for mode in MODES:
        m, c, tie = ct.hip.decode(np.zeros((0, 8), np.uint8), mode)
        assert m.shape == (0, 4)
        for B in (1, 3, 63, 65, 257):
            S = GOLDEN["d3_p005_syn"][:B]
            m, c, tie = ct.hip.decode(S, mode)
            check_marginals(m, GOLDEN["d3_p005_mar"][:B], mode)
