"""Full-size oracle parity of the configurations the throughput benchmark
runs, through the C ABI (reference call: `decode(ct, syndrome)`,
src/decoding/tndecoder.jl:160-165, evaluated for every syndrome).

  * d=7 rotated surface code, code-capacity depolarizing p in {0.02, 0.05,
    0.08}: whole 65,536-syndrome batches through the register-resident
    plan-specialised FP32 kernel in the benchmark's configurations (tree
    seeds 11 and 14; subtrees on up to 6, 7 or 8 syndrome bits tabulated;
    256 threads; 2 waves per SIMD), and through the exact FP64
    register-resident kernel, against the CPU oracle's FP64 executor on
    the same plan and the same syndromes.
  * d=9 on 2,048 syndromes: every fused executor the plan fits.
  * toric code L=5 (16 classes), nested-dissection network of width 2^21:
    the plain and specialised FP64, wide FP32 and specialised FP32 (with and
    without FP32-MFMA steps) executors on the 16
    syndromes of tests/golden/toric5.npz (the C oracle's marginals, written
    by tests/golden/make_toric5.py; about a CPU-minute per syndrome).
Tolerances: tests/_parity.py (FP64 1e-12 relative with exact zeros, tie
flags equal; FP32 2e-5 of the row maximum; classes bit-exact on every
non-tie syndrome).
"""
import sys

import numpy as np
import pytest

from _parity import check_f32, check_f64
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(torch, ct, dS, n, mode):
    dev = torch.device("cuda:0")
    nout = 1 << ct.hip.n_out
    mar = torch.full((n, nout), -1.0, dtype=torch.float64, device=dev)
    cls = torch.full((n,), -1, dtype=torch.int32, device=dev)
    tie = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    ct.hip.decode_device(dS.data_ptr(), n, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr())
    torch.cuda.synchronize()
    return mar.cpu().numpy(), cls.cpu().numpy(), tie.cpu().numpy()


@pytest.mark.parametrize("p,seed,tab", [(0.02, 11, 7), (0.05, 11, 7), (0.08, 11, 7),
                                        (0.05, 14, 6), (0.05, 14, 7), (0.05, 11, 8)])
def test_d7_benchmark_configuration_full_batch(p, seed, tab):
    torch = pytest.importorskip("torch")
    n = 65536
    say(f"case d7 full batch p={p} tree seed {seed} tabulate {tab}")
    t = CSSTannerGraph.from_code(SurfaceCode(7, 7))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="spec_f32", optimizer=D.TreeSA(seed=seed), tabulate_bits=tab), t, em)
    assert ct.hip.stats()["tabulate_bits"] == tab
    S = css_syndrome_rows(sample_depolarizing(em, seed=int(p * 1e4) + 31 * seed + tab, count=n), t)
    ref, _, _ = OraclePlan(ct.plan).decode(S)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    ct.hip.set_spec(64, 256, False, 2, reg=True)
    cfg = ct.hip.spec_config()
    assert cfg["reg"] and cfg["threads"] == 256 and cfg["waves"] == 2
    m, c, _ = run(torch, ct, dS, n, "spec_f32")
    err, ties = check_f32(m, c, ref, f"spec_f32 p={p}")
    say(f"  spec_f32 (register-resident): max err {err:.2e} of the row maximum, {ties} ties, "
        f"{ct.hip.last_recheck_count()} FP64 re-checks")
    m, c, tie = run(torch, ct, dS, n, "spec_f64")
    err = check_f64(m, c, tie, ref, f"spec_f64 p={p}")
    say(f"  spec_f64: max rel err {err:.2e}")


def test_d9_vs_oracle():
    torch = pytest.importorskip("torch")
    n = 2048
    t = CSSTannerGraph.from_code(SurfaceCode(9, 9))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
    S = css_syndrome_rows(sample_depolarizing(em, seed=909, count=n), t)
    ref, _, _ = OraclePlan(ct.plan).decode(S)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    for mode in ("spec_f32", "spec_f64", "fused_f32", "fused_f64"):
        say(f"case d9 {mode}")
        try:
            m, c, tie = run(torch, ct, dS, n, mode)
        except L.TqecError as e:
            assert e.code == L.TQEC_EUNSUPPORTED and mode.startswith("fused")
            say(f"  {mode}: plan does not fit the generic fused executor")
            continue
        if mode.endswith("f32"):
            err, ties = check_f32(m, c, ref, f"d9 {mode}")
        else:
            err = check_f64(m, c, tie, ref, f"d9 {mode}")
        say(f"  max err {err:.2e}")


@pytest.mark.parametrize("mode", ["plain_f64", "spec_f64", "wide_f32", "spec_f32", "spec_f32_mfma32"])
def test_toric5_vs_oracle_fixture(mode):
    import os
    torch = pytest.importorskip("torch")
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "toric5.npz"))
    S, ref, p = g["syn"], g["mar"], float(g["p"])
    say(f"case toric L=5 {mode}: {len(S)} syndromes")
    t = CSSTannerGraph.from_code(ToricCode(5, 5))
    em = iid_error(p, p, p, t.nq)
    mfma32 = mode.endswith("_mfma32")
    mode = mode.replace("_mfma32", "")
    ct = D.compile(D.TNMMAP(mode=mode, optimizer=D.Dissection(), tabulate_bits=0), t, em)
    assert ct.hip.stats()["sc"] <= 22
    if mfma32:
        # the 512 x 4096 x 512 step (81% of the multiply-adds) on FP32 MFMA
        ct.hip.set_spec(1, 256, False, 0, slab=True, unroll=4, mfma32=True)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    mar, cls, tie = run(torch, ct, dS, len(S), mode)
    if mode.endswith("_f64"):
        check_f64(mar, cls, tie, ref, mode)
    else:
        # the plan's last step contracts 18 labels: every marginal is one
        # FP32 accumulation chain of 2^18 products, whose rounding reaches
        # ~2.5e-5 of the row maximum, so this plan is held to 1e-4; classes
        # stay exact on every non-tie syndrome (FP64 re-check below a 1e-3
        # margin)
        check_f32(mar, cls, ref, mode, tol=1e-4)


@pytest.mark.parametrize("mode", ["plain_f64", "spec_f32"])
def test_toric5_sliced_vs_oracle_fixture(mode):
    """Label slicing on the GPU (tensorqec/slicing.py; the reference's slicer
    hook is src/nonclifford/tensornetwork.jl:174-192): the toric L=5 plan with
    its widest intermediates cut by two ranks, each syndrome decoded as 2^k
    rows with the sliced labels fixed and the rows' marginals summed, against
    the oracle's marginals in tests/golden/toric5.npz (FP64: 1e-12 relative;
    FP32: 1e-4 of the row maximum)."""
    import os
    from tensorqec.slicing import SlicedDecoder, choose_slices
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "toric5.npz"))
    S, ref, p = g["syn"], g["mar"], float(g["p"])
    t = CSSTannerGraph.from_code(ToricCode(5, 5))
    em = iid_error(p, p, p, t.nq)
    ct = D.compile(D.TNMMAP(mode=mode, optimizer=D.Dissection(), tabulate_bits=0), t, em)
    sc = ct.plan.stats()["sc_dep"]
    labels = choose_slices(ct.plan, sc - 2)
    dec = SlicedDecoder(ct.plan, labels)
    assert dec.plan.stats()["sc_dep"] <= sc - 2
    say(f"case toric L=5 sliced {mode}: {len(labels)} labels, width 2^{dec.plan.stats()['sc_dep']}")
    mar, cls, tie = dec.decode(S, mode)
    if mode.endswith("_f64"):
        check_f64(mar, cls, tie, ref, mode)
    else:
        check_f32(mar, cls, ref, mode, tol=1e-4)


@pytest.mark.parametrize("mode", ["spec_f64", "spec_f32"])
def test_toric6_sliced_vs_oracle_fixture(mode):
    """Toric L=6 (2^36.3 multiply-adds per syndrome, widest tensor 2^27
    unsliced), decoded with the tree and the sliced labels annealed together
    for rows of width 2^21 (tensorqec.slicing.sliced_dissection_plan),
    against the oracle's marginals in tests/golden/toric6.npz (FP64: 1e-12
    relative; FP32: 1e-4 of the row maximum)."""
    import os
    from tensorqec.gf2 import logical_operator
    from tensorqec.network import dissected_css_network
    from tensorqec.slicing import SlicedDecoder, sliced_dissection_plan
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "toric6.npz"))
    S, ref, p = g["syn"], g["mar"], float(g["p"])
    t = CSSTannerGraph.from_code(ToricCode(6, 6))
    em = iid_error(p, p, p, t.nq)
    lx, lz = logical_operator(t)
    plan, labels, _, sc = sliced_dissection_plan(lambda s: dissected_css_network(t, em, lxlz=(lx, lz), seed=s), 21)
    assert sc <= 21
    dec = SlicedDecoder(plan, labels)
    say(f"case toric L=6 sliced {mode}: {len(labels)} labels, width 2^{sc}")
    mar, cls, tie = dec.decode(S, mode)
    if mode.endswith("_f64"):
        check_f64(mar, cls, tie, ref, mode)
    else:
        check_f32(mar, cls, ref, mode, tol=1e-4)


@pytest.mark.parametrize("mode", ["plain_f64", "spec_f64", "spec_f32"])
def test_sliced_one_hot_leaves_vs_oracle(mode):
    """The drop-in network (the reference's, evidence as one-hot leaves
    selected by syndrome bits, tndecoder.jl:148-158) sliced on two labels and
    decoded on the GPU (the sliced labels become syndrome bits after the
    leaves' own), against the oracle on the unsliced plan, d=5."""
    from tensorqec.slicing import SlicedDecoder, choose_slices
    t = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    em = iid_error(0.05 / 3, 0.05 / 3, 0.05 / 3, t.nq)
    plan = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="plain_f64", tabulate_bits=0), t, em).plan
    labels = choose_slices(plan, plan.stats()["sc_dep"] - 2)
    assert labels
    S = css_syndrome_rows(sample_depolarizing(em, seed=8, count=256), t)
    ref, _, _ = OraclePlan(plan).decode(S)
    say(f"case sliced one-hot leaves d=5 {mode}: {len(labels)} labels")
    mar, cls, tie = SlicedDecoder(plan, labels).decode(S, mode)
    if mode.endswith("_f64"):
        check_f64(mar, cls, tie, ref, mode)
    else:
        check_f32(mar, cls, ref, mode, tol=2e-5)
