"""Plan-specialised kernels (TQEC_MODE_SPEC_F32 / SPEC_F64) against the
generic fused kernels of the same precision on the same syndromes:
marginals, classes and tie flags must be bit-identical — the generated kernel performs the same operations in
the same order, with the descriptor values substituted.  This holds for the
LDS-tile kernels, the FP64 kernels and the register-resident FP32 kernel
with per-step scaling (TQEC_SPEC_DYNSCALE).  The register-resident FP32
kernel's default, plan-time static exponents, changes the arithmetic: it is
checked against the CPU oracle's FP64 contraction of the same plan on the
same syndromes instead.  Several tile configurations and the list (gated
fall-back) entry point are covered.  Deep tabulation, plans too wide for LDS
and the BF16-MFMA mode are also checked against the oracle (tolerances in
tests/_parity.py; BF16 steps 3e-2 of the row maximum)."""
import sys

import numpy as np
import pytest

from _parity import check_f32
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, ShorCode, SteaneCode, SurfaceCode, ToricCode
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


@pytest.mark.parametrize("name,code,p,n", [("surface3", SurfaceCode(3, 3), 0.05, 777),
                                           ("steane", SteaneCode(), 0.08, 300), ("shor", ShorCode(), 0.03, 129),
                                           ("surface5", SurfaceCode(5, 5), 0.06, 4097),
                                           ("surface7", SurfaceCode(7, 7), 0.05, 20000),
                                           ("toric3", ToricCode(3, 3), 0.07, 65)])
def test_spec_bit_identical_to_generic(name, code, p, n):
    torch = pytest.importorskip("torch")
    say(f"case spec {name} n={n}")
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
    S = css_syndrome_rows(sample_depolarizing(em, seed=4242 + t.nq, count=n), t)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    try:
        ref = run(torch, ct, dS, n, "fused_f32")
    except L.TqecError as e:
        assert e.code == L.TQEC_EUNSUPPORTED
        pytest.skip(f"{name}: plan does not fit the fused executors")
    say("  spec_f64 vs fused_f64")
    try:
        ref64 = run(torch, ct, dS, n, "fused_f64")
    except L.TqecError as e:
        assert e.code == L.TQEC_EUNSUPPORTED
        ref64 = None
    if ref64 is not None:
        got64 = run(torch, ct, dS, n, "spec_f64")
        for a, b in zip(got64, ref64):
            assert np.array_equal(a, b), (name, "f64")
    configs = [(1, 64, 1), (4, 128, 1), (16, 256, 1), (4, 64, 2), (2, 64, 4), (8, 256, 4)]
    for cfg in configs:
        try:
            ct.hip.set_spec(cfg[0], cfg[1], False, 0, False, cfg[2])
        except L.TqecError as e:
            assert e.code == L.TQEC_EUNSUPPORTED
            continue
        say(f"  tile {cfg}")
        got = run(torch, ct, dS, n, "spec_f32")
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (name, cfg)


@pytest.mark.parametrize("name,code,p,n", [("surface3", SurfaceCode(3, 3), 0.05, 777),
                                           ("steane", SteaneCode(), 0.08, 300), ("shor", ShorCode(), 0.03, 129),
                                           ("surface5", SurfaceCode(5, 5), 0.06, 4097),
                                           ("surface7", SurfaceCode(7, 7), 0.05, 20001),
                                           ("toric3", ToricCode(3, 3), 0.07, 65)])
def test_spec_reg_bit_identical_to_generic(name, code, p, n):
    """Register-resident kernel (one syndrome per quad of lanes, DPP
    exchanges).  With per-step scaling: the same operations in the same
    order as the generic FP32 kernel, so bit-identical, with the same FP64
    re-checks, for every thread count, table placement and register cap.
    With static exponents (the default): within the FP32 tolerance of the
    CPU oracle's FP64 contraction, classes equal on every non-tie syndrome.
    Also through the list entry point."""
    torch = pytest.importorskip("torch")
    say(f"case spec reg {name} n={n}")
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
    S = css_syndrome_rows(sample_depolarizing(em, seed=777 + t.nq, count=n), t)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
    try:
        ref = run(torch, ct, dS, n, "fused_f32")
    except L.TqecError as e:
        assert e.code == L.TQEC_EUNSUPPORTED
        pytest.skip(f"{name}: plan does not fit the fused executors")
    ref_rechecks = ct.hip.last_recheck_count()
    ref_generic = ref
    m64, _, _ = OraclePlan(ct.plan).decode(S)
    done = 0
    for thr, tab, wv in ((64, False, 0), (128, True, 0), (256, True, 0), (256, False, 2)):
        for dyn in (True, False):
            try:
                ct.hip.set_spec(thr // 4, thr, tab, wv, reg=True, dyn=dyn)
            except L.TqecError as e:
                assert e.code == L.TQEC_EUNSUPPORTED
                continue
            cfg = ct.hip.spec_config()
            assert cfg["reg"] and cfg["dyn"] == dyn
            say(f"  threads {thr} tables_in_lds {tab} waves {wv} per-step scaling {dyn}")
            got = run(torch, ct, dS, n, "spec_f32")
            if dyn:
                for a, b in zip(got, ref_generic):
                    assert np.array_equal(a, b), (name, thr, tab, wv)
                # the same syndromes were re-checked in FP64 (per-tile re-check segments)
                assert ct.hip.last_recheck_count() == ref_rechecks, (name, thr, ref_rechecks)
            else:
                err, ties = check_f32(got[0], got[1], m64, f"{name} static {thr}")
                say(f"    static exponents: max err {err:.2e} of the row maximum, {ties} ties, "
                    f"{ct.hip.last_recheck_count()} FP64 re-checks (per-step scaling: {ref_rechecks})")
                ref = got
            done += 1
    if done == 0:
        pytest.skip(f"{name}: too many live entries per lane for registers")
    say("  list entry point")
    rows = np.arange(1, n, 3, dtype=np.int32)
    dl = torch.from_numpy(rows).to(dev)
    dc = torch.tensor([len(rows)], dtype=torch.int32, device=dev)
    nout = 1 << ct.hip.n_out
    mar = torch.zeros((n, nout), dtype=torch.float64, device=dev)
    cls = torch.full((n,), -1, dtype=torch.int32, device=dev)
    tie = torch.zeros(n, dtype=torch.uint8, device=dev)
    ct.hip.decode_list_device(dS.data_ptr(), dl.data_ptr(), dc.data_ptr(), len(rows), "spec_f32", mar.data_ptr(),
                              cls.data_ptr(), tie.data_ptr())
    torch.cuda.synchronize()
    c = cls.cpu().numpy()
    assert np.array_equal(c[rows], ref[1][rows])
    untouched = np.ones(n, bool)
    untouched[rows] = False
    assert np.all(c[untouched] == -1)
    assert np.array_equal(mar.cpu().numpy()[rows], ref[0][rows])


@pytest.mark.parametrize("name,code,tab,factorize", [("surface3x5", SurfaceCode(3, 5), 0, False),
                                                     ("steane", SteaneCode(), 8, True),
                                                     ("toric2", ToricCode(2, 2), 2, True),
                                                     ("surface5", SurfaceCode(5, 5), 8, False),
                                                     ("shor", ShorCode(), 1, False)])
def test_spec_reg_plan_variants_bit_identical(name, code, tab, factorize):
    """Register-resident kernels on plans of other shapes (rectangular code,
    no chain factorization, tabulation depth 0-8, 16 classes for the toric
    code): FP32 with per-step scaling bit-identical to the generic FP32
    kernel, FP64 to the generic FP64 kernel; FP32 with static exponents
    within the FP32 tolerance of the CPU oracle."""
    torch = pytest.importorskip("torch")
    say(f"case spec reg variant {name} tab={tab} factorize={factorize}")
    t = CSSTannerGraph.from_code(code)
    em = iid_error(0.02, 0.03, 0.01, t.nq)
    ct = D.compile(D.TNMMAP(mode="fused_f32", factorize=factorize, tabulate_bits=tab), t, em)
    n = 2500
    S = css_syndrome_rows(sample_depolarizing(em, seed=1234 + t.nq + tab, count=n), t)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    ct.hip.set_spec(16, 64, False, 0, reg=True, dyn=True)
    assert ct.hip.spec_config()["reg"]
    for mode, ref_mode in (("spec_f32", "fused_f32"), ("spec_f64", "fused_f64")):
        ref = run(torch, ct, dS, n, ref_mode)
        got = run(torch, ct, dS, n, mode)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (name, mode)
    ct.hip.set_spec(16, 64, False, 0, reg=True)
    got = run(torch, ct, dS, n, "spec_f32")
    m64, _, _ = OraclePlan(ct.plan).decode(S)
    err, ties = check_f32(got[0], got[1], m64, f"{name} static")
    say(f"  static exponents: max err {err:.2e}, {ties} ties")


@pytest.mark.parametrize("d,tab,n", [(5, 5, 3001), (5, 6, 3001), (7, 6, 8192), (7, 7, 8192), (7, 8, 8192)])
def test_deep_tabulation_matches_oracle(d, tab, n):
    """Subtrees on up to 8 syndrome bits tabulated at plan creation (up to
    256 variants each): the register-resident kernel with per-step scaling
    stays bit-identical to the generic FP32 kernel, and both it and the
    static-exponent default agree with the CPU oracle's FP64 contraction
    (classes on every non-tie syndrome, marginals to 2e-5 of the row
    maximum)."""
    torch = pytest.importorskip("torch")
    say(f"case tabulate_bits={tab} d={d} n={n}")
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="fused_f32", optimizer=D.TreeSA(seed=14), tabulate_bits=tab), t, em)
    assert ct.hip.stats()["tabulate_bits"] == tab
    S = css_syndrome_rows(sample_depolarizing(em, seed=606 + d + tab, count=n), t)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    ref = run(torch, ct, dS, n, "fused_f32")
    ct.hip.set_spec(64, 256, False, 2, reg=True, dyn=True)
    got = run(torch, ct, dS, n, "spec_f32")
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    m64, _, _ = OraclePlan(ct.plan).decode(S)
    err, _ = check_f32(got[0], got[1], m64, f"tab {tab}")
    ct.hip.set_spec(64, 256, False, 2, reg=True)
    got = run(torch, ct, dS, n, "spec_f32")
    err_s, _ = check_f32(got[0], got[1], m64, f"tab {tab} static")
    say(f"  max err {err:.2e} (per-step scaling), {err_s:.2e} (static exponents)")


@pytest.mark.parametrize("reg", [True, False])
def test_spec_call_replays_from_a_captured_graph(reg):
    """A SPEC_F32 call captured into a HIP graph (torch.cuda.graph on a side
    stream) replays correctly several times in a row, with eager work on
    other batches between the replays: the register-resident kernel's
    per-tile re-check segments need no clearing, and the LDS kernel's list
    length is cleared by a kernel, not a memset node."""
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(7, 7))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
    if reg:
        ct.hip.set_spec(64, 256, False, 0, reg=True)
    else:
        ct.hip.set_spec(8, 128, True, 0)
    n = 16384
    dev = torch.device("cuda:0")
    S = [torch.from_numpy(np.ascontiguousarray(css_syndrome_rows(sample_depolarizing(em, seed=90 + i, count=n), t)))
         .to(dev) for i in range(2)]
    s = torch.cuda.Stream()
    mar = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    cls = torch.zeros(n, dtype=torch.int32, device=dev)
    tie = torch.zeros(n, dtype=torch.uint8, device=dev)

    def call(i):
        ct.hip.decode_device(S[i].data_ptr(), n, "spec_f32", mar.data_ptr(), cls.data_ptr(), tie.data_ptr(),
                             s.cuda_stream)

    with torch.cuda.stream(s):
        call(0)
    torch.cuda.synchronize()
    ref = (mar.cpu().numpy().copy(), cls.cpu().numpy().copy(), tie.cpu().numpy().copy())
    ref_rechecks = ct.hip.last_recheck_count(s.cuda_stream)
    assert ref_rechecks > 0            # the re-check list is exercised
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call(0)
    torch.cuda.synchronize()
    for rep in range(3):
        say(f"case graph replay reg={reg} rep {rep}")
        mar.zero_()
        cls.fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        got = (mar.cpu().numpy(), cls.cpu().numpy(), tie.cpu().numpy())
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (reg, rep)
        with torch.cuda.stream(s):
            call(1)                    # eager work on another batch between replays
        torch.cuda.synchronize()


def test_spec_list_decode_and_autotune():
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
    n = 8192
    S = css_syndrome_rows(sample_depolarizing(em, seed=99, count=n), t)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
    say("case spec autotune")
    tile, thr = ct.hip.autotune("spec_f32", dS.data_ptr(), n)
    cfg = ct.hip.spec_config()
    say(f"  chosen {cfg}")
    got = run(torch, ct, dS, n, "spec_f32")
    if cfg["reg"] and not cfg["dyn"]:
        m64, _, _ = OraclePlan(ct.plan).decode(S)
        check_f32(got[0], got[1], m64, "autotuned static")
    else:
        for a, b in zip(got, run(torch, ct, dS, n, "fused_f32")):
            assert np.array_equal(a, b)
    ref = got
    say("case spec list decode")
    rows = np.arange(3, n, 7, dtype=np.int32)
    dl = torch.from_numpy(rows).to(dev)
    dc = torch.tensor([len(rows)], dtype=torch.int32, device=dev)
    mar = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    cls = torch.full((n,), -1, dtype=torch.int32, device=dev)
    tie = torch.zeros(n, dtype=torch.uint8, device=dev)
    ct.hip.decode_list_device(dS.data_ptr(), dl.data_ptr(), dc.data_ptr(), len(rows), "spec_f32", mar.data_ptr(),
                              cls.data_ptr(), tie.data_ptr())
    torch.cuda.synchronize()
    c = cls.cpu().numpy()
    assert np.array_equal(c[rows], ref[1][rows])
    untouched = np.ones(n, bool)
    untouched[rows] = False
    assert np.all(c[untouched] == -1)
    assert np.array_equal(mar.cpu().numpy()[rows], ref[0][rows])


def test_spec_slab_mode_bit_identical():
    """Intermediates in a per-workgroup global slab instead of LDS: same
    arithmetic, so still bit-identical to the generic fused kernel."""
    torch = pytest.importorskip("torch")
    for code, n in ((SurfaceCode(5, 5), 3000), (SurfaceCode(7, 7), 5000)):
        t = CSSTannerGraph.from_code(code)
        p = 0.05
        em = iid_error(p / 3, p / 3, p / 3, t.nq)
        ct = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
        S = css_syndrome_rows(sample_depolarizing(em, seed=31 + t.nq, count=n), t)
        dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
        ref = run(torch, ct, dS, n, "fused_f32")
        for tile, thr in ((1, 256), (4, 128), (16, 256)):
            say(f"case spec slab {t.nq} tile {tile} threads {thr}")
            ct.hip.set_spec(tile, thr, False, 0, slab=True)
            assert ct.hip.spec_config()["slab"]
            got = run(torch, ct, dS, n, "spec_f32")
            for a, b in zip(got, ref):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["toric4", "dem_d3_r3"])
def test_spec_wide_plans_vs_oracle(case):
    """Plans too wide for LDS (toric L=4 code capacity, circuit-level d=3
    memory experiment over 3 rounds) run the specialised kernel from slabs;
    marginals within 2e-5 of the row maximum of the CPU oracle's FP64
    contraction and classes equal on every non-tie syndrome."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    if case == "toric4":
        t = CSSTannerGraph.from_code(ToricCode(4, 4))
        em = iid_error(0.05 / 3, 0.05 / 3, 0.05 / 3, t.nq)
        ct = D.compile(D.TNMMAP(mode="spec_f32", optimizer=D.TreeSA(seed=1, ntrials=8, sweeps=1000, restarts=8)), t, em)
        S = css_syndrome_rows(sample_depolarizing(em, seed=5, count=96), t)
    else:
        from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
        dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
        ct = D.compile(D.TNMMAP(mode="spec_f32"), dem)
        smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
        dS0 = torch.empty((2000, smp.nsyn), dtype=torch.uint8, device=dev)
        smp.sample(17, 0, 2000, dS0.data_ptr())
        S = dS0.cpu().numpy()
    n = len(S)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev)
    say(f"case spec wide {case}: config {ct.hip.spec_config()}")
    ref, _, _ = OraclePlan(ct.plan).decode(S)
    got = run(torch, ct, dS, n, "spec_f32")
    err, _ = check_f32(got[0], got[1], ref, case)
    say(f"  max rel marginal err {err:.2e}")
    plain = run(torch, ct, dS, n, "plain_f64")
    assert np.max(np.abs(plain[0] - ref) / np.maximum(ref, 1e-300)) < 1e-12


@pytest.mark.parametrize("d,rounds,slab,n", [(3, 3, False, 4096), (5, 2, True, 1024)])
def test_spec_fp32_mfma_dem_vs_oracle(d, rounds, slab, n):
    """Circuit-level DEMs with their GEMM-shaped steps (M, N >= 32, K >= 2 per
    syndrome) on FP32 MFMA (TQEC_SPEC_MFMA32), in LDS tiles (d=3, 3 rounds)
    and per-workgroup slabs (d=5, 2 rounds, nested-dissection network):
    the FP32 contract, marginals within 2e-5 of the row maximum of the CPU
    oracle's FP64 contraction and classes equal on every non-tie syndrome;
    the generated kernel does use the MFMA steps."""
    torch = pytest.importorskip("torch")
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    dev = torch.device("cuda:0")
    dem = detector_error_model_from_stim(rotated_memory_z(d, rounds, 0.001))
    ct = D.compile(D.TNMMAP(mode="spec_f32"), dem)
    ct.hip.set_spec(1, 256, False, 0, slab=slab, unroll=2, mfma32=True)
    assert ct.hip.spec_config()["mfma32"]
    assert "__builtin_amdgcn_mfma_f32_32x32x2f32" in ct.hip.spec_source(1, 256, False, slab=slab, unroll=2,
                                                                           mfma32=True)
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    dS = torch.empty((n, smp.nsyn), dtype=torch.uint8, device=dev)
    smp.sample(31, 0, n, dS.data_ptr())
    say(f"case spec fp32 mfma dem d{d} r{rounds} slab {slab}")
    ref, _, _ = OraclePlan(ct.plan).decode(dS.cpu().numpy())
    got = run(torch, ct, dS, n, "spec_f32")
    err, ties = check_f32(got[0], got[1], ref, "fp32 mfma")
    say(f"  max rel marginal err {err:.2e}, ties {ties}, rechecks {ct.hip.last_recheck_count()}")


def test_spec_bf16_mfma_dem_vs_oracle():
    """Circuit-level d=3, 3-round DEM with its GEMM-shaped steps (M, N >= 32,
    K >= 16 per syndrome) on BF16 MFMA: marginals within 3e-2 of the row
    maximum of the CPU oracle's FP64 contraction (BF16 operands: 2^-9
    relative each), classes equal on every non-tie syndrome (margins below
    5% are re-checked in FP64)."""
    torch = pytest.importorskip("torch")
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    dev = torch.device("cuda:0")
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    ct = D.compile(D.TNMMAP(mode="spec_bf16"), dem)
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    n = 4096
    dS = torch.empty((n, smp.nsyn), dtype=torch.uint8, device=dev)
    smp.sample(29, 0, n, dS.data_ptr())
    say("case spec bf16 dem d3 r3")
    ref, _, _ = OraclePlan(ct.plan).decode(dS.cpu().numpy())
    got = run(torch, ct, dS, n, "spec_bf16")
    f32 = run(torch, ct, dS, n, "spec_f32")
    err, _ = check_f32(got[0], got[1], ref, "bf16", tol=3e-2)
    err32, _ = check_f32(f32[0], f32[1], ref, "f32")
    say(f"  max rel marginal err bf16 {err:.2e} (fp32 {err32:.2e}), rechecks {ct.hip.last_recheck_count()}")
    assert err > err32 or err == 0.0          # the MFMA steps did run in BF16
