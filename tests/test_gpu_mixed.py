"""Mixed-code executor: batches of several codes decoded by one persistent
launch (tqec_mixed_*), against the CPU oracle and the single-plan executors
on the same syndromes.

  * Every batch: marginals within 2e-5 of the row maximum of the CPU
    oracle's FP64 contraction of the same plan, classes equal on every row
    whose FP64 top-2 gap is above 1e-12 (the fused_f32 contract; reference
    call `decode(ct, syndrome)`, src/decoding/tndecoder.jl:160-165).
  * Plans whose tiles fit the LDS, in addition: marginals, classes and tie
    flags bit-identical to tqec_decode_batch_device(..., fused_f32) — every
    syndrome goes through the same per-syndrome FP32 arithmetic and the same
    FP64 re-check rule, only the tiling differs.
Graph replay and direct enqueue give identical results, and a second replay
with different (smaller) batches reuses the captured graph.  Both kernels of
the executor are covered: the generic interpreter and the kernel generated
from every plan's specialised code.
"""
import sys
import time

import numpy as np
import pytest

from _parity import check_f32
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, ShorCode, SteaneCode, SurfaceCode, ToricCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing

pytestmark = pytest.mark.gpu

CODES = [("surface3", SurfaceCode(3, 3), 0.05), ("steane", SteaneCode(), 0.08), ("shor", ShorCode(), 0.03),
         ("surface5", SurfaceCode(5, 5), 0.06), ("surface7", SurfaceCode(7, 7), 0.05),
         ("surface9", SurfaceCode(9, 9), 0.02), ("toric3", ToricCode(3, 3), 0.07), ("toric4", ToricCode(4, 4), 0.04)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    torch = pytest.importorskip("torch")
    plans = []
    for name, code, p in CODES:
        t = CSSTannerGraph.from_code(code)
        em = iid_error(p / 3, p / 3, p / 3, t.nq)
        opt = D.TreeSA(seed=1, ntrials=8, sweeps=1000, restarts=8) if name == "toric4" else D.TreeSA(seed=1)
        ct = D.compile(D.TNMMAP(optimizer=opt, mode="fused_f32"), t, em)
        plans.append((name, t, em, ct, OraclePlan(ct.plan)))
    return torch, plans


def make_batches(torch, plans, spec, seed):
    dev = torch.device("cuda:0")
    out = []
    for k, (pi, n) in enumerate(spec):
        name, t, em, ct, _ = plans[pi]
        S = css_syndrome_rows(sample_depolarizing(em, seed=seed + 101 * k, count=max(n, 1)), t)[:n]
        nout = 1 << ct.hip.n_out
        dS = torch.from_numpy(np.ascontiguousarray(S)).to(dev) if n else torch.empty((0, t.ns), dtype=torch.uint8, device=dev)
        out.append(dict(plan=pi, n=n, S=S, dS=dS,
                        mar=torch.full((max(n, 1), nout), -1.0, dtype=torch.float64, device=dev),
                        cls=torch.full((max(n, 1),), -1, dtype=torch.int32, device=dev),
                        tie=torch.full((max(n, 1),), 7, dtype=torch.uint8, device=dev)))
    return out


def table(batches):
    return [(b["plan"], b["n"], b["dS"].data_ptr() if b["n"] else 0, b["mar"].data_ptr(), b["cls"].data_ptr(),
             b["tie"].data_ptr()) for b in batches]


def check_against_single(torch, plans, mx, batches):
    dev = torch.device("cuda:0")
    for b in batches:
        name, t, em, ct, op = plans[b["plan"]]
        n = b["n"]
        if n == 0:
            continue
        tile, slab = mx.plan_config(b["plan"])
        mar = b["mar"][:n].cpu().numpy()
        cls = b["cls"][:n].cpu().numpy()
        tie = b["tie"][:n].cpu().numpy()
        nout = mar.shape[1]
        ref, _, _ = op.decode(b["S"])
        check_f32(mar, cls, ref, name)
        if slab & 1 == 0:
            rm = torch.empty((n, nout), dtype=torch.float64, device=dev)
            rc = torch.empty(n, dtype=torch.int32, device=dev)
            rt = torch.empty(n, dtype=torch.uint8, device=dev)
            ct.hip.decode_device(b["dS"].data_ptr(), n, "fused_f32", rm.data_ptr(), rc.data_ptr(), rt.data_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(mar, rm.cpu().numpy()), name
            assert np.array_equal(cls, rc.cpu().numpy()), name
            assert np.array_equal(tie, rt.cpu().numpy()), name


SPEC1 = [(0, 1), (4, 300), (1, 7), (6, 33), (3, 64), (2, 0), (5, 17), (4, 5), (7, 9), (0, 1000), (2, 129),
         (6, 1), (3, 4096)]
SPEC2 = [(4, 40), (0, 3), (7, 2), (5, 1), (6, 10)]


@pytest.mark.parametrize("kernel", ["generic", "specialised"])
def test_mixed_stream_matches_single_plan_executors(setup, kernel):
    torch, plans = setup
    say(f"case mixed ({kernel}): create")
    mx = L.HipMixed([pl[3].hip for pl in plans], max_batches=16, max_rows=8192)
    if kernel == "specialised":
        t0 = time.time()
        mx.set_kernel(True)
        say(f"  generated kernel compiled in {time.time() - t0:.1f}s")
    for i, (name, *_rest) in enumerate(plans):
        say(f"  plan {name}: tile/slab {mx.plan_config(i)}")
    assert mx.plan_config(6)[1] & 1 == 1          # toric3 needs the slab
    assert mx.plan_config(4)[1] == 0
    stream = torch.cuda.Stream()
    b1 = make_batches(torch, plans, SPEC1, seed=11)
    say("case mixed: graph call 1")
    mx.decode_device(table(b1), stream=stream.cuda_stream)
    lat, nre = mx.last_call()
    assert len(lat) == len(SPEC1)
    for (pi, n), us in zip(SPEC1, lat):
        assert (us > 0) == (n > 0)
    say(f"  latencies us {np.round(lat, 1).tolist()} rechecks {nre}")
    check_against_single(torch, plans, mx, b1)
    say("case mixed: direct enqueue gives the same results")
    b1d = make_batches(torch, plans, SPEC1, seed=11)
    mx.set_graph(False)
    mx.decode_device(table(b1d), stream=stream.cuda_stream)
    mx.last_call()
    for x, y in zip(b1, b1d):
        for k in ("mar", "cls", "tie"):
            assert torch.equal(x[k], y[k])
    mx.set_graph(True)
    say("case mixed: graph calls 2 and 3 (smaller batches, both slots)")
    b2 = make_batches(torch, plans, SPEC2, seed=23)
    mx.decode_device(table(b2), stream=stream.cuda_stream)
    b3 = make_batches(torch, plans, SPEC2[::-1], seed=37)
    mx.decode_device(table(b3), stream=stream.cuda_stream)
    lat3, _ = mx.last_call()
    assert len(lat3) == len(SPEC2)
    check_against_single(torch, plans, mx, b2)
    check_against_single(torch, plans, mx, b3)
    say("case mixed: replay of slot 1 with the first table again")
    b4 = make_batches(torch, plans, SPEC1, seed=11)
    mx.decode_device(table(b4), stream=stream.cuda_stream)
    mx.last_call()
    for x, y in zip(b1, b4):
        for k in ("mar", "cls", "tie"):
            assert torch.equal(x[k], y[k])


def test_mixed_rejects_bad_tables(setup):
    torch, plans = setup
    mx = L.HipMixed([plans[0][3].hip, plans[1][3].hip], max_batches=2, max_rows=10)
    dev = torch.device("cuda:0")
    S = torch.zeros((8, 8), dtype=torch.uint8, device=dev)
    with pytest.raises(L.TqecError) as e:
        mx.decode_device([(2, 1, S.data_ptr(), 0, 0, 0)])
    assert e.value.code == L.TQEC_EINVAL
    with pytest.raises(L.TqecError) as e:
        mx.decode_device([(0, 8, S.data_ptr(), 0, 0, 0), (0, 8, S.data_ptr(), 0, 0, 0)])
    assert e.value.code == L.TQEC_EINVAL
    with pytest.raises(L.TqecError) as e:
        mx.decode_device([(0, 1, S.data_ptr(), 0, 0, 0)] * 3)
    assert e.value.code == L.TQEC_EINVAL
    with pytest.raises(L.TqecError) as e:
        mx.decode_device([(0, 1, 0, 0, 0, 0)])
    assert e.value.code == L.TQEC_ENULL


def test_mixed_specialised_with_toric5():
    """The generated mixed-code kernel with a plan of width 2^21 (toric code
    L=5, nested-dissection network: per-workgroup slabs, GEMM-shaped steps
    on FP32 MFMA) beside a small one: the toric L=5 batch against the
    oracle's marginals in tests/golden/toric5.npz (1e-4 of the row maximum:
    its last step sums 2^18 products per marginal in FP32), the surface-code
    batch against the CPU oracle at 2e-5; classes exact on non-tie rows."""
    import os
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "toric5.npz"))
    p5 = float(g["p"])
    t5 = CSSTannerGraph.from_code(ToricCode(5, 5))
    ct5 = D.compile(D.TNMMAP(mode="fused_f32", optimizer=D.Dissection(), tabulate_bits=0), t5,
                    iid_error(p5, p5, p5, t5.nq))
    t3 = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    em3 = iid_error(0.05 / 3, 0.05 / 3, 0.05 / 3, t3.nq)
    ct3 = D.compile(D.TNMMAP(mode="fused_f32"), t3, em3)
    say("case mixed with toric5: create and generate")
    mx = L.HipMixed([ct3.hip, ct5.hip], max_batches=4, max_rows=4096)
    mx.set_kernel(True)
    assert mx.plan_config(1)[1] & 1 == 1           # the toric L=5 tiles work from slabs
    S3 = css_syndrome_rows(sample_depolarizing(em3, seed=77, count=500), t3)
    S5 = g["syn"]
    outs = []
    for S, nout in ((S3, 4), (S5, 16)):
        n = len(S)
        outs.append(dict(n=n, S=S, dS=torch.from_numpy(np.ascontiguousarray(S)).to(dev),
                         mar=torch.full((n, nout), -1.0, dtype=torch.float64, device=dev),
                         cls=torch.full((n,), -1, dtype=torch.int32, device=dev),
                         tie=torch.full((n,), 7, dtype=torch.uint8, device=dev)))
    tab = [(i, o["n"], o["dS"].data_ptr(), o["mar"].data_ptr(), o["cls"].data_ptr(), o["tie"].data_ptr())
           for i, o in enumerate(outs)]
    mx.decode_device(tab)
    lat, nre = mx.last_call()
    say(f"  latencies us {np.round(lat, 1).tolist()} rechecks {nre}")
    ref3, _, _ = OraclePlan(ct3.plan).decode(S3)
    check_f32(outs[0]["mar"].cpu().numpy(), outs[0]["cls"].cpu().numpy(), ref3, "surface3")
    err, _ = check_f32(outs[1]["mar"].cpu().numpy(), outs[1]["cls"].cpu().numpy(), g["mar"], "toric5", tol=1e-4)
    say(f"  toric5 max rel marginal err {err:.2e}")
