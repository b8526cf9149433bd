"""The decode entry point that allocates nothing
(tqec_decode_batch_device_noalloc) against tqec_decode_batch_device on the
same plan, mode and syndromes, through the C ABI.

Both run the same kernels, so marginals, classes and tie flags must be
bit-identical.  Each case makes two calls in a row, the second on a smaller
(ragged) batch, so stale re-check lists or counts left in the shared static
scratch by the first call would show.  Covered: the register-resident FP32
kernel with its segmented FP64 re-check (d=7, p=0.08, where re-checks
occur), the LDS-tile FP32 kernel with its single list (d=5), and the FP64
kernel.  Host-side refusals (batch above the limit, unprepared plan, other
modes) are checked too; none of them launches anything.
"""
import sys

import numpy as np
import pytest

from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def outputs(torch, n, nout):
    dev = torch.device("cuda:0")
    return (torch.full((n, nout), -1.0, dtype=torch.float64, device=dev),
            torch.full((n,), -1, dtype=torch.int32, device=dev),
            torch.full((n,), 7, dtype=torch.uint8, device=dev))


def both(torch, ct, dS, n, mode):
    nout = 1 << ct.hip.n_out
    ref = outputs(torch, n, nout)
    got = outputs(torch, n, nout)
    ct.hip.decode_device(dS.data_ptr(), n, mode, *(t.data_ptr() for t in ref))
    torch.cuda.synchronize()
    ct.hip.decode_device_noalloc(dS.data_ptr(), n, mode, *(t.data_ptr() for t in got))
    torch.cuda.synchronize()
    for r, g, what in zip(ref, got, ("marginals", "classes", "tie flags")):
        assert torch.equal(r, g), f"{mode} n={n}: {what} differ"
    return ref


@pytest.mark.parametrize("d,p,n1,n2,lds_tile", [(7, 0.08, 65536, 10001, False), (5, 0.08, 20000, 3001, True)])
def test_noalloc_matches_device_entry_point(d, p, n1, n2, lds_tile):
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
    if lds_tile:
        ct.hip.set_spec(8, 128, False)
    S = css_syndrome_rows(sample_depolarizing(em, seed=7000 + d, count=n1), t)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    for mode in ("spec_f32", "spec_f64"):
        # unprepared on the device: refused before anything is enqueued
        mar, cls, tie = outputs(torch, n2, 1 << ct.hip.n_out)
        fresh = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
        with pytest.raises(L.TqecError) as ei:
            fresh.hip.decode_device_noalloc(dS.data_ptr(), n2, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr())
        assert ei.value.code == L.TQEC_EINVAL
        for n in (n1, n2):
            say(f"case noalloc d={d} p={p} {mode} n={n} lds_tile={lds_tile}")
            ref = both(torch, ct, dS, n, mode)
            assert int((ref[1] < 0).sum()) == 0
            if mode == "spec_f32":
                say(f"  configuration {ct.hip.spec_config()}, FP64 re-checks {ct.hip.last_recheck_count()}")
        if mode == "spec_f32":
            assert ct.hip.spec_config()["reg"] == (not lds_tile)


def test_noalloc_refusals():
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    em = iid_error(0.01, 0.01, 0.01, t.nq)
    ct = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
    S = css_syndrome_rows(sample_depolarizing(em, seed=5, count=64), t)
    dS = torch.from_numpy(np.ascontiguousarray(S)).to("cuda:0")
    mar, cls, tie = outputs(torch, 64, 4)
    ptrs = (mar.data_ptr(), cls.data_ptr(), tie.data_ptr())
    ct.hip.decode_device(dS.data_ptr(), 64, "spec_f32", *ptrs)
    torch.cuda.synchronize()
    with pytest.raises(L.TqecError) as ei:
        ct.hip.decode_device_noalloc(dS.data_ptr(), L.NOALLOC_MAX_ROWS + 1, "spec_f32", *ptrs)
    assert ei.value.code == L.TQEC_EINVAL and "limit" in str(ei.value)
    for mode in ("plain_f64", "fused_f32", "spec_bf16"):
        with pytest.raises(L.TqecError) as ei:
            ct.hip.decode_device_noalloc(dS.data_ptr(), 64, mode, *ptrs)
        assert ei.value.code == L.TQEC_EUNSUPPORTED
    ct.hip.decode_device_noalloc(dS.data_ptr(), 0, "spec_f32", *ptrs)
