"""On-device sampling and error-pattern recovery through the C ABI.

The sampler must be bit-identical to the host sampler (same Philox counters);
the recovery must equal the host restatement of tndecoder.jl:167-174 and
produce syndrome-consistent patterns in the decided class."""
import numpy as np
import pytest

from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing, sample_flips, syndrome_extraction
from tensorqec.gf2 import gf2_right_inverse, logical_operator

pytestmark = pytest.mark.gpu


def true_classes(ep, lx, lz):
    k = lx.shape[0]
    c = np.zeros(ep.xerror.shape[0], np.int64)
    for i in range(k):
        c |= ((ep.zerror.astype(np.int64) @ lx[i]) & 1) << i
        c |= ((ep.xerror.astype(np.int64) @ lz[i]) & 1) << (k + i)
    return c


@pytest.mark.parametrize("code,p", [(SurfaceCode(7, 7), 0.05), (ToricCode(5, 5), 0.08), (SurfaceCode(9, 9), 0.02)])
def test_css_sampler_bit_identical(code, p):
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(code)
    lx, lz = logical_operator(t)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    n, first, seed = 20000, 12345, 987654321012
    s = L.HipSampler.css(t, em, lx, lz)
    dev = torch.device("cuda:0")
    dS = torch.empty((n, t.ns), dtype=torch.uint8, device=dev)
    dC = torch.empty(n, dtype=torch.int32, device=dev)
    dE = torch.empty((n, 2 * t.nq), dtype=torch.uint8, device=dev)
    s.sample(seed, first, n, dS.data_ptr(), dC.data_ptr(), dE.data_ptr())
    torch.cuda.synchronize()
    ep = sample_depolarizing(em, seed, n, first=first)
    assert np.array_equal(dE.cpu().numpy(), np.concatenate([ep.xerror, ep.zerror], axis=1))
    assert np.array_equal(dS.cpu().numpy(), css_syndrome_rows(ep, t))
    assert np.array_equal(dC.cpu().numpy(), true_classes(ep, lx, lz))


def test_flip_sampler_bit_identical():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    nm, nd, nl = 300, 70, 2
    p = rng.uniform(0.001, 0.05, nm)
    Dm = (rng.random((nd, nm)) < 0.03).astype(np.uint8)
    Lm = (rng.random((nl, nm)) < 0.05).astype(np.uint8)
    s = L.HipSampler.flips(p, Dm, Lm)
    n = 10000
    dev = torch.device("cuda:0")
    dS = torch.empty((n, nd), dtype=torch.uint8, device=dev)
    dC = torch.empty(n, dtype=torch.int32, device=dev)
    dE = torch.empty((n, nm), dtype=torch.uint8, device=dev)
    s.sample(77, 0, n, dS.data_ptr(), dC.data_ptr(), dE.data_ptr())
    torch.cuda.synchronize()
    e = sample_flips(type("E", (), {"p": p})(), 77, n)
    assert np.array_equal(dE.cpu().numpy(), e)
    assert np.array_equal(dS.cpu().numpy(), (e.astype(np.int64) @ Dm.T) & 1)
    cls = ((e.astype(np.int64) @ Lm.T) & 1) @ (1 << np.arange(nl))
    assert np.array_equal(dC.cpu().numpy(), cls)


@pytest.mark.parametrize("code", [SurfaceCode(7, 7), ToricCode(4, 4)])
def test_css_recovery_matches_host(code):
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(code)
    em = iid_error(0.03, 0.03, 0.03, t.nq)
    ct = D.compile(D.TNMMAP(), t, em)
    S = css_syndrome_rows(sample_depolarizing(em, 5, 4096), t)
    rng = np.random.default_rng(2)
    cls = rng.integers(0, 1 << (2 * ct.lx.shape[0]), size=len(S)).astype(np.int32)
    host = D.css_error_patterns(ct, S, cls)
    rec = L.HipRecovery.css(t, ct.pure_x, ct.pure_z, ct.lx, ct.lz)
    dev = torch.device("cuda:0")
    dS, dC = torch.from_numpy(S).to(dev), torch.from_numpy(cls).to(dev)
    dA = torch.empty((len(S), t.nq), dtype=torch.uint8, device=dev)
    dB = torch.empty((len(S), t.nq), dtype=torch.uint8, device=dev)
    rec.recover(dS.data_ptr(), dC.data_ptr(), len(S), dA.data_ptr(), dB.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dA.cpu().numpy(), host.xerror)
    assert np.array_equal(dB.cpu().numpy(), host.zerror)
