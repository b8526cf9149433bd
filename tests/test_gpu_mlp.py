"""Learned front-end on the GPU: forward logits and all parameter gradients
against the float64 CPU oracle (oracle/mlp_oracle.py), training on
TN-labelled syndromes, and the confidence gate in front of the TN path.

Tolerances: logits within 1e-2 of the largest |logit|; every gradient tensor
within 2e-2 relative (Frobenius norm).  The GPU accumulates in FP32 and
rounds activations and back-propagated gradients to BF16 where the oracle
does the same rounding in float64 arithmetic, so a value sitting on a BF16
rounding boundary can land one BF16 ulp (0.4%) apart.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing
from tensorqec.gf2 import logical_operator
from tensorqec.learned import GatedDecoder, HipMLP, train_on_tn

pytestmark = pytest.mark.gpu


def setup(n_in=48, n_out=4, B=256, seed=0):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(seed)
    mlp = HipMLP(n_in, n_out, seed=seed)
    p = mlp.get_params()
    p += rng.normal(0, 0.01, size=p.shape).astype(np.float32)     # non-zero biases too
    mlp.set_params(p)
    S = (rng.random((B, n_in)) < 0.2).astype(np.uint8)
    Q = rng.dirichlet(np.ones(n_out), size=B).astype(np.float32)
    return torch, mlp, S, Q


def features(S, kin):
    X = np.zeros((S.shape[0], kin))
    X[:, :S.shape[1]] = S
    return X


def test_forward_matches_oracle():
    torch, mlp, S, Q = setup()
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dz = torch.empty((len(S), 4), dtype=torch.float32, device=dev)
    mlp.forward(dS.data_ptr(), len(S), dz.data_ptr())
    torch.cuda.synchronize()
    params = mlp.unflatten(mlp.get_params())
    _, z, _ = MO.forward_backward(params, features(S, mlp.kin), Q)
    got = dz.cpu().numpy()
    assert np.max(np.abs(got - z)) < 1e-2 * np.max(np.abs(z))


@pytest.mark.parametrize("n_in", [48, 120])
def test_gradients_match_oracle(n_in):
    torch, mlp, S, Q = setup(n_in=n_in, B=512, seed=3)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    dQ = torch.from_numpy(Q).to(dev)
    dG = torch.empty(mlp.nparams, dtype=torch.float32, device=dev)
    loss = mlp.gradients(dS.data_ptr(), dQ.data_ptr(), len(S), dG.data_ptr())
    params = mlp.unflatten(mlp.get_params())
    ref_loss, _, ref = MO.forward_backward(params, features(S, mlp.kin), Q)
    assert abs(loss - ref_loss) < 1e-3 * max(1.0, abs(ref_loss))
    got = mlp.unflatten(dG.cpu().numpy())
    for k, g in ref.items():
        err = np.linalg.norm(got[k] - g) / max(np.linalg.norm(g), 1e-12)
        assert err < 2e-2, (k, err)
    # padded input rows of W1 carry no gradient
    assert not np.any(got["W1"][n_in:, :])


def test_train_and_gate_d5():
    """Train on TN posteriors at d=5, p=0.05; the gate must hand the TN path
    only a minority of syndromes and the gated decisions must keep the
    logical error rate of the TN decoder (within 3 sigma)."""
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    lx, lz = logical_operator(t)
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, em)
    sampler = L.HipSampler.css(t, em, lx, lz)
    mlp = HipMLP(t.ns, 4, seed=1)
    rep = train_on_tn(mlp, ct, sampler, n_samples=8192 * 150, batch=8192, lr=2e-3, lr_final=2e-4, seed=99,
                      torch=torch)
    first, last = rep.losses[0][1], rep.losses[-1][1]
    assert last < 0.5 * first, rep.losses
    # held-out evaluation: different seed
    n = 65536
    dev = torch.device("cuda:0")
    S = torch.empty((n, t.ns), dtype=torch.uint8, device=dev)
    truth = torch.empty(n, dtype=torch.int32, device=dev)
    sampler.sample(7777, 0, n, S.data_ptr(), truth.data_ptr())
    cls_tn = torch.empty(n, dtype=torch.int32, device=dev)
    mar = torch.empty((n, 4), dtype=torch.float64, device=dev)
    tie = torch.empty(n, dtype=torch.uint8, device=dev)
    ct.hip.decode_device(S.data_ptr(), n, "fused_f32", mar.data_ptr(), cls_tn.data_ptr(), tie.data_ptr())
    gd = GatedDecoder(mlp, ct, threshold=0.99, capacity=n, torch=torch)
    cls_g = torch.empty(n, dtype=torch.int32, device=dev)
    conf = torch.empty(n, dtype=torch.float32, device=dev)
    mar_g = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    tie_g = torch.zeros(n, dtype=torch.uint8, device=dev)
    gd.decode_device(S.data_ptr(), n, cls_g.data_ptr(), conf.data_ptr(), mar_g.data_ptr(), tie_g.data_ptr())
    torch.cuda.synchronize()
    fall = gd.fallbacks() / n
    a, b, tr = cls_tn.cpu().numpy(), cls_g.cpu().numpy(), truth.cpu().numpy()
    ler_tn, ler_g = (a != tr).mean(), (b != tr).mean()
    sigma = np.sqrt(ler_tn * (1 - ler_tn) / n)
    print(f"d5 gate: fallback {fall:.4f}, disagreement {(a != b).mean():.5f}, LER tn {ler_tn:.5f} gated {ler_g:.5f}")
    assert fall < 0.5
    assert abs(ler_g - ler_tn) < 3 * sigma + 1e-4


@pytest.mark.parametrize("n_in,n_out,B", [(48, 4, 1000), (120, 2, 4096), (24, 2, 77), (80, 4, 65536)])
def test_fused_forward_matches_layered_and_oracle(n_in, n_out, B):
    """The fused inference kernel (all four layers per 128-row tile, the
    activations in LDS) against the layer-by-layer kernels and the float64
    oracle: logits within 1e-2 of the largest |logit| (BF16 activations on
    both sides; the fused kernel adds the bias before the products, which can
    move an activation across a BF16 rounding boundary), gate classes equal
    wherever the top two logits are apart.  Ragged tiles (B not a multiple
    of 128) and every supported input width class are covered."""
    torch, mlp, S, Q = setup(n_in=n_in, n_out=n_out, B=B, seed=11)
    dev = torch.device("cuda:0")
    dS = torch.from_numpy(S).to(dev)
    assert mlp.inference() == 1
    zf = torch.full((B, n_out), float("nan"), dtype=torch.float32, device=dev)
    mlp.forward(dS.data_ptr(), B, zf.data_ptr())
    mlp.set_inference(0)
    assert mlp.inference() == 0
    zl = torch.full((B, n_out), float("nan"), dtype=torch.float32, device=dev)
    mlp.forward(dS.data_ptr(), B, zl.data_ptr())
    mlp.set_inference(1)
    torch.cuda.synchronize()
    zf, zl = zf.cpu().numpy(), zl.cpu().numpy()
    assert np.isfinite(zf).all()
    scale = np.max(np.abs(zl))
    assert np.max(np.abs(zf - zl)) < 1e-2 * scale
    m = min(B, 2048)
    params = mlp.unflatten(mlp.get_params())
    _, z, _ = MO.forward_backward(params, features(S[:m], mlp.kin), Q[:m])
    assert np.max(np.abs(zf[:m] - z)) < 1e-2 * np.max(np.abs(z))
    srt = np.sort(zl, axis=1)
    apart = (srt[:, -1] - srt[:, -2]) > 2e-2 * scale
    assert np.array_equal(np.argmax(zf, axis=1)[apart], np.argmax(zl, axis=1)[apart])
    # the gate runs the same kernel
    cls = torch.empty(B, dtype=torch.int32, device=dev)
    conf = torch.empty(B, dtype=torch.float32, device=dev)
    mlp.gate(dS.data_ptr(), B, 0.9, cls.data_ptr(), conf.data_ptr(), 0, 0)
    torch.cuda.synchronize()
    assert np.array_equal(cls.cpu().numpy(), np.argmax(zf, axis=1))
