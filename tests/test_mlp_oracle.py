"""Pins of the learned front-end's CPU oracle (oracle/mlp_oracle.py).

The oracle has no counterpart in the reference, so it is pinned against
independent statements of the same mathematics:
  * ``bf16`` against torch's float32 -> bfloat16 conversion (round to
    nearest, ties to even), on random values, exact halfway points and
    values next to them, signed zeros, subnormals and large magnitudes;
  * ``forward_backward`` with BF16 rounding switched off against torch CPU
    autograd in float64 (loss, logits, every gradient to 1e-10 relative)
    and against central finite differences of the loss;
  * the rounding itself: with rounding on, the forward pass equals the
    float64 network evaluated on BF16-rounded weights and activations;
  * planted mistakes (a dropped ReLU mask in one backward layer, the mask
    of the wrong layer, a weight gradient paired with the wrong samples)
    that the autograd check must catch.
"""
import inspect

import numpy as np
import pytest

from oracle import mlp_oracle as MO

torch = pytest.importorskip("torch")


def make_params(n_in=20, kin=32, hidden=16, n_out=4, seed=0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, 0.4, (kin, hidden)), "b1": rng.normal(0, 0.1, hidden),
         "W2": rng.normal(0, 0.4, (hidden, hidden)), "b2": rng.normal(0, 0.1, hidden),
         "W3": rng.normal(0, 0.4, (hidden, hidden)), "b3": rng.normal(0, 0.1, hidden),
         "W4": rng.normal(0, 0.4, (hidden, n_out)), "b4": rng.normal(0, 0.1, n_out)}
    p["W1"][n_in:, :] = 0.0
    B = 24
    X = np.zeros((B, kin))
    X[:, :n_in] = rng.random((B, n_in)) < 0.3
    Q = rng.dirichlet(np.ones(n_out), size=B)
    return p, X, Q


def torch_reference(p, X, Q):
    """The same network written with torch modules' arithmetic in float64
    and differentiated by autograd."""
    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in p.items()}
    x = torch.tensor(X, dtype=torch.float64)
    h = torch.relu(x @ T["W1"] + T["b1"])
    h = torch.relu(h @ T["W2"] + T["b2"])
    h = torch.relu(h @ T["W3"] + T["b3"])
    z = h @ T["W4"] + T["b4"]
    loss = -(torch.tensor(Q) * torch.log_softmax(z, dim=1)).sum() / X.shape[0]
    loss.backward()
    return float(loss.detach()), z.detach().numpy(), {k: t.grad.numpy() for k, t in T.items()}


def assert_matches_torch(fb, p, X, Q, tol=1e-10):
    loss, z, g = fb(p, X, Q, round_bf16=False)
    rl, rz, rg = torch_reference(p, X, Q)
    assert abs(loss - rl) <= tol * abs(rl)
    assert np.max(np.abs(z - rz)) <= tol * np.max(np.abs(rz))
    for k in rg:
        den = max(np.linalg.norm(rg[k]), 1e-300)
        assert np.linalg.norm(g[k] - rg[k]) <= tol * den, k


def test_bf16_matches_torch_round_to_nearest_even():
    rng = np.random.default_rng(1)
    vals = [rng.normal(0, 1, 4096), rng.normal(0, 1e-30, 256), rng.normal(0, 1e30, 256),
            np.array([0.0, -0.0, 1.0, -1.0, 3.0e38, -3.0e38, 1e-40, -1e-40, 1e-45])]
    # exact halfway points between neighbouring BF16 values: low 16 bits 0x8000,
    # with an even and an odd BF16 mantissa below; and one ulp either side
    base = (rng.integers(0x0080, 0x7F00, 512).astype(np.uint32) << 16)
    for off in (0x8000, 0x7FFF, 0x8001, 0x0000, 0xFFFF):
        u = (base | np.uint32(off)).astype(np.uint32)
        for sign in (0, 0x80000000):
            vals.append((u | np.uint32(sign)).view(np.float32).astype(np.float64))
    x = np.concatenate(vals).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = MO.bf16(x)
    assert np.array_equal(np.signbit(got), np.signbit(ref))
    assert np.array_equal(got, ref)
    # ties really go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert MO.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert MO.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_forward_backward_matches_torch_autograd(seed):
    p, X, Q = make_params(seed=seed)
    assert_matches_torch(MO.forward_backward, p, X, Q)


def test_forward_backward_matches_finite_differences():
    p, X, Q = make_params(n_in=10, kin=32, hidden=8, seed=5)
    _, _, g = MO.forward_backward(p, X, Q, round_bf16=False)
    rng = np.random.default_rng(9)
    h = 1e-6
    for k, v in p.items():
        idx = [tuple(rng.integers(0, s) for s in v.shape) for _ in range(6)]
        if k == "W1":
            idx = [(i % 10,) + j[1:] for i, j in zip(range(6), idx)]   # rows that carry input
        for ix in idx:
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[k][ix] += h
            lp = MO.forward_backward(q, X, Q, round_bf16=False)[0]
            q[k][ix] -= 2 * h
            lm = MO.forward_backward(q, X, Q, round_bf16=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k][ix]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (k, ix, fd, g[k][ix])


def test_rounding_is_applied_at_storage_points():
    p, X, Q = make_params(seed=3)
    _, z, _ = MO.forward_backward(p, X, Q, round_bf16=True)
    r = lambda a: torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    h = r(np.maximum(X @ r(p["W1"]) + p["b1"], 0))
    h = r(np.maximum(h @ r(p["W2"]) + p["b2"], 0))
    h = r(np.maximum(h @ r(p["W3"]) + p["b3"], 0))
    zz = h @ r(p["W4"]) + p["b4"]
    assert np.allclose(z, zz, rtol=1e-12, atol=1e-12)
    # and it changes the result (the check above is not vacuous)
    _, z0, _ = MO.forward_backward(p, X, Q, round_bf16=False)
    assert np.max(np.abs(z - z0)) > 1e-4


@pytest.mark.parametrize("mutation", [
    ('bf16((dA3 @ W["W3"].T) * (H2 > 0))', 'bf16((dA3 @ W["W3"].T))'),      # dropped ReLU mask
    ('bf16((dz @ W["W4"].T) * (H3 > 0))', 'bf16((dz @ W["W4"].T) * (H2 > 0))'),  # mask of the wrong layer
    ('g["W2"] = H1.T @ dA2', 'g["W2"] = H1.T @ dA2[::-1]'),                     # rows of the wrong sample
])
def test_planted_mistakes_are_caught(mutation):
    """A copy of the oracle with one backward-pass line changed must fail
    the autograd comparison (mutation made on the CPU, in memory)."""
    src = inspect.getsource(MO)
    old, new = mutation
    assert old in src
    ns = {}
    exec(compile(src.replace(old, new), "mlp_oracle_mutant", "exec"), ns)
    p, X, Q = make_params(seed=0)
    with pytest.raises(AssertionError):
        assert_matches_torch(ns["forward_backward"], p, X, Q)
