"""CPU oracle of the learned front-end's forward and backward pass — TEST
INFRASTRUCTURE ONLY (tests/ may use it, the product never does).

Float64 restatement of the network the GPU kernels implement, with BF16
rounding (round to nearest even) applied where the GPU stores BF16: the
weights it multiplies, the hidden activations, and the back-propagated
pre-activation gradients.  Products are summed in float64 here and in FP32
on the GPU.  Loss: mean soft cross-entropy -sum_c q_c log softmax(z)_c.
"""
import numpy as np


def bf16(x):
    """Round to the nearest BF16 value, ties to even (finite inputs)."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


_bf16 = bf16


def forward_backward(params: dict, X: np.ndarray, Q: np.ndarray, round_bf16: bool = True):
    """params: W1 b1 W2 b2 W3 b3 W4 b4 (float arrays, W1 rows = padded input
    width); X: (B, kin) 0/1 features; Q: (B, n_out) target probabilities.
    Returns (loss, logits, grads dict).  With ``round_bf16=False`` every
    value stays float64 (the exact gradient of the float64 network); with
    rounding, the gradient is taken through BF16 storage as through the
    identity (what the GPU's backward pass does).  Shapes are checked."""
    B = X.shape[0]
    assert Q.shape[0] == B and params["W4"].shape[1] == Q.shape[1]
    assert params["W1"].shape[0] == X.shape[1]
    bf16 = _bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    W = {k: bf16(params[k]) for k in ("W1", "W2", "W3", "W4")}
    b = {k: np.asarray(params[k], dtype=np.float64) for k in ("b1", "b2", "b3", "b4")}
    H0 = np.asarray(X, dtype=np.float64)
    H1 = bf16(np.maximum(H0 @ W["W1"] + b["b1"], 0))
    H2 = bf16(np.maximum(H1 @ W["W2"] + b["b2"], 0))
    H3 = bf16(np.maximum(H2 @ W["W3"] + b["b3"], 0))
    z = H3 @ W["W4"] + b["b4"]
    m = z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(z - m).sum(axis=1, keepdims=True)) + m
    logp = z - lse
    loss = -(Q * logp).sum() / B
    dz = (np.exp(logp) - Q) / B
    g = {"W4": H3.T @ dz, "b4": dz.sum(axis=0)}
    dA3 = bf16((dz @ W["W4"].T) * (H3 > 0))
    g["W3"] = H2.T @ dA3
    g["b3"] = dA3.sum(axis=0)
    dA2 = bf16((dA3 @ W["W3"].T) * (H2 > 0))
    g["W2"] = H1.T @ dA2
    g["b2"] = dA2.sum(axis=0)
    dA1 = bf16((dA2 @ W["W2"].T) * (H1 > 0))
    g["W1"] = H0.T @ dA1
    g["b1"] = dA1.sum(axis=0)
    return loss, z, g
