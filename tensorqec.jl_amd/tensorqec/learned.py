"""Learned front-end: a syndrome -> logical-class MLP trained on the
tensor-network decoder's class posteriors, with a confidence gate that sends
uncertain syndromes back to the TN path.

There is no learned decoder in the reference; this module sits in front of
the TNMMAP path (src/decoding/tndecoder.jl:160-165).  Device memory is held
in torch tensors (plumbing only); every kernel is in libtqec_hip.so.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import lib as L


def _vp(x):
    return ctypes.c_void_p(x) if x else None


class HipMLP:
    """``n_in -> hidden x 3 (ReLU) -> n_out`` with BF16 weights, FP32 accumulation."""

    def __init__(self, n_in: int, n_out: int, hidden: int = 512, seed: int = 0):
        lib = L.load()
        h = ctypes.c_void_p()
        L.check(lib.tqec_mlp_create(int(n_in), int(hidden), int(n_out), ctypes.byref(h)))
        self._h = h
        self.n_in, self.n_out, self.hidden = n_in, n_out, hidden
        self.kin = (n_in + 31) // 32 * 32
        self.nparams = int(lib.tqec_mlp_num_params(h))
        self.set_params(self.init_params(seed))

    def layout(self):
        """(name, shape) of the flat parameter vector."""
        k, h, o = self.kin, self.hidden, self.n_out
        return [("W1", (k, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (h, h)), ("b3", (h,)),
                ("W4", (h, o)), ("b4", (o,))]

    def unflatten(self, flat: np.ndarray):
        out, o = {}, 0
        for name, shp in self.layout():
            n = int(np.prod(shp))
            out[name] = flat[o:o + n].reshape(shp)
            o += n
        return out

    def init_params(self, seed: int) -> np.ndarray:
        """He-uniform weights for the real fan-in, zero biases (seeded)."""
        rng = np.random.default_rng(seed)
        parts = []
        fan = {"W1": self.n_in, "W2": self.hidden, "W3": self.hidden, "W4": self.hidden}
        for name, shp in self.layout():
            if name.startswith("W"):
                lim = np.sqrt(6.0 / fan[name])
                w = rng.uniform(-lim, lim, size=shp)
                if name == "W1":
                    w[self.n_in:, :] = 0.0
                parts.append(w.ravel())
            else:
                parts.append(np.zeros(shp).ravel())
        return np.concatenate(parts).astype(np.float32)

    def set_params(self, flat: np.ndarray) -> None:
        a = np.ascontiguousarray(np.asarray(flat, dtype=np.float32))
        assert a.size == self.nparams
        L.check(L.load().tqec_mlp_set_params(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get_params(self) -> np.ndarray:
        a = np.empty(self.nparams, dtype=np.float32)
        L.check(L.load().tqec_mlp_get_params(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return a

    def save(self, path: str) -> None:
        np.savez(path, params=self.get_params(), n_in=self.n_in, n_out=self.n_out, hidden=self.hidden)

    @classmethod
    def load(cls, path: str) -> "HipMLP":
        z = np.load(path)
        m = cls(int(z["n_in"]), int(z["n_out"]), int(z["hidden"]))
        m.set_params(z["params"])
        return m

    def forward(self, d_syn: int, n: int, d_logits: int, stream: int = 0) -> None:
        L.check(L.load().tqec_mlp_forward(self._h, _vp(d_syn), int(n), _vp(d_logits), _vp(stream)))

    def gradients(self, d_syn: int, d_q: int, n: int, d_grad: int = 0, stream: int = 0) -> float:
        loss = ctypes.c_float()
        L.check(L.load().tqec_mlp_gradients(self._h, _vp(d_syn), _vp(d_q), int(n), _vp(d_grad),
                                            ctypes.byref(loss), _vp(stream)))
        return loss.value

    def train_step(self, d_syn: int, d_q: int, n: int, lr: float, want_loss: bool = False, stream: int = 0):
        loss = ctypes.c_float()
        L.check(L.load().tqec_mlp_train_step(self._h, _vp(d_syn), _vp(d_q), int(n), float(lr),
                                             ctypes.byref(loss) if want_loss else None, _vp(stream)))
        return loss.value if want_loss else None

    def gate(self, d_syn: int, n: int, threshold: float, d_cls: int, d_conf: int, d_reject: int, d_count: int,
             stream: int = 0) -> None:
        L.check(L.load().tqec_mlp_gate(self._h, _vp(d_syn), int(n), float(threshold), _vp(d_cls), _vp(d_conf),
                                       _vp(d_reject), _vp(d_count), _vp(stream)))

    def set_inference(self, kernel: int) -> None:
        """Kernel of forward / gate: 1 fused (default where it applies), 0 the
        layer-by-layer training kernels."""
        L.check(L.load().tqec_mlp_set_inference(self._h, int(kernel)))

    def inference(self) -> int:
        k = ctypes.c_int32()
        L.check(L.load().tqec_mlp_inference(self._h, ctypes.byref(k)))
        return k.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().tqec_mlp_destroy(h)
            except Exception:
                pass
            self._h = None


def posteriors(d_mar: int, n: int, n_classes: int, d_q: int, stream: int = 0) -> None:
    L.check(L.load().tqec_posteriors_from_marginals(_vp(d_mar), int(n), int(n_classes), _vp(d_q), _vp(stream)))


@dataclass
class TrainReport:
    samples: int
    steps: int
    seconds: float
    losses: list


def train_on_tn(mlp: HipMLP, ct, sampler, n_samples: int, batch: int = 8192, lr: float = 1e-3,
                seed: int = 1, first: int = 0, lr_final: float = 1e-4, log_every: int = 50,
                torch=None, mode: str = "fused_f32", max_seconds: Optional[float] = None) -> TrainReport:
    """Train on TN-labelled syndromes: each batch is sampled on the device
    (seeded), labelled by the fused TN decoder (FP32 marginals normalised to
    class posteriors) and used for one Adam step of soft cross-entropy."""
    if torch is None:
        import torch
    dev = torch.device("cuda:0")
    nclass = 1 << ct.hip.n_out
    S = torch.empty((batch, ct.hip.n_syn), dtype=torch.uint8, device=dev)
    M = torch.empty((batch, nclass), dtype=torch.float64, device=dev)
    C = torch.empty(batch, dtype=torch.int32, device=dev)
    T = torch.empty(batch, dtype=torch.uint8, device=dev)
    Q = torch.empty((batch, nclass), dtype=torch.float32, device=dev)
    steps = max(1, n_samples // batch)
    losses = []
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    done = 0
    for s in range(steps):
        if max_seconds is not None and s % 64 == 63:
            torch.cuda.synchronize()
            if time.perf_counter() - t0 > max_seconds:
                break
        sampler.sample(seed, first + s * batch, batch, S.data_ptr(), stream=stream)
        ct.hip.decode_device(S.data_ptr(), batch, mode, M.data_ptr(), C.data_ptr(), T.data_ptr(), stream)
        posteriors(M.data_ptr(), batch, nclass, Q.data_ptr(), stream)
        frac = s / max(1, steps - 1)
        rate = lr * (lr_final / lr) ** frac            # geometric decay
        want = (s % log_every == 0) or s == steps - 1
        loss = mlp.train_step(S.data_ptr(), Q.data_ptr(), batch, rate, want_loss=want, stream=stream)
        if want:
            losses.append((s, loss))
        done = s + 1
    torch.cuda.synchronize()
    return TrainReport(done * batch, done, time.perf_counter() - t0, losses)


class GatedDecoder:
    """MLP classes where the network is confident, TN (fused FP32 with FP64
    re-check) for the rest — all on the device, no host round trip."""

    def __init__(self, mlp: HipMLP, ct, threshold: float, capacity: int, torch=None, mode: str = "fused_f32"):
        if torch is None:
            import torch
        dev = torch.device("cuda:0")
        self.mlp, self.ct, self.threshold, self.mode = mlp, ct, float(threshold), mode
        self.reject = torch.empty(capacity, dtype=torch.int32, device=dev)
        self.count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.capacity = capacity

    def decode_device(self, d_syn: int, n: int, d_cls: int, d_conf: int, d_mar: int, d_tie: int,
                      stream: int = 0) -> None:
        if n > self.capacity:
            raise ValueError("batch larger than the gate's capacity")
        self.mlp.gate(d_syn, n, self.threshold, d_cls, d_conf, self.reject.data_ptr(), self.count.data_ptr(), stream)
        self.ct.hip.decode_list_device(d_syn, self.reject.data_ptr(), self.count.data_ptr(), n, self.mode,
                                       d_mar, d_cls, d_tie, stream)

    def fallbacks(self) -> int:
        return int(self.count.item())
