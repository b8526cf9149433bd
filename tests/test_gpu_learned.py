"""Learned front-end end to end on the GPU: trained by the library's own
forward/backward/Adam kernels on syndromes labelled by the TN decoder, gated
by confidence (uncertain syndromes go to the TN path on the device), and
judged against the CPU oracle's FP64 contraction of the same plan
(`decode(ct, syndrome)`, src/decoding/tndecoder.jl:160-165 and :248-253).

  * d=7 rotated surface code, code capacity p=0.05: 48 inputs, 4 classes,
    1.3 M training syndromes.
  * circuit-level memory-Z, d=3, 3 rounds, p=0.001: 24 detector inputs,
    2 classes (logical Z flip or not), 1 M training syndromes.
The gate threshold is the smallest of a fixed ladder whose decisions differ
from the TN decoder's on at most 1e-4 of a calibration batch; evaluation
batches are held out (other seeds).  Required: the training loss falls, the
gated logical error rate lies inside the 95% Wilson interval of the FP64
oracle's, and the gated decisions differ from the oracle's on at most 1e-3
of syndromes; the fallback fraction is reported.
"""
import math
import sys

import numpy as np
import pytest

from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode
from tensorqec.errors import iid_error
from tensorqec.learned import GatedDecoder, HipMLP, train_on_tn

pytestmark = pytest.mark.gpu

LADDER = [0.9, 0.95, 0.98, 0.99, 0.995, 0.998, 0.999, 0.9995, 0.9999]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def wilson(k, n, z=1.96):
    p = k / n
    den = 1 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return c - h, c + h


def train_gate_evaluate(torch, ct, smp, n_in, n_classes, n_train, n_eval, mode, seed):
    dev = torch.device("cuda:0")
    mlp = HipMLP(n_in, n_classes, seed=seed)
    rep = train_on_tn(mlp, ct, smp, n_samples=n_train, batch=8192, lr=2e-3, lr_final=2e-4, seed=seed + 1,
                      torch=torch, mode=mode)
    first, last = rep.losses[0][1], rep.losses[-1][1]
    say(f"  trained on {rep.samples} syndromes in {rep.seconds:.1f}s, loss {first:.4f} -> {last:.4f}")
    assert last < 0.5 * first
    nb = 65536
    S = torch.empty((nb, smp.nsyn), dtype=torch.uint8, device=dev)
    truth = torch.empty(nb, dtype=torch.int32, device=dev)
    mar = torch.empty((nb, n_classes), dtype=torch.float64, device=dev)
    c_tn = torch.empty(nb, dtype=torch.int32, device=dev)
    c_g = torch.empty(nb, dtype=torch.int32, device=dev)
    conf = torch.empty(nb, dtype=torch.float32, device=dev)
    tie = torch.empty(nb, dtype=torch.uint8, device=dev)
    smp.sample(seed + 2, 0, nb, S.data_ptr(), truth.data_ptr())
    ct.hip.decode_device(S.data_ptr(), nb, mode, mar.data_ptr(), c_tn.data_ptr(), tie.data_ptr())
    chosen = LADDER[-1]
    for th in LADDER:
        gd = GatedDecoder(mlp, ct, th, nb, torch=torch, mode=mode)
        gd.decode_device(S.data_ptr(), nb, c_g.data_ptr(), conf.data_ptr(), mar.data_ptr(), tie.data_ptr())
        torch.cuda.synchronize()
        if float((c_g != c_tn).float().mean().item()) <= 1e-4:
            chosen = th
            break
    gd = GatedDecoder(mlp, ct, chosen, nb, torch=torch, mode=mode)
    op = OraclePlan(ct.plan)
    tot = dict(n=0, fall=0, err_g=0, err_o=0, dis=0)
    for k in range(0, n_eval, nb):
        m = min(nb, n_eval - k)
        smp.sample(seed + 3, k, m, S.data_ptr(), truth.data_ptr())
        gd.decode_device(S.data_ptr(), m, c_g.data_ptr(), conf.data_ptr(), mar.data_ptr(), tie.data_ptr())
        torch.cuda.synchronize()
        _, c_o, _ = op.decode(S[:m].cpu().numpy())
        g, tr = c_g[:m].cpu().numpy(), truth[:m].cpu().numpy()
        tot["n"] += m
        tot["fall"] += gd.fallbacks()
        tot["err_g"] += int((g != tr).sum())
        tot["err_o"] += int((c_o != tr).sum())
        tot["dis"] += int((g != c_o).sum())
    n = tot["n"]
    lo, hi = wilson(tot["err_o"], n)
    ler_g = tot["err_g"] / n
    say(f"  gate {chosen}: fallback {tot['fall'] / n:.4f}, disagreement with FP64 oracle {tot['dis'] / n:.2e}, "
        f"LER gated {ler_g:.5f}, FP64 oracle {tot['err_o'] / n:.5f} [{lo:.5f}, {hi:.5f}] on {n} syndromes")
    assert lo <= ler_g <= hi
    assert tot["dis"] <= 1e-3 * n
    return tot


def test_d7_code_capacity_front_end():
    torch = pytest.importorskip("torch")
    t = CSSTannerGraph.from_code(SurfaceCode(7, 7))
    p = 0.05
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
    smp = L.HipSampler.css(t, em, ct.lx, ct.lz)
    say("case learned d=7 code capacity p=0.05")
    train_gate_evaluate(torch, ct, smp, t.ns, 4, n_train=8192 * 160, n_eval=4 * 65536, mode="spec_f32", seed=71)


def test_circuit_level_dem_front_end():
    torch = pytest.importorskip("torch")
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    ct = D.compile(D.TNMMAP(mode="spec_f32"), dem)
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    say(f"case learned circuit-level d=3 3 rounds p=0.001: {smp.nsyn} detectors, {len(dem.error_rates)} mechanisms")
    train_gate_evaluate(torch, ct, smp, smp.nsyn, 2, n_train=8192 * 128, n_eval=16384, mode="spec_f32", seed=33)
