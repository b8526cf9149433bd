"""DEM TNMMAP on the GPU: the reference's dem.dem fixture against brute force,
the device recovery against the host restatement of tndecoder.jl:255-271,
and a circuit-level rotated memory DEM against the CPU oracle."""
import os

import numpy as np
import pytest

from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
from tensorqec.dem import dem_error_patterns, parse_dem_file, random_error_pattern

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden")
GOLDEN = np.load(os.path.join(HERE, "golden.npz"))


@pytest.mark.parametrize("mode", ["plain_f64", "fused_f64", "fused_f32"])
@pytest.mark.parametrize("factorize,evidence", [(True, "slices"), (False, "leaves")])
def test_dem_fixture_all_syndromes(mode, factorize, evidence):
    dem = parse_dem_file(os.path.join(HERE, "dem.dem"))
    ct = D.compile(D.TNMMAP(mode=mode, factorize=factorize, evidence=evidence), dem)
    S, ref = GOLDEN["dem_syn"], GOLDEN["dem_mar"]
    r = D.decode_batch(ct, S)
    valid = ref.max(axis=1) > 0
    if mode == "fused_f32":
        err = np.max(np.abs(r.marginals[valid] - ref[valid]) / ref[valid].max(axis=1, keepdims=True))
        assert err < 2e-5
    else:
        nz = ref > 0
        assert np.all(r.marginals[~nz] == 0)
        assert np.max(np.abs(r.marginals[nz] - ref[nz]) / ref[nz]) < 1e-12
    gap_ok = np.abs(ref[:, 0] - ref[:, 1]) > 1e-12 * ref.max(axis=1)
    assert np.array_equal(r.classes[valid & gap_ok], np.argmax(ref, axis=1)[valid & gap_ok])


def test_dem_device_recovery_matches_host():
    torch = pytest.importorskip("torch")
    dem = parse_dem_file(os.path.join(HERE, "dem.dem"))
    ct = D.compile(D.TNMMAP(), dem)
    e = random_error_pattern(dem, 5, 4096)
    S = ((e.astype(np.int64) @ ct.tanner.H.T.astype(np.int64)) & 1).astype(np.uint8)
    cls = (np.arange(len(S)) % 2).astype(np.int32)
    host = dem_error_patterns(ct, S, cls)
    rec = L.HipRecovery.dem(ct.pure, ct.l2q_mask)
    dev = torch.device("cuda:0")
    dS, dC = torch.from_numpy(S).to(dev), torch.from_numpy(cls).to(dev)
    dA = torch.empty((len(S), ct.pure.shape[0]), dtype=torch.uint8, device=dev)
    rec.recover(dS.data_ptr(), dC.data_ptr(), len(S), dA.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(dA.cpu().numpy(), host)


def test_circuit_level_d3_dem_vs_oracle():
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    ct = D.compile(D.TNMMAP(mode="plain_f64", optimizer=D.TreeSA(sweeps=100, restarts=2)), dem)
    e = random_error_pattern(dem, 11, 256)
    S = ((e.astype(np.int64) @ ct.tanner.H.T.astype(np.int64)) & 1).astype(np.uint8)
    ref, _, _ = OraclePlan(ct.plan).decode(S)
    r = D.decode_batch(ct, S, error_patterns=False)
    nz = ref > 0
    assert np.all(r.marginals[~nz] == 0)
    assert np.max(np.abs(r.marginals[nz] - ref[nz]) / ref[nz]) < 1e-12
    # the wide FP32 executor (too wide for LDS) with its FP64 re-check
    w = D.decode_batch(ct, S, mode="wide_f32", error_patterns=False)
    assert np.max(np.abs(w.marginals - ref) / ref.max(axis=1, keepdims=True)) < 2e-5
    ok = np.abs(np.diff(np.sort(ref, axis=1)[:, -2:], axis=1))[:, 0] > 1e-12 * ref.max(axis=1)
    assert np.array_equal(w.classes[ok], np.argmax(ref, axis=1)[ok])


@pytest.fixture(scope="module")
def d5_two_rounds():
    """Circuit-level memory-Z, rotated surface code d=5, 2 rounds, p=0.001
    (48 detectors, about 600 mechanisms), compiled once, with 96 syndromes
    drawn at 10x the mechanism rates (most rows carry several flips) and the
    C oracle's marginals on the same plan."""
    dem = detector_error_model_from_stim(rotated_memory_z(5, 2, 0.001))
    ct = D.compile(D.TNMMAP(), dem)
    rng = np.random.default_rng(13)
    e = (rng.random((96, len(dem.error_rates))) < 10 * np.asarray(dem.error_rates)[None, :]).astype(np.int64)
    S = ((e @ ct.tanner.H.T.astype(np.int64)) & 1).astype(np.uint8)
    ref, _, _ = OraclePlan(ct.plan).decode(S)
    return ct, S, ref


@pytest.mark.parametrize("mode", ["plain_f64", "spec_f32", "spec_bf16"])
def test_circuit_level_d5_two_rounds_vs_oracle(d5_two_rounds, mode):
    """The compile picks the nested-dissection network (width 2^19, ~23 M
    multiply-adds per syndrome) and every executor matches the C oracle on
    the same plan (tndecoder.jl:248-253 decodes this network one syndrome at
    a time)."""
    from _parity import check_f32, check_f64
    ct, S, ref = d5_two_rounds
    assert ct.graph.meta.get("dissected")
    print(f"case d5r2 dem {mode}: {len(S)} syndromes", flush=True)
    r = D.decode_batch(ct, S, mode=mode, error_patterns=False)
    if mode == "plain_f64":
        check_f64(r.marginals, r.classes, r.ties, ref, mode)
    else:
        check_f32(r.marginals, r.classes, ref, mode, tol=3e-2 if mode == "spec_bf16" else 2e-5)


@pytest.mark.parametrize("mode", ["spec_f64", "spec_f32"])
def test_circuit_level_d5_two_rounds_sliced_vs_oracle(d5_two_rounds, mode):
    """Label slicing on a detector error model, with the subtrees that depend
    on at most 4 syndrome bits tabulated: the d=5, 2-round network with its
    tree and sliced labels annealed together, each syndrome decoded as 2^k
    rows and the rows summed, against the oracle marginals of the unsliced
    plan.  The rows are 2^14 wide (the unsliced plan: 2^19), narrow enough
    that annealing alone cannot reach them and labels must be sliced."""
    from _parity import check_f32, check_f64
    from tensorqec.network import dissected_dem_network
    from tensorqec.slicing import SlicedDecoder, sliced_dissection_plan
    ct, S, ref = d5_two_rounds
    dem = detector_error_model_from_stim(rotated_memory_z(5, 2, 0.001))
    target = 14
    plan, labels, _, sc = sliced_dissection_plan(lambda s: dissected_dem_network(dem, seed=s), target,
                                                 seeds=1, sweeps=100, iters=3)
    assert sc <= target and labels
    dec = SlicedDecoder(plan, labels, tabulate_bits=4)
    print(f"case d5r2 dem sliced {mode}: {len(labels)} labels, width 2^{sc}, {len(S)} syndromes", flush=True)
    mar, cls, tie = dec.decode(S, mode)
    if mode == "spec_f64":
        check_f64(mar, cls, tie, ref, mode)
    else:
        check_f32(mar, cls, ref, mode, tol=2e-5)
