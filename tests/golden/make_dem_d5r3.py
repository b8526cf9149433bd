"""Regenerate tests/golden/dem_d5r3.npz: exact TNMMAP marginals of the
circuit-level memory-Z experiment on the rotated surface code d=5 with 3
rounds (72 detectors, 821 mechanisms, 1 logical; depolarizing p=0.001 after
every gate, reset and measurement flips p, as tensorqec.circuits generates
it) for 2 syndromes.

The marginals are the C oracle's FP64 contraction (oracle/cexec.c) of the
nested-dissection DEM network sliced on the labels that
tensorqec.slicing.sliced_dissection_plan picks for rows of 2^21 entries:
each syndrome is contracted once per assignment of the 9 sliced labels and
the slices are summed in FP64 (2^40.5 multiply-adds per syndrome, about 8
minutes per syndrome on 7 threads, ~42 GB of host memory).  Slicing is an
exact rewriting (the CPU tests compare sliced and unsliced plans), so the
marginals are a property of the model: any exact plan of its network must
reproduce them.

The syndromes come from mechanism flips drawn at 10 times their rates
(numpy default_rng, seed 31), so that each carries several detection
events.

Contents
  syn    (2, 72)  uint8   detector rows
  mar    (2, 2)   float64 P(syndrome, logical parity)
  p      ()       float64 0.001

Run: python tests/golden/make_dem_d5r3.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tensorqec.jl_amd"))

from oracle.cexec import OraclePlan  # noqa: E402
from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z  # noqa: E402
from tensorqec.network import dissected_dem_network  # noqa: E402
from tensorqec.slicing import combine, expand_rows, slice_plan, sliced_dissection_plan  # noqa: E402

P, SEED, COUNT, WIDTH = 0.001, 31, 2, 21


def syndromes(dem):
    rng = np.random.default_rng(SEED)
    e = (rng.random((COUNT, len(dem.error_rates))) < 10 * np.asarray(dem.error_rates)[None, :]).astype(np.int64)
    return ((e @ dem.tanner().H.T.astype(np.int64)) & 1).astype(np.uint8)


def main():
    dem = detector_error_model_from_stim(rotated_memory_z(5, 3, P))
    plan, labels, tc, sc = sliced_dissection_plan(lambda s: dissected_dem_network(dem, seed=s), WIDTH)
    k = len(labels)
    print(f"plan: {k} sliced labels, width 2^{sc}, {tc:.3g} multiply-adds per syndrome", flush=True)
    S = syndromes(dem)
    t0 = time.time()
    part, _, _ = OraclePlan(slice_plan(plan, labels)).decode(expand_rows(S, k))
    mar, _, _ = combine(part, len(S), k)
    print(f"oracle: {len(S)} syndromes in {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "dem_d5r3.npz"), syn=S, mar=mar, p=np.float64(P))


if __name__ == "__main__":
    main()
