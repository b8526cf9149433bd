"""Regenerate tests/golden/toric6.npz: exact TNMMAP marginals of the toric
code L=6 (72 qubits, 2 logical qubits, 16 classes) under code-capacity
depolarizing noise px = py = pz = 0.03, for 8 sampled syndromes.

The marginals are the C oracle's FP64 contraction (oracle/cexec.c) of the
nested-dissection network of that code, sliced on the labels that
tensorqec.slicing.sliced_dissection_plan picks for a widest per-row tensor
of 2^21 entries: each syndrome is contracted once per assignment of the
sliced labels and the slices' marginals are summed in FP64 (2^36.3
multiply-adds per syndrome; about 6 minutes on 7 threads).  Slicing is an exact
rewriting (the CPU tests compare sliced and unsliced plans and brute
force), so the marginals are a property of the code and the noise: any
exact plan of either network must reproduce them.

Contents
  syn    (8, 70)  uint8   syndrome rows [sx | sz] (35 + 35 independent checks),
                          Philox sampler seed 6
  mar    (8, 16)  float64 P(syndrome, class), classes ordered as
                          tensorqec.gf2.logical_operator numbers them
  p      ()       float64 0.03

Run: python tests/golden/make_toric6.py   (~32 GB of host memory)
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
from tensorqec.codes import CSSTannerGraph, ToricCode  # noqa: E402
from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing  # noqa: E402
from tensorqec.gf2 import logical_operator  # noqa: E402
from tensorqec.network import dissected_css_network  # noqa: E402
from tensorqec.slicing import combine, expand_rows, slice_plan, sliced_dissection_plan  # noqa: E402

P, SEED, COUNT, WIDTH = 0.03, 6, 8, 21


def main():
    t = CSSTannerGraph.from_code(ToricCode(6, 6))
    em = iid_error(P, P, P, t.nq)
    lx, lz = logical_operator(t)
    plan, labels, tc, sc = sliced_dissection_plan(
        lambda s: dissected_css_network(t, em, lxlz=(lx, lz), seed=s), WIDTH)
    k = len(labels)
    print(f"plan: {k} sliced labels, width 2^{sc}, {tc:.3g} multiply-adds per syndrome", flush=True)
    S = css_syndrome_rows(sample_depolarizing(em, seed=SEED, count=COUNT), t)
    t0 = time.time()
    part, _, _ = OraclePlan(slice_plan(plan, labels)).decode(expand_rows(S, k))
    mar, _, _ = combine(part, len(S), k)
    print(f"oracle: {len(S)} syndromes in {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "toric6.npz"), syn=S.astype(np.uint8), mar=mar, p=np.float64(P))


if __name__ == "__main__":
    main()
