"""Regenerate tests/golden/toric5.npz: exact TNMMAP marginals of the toric
code L=5 (50 qubits, 2 logical qubits, 16 classes) under code-capacity
depolarizing noise px = py = pz = 0.03, for 16 sampled syndromes.

The marginals are the C oracle's FP64 contraction (oracle/cexec.c) of the
nested-dissection network of that code (width 2^21, 1.33 G multiply-adds per
syndrome, 46-72 s per syndrome on one core), which the CPU tests
pin against the reference network (src/decoding/tndecoder.jl:97-146) on
smaller codes.  They are a property of the code and the noise, not of the
contraction order: any exact plan of either network must reproduce them.

Contents
  syn    (16, 48) uint8   syndrome rows [sx | sz] (24 + 24 independent checks),
                          Philox sampler seed 5
  mar    (16, 16) float64 P(syndrome, class), classes ordered as
                          tensorqec.gf2.logical_operator numbers them
  p      ()       float64 0.03

Run: python tests/golden/make_toric5.py   (about 12 minutes; ~32 GB of host memory)
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
from tensorqec.planner import flatten  # noqa: E402

P, SEED, COUNT = 0.03, 5, 16


def main():
    t = CSSTannerGraph.from_code(ToricCode(5, 5))
    em = iid_error(P, P, P, t.nq)
    lx, lz = logical_operator(t)
    g, tree = dissected_css_network(t, em, lxlz=(lx, lz), seed=0)
    plan = flatten(g.without_evidence_leaves(), tree, True)
    S = css_syndrome_rows(sample_depolarizing(em, seed=SEED, count=COUNT), t)
    t0 = time.time()
    mar, _, _ = OraclePlan(plan).decode(S)
    print(f"oracle: {len(S)} syndromes in {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "toric5.npz"), syn=S.astype(np.uint8), mar=mar, p=np.float64(P))


if __name__ == "__main__":
    main()
