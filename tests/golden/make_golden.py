"""Regenerate tests/golden/golden.npz.

The vectors are computed by the CPU oracle's exact brute-force enumerator
(oracle/tn_oracle.py) — not by the reference, which is Julia and is not
runnable in this environment — and by the host package's logical-operator
restatement (pinned separately to test/codes/code_distance.jl:57-58).

Contents
  d3_p005_syn / d3_p005_mar   all 256 syndromes of SurfaceCode(3,3) under
                              depolarizing px=py=pz=0.05/3: P(s, class)
  d3_px01_mar                 same syndromes, px=0.1 only (the setting of
                              test/decoding/tndecoder.jl:101-110)
  steane_syn / steane_mar     SteaneCode, px=py=pz=0.05/3
  shor_syn / shor_mar         ShorCode, px=py=pz=0.05/3
  toric2_syn / toric2_mar     ToricCode(2,2) (k=2, 16 classes), px=py=pz=0.05/3
  dem_syn / dem_mar           test/decoding/test_circuits/dem.dem
  lx_<code> / lz_<code>       logical operators of every code the configs use

Run: python tests/golden/make_golden.py
"""
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tensorqec.jl_amd"))

from oracle import tn_oracle as O  # noqa: E402
from tensorqec.codes import CSSTannerGraph, ShorCode, SteaneCode, SurfaceCode, ToricCode  # noqa: E402
from tensorqec.gf2 import logical_operator  # noqa: E402


def table(code, px, py, pz):
    t = CSSTannerGraph.from_code(code)
    lx, lz = logical_operator(t)
    n = t.nq
    bf = O.brute_force_css(t.stgx.H, t.stgz.H, lx, lz, np.full(n, px), np.full(n, py), np.full(n, pz))
    ns = t.ns
    syn = np.array(list(itertools.product([0, 1], repeat=ns)), dtype=np.uint8)[:, ::-1]
    mar = np.array([bf.get(tuple(s), np.zeros(1 << (2 * lx.shape[0]))) for s in syn])
    return syn, mar


def main():
    out = {}
    p = 0.05 / 3
    out["d3_p005_syn"], out["d3_p005_mar"] = table(SurfaceCode(3, 3), p, p, p)
    _, out["d3_px01_mar"] = table(SurfaceCode(3, 3), 0.1, 0.0, 0.0)
    out["steane_syn"], out["steane_mar"] = table(SteaneCode(), p, p, p)
    out["shor_syn"], out["shor_mar"] = table(ShorCode(), p, p, p)
    out["toric2_syn"], out["toric2_mar"] = table(ToricCode(2, 2), p, p, p)
    rates, flipped, dl, ll = O.parse_dem_text(open(os.path.join(HERE, "dem.dem")).read())
    bf = O.brute_force_dem(rates, flipped, dl, ll)
    syn = np.array(list(itertools.product([0, 1], repeat=len(dl))), dtype=np.uint8)[:, ::-1]
    out["dem_syn"] = syn
    out["dem_mar"] = np.array([bf.get(tuple(s), np.zeros(1 << len(ll))) for s in syn])
    codes = {f"surface{d}": SurfaceCode(d, d) for d in (3, 5, 7, 9)}
    codes.update({f"toric{L}": ToricCode(L, L) for L in (2, 3, 4, 5, 6)})
    codes.update({"steane": SteaneCode(), "shor": ShorCode()})
    for name, code in codes.items():
        lx, lz = logical_operator(CSSTannerGraph.from_code(code))
        out[f"lx_{name}"] = lx
        out[f"lz_{name}"] = lz
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
