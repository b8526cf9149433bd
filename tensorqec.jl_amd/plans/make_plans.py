"""Regenerate the plan files of the d=7 code-capacity workload: the
rotated surface code d=7 under depolarizing noise p/3 per Pauli for
p = 0.02, 0.05, 0.08, contracted along the tree of the seeded tree search
with seed 11 (the tree the benchmark's tuner picks at p=0.05), with the
library's per-syndrome work at the two tabulation depths the tuner chooses
between (7 and 8 syndrome bits).  The tree depends on the network's labels
only, so the three files differ in their leaf tensors alone.

Run: python tensorqec.jl_amd/plans/make_plans.py   (seconds)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from tensorqec import decoders as D  # noqa: E402
from tensorqec.codes import CSSTannerGraph, SurfaceCode  # noqa: E402
from tensorqec.errors import iid_error  # noqa: E402
from tensorqec.planio import save_plan  # noqa: E402

SEED, TABS = 11, (7, 8)


def d7_plan(p):
    t = CSSTannerGraph.from_code(SurfaceCode(7, 7))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    return D.compile(D.TNMMAP(mode="fused_f32", optimizer=D.TreeSA(seed=SEED), tabulate_bits=TABS[0]), t, em).plan


def path_for(p):
    return os.path.join(HERE, f"surface_d7_p{int(round(p * 1000)):03d}_seed{SEED}.json")


def main():
    for p in (0.02, 0.05, 0.08):
        save_plan(d7_plan(p), path_for(p), tabulate_bits=TABS,
                  meta={"code": "rotated surface code d=7", "noise": f"code-capacity depolarizing, p={p} (p/3 each)",
                        "tree": f"TreeSA(seed={SEED}, sc_target=None)", "evidence": "slices"})
        print(path_for(p))


if __name__ == "__main__":
    main()
