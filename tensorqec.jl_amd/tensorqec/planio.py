"""Plans as files: a flattened contraction plan (``planner.Plan``) written
to and read back from JSON, exactly (leaf tensors as decimal floats that
round-trip bit for bit), with the work it implies per syndrome.

A plan file fixes what a benchmark executes: the tree (steps, their label
orders), the leaf tensors and how syndrome bits enter.  Its ``work`` block
gives, per tabulation depth, the library's own count of the per-syndrome
steps (``tqec_plan_stats``: ``dep_tc`` multiply-adds, ``dep_rw`` tensor
entries read and written, ``sc`` log2 of the widest tensor), from which
W_flops = 2 x dep_tc and W_bytes = 4 x dep_rw (FP32) follow; the same
counts as ``contraction_complexity`` (src/decoding/tndecoder.jl:273-279)
restricted to the steps that run per syndrome.
"""
from __future__ import annotations

import json
from typing import Dict, Iterable, Optional

import numpy as np

from .planner import Plan

FORMAT = "tensorqec-plan-1"


def plan_to_dict(plan: Plan) -> dict:
    return {
        "format": FORMAT,
        "n_leaves": plan.n_leaves,
        "leaf_labels": [list(map(int, l)) for l in plan.leaf_labels],
        "leaf_data": [None if d is None else [float(x) for x in np.asarray(d, np.float64).reshape(-1)]
                      for d in plan.leaf_data],
        "leaf_slot": [int(s) for s in plan.leaf_slot],
        "lhs": [int(x) for x in plan.lhs],
        "rhs": [int(x) for x in plan.rhs],
        "out_labels": [list(map(int, l)) for l in plan.out_labels],
        "out_labels_final": [int(x) for x in plan.out_labels_final],
        "n_syndrome_bits": int(plan.n_syndrome_bits),
        "evidence_label_bits": {str(int(k)): int(v) for k, v in plan.evidence_label_bits.items()},
    }


def plan_from_dict(d: dict) -> Plan:
    if d.get("format") != FORMAT:
        raise ValueError(f"not a {FORMAT} plan")
    return Plan(n_leaves=int(d["n_leaves"]), leaf_labels=[list(l) for l in d["leaf_labels"]],
                leaf_data=[None if x is None else np.asarray(x, np.float64) for x in d["leaf_data"]],
                leaf_slot=list(d["leaf_slot"]), lhs=list(d["lhs"]), rhs=list(d["rhs"]),
                out_labels=[list(l) for l in d["out_labels"]], out_labels_final=list(d["out_labels_final"]),
                n_syndrome_bits=int(d["n_syndrome_bits"]),
                evidence_label_bits={int(k): int(v) for k, v in d["evidence_label_bits"].items()})


def plan_work(plan: Plan, tabulate_bits: Iterable[int]) -> Dict[str, dict]:
    """Per tabulation depth, the library's per-syndrome counts and the
    W_flops / W_bytes they give (FP32 entries)."""
    from . import lib as L
    out = {}
    for tab in tabulate_bits:
        st = L.HipPlan(plan, tabulate_bits=int(tab)).stats()
        out[str(int(tab))] = {"dep_tc": st["dep_tc"], "dep_rw": st["dep_rw"], "sc": st["sc"],
                              "n_dep_steps": st["n_dep_steps"], "table_entries": st["table_entries"],
                              "w_flops_per_syndrome": 2.0 * st["dep_tc"],
                              "w_bytes_per_syndrome_fp32": 4.0 * st["dep_rw"]}
    return out


def save_plan(plan: Plan, path: str, meta: Optional[dict] = None, tabulate_bits: Iterable[int] = ()) -> None:
    d = plan_to_dict(plan)
    d["meta"] = dict(meta or {})
    d["work"] = plan_work(plan, tabulate_bits)
    with open(path, "w") as f:
        json.dump(d, f, separators=(",", ":"))
        f.write("\n")


def load_plan(path: str):
    """(plan, meta, work) from a plan file."""
    with open(path) as f:
        d = json.load(f)
    return plan_from_dict(d), d.get("meta", {}), d.get("work", {})
