"""Label slicing: contractions wider than one workgroup's slab are split into
2^k slices, each with k inner labels fixed, decoded as separate rows of one
batch and summed.

This is the slicing step of the reference's planner
(`TreeSASlicer`, src/nonclifford/tensornetwork.jl:174-192: fix a few labels,
contract every assignment, add the results), applied to a fixed contraction
tree: the labels are chosen greedily among those of the widest
intermediates, each becomes a *virtual syndrome bit* of the plan (the plan
compiler already fixes labels by syndrome bits: that is how evidence is
sliced), and a syndrome row is expanded into 2^k rows carrying its slice
index in those bits.  The marginals of the slices are summed in FP64; the
class is the first maximum of the sum (`findmax`, tndecoder.jl:163) and the
tie flag uses the executors' rule.  Work grows by the slicing overhead
(shared subtrees are recomputed per slice); the widest tensor shrinks by a
factor 2 per sliced label.
"""
from __future__ import annotations

from dataclasses import replace
from typing import List, Optional, Sequence

import numpy as np

from .planner import Plan


def choose_slices(plan: Plan, target_rank: int, max_labels: int = 12) -> List[int]:
    """Labels to slice so that no per-syndrome intermediate of ``plan`` has
    more than ``target_rank`` labels.  Slicing a label halves every step
    whose operands carry it and doubles the number of slices, so the total
    work grows by sum(steps without it) / sum(all steps); each round takes,
    among the labels of the current widest intermediates, the one carried by
    the costliest steps (ties: the smallest label).  Output labels and labels
    already fixed by syndrome bits are never chosen."""
    dep = plan.dependent_nodes()
    fixed = set(plan.evidence_label_bits) | set(plan.out_labels_final)
    L = plan.n_leaves

    def eff(node):
        labs = plan.node_labels(node)
        return [l for l in labs if l not in plan.evidence_label_bits] if node < L else labs

    steps = [(set(eff(plan.lhs[i])) | set(eff(plan.rhs[i])), set(plan.out_labels[i]))
             for i in range(plan.n_steps) if dep[L + i]]
    chosen: List[int] = []
    while steps and len(chosen) < max_labels:
        top = max(len(out) for _, out in steps)
        if top <= target_rank:
            break
        cand = {l for _, out in steps if len(out) == top for l in out if l not in fixed}
        if not cand:
            break
        weight = {l: 0.0 for l in cand}
        for w, _ in steps:
            c = 2.0 ** len(w)
            for l in w & cand:
                weight[l] += c
        best = min(cand, key=lambda l: (-weight[l], l))
        chosen.append(best)
        fixed.add(best)
        for w, out in steps:
            w.discard(best)
            out.discard(best)
    return chosen


def slice_plan(plan: Plan, labels: Sequence[int]) -> Plan:
    """``plan`` with ``labels`` fixed by virtual syndrome bits
    ``n_syndrome_bits + j``: the labels leave every intermediate, the leaves
    that carry them become variants selected by those bits."""
    labels = list(labels)
    bad = [l for l in labels if l in plan.evidence_label_bits or l in plan.out_labels_final]
    if bad:
        raise ValueError(f"labels {bad} are output labels or already fixed by syndrome bits")
    ns = plan.n_syndrome_bits
    ev = dict(plan.evidence_label_bits)
    for j, l in enumerate(labels):
        ev[l] = ns + j
    drop = set(labels)
    outs = [[l for l in out if l not in drop] for out in plan.out_labels]
    return replace(plan, out_labels=outs, n_syndrome_bits=ns + len(labels), evidence_label_bits=ev)


def expand_rows(syndromes: np.ndarray, k: int) -> np.ndarray:
    """Each row followed by its 2^k slices: row r, slice s -> row r * 2^k + s
    with bit j of s appended as syndrome bit ``n + j``."""
    S = np.ascontiguousarray(np.atleast_2d(np.asarray(syndromes, dtype=np.uint8)))
    n, ns = S.shape
    m = 1 << k
    out = np.empty((n * m, ns + k), dtype=np.uint8)
    out[:, :ns] = np.repeat(S, m, axis=0)
    s = np.tile(np.arange(m, dtype=np.int64), n)
    for j in range(k):
        out[:, ns + j] = (s >> j) & 1
    return out


def combine(marginals: np.ndarray, n: int, k: int, tie_tol: float = 1e-12):
    """Sum the slices' marginals of each row; classes = first maximum,
    ties = top-2 gap at most ``tie_tol`` of the maximum."""
    nout = marginals.shape[1]
    mar = np.asarray(marginals, dtype=np.float64).reshape(n, 1 << k, nout).sum(axis=1)
    cls = np.argmax(mar, axis=1).astype(np.int32)
    if nout > 1:
        top2 = np.sort(mar, axis=1)[:, -2:]
        tie = ((top2[:, 1] - top2[:, 0]) <= tie_tol * top2[:, 1]).astype(np.uint8)
    else:
        tie = np.zeros(n, dtype=np.uint8)
    return mar, cls, tie


class SlicedDecoder:
    """A compiled plan decoded in 2^k slices (see the module docstring).

    ``hip`` is the library plan of the sliced plan (``tensorqec.lib.HipPlan``);
    ``decode`` takes host syndrome rows of the original width."""

    def __init__(self, plan: Plan, labels: Sequence[int], tabulate_bits: int = 0):
        from . import lib as L
        self.labels = list(labels)
        self.k = len(self.labels)
        self.plan = slice_plan(plan, self.labels)
        self.n_syn = plan.n_syndrome_bits
        self.hip = L.HipPlan(self.plan, tabulate_bits=tabulate_bits)

    def decode(self, syndromes, mode: str = "spec_f32", tie_tol: float = 1e-12):
        S = np.atleast_2d(np.asarray(syndromes, dtype=np.uint8))
        if S.shape[1] != self.n_syn:
            raise ValueError(f"syndrome rows have {S.shape[1]} bits, the plan {self.n_syn}")
        mar, _, _ = self.hip.decode(expand_rows(S, self.k), mode, True)
        return combine(mar, len(S), self.k, tie_tol)

    def decode_device(self, torch, d_syn, n: int, mode: str = "spec_f32", stream: Optional[int] = None,
                      tie_tol: float = 1e-12):
        """Device rows in, device (marginals, classes, ties) out; the slices
        are expanded and summed with torch on the same device."""
        dev = d_syn.device
        m = 1 << self.k
        ext = torch.empty((n * m, self.n_syn + self.k), dtype=torch.uint8, device=dev)
        ext[:, :self.n_syn] = d_syn[:n].repeat_interleave(m, dim=0)
        s = torch.arange(m, device=dev).repeat(n)
        for j in range(self.k):
            ext[:, self.n_syn + j] = ((s >> j) & 1).to(torch.uint8)
        nout = 1 << self.hip.n_out
        mar = torch.empty((n * m, nout), dtype=torch.float64, device=dev)
        cls = torch.empty(n * m, dtype=torch.int32, device=dev)
        tie = torch.empty(n * m, dtype=torch.uint8, device=dev)
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        self.hip.decode_device(ext.data_ptr(), n * m, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), st)
        tot = mar.view(n, m, nout).sum(dim=1)
        top2 = torch.topk(tot, 2, dim=1).values if nout > 1 else None
        c = torch.argmax(tot, dim=1).to(torch.int32)
        t = ((top2[:, 0] - top2[:, 1]) <= tie_tol * top2[:, 0]).to(torch.uint8) if nout > 1 else \
            torch.zeros(n, dtype=torch.uint8, device=dev)
        return tot, c, t
