"""Label slicing: contractions wider than one workgroup's slab are split into
2^k slices, each with k inner labels fixed, decoded as separate rows of one
batch and summed.

This is the slicing step of the reference's planner
(`TreeSASlicer`, src/nonclifford/tensornetwork.jl:174-192: fix a few labels,
contract every assignment, add the results).  The labels are chosen greedily
among those of the widest intermediates (:func:`choose_slices`, on a fixed
tree), or together with the tree, which is re-annealed around them
(:func:`anneal_sliced`, :func:`sliced_dissection_plan`).  Each becomes a
*virtual syndrome bit* of the plan (the plan compiler already fixes labels by
syndrome bits: that is how evidence is sliced), and a syndrome row is
expanded into 2^k rows carrying its slice index in those bits.  The
marginals of the slices are summed in FP64; the class is the first maximum
of the sum (`findmax`, tndecoder.jl:163) and the tie flag uses the
executors' rule; in reduced precision, syndromes whose sum is close to a
tie are decided again in FP64 (:class:`SlicedDecoder`).  Work grows by the
slicing overhead (subtrees that carry no sliced label are recomputed per
slice); the widest tensor shrinks by a factor 2 per sliced label.
"""
from __future__ import annotations

from dataclasses import replace
from typing import List, Optional, Sequence

import numpy as np

from .planner import Plan


def _step_masks(plan: Plan):
    """Per step: bitmask of the labels it multiplies over (both operands,
    evidence labels of leaves excluded) and of its output labels; per leaf:
    whether it depends on the syndrome, and its label bitmask."""
    L = plan.n_leaves
    ev = plan.evidence_label_bits

    def mask(labs):
        m = 0
        for l in labs:
            m |= 1 << l
        return m

    leaf_dep = [plan.leaf_slot[t] >= 0 or any(l in ev for l in plan.leaf_labels[t]) for t in range(L)]
    leaf_mask = [mask(plan.leaf_labels[t]) for t in range(L)]
    eff = [mask(l for l in plan.leaf_labels[t] if l not in ev) for t in range(L)] + \
          [mask(out) for out in plan.out_labels]
    w = [eff[plan.lhs[i]] | eff[plan.rhs[i]] for i in range(plan.n_steps)]
    out = eff[L:]
    return leaf_dep, leaf_mask, w, out


def _sliced_cost(plan: Plan, masks, sliced: int):
    """(per-syndrome multiply-adds summed over the 2^k slices, widest
    per-row intermediate) of ``plan`` with the labels of bitmask ``sliced``
    fixed: leaves carrying a sliced label become row-dependent, and so does
    every step above them."""
    leaf_dep, leaf_mask, w, out = masks
    L = plan.n_leaves
    dep = [leaf_dep[t] or bool(leaf_mask[t] & sliced) for t in range(L)]
    keep = ~sliced
    tc, sc = 0.0, 0
    for i in range(plan.n_steps):
        d = dep[plan.lhs[i]] or dep[plan.rhs[i]]
        dep.append(d)
        if d:
            tc += 2.0 ** (w[i] & keep).bit_count()
            sc = max(sc, (out[i] & keep).bit_count())
    return tc * 2.0 ** sliced.bit_count(), sc


def choose_slices(plan: Plan, target_rank: int, max_labels: int = 12) -> List[int]:
    """Labels to slice so that no per-row intermediate of ``plan`` has more
    than ``target_rank`` labels.  Fixing a label removes it from every step
    that carries it, doubles the number of rows, and makes every constant
    (syndrome-independent) subtree that carries it row-dependent; each round
    takes, among the labels of the current widest per-row intermediates, the
    one that gives the least total work per syndrome after slicing (ties:
    the narrower result, then the smallest label).  Output labels and labels
    already fixed by syndrome bits are never chosen."""
    masks = _step_masks(plan)
    _, _, _, out = masks
    fixed = set(plan.evidence_label_bits) | set(plan.out_labels_final)
    chosen: List[int] = []
    sliced = 0
    tc, sc = _sliced_cost(plan, masks, 0)
    while sc > target_rank and len(chosen) < max_labels:
        L = plan.n_leaves
        dep = _sliced_dep(plan, masks, sliced)
        keep = ~sliced
        cand = set()
        for i in range(plan.n_steps):
            if dep[L + i] and (out[i] & keep).bit_count() == sc:
                m = out[i] & keep
                while m:
                    b = m & -m
                    cand.add(b.bit_length() - 1)
                    m ^= b
        cand -= fixed
        if not cand:
            break
        best = min(cand, key=lambda l: (*_sliced_cost(plan, masks, sliced | (1 << l)), l))
        chosen.append(best)
        fixed.add(best)
        sliced |= 1 << best
        tc, sc = _sliced_cost(plan, masks, sliced)
    return chosen


def _sliced_dep(plan: Plan, masks, sliced: int) -> List[bool]:
    leaf_dep, leaf_mask, _, _ = masks
    dep = [leaf_dep[t] or bool(leaf_mask[t] & sliced) for t in range(plan.n_leaves)]
    for i in range(plan.n_steps):
        dep.append(dep[plan.lhs[i]] or dep[plan.rhs[i]])
    return dep


def anneal_sliced(graph, tree, target_rank: int, iters: int = 4, sweeps: int = 100, seed: int = 0,
                  max_labels: int = 12):
    """Slices and contraction tree chosen together, for ``graph`` (a network
    whose evidence labels are fixed by syndrome bits; ``tree`` over its
    non-evidence tensors): alternately pick the labels to slice on the
    current tree (:func:`choose_slices`) and re-anneal the tree with those
    labels fixed and a penalty on intermediates wider than ``target_rank``
    (``planner.anneal_tree``), keeping the cheapest sliced plan whose widest
    per-row intermediate fits.  Slicing a fixed tree can cost orders of
    magnitude where many distinct intermediates are at the top width and
    share few labels; the re-annealed tree moves those labels onto few
    intermediates.  The reference interleaves slicing with its TreeSA
    annealing the same way (`TreeSASlicer`,
    src/nonclifford/tensornetwork.jl:174-192).

    Returns ``(plan, labels, tc, sc)``: the flattened plan (evidence as
    slices), the labels to slice, the per-syndrome multiply-adds over all
    slices and the per-row width."""
    from .planner import ContractionTree, anneal_tree, flatten, recompute_masks
    gw = graph.without_evidence_leaves()
    L = tree.n_leaves
    ev = set(graph.evidence_labels)
    ixs = gw.ixs[:L]
    best = None
    labels: List[int] = []
    for it in range(max(1, iters)):
        plan = flatten(gw, tree, True)
        labels = choose_slices(plan, target_rank, max_labels=max_labels)
        tc, sc = _sliced_cost(plan, _step_masks(plan), sum(1 << l for l in labels))
        if best is None or (sc > target_rank, tc) < (best[3] > target_rank, best[2]):
            best = (plan, list(labels), tc, sc)
        if it + 1 == iters:
            break
        fixed = ev | set(labels)
        t = ContractionTree(L, list(tree.children), [])
        t.masks = recompute_masks(t, [[l for l in ix if l not in fixed] for ix in ixs], graph.iy)
        t = anneal_tree(t, graph.iy, seed=seed * 1009 + it, sweeps=sweeps, sc_target=target_rank)
        tree = ContractionTree(L, list(t.children), [])
        tree.masks = recompute_masks(tree, [[l for l in ix if l not in ev] for ix in ixs], graph.iy)
    return best


def sliced_dissection_plan(build, target_rank: int, seeds: int = 2, sweeps: int = 200, iters: int = 4,
                           slice_sweeps: int = 100, max_labels: int = 12):
    """:func:`anneal_sliced` over ``seeds`` dissected networks
    ``build(seed) -> (graph, tree)`` (each first annealed as
    ``decoders.dissection_plan`` does): the cheapest result whose widest
    per-row intermediate fits ``target_rank``, as ``(plan, labels, tc,
    sc)``."""
    from .planner import anneal_tree, tree_cost
    best = None
    for seed in range(max(1, seeds)):
        g, tree = build(seed)
        if sweeps > 0:
            ref = anneal_tree(tree, g.iy, seed=seed, sweeps=sweeps)
            if tree_cost(ref)[0] < tree_cost(tree)[0]:
                tree = ref
        r = anneal_sliced(g, tree, target_rank, iters=iters, sweeps=slice_sweeps, seed=seed,
                          max_labels=max_labels)
        if best is None or (r[3] > target_rank, r[2]) < (best[3] > target_rank, best[2]):
            best = r
    return best


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


# FP64 counterpart of each reduced-precision mode, for the re-check of
# syndromes whose summed marginals are too close to call.
_FP64_OF = {"spec_f32": "spec_f64", "spec_bf16": "spec_f64", "fused_f32": "fused_f64", "wide_f32": "plain_f64"}


def needs_fp64(totals: np.ndarray, recheck_tol: float) -> np.ndarray:
    """Rows of summed marginals to decide again in FP64: no positive entry,
    or a top-2 gap within ``recheck_tol`` of the maximum (the executors'
    per-syndrome rule, applied to the sum of the slices)."""
    t = np.asarray(totals, dtype=np.float64)
    best = t.max(axis=1)
    flag = ~(best > 0)
    if t.shape[1] > 1:
        second = np.sort(t, axis=1)[:, -2]
        flag |= (best - second) <= recheck_tol * best
    return flag


class SlicedDecoder:
    """A compiled plan decoded in 2^k slices (see the module docstring).

    ``hip`` is the library plan of the sliced plan (``tensorqec.lib.HipPlan``);
    ``decode`` takes host syndrome rows of the original width, and
    ``decode_device`` device rows.  Precision follows the FP32 executors'
    contract at the level of the summed marginals: in a reduced-precision
    mode, a syndrome whose sum has no positive entry or a top-2 gap within
    ``recheck_tol`` of its maximum has all its slices decoded again in the
    FP64 counterpart mode.  The executors' own per-row margin re-check is
    off for slice rows (``slice_recheck_tol`` = 0: a slice's margin decides
    nothing), while rows that underflowed or have no positive entry are
    still re-checked in FP64 row by row.  ``last_rechecked`` counts the
    syndromes the last call decided again."""

    def __init__(self, plan: Plan, labels: Sequence[int], tabulate_bits: int = 0,
                 slice_recheck_tol: float = 0.0, recheck_tol: float = 1e-3):
        from . import lib as L
        self.labels = list(labels)
        self.k = len(self.labels)
        self.plan = slice_plan(plan, self.labels)
        self.n_syn = plan.n_syndrome_bits
        self.hip = L.HipPlan(self.plan, tabulate_bits=tabulate_bits)
        self.hip.set_tolerances(1e-12, float(slice_recheck_tol))
        self.recheck_tol = float(recheck_tol)
        self.last_rechecked = 0

    def decode(self, syndromes, mode: str = "spec_f32", tie_tol: float = 1e-12):
        S = np.atleast_2d(np.asarray(syndromes, dtype=np.uint8))
        if S.shape[1] != self.n_syn:
            raise ValueError(f"syndrome rows have {S.shape[1]} bits, the plan {self.n_syn}")
        mar, _, _ = self.hip.decode(expand_rows(S, self.k), mode, True)
        tot, cls, tie = combine(mar, len(S), self.k, tie_tol)
        self.last_rechecked = 0
        if mode in _FP64_OF:
            redo = needs_fp64(tot, self.recheck_tol)
            self.last_rechecked = int(redo.sum())
            if redo.any():
                m64, _, _ = self.hip.decode(expand_rows(S[redo], self.k), _FP64_OF[mode], True)
                tot[redo], cls[redo], tie[redo] = combine(m64, self.last_rechecked, self.k, tie_tol)
        return tot, cls, tie

    def _decode_rows(self, torch, d_syn, n, mode, stream):
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
        self.hip.decode_device(ext.data_ptr(), n * m, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), stream)
        return mar.view(n, m, nout).sum(dim=1)

    def decode_device(self, torch, d_syn, n: int, mode: str = "spec_f32", stream: Optional[int] = None,
                      tie_tol: float = 1e-12):
        """Device rows in, device (marginals, classes, ties) out; the slices
        are expanded and summed with torch on the same device and stream as
        the library's launches (the FP64 re-check reads the flagged
        syndromes' count back to the host)."""
        if stream is not None and stream != torch.cuda.current_stream().cuda_stream:
            with torch.cuda.stream(torch.cuda.ExternalStream(stream)):
                return self.decode_device(torch, d_syn, n, mode, None, tie_tol)
        st = torch.cuda.current_stream().cuda_stream
        tot = self._decode_rows(torch, d_syn, n, mode, st)
        self.last_rechecked = 0
        if mode in _FP64_OF and n > 0:
            best = tot.max(dim=1).values
            flag = ~(best > 0)
            if tot.shape[1] > 1:
                second = torch.topk(tot, 2, dim=1).values[:, 1]
                flag |= (best - second) <= self.recheck_tol * best
            idx = torch.nonzero(flag).flatten()
            self.last_rechecked = int(idx.numel())
            if self.last_rechecked:
                sub = d_syn[:n].index_select(0, idx).contiguous()
                tot[idx] = self._decode_rows(torch, sub, self.last_rechecked, _FP64_OF[mode], st)
        nout = tot.shape[1]
        c = torch.argmax(tot, dim=1).to(torch.int32)
        if nout > 1:
            top2 = torch.topk(tot, 2, dim=1).values
            t = ((top2[:, 0] - top2[:, 1]) <= tie_tol * top2[:, 0]).to(torch.uint8)
        else:
            t = torch.zeros(n, dtype=torch.uint8, device=tot.device)
        return tot, c, t
