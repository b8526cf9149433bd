"""Contraction-order search and the flat contraction plan.

The reference asks OMEinsum for a contraction tree
(``optimize_code(code, uniformsize(code, 2), TreeSA())``,
src/decoding/tndecoder.jl:140-144) and evaluates the resulting NestedEinsum.
This module plays that role: it searches a binary contraction tree over the
network (seeded greedy, seeded variable-elimination orders and a seeded
simulated-annealing refinement of the tree by local rotations, in the spirit
of TreeSA) and flattens it into a post-order list of pairwise steps — the
form the HIP library consumes.

All variables are bits, so a label set is a Python int bitmask and the size
of a tensor is 2^popcount.
"""
from __future__ import annotations

import math
import random
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple


def _bits(mask: int) -> List[int]:
    out = []
    while mask:
        low = mask & -mask
        out.append(low.bit_length() - 1)
        mask ^= low
    return out


# population count of a label mask
_pc = int.bit_count


@dataclass
class ContractionTree:
    """Binary tree over ``n_leaves`` leaves.  Node ids: leaves ``0..L-1``,
    internal nodes ``L..2L-2`` in post-order; ``children[i]`` is the pair of
    internal node ``L + i``; ``masks[node]`` its open labels."""
    n_leaves: int
    children: List[Tuple[int, int]]
    masks: List[int]

    @property
    def root(self) -> int:
        return self.n_leaves + len(self.children) - 1 if self.children else 0


def _out_mask(a: int, b: int, count: Dict[int, int], iy: int) -> int:
    """Labels of a|b still needed: in the output, or in another tensor."""
    u = a | b
    keep = u & iy
    for l in _bits(u & ~iy):
        c = count[l] - ((a >> l) & 1) - ((b >> l) & 1)
        if c > 0:
            keep |= 1 << l
    return keep


def tree_cost(tree: ContractionTree) -> Tuple[float, float, int]:
    """``(tc, rw, sc)``: multiply-adds (sum over steps of 2^|La u Lb|),
    element reads+writes (sum of |A|+|B|+|C|) and log2 of the largest
    tensor.  Mirrors ``contraction_complexity`` (tndecoder.jl:273-279)."""
    tc = 0.0
    rw = 0.0
    sc = max((_pc(m) for m in tree.masks[:tree.n_leaves]), default=0)
    for i, (a, b) in enumerate(tree.children):
        la, lb, lc = tree.masks[a], tree.masks[b], tree.masks[tree.n_leaves + i]
        tc += 2.0 ** _pc(la | lb)
        rw += 2.0 ** _pc(la) + 2.0 ** _pc(lb) + 2.0 ** _pc(lc)
        sc = max(sc, _pc(lc))
    return tc, rw, sc


def greedy_tree(ixs: Sequence[Sequence[int]], iy: Sequence[int], seed: int = 0,
                alpha: float = 0.0, temperature: float = 0.0) -> ContractionTree:
    """Greedy pairwise order: contract the adjacent pair minimising
    ``|C| - alpha (|A| + |B|)``, ties (or a Boltzmann draw when
    ``temperature > 0``) broken by the seeded RNG."""
    rng = random.Random(seed)
    L = len(ixs)
    masks = [sum(1 << l for l in set(ix)) for ix in ixs]
    iym = sum(1 << l for l in iy)
    count: Dict[int, int] = {}
    holders: Dict[int, set] = {}
    for t, m in enumerate(masks):
        for l in _bits(m):
            count[l] = count.get(l, 0) + 1
            holders.setdefault(l, set()).add(t)
    active = set(range(L))
    children: List[Tuple[int, int]] = []
    node_of = {t: t for t in range(L)}  # active tensor id -> node id (same ids)
    next_id = L
    while len(active) > 1:
        cand = {}
        for l, hs in holders.items():
            hs2 = sorted(hs)
            for i in range(len(hs2)):
                for j in range(i + 1, len(hs2)):
                    p = (hs2[i], hs2[j])
                    if p in cand:
                        continue
                    a, b = masks[p[0]], masks[p[1]]
                    c = _out_mask(a, b, count, iym)
                    cand[p] = 2.0 ** _pc(c) - alpha * (2.0 ** _pc(a) + 2.0 ** _pc(b))
        if not cand:
            # disconnected components: outer product of the two smallest
            srt = sorted(active, key=lambda t: (_pc(masks[t]), t))
            cand[(srt[0], srt[1])] = 0.0
        if temperature > 0:
            items = sorted(cand.items())
            best_v = min(v for _, v in items)
            ws = [math.exp(-(v - best_v) / (temperature * max(1.0, abs(best_v)))) for _, v in items]
            r = rng.random() * sum(ws)
            acc = 0.0
            pick = items[-1][0]
            for (p, _), w in zip(items, ws):
                acc += w
                if acc >= r:
                    pick = p
                    break
        else:
            best_v = min(cand.values())
            ties = sorted(p for p, v in cand.items() if v == best_v)
            pick = ties[rng.randrange(len(ties))]
        a, b = pick
        ma, mb = masks[a], masks[b]
        mc = _out_mask(ma, mb, count, iym)
        for l in _bits(ma):
            count[l] -= 1
            holders[l].discard(a)
        for l in _bits(mb):
            count[l] -= 1
            holders[l].discard(b)
        for l in list(holders):
            if not holders[l] and count.get(l, 0) == 0:
                del holders[l]
        t = next_id
        next_id += 1
        masks.append(mc)
        for l in _bits(mc):
            count[l] = count.get(l, 0) + 1
            holders.setdefault(l, set()).add(t)
        active.discard(a)
        active.discard(b)
        active.add(t)
        children.append((a, b))
    return _renumber(L, children, masks)


def _renumber(L: int, children: List[Tuple[int, int]], masks: List[int]) -> ContractionTree:
    """Children are already in creation order, which is a valid post-order."""
    return ContractionTree(L, list(children), list(masks))


def elimination_tree(ixs: Sequence[Sequence[int]], iy: Sequence[int], order: Sequence[int]) -> ContractionTree:
    """Bucket elimination: for each variable in ``order`` (skipping output
    variables), merge all tensors holding it pairwise (smallest first) and
    sum it out; finally merge what is left."""
    L = len(ixs)
    masks = [sum(1 << l for l in set(ix)) for ix in ixs]
    iym = sum(1 << l for l in iy)
    count: Dict[int, int] = {}
    for m in masks:
        for l in _bits(m):
            count[l] = count.get(l, 0) + 1
    active = list(range(L))
    children: List[Tuple[int, int]] = []

    def merge(a: int, b: int) -> int:
        ma, mb = masks[a], masks[b]
        mc = _out_mask(ma, mb, count, iym)
        for l in _bits(ma):
            count[l] -= 1
        for l in _bits(mb):
            count[l] -= 1
        for l in _bits(mc):
            count[l] = count.get(l, 0) + 1
        masks.append(mc)
        children.append((a, b))
        return len(masks) - 1

    for v in order:
        if (iym >> v) & 1:
            continue
        hold = [t for t in active if (masks[t] >> v) & 1]
        if len(hold) < 2:
            continue
        rest = [t for t in active if not (masks[t] >> v) & 1]
        hold.sort(key=lambda t: (_pc(masks[t]), t))
        while len(hold) > 1:
            a = hold.pop(0)
            b = hold.pop(0)
            c = merge(a, b)
            hold.append(c)
            hold.sort(key=lambda t: (_pc(masks[t]), t))
        active = rest + hold
    active.sort(key=lambda t: (_pc(masks[t]), t))
    while len(active) > 1:
        a = active.pop(0)
        b = active.pop(0)
        active.append(merge(a, b))
        active.sort(key=lambda t: (_pc(masks[t]), t))
    return ContractionTree(L, children, masks)


def min_fill_order(ixs: Sequence[Sequence[int]], iy: Sequence[int], seed: int = 0) -> List[int]:
    """Greedy min-fill variable order on the interaction graph (ties by
    degree, then the seeded RNG).  Output variables are never eliminated."""
    rng = random.Random(seed)
    adj: Dict[int, set] = {}
    for ix in ixs:
        s = set(ix)
        for u in s:
            adj.setdefault(u, set()).update(s - {u})
    out = set(iy)
    order = []
    remaining = set(v for v in adj if v not in out)
    while remaining:
        best = None
        best_key = None
        for v in remaining:
            nb = list(adj[v])
            fill = 0
            for i in range(len(nb)):
                ai = adj[nb[i]]
                for j in range(i + 1, len(nb)):
                    if nb[j] not in ai:
                        fill += 1
            key = (fill, len(nb), rng.random())
            if best_key is None or key < best_key:
                best, best_key = v, key
        v = best
        nb = adj[v]
        for u in nb:
            adj[u].update(nb - {u})
            adj[u].discard(v)
        del adj[v]
        remaining.discard(v)
        order.append(v)
    return order


def anneal_tree(tree: ContractionTree, iy: Sequence[int], seed: int = 0, sweeps: int = 200,
                sc_target: Optional[int] = None, beta0: float = 1.0, beta1: float = 20.0) -> ContractionTree:
    """Simulated annealing over tree rotations ((A,B),C) -> ((A,C),B) or
    ((C,B),A), accepting by the change of log2(tc) plus a penalty on steps
    wider than ``sc_target``.  Deterministic for a given seed."""
    rng = random.Random(seed)
    L = tree.n_leaves
    iym = sum(1 << l for l in iy)
    # Build explicit parent/children structure.
    left: Dict[int, int] = {}
    right: Dict[int, int] = {}
    for i, (a, b) in enumerate(tree.children):
        left[L + i], right[L + i] = a, b
    root = tree.root
    leafmask = tree.masks[:L]
    # label multiplicity over leaves (plus output)
    total: Dict[int, int] = {}
    for m in leafmask:
        for l in _bits(m):
            total[l] = total.get(l, 0) + 1
    for l in _bits(iym):
        total[l] = total.get(l, 0) + 1

    # subtree label counts: cnt[node][l] would be heavy; keep union and "inner" masks:
    # a label is open at node iff it occurs in the subtree and (in iy or occurs outside).
    # occurs-outside <=> subtree count < total count.
    def sub_counts(node, memo):
        if node < L:
            d = {l: 1 for l in _bits(leafmask[node])}
            memo[node] = d
            return d
        da = sub_counts(left[node], memo)
        db = sub_counts(right[node], memo)
        d = dict(da)
        for l, c in db.items():
            d[l] = d.get(l, 0) + c
        memo[node] = d
        return d

    import sys
    sys.setrecursionlimit(max(10000, sys.getrecursionlimit()))

    def open_mask(d):
        m = 0
        for l, c in d.items():
            if c < total[l]:
                m |= 1 << l
        return m

    memo: Dict[int, Dict[int, int]] = {}
    sub_counts(root, memo)
    omask = {node: open_mask(d) for node, d in memo.items()}

    def step_cost(node):
        a, b = left[node], right[node]
        w = _pc(omask[a] | omask[b])
        c = 2.0 ** w
        if sc_target is not None and _pc(omask[node]) > sc_target:
            c *= 2.0 ** (4 * (_pc(omask[node]) - sc_target))
        return c

    internal = [n for n in left]
    costs = {n: step_cost(n) for n in internal}
    total_cost = sum(costs.values())
    parent: Dict[int, int] = {}
    for n in internal:
        parent[left[n]] = n
        parent[right[n]] = n
    nsteps = sweeps * len(internal)
    for it in range(nsteps):
        beta = beta0 + (beta1 - beta0) * it / max(1, nsteps - 1)
        n = internal[rng.randrange(len(internal))]
        # n = (X, Y); pick a child that is internal: X = (A, B)
        side = rng.randrange(2)
        x = left[n] if side == 0 else right[n]
        y = right[n] if side == 0 else left[n]
        if x < L:
            x, y = y, x
            side = 1 - side
            if x < L:
                continue
        a, b = left[x], right[x]
        # rotate: (a, b), y  ->  (a, y), b   or  (b, y), a
        swap_a = rng.randrange(2)
        keep, move = (a, b) if swap_a == 0 else (b, a)
        # new x' = (keep, y), n' = (x', move)
        dk, dy = memo[keep], memo[y]
        dnew = dict(dk)
        for l, c in dy.items():
            dnew[l] = dnew.get(l, 0) + c
        new_open = open_mask(dnew)
        old = costs[n] + costs[x]
        w1 = _pc(omask[keep] | omask[y])
        c1 = 2.0 ** w1
        if sc_target is not None and _pc(new_open) > sc_target:
            c1 *= 2.0 ** (4 * (_pc(new_open) - sc_target))
        w2 = _pc(new_open | omask[move])
        c2 = 2.0 ** w2
        if sc_target is not None and _pc(omask[n]) > sc_target:
            c2 *= 2.0 ** (4 * (_pc(omask[n]) - sc_target))
        new_total = total_cost - old + c1 + c2
        d = math.log2(new_total) - math.log2(total_cost)
        if d <= 0 or rng.random() < math.exp(-beta * d * 50):
            left[x], right[x] = keep, y
            if side == 0:
                left[n], right[n] = x, move
            else:
                right[n], left[n] = x, move
            parent[keep] = x
            parent[y] = x
            parent[move] = n
            parent[x] = n
            memo[x] = dnew
            omask[x] = new_open
            costs[x] = c1
            costs[n] = c2
            total_cost = new_total
    # rebuild post-order
    children: List[Tuple[int, int]] = []
    masks = list(leafmask)
    newid: Dict[int, int] = {}

    def emit(node):
        stack = [(node, False)]
        while stack:
            nd, done = stack.pop()
            if nd < L:
                newid[nd] = nd
                continue
            if done:
                children.append((newid[left[nd]], newid[right[nd]]))
                masks.append(omask[nd] if nd != root else (omask[nd] & iym) | omask[nd])
                newid[nd] = L + len(children) - 1
            else:
                stack.append((nd, True))
                stack.append((right[nd], False))
                stack.append((left[nd], False))

    emit(root)
    return ContractionTree(L, children, masks)


def sweep_order(ixs: Sequence[Sequence[int]], iy: Sequence[int], primary: Sequence[int]) -> List[int]:
    """Eliminate the ``primary`` variables in the given order; every other
    variable is eliminated as soon as all primary variables of the tensors
    holding it are gone.  For a detector error model with mechanisms sorted
    by their detectors (i.e. by time) this is a sweep through the circuit
    whose state is one partial-parity bit per open detector chain."""
    holders: Dict[int, List[int]] = {}
    for t, ix in enumerate(ixs):
        for l in ix:
            holders.setdefault(l, []).append(t)
    prim = list(primary)
    pset = set(prim)
    out = set(iy)
    rest = [l for l in holders if l not in pset and l not in out]
    need = {a: set(m for t in holders[a] for m in ixs[t] if m in pset) for a in rest}
    waiting: Dict[int, List[int]] = {}
    pos = {m: i for i, m in enumerate(prim)}
    for a, ms in need.items():
        last = max((pos[m] for m in ms), default=-1)
        waiting.setdefault(last, []).append(a)
    order = list(sorted(waiting.get(-1, [])))
    for i, m in enumerate(prim):
        order.append(m)
        order.extend(sorted(waiting.get(i, [])))
    return order


def dissect(n: int, edges: Sequence[Sequence[int]], is_output: Sequence[bool], seed: int = 0,
            balance: float = 0.2, refine: int = 2) -> List[Tuple[int, int]]:
    """Binary dissection tree over items ``0..n-1`` (internal node ``n + i`` is
    ``children[i]``, post-order) for the hypergraph ``edges`` (item lists).
    A part is bisected along the Fiedler vector of its items' graph (items
    sharing an edge joined, an edge of k items weighing 1/(k-1) per pair),
    then improved by ``refine`` passes of single-item moves; the cut
    minimises the larger side's *open* edges — edges with items on both
    sides of that side's boundary, or with ``is_output`` set — with at least
    ``balance`` of the part on each side."""
    import numpy as np
    rng = random.Random(seed)
    item_edges: List[List[int]] = [[] for _ in range(n)]
    for e, it in enumerate(edges):
        for i in set(it):
            item_edges[i].append(e)
    total = [len(set(it)) for it in edges]
    children: List[Tuple[int, int]] = []

    def fiedler(part: List[int]) -> List[int]:
        pos = {t: i for i, t in enumerate(part)}
        m = len(part)
        rows, cols, vals = [], [], []
        seen = set()
        for t in part:
            for e in item_edges[t]:
                if e in seen:
                    continue
                seen.add(e)
                hs = sorted(set(pos[u] for u in edges[e] if u in pos))
                k = len(hs)
                if k < 2:
                    continue
                w = 1.0 / (k - 1)
                for a in range(k):
                    for b in range(a + 1, k):
                        rows += [hs[a], hs[b]]
                        cols += [hs[b], hs[a]]
                        vals += [w, w]
        from scipy.sparse import coo_matrix
        from scipy.sparse.csgraph import connected_components
        W = coo_matrix((vals, (rows, cols)), shape=(m, m)).tocsr()
        jit = np.array([rng.random() for _ in range(m)])
        ncomp, lab = connected_components(W, directed=False)
        if ncomp > 1:
            sizes = np.bincount(lab)
            return [part[i] for i in sorted(range(m), key=lambda i: (-sizes[lab[i]], lab[i], jit[i]))]
        deg = np.asarray(W.sum(axis=1)).ravel()
        if m <= 2000:
            _, vec = np.linalg.eigh(np.diag(deg) - W.toarray())
            f = vec[:, 1]
        else:
            from scipy.sparse import diags
            from scipy.sparse.linalg import eigsh
            _, vec = eigsh(diags(deg) - W, k=2, sigma=-1e-6, which="LM", v0=jit)
            f = vec[:, 1]
        return [part[i] for i in np.lexsort((jit * 1e-9, f))]

    def widths(side_of: Dict[int, int], part: List[int]):
        """Open-edge counts of side 0 and side 1 of the part, and per edge
        the items of the part on each side."""
        cnt: Dict[int, List[int]] = {}
        for t in part:
            s = side_of[t]
            for e in item_edges[t]:
                c = cnt.setdefault(e, [0, 0])
                c[s] += 1
        w = [0, 0]
        for e, c in cnt.items():
            for s in (0, 1):
                if c[s] > 0 and (c[s] < total[e] or is_output[e]):
                    w[s] += 1
        return w, cnt

    def split(part: List[int]) -> Tuple[List[int], List[int]]:
        order = fiedler(part)
        m = len(order)
        lo, hi = max(1, int(math.floor(balance * m))), min(m - 1, int(math.ceil((1 - balance) * m)))
        side_of = {t: 1 for t in order}
        w, cnt = widths(side_of, order)
        best, best_i = None, lo
        for i, t in enumerate(order[:hi], start=1):
            for e in item_edges[t]:
                c = cnt[e]
                for s in (0, 1):
                    w[s] -= c[s] > 0 and (c[s] < total[e] or is_output[e])
                c[1] -= 1
                c[0] += 1
                for s in (0, 1):
                    w[s] += c[s] > 0 and (c[s] < total[e] or is_output[e])
            side_of[t] = 0
            if i >= lo:
                key = (max(w), w[0] + w[1], abs(2 * i - m))
                if best is None or key < best:
                    best, best_i = key, i
        side_of = {t: (0 if j < best_i else 1) for j, t in enumerate(order)}
        # single-item moves that lower (max width, total width), keeping balance
        for _ in range(refine):
            moved = False
            for t in order:
                s = side_of[t]
                nside = sum(1 for u in order if side_of[u] == s)
                if nside - 1 < max(1, int(balance * m)):
                    continue
                w0, _ = widths(side_of, order)
                side_of[t] = 1 - s
                w1, _ = widths(side_of, order)
                if (max(w1), w1[0] + w1[1]) < (max(w0), w0[0] + w0[1]):
                    moved = True
                else:
                    side_of[t] = s
            if not moved:
                break
        left = [t for t in order if side_of[t] == 0]
        right = [t for t in order if side_of[t] == 1]
        return left, right

    def build(part: List[int]) -> int:
        if len(part) == 1:
            return part[0]
        if len(part) == 2:
            a, b = part
        else:
            left, right = split(part)
            a, b = build(left), build(right)
        children.append((a, b))
        return n + len(children) - 1

    import sys
    limit = sys.getrecursionlimit()
    sys.setrecursionlimit(max(limit, 4 * n + 100))
    try:
        if n > 0:
            build(list(range(n)))
    finally:
        sys.setrecursionlimit(limit)
    return children


def search_tree(ixs: Sequence[Sequence[int]], iy: Sequence[int], seed: int = 0, ntrials: int = 8,
                sa_sweeps: int = 0, sc_target: Optional[int] = None,
                objective: str = "tc", primary: Optional[Sequence[int]] = None,
                greedy_limit: int = 1200) -> ContractionTree:
    """Best of several seeded greedy and min-fill elimination trees,
    optionally refined by :func:`anneal_tree`.  ``objective`` ranks trees by
    multiply-adds ("tc") or element traffic ("rw")."""
    cands: List[ContractionTree] = []
    if primary is not None:
        cands.append(elimination_tree(ixs, iy, sweep_order(ixs, iy, primary)))
    for t in range(ntrials if len(ixs) <= greedy_limit else 0):
        cands.append(greedy_tree(ixs, iy, seed=seed * 1000 + t, alpha=0.0, temperature=0.0 if t == 0 else 0.3))
        cands.append(greedy_tree(ixs, iy, seed=seed * 1000 + t, alpha=1.0, temperature=0.0 if t == 0 else 0.3))
        cands.append(elimination_tree(ixs, iy, min_fill_order(ixs, iy, seed=seed * 1000 + t)))

    def key(tr):
        tc, rw, sc = tree_cost(tr)
        pen = 0 if sc_target is None or sc <= sc_target else sc - sc_target
        return (pen, tc if objective == "tc" else rw, sc)

    best = min(cands, key=key)
    if sa_sweeps > 0:
        ref = anneal_tree(best, iy, seed=seed, sweeps=sa_sweeps, sc_target=sc_target)
        if key(ref) < key(best):
            best = ref
    return best


# ----------------------------------------------------------------------------
# Flat plan
# ----------------------------------------------------------------------------

@dataclass
class Plan:
    """Post-order list of pairwise steps over a network.

    ``leaf_labels[t]`` / ``leaf_data[t]`` are the leaves (data may be None for
    one-hot evidence leaves, ``leaf_slot[t] >= 0``).  Step i produces node
    ``n_leaves + i`` from nodes ``lhs[i]`` and ``rhs[i]`` with labels
    ``out_labels[i]`` (its memory order: first label = least significant
    bit).  ``evidence_label_bits`` maps a sliced label to its syndrome bit.
    ``out_labels_final`` is the output order (the reference's ``iy``).
    """
    n_leaves: int
    leaf_labels: List[List[int]]
    leaf_data: List
    leaf_slot: List[int]
    lhs: List[int]
    rhs: List[int]
    out_labels: List[List[int]]
    out_labels_final: List[int]
    n_syndrome_bits: int
    evidence_label_bits: Dict[int, int] = field(default_factory=dict)

    @property
    def n_steps(self) -> int:
        return len(self.lhs)

    def node_labels(self, node: int) -> List[int]:
        return self.leaf_labels[node] if node < self.n_leaves else self.out_labels[node - self.n_leaves]

    def stats(self) -> Dict[str, float]:
        """tc (multiply-adds), rw (elements touched), sc (log2 widest tensor),
        counted over every step, and the same restricted to syndrome-dependent
        steps (the ones executed per syndrome)."""
        dep = self.dependent_nodes()
        tc = rw = tcd = rwd = 0.0
        sc = scd = 0
        for i in range(self.n_steps):
            a, b = self.lhs[i], self.rhs[i]
            la, lb, lc = (len(self._eff(a)), len(self._eff(b)), len(self.out_labels[i]))
            w = len(set(self._eff(a)) | set(self._eff(b)))
            tc += 2.0 ** w
            rw += 2.0 ** la + 2.0 ** lb + 2.0 ** lc
            sc = max(sc, lc)
            if dep[self.n_leaves + i]:
                tcd += 2.0 ** w
                rwd += (2.0 ** la if dep[a] else 0) + (2.0 ** lb if dep[b] else 0) + 2.0 ** lc
                scd = max(scd, lc)
        return {"tc": tc, "rw": rw, "sc": sc, "tc_dep": tcd, "rw_dep": rwd, "sc_dep": scd,
                "steps": self.n_steps, "steps_dep": sum(dep[self.n_leaves:])}

    def _eff(self, node: int) -> List[int]:
        labs = self.node_labels(node)
        if node < self.n_leaves:
            return [l for l in labs if l not in self.evidence_label_bits]
        return labs

    def dependent_nodes(self) -> List[bool]:
        dep = []
        for t in range(self.n_leaves):
            dep.append(self.leaf_slot[t] >= 0 or any(l in self.evidence_label_bits for l in self.leaf_labels[t]))
        for i in range(self.n_steps):
            dep.append(dep[self.lhs[i]] or dep[self.rhs[i]])
        return dep


def _labels_of(mask: int) -> List[int]:
    return _bits(mask)


def flatten(graph, tree: ContractionTree, evidence_as_slices: bool) -> Plan:
    """Turn a tree over ``graph``'s tensors into a :class:`Plan`.

    Intermediate label order: the kept labels of the left operand in its
    order, then the new labels of the right operand in its order; the root
    takes the output order ``graph.iy``.
    """
    L = tree.n_leaves
    leaf_labels = [list(ix) for ix in graph.ixs]
    slot_of_leaf = {t: i for i, t in enumerate(graph.evidence_leaves) if t >= 0}
    leaf_slot = [slot_of_leaf.get(t, -1) for t in range(L)]
    leaf_data = [None if leaf_slot[t] >= 0 else graph.tensors[t] for t in range(L)]
    ev_bits = {}
    if evidence_as_slices:
        ev_bits = {l: i for i, l in enumerate(graph.evidence_labels)}
    node_labels: Dict[int, List[int]] = {t: [l for l in leaf_labels[t] if l not in ev_bits] for t in range(L)}
    lhs, rhs, outs = [], [], []
    root = tree.root
    for i, (a, b) in enumerate(tree.children):
        node = L + i
        keep = tree.masks[node]
        la, lb = node_labels[a], node_labels[b]
        if node == root:
            out = list(graph.iy)
        else:
            out = [l for l in la if (keep >> l) & 1]
            out += [l for l in lb if (keep >> l) & 1 and l not in out]
        node_labels[node] = out
        lhs.append(a)
        rhs.append(b)
        outs.append(out)
    return Plan(L, leaf_labels, leaf_data, leaf_slot, lhs, rhs, outs, list(graph.iy), graph.n_syndrome, ev_bits)


def recompute_masks(tree: ContractionTree, ixs: Sequence[Sequence[int]], iy: Sequence[int]) -> List[int]:
    """Open-label masks of every node computed from scratch (a label is open
    at a node iff it is an output label or also occurs outside the subtree)."""
    L = tree.n_leaves
    iym = sum(1 << l for l in iy)
    total: Dict[int, int] = {}
    leafm = [sum(1 << l for l in set(ix)) for ix in ixs]
    for m in leafm:
        for l in _bits(m):
            total[l] = total.get(l, 0) + 1
    sub: List[Dict[int, int]] = [{l: 1 for l in _bits(m)} for m in leafm]
    masks = list(leafm)
    for (a, b) in tree.children:
        d = dict(sub[a])
        for l, c in sub[b].items():
            d[l] = d.get(l, 0) + c
        sub.append(d)
        m = 0
        for l, c in d.items():
            if (iym >> l) & 1 or c < total[l]:
                m |= 1 << l
        masks.append(m)
    return masks


def check_tree(tree: ContractionTree, ixs: Sequence[Sequence[int]], iy: Sequence[int]) -> None:
    """Raise if ``tree`` is not a full binary tree over all leaves or if its
    masks disagree with a from-scratch recomputation."""
    L = tree.n_leaves
    if len(tree.children) != max(0, L - 1):
        raise ValueError("tree must have L-1 steps")
    used = [0] * (L + len(tree.children))
    for i, (a, b) in enumerate(tree.children):
        for c in (a, b):
            if c >= L + i or c < 0:
                raise ValueError("child refers to a later node")
            used[c] += 1
    if any(u != 1 for u in used[:-1]) or (L > 1 and used[-1] != 0):
        raise ValueError("every node except the root must be used exactly once")
    if recompute_masks(tree, ixs, iy) != list(tree.masks):
        raise ValueError("tree label masks are inconsistent")
