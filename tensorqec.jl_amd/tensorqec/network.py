"""Factor graphs of the TNMMAP decoder.

Every variable is a bit (size 2).  A tensor of rank r is a float64 vector of
length 2^r whose flat index has bit j equal to the value of its j-th label —
the column-major layout of a Julia ``Array{Float64}`` with all sizes 2, so a
tensor built by the reference can be handed to the library unchanged.

Reference: src/decoding/tndecoder.jl:42-44 (parity tensor), :94-146 (CSS
network and variable order), :186-238 (DEM network and check factorization).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .gf2 import logical_operator


def parity_check_tensor(w: int) -> np.ndarray:
    """Even-parity indicator on w+1 bits: ``t[i] = 1 - popcount(i) mod 2``
    (tndecoder.jl:42-44)."""
    i = np.arange(1 << (w + 1), dtype=np.uint64)
    pc = np.zeros_like(i)
    for b in range(w + 1):
        pc += (i >> np.uint64(b)) & np.uint64(1)
    return (1 - (pc & np.uint64(1))).astype(np.float64)


def qubit_prior(px: float, py: float, pz: float) -> np.ndarray:
    """``[1-px-py-pz pz; px py]`` indexed ``[x, z]``, flattened column-major:
    entry (x, z) at x + 2 z (tndecoder.jl:118-121)."""
    return np.array([1.0 - px - py - pz, px, pz, py], dtype=np.float64)


@dataclass
class FactorGraph:
    """A tensor network with evidence.

    ``ixs[t]`` labels of tensor t (0-based variable ids), ``tensors[t]`` its
    data, ``iy`` the open output labels.  ``evidence_labels[i]`` is the
    variable fixed by syndrome bit i; ``evidence_leaves[i]`` is the index of
    the one-hot evidence tensor ``[1, 0]`` attached to it (or -1 when the
    network was built with evidence folded into the labels).
    """
    nvars: int
    ixs: List[List[int]]
    tensors: List[np.ndarray]
    iy: List[int]
    evidence_labels: List[int]
    evidence_leaves: List[int]
    meta: Dict = field(default_factory=dict)

    def copy(self) -> "FactorGraph":
        return FactorGraph(self.nvars, [list(x) for x in self.ixs], [t.copy() for t in self.tensors],
                           list(self.iy), list(self.evidence_labels), list(self.evidence_leaves), dict(self.meta))

    @property
    def n_syndrome(self) -> int:
        return len(self.evidence_labels)

    def without_evidence_leaves(self) -> "FactorGraph":
        """Drop the one-hot evidence tensors; the evidence variables stay as
        labels of the tensors that carry them and are fixed per syndrome by
        slicing (contracting a one-hot vector over a label selects a slice)."""
        drop = set(i for i in self.evidence_leaves if i >= 0)
        keep = [t for t in range(len(self.ixs)) if t not in drop]
        g = FactorGraph(self.nvars, [list(self.ixs[t]) for t in keep], [self.tensors[t].copy() for t in keep],
                        list(self.iy), list(self.evidence_labels), [-1] * len(self.evidence_labels), dict(self.meta))
        return g


def _push_check(ixs, tensors, members: List[int], check_var: int, nvars: int, factorize: bool) -> int:
    """Parity constraint on ``members`` + ``check_var``; with ``factorize`` a
    check of 3+ members becomes a chain of rank-3 parity tensors linked by new
    auxiliary variables (tndecoder.jl:221-238).  Returns the new ``nvars``."""
    c = list(members)
    w = len(c)
    if not factorize or w <= 2:
        ixs.append(c + [check_var])
        tensors.append(parity_check_tensor(w))
        return nvars
    p2 = parity_check_tensor(2)
    ixs.append([c[0], c[1], nvars])
    tensors.append(p2.copy())
    for j in range(2, w - 1):
        ixs.append([nvars + j - 1, c[j], nvars + j - 2])
        tensors.append(p2.copy())
    ixs.append([check_var, c[w - 1], nvars + w - 3])
    tensors.append(p2.copy())
    return nvars + w - 2


def dissected_parity_network(nvars: int, units: List[List[int]], priors: List[np.ndarray],
                             checks: List[tuple], iy: List[int], ev_labels: List[int], meta: Dict,
                             seed: int = 0):
    """The same joint distribution as a parity-check network, rewritten
    along a dissection of its error units, with the contraction tree that
    follows it.

    ``units[u]`` are the variables of error unit u (a DEM mechanism, or the
    x- and z-error of a qubit) and ``priors[u]`` their joint prior; each
    check ``(members, closing)`` asks that the XOR of its member variables
    equal the ``closing`` variable (a syndrome bit, fixed by evidence, or a
    logical parity in ``iy``).  The units are bisected recursively
    (``planner.dissect``, cutting as few checks as it can); for a subtree S
    and a check c with units inside and outside S, one auxiliary label holds
    the XOR of c's members inside S, and where two subtrees meet one rank-3
    parity tensor [xor in S1, xor in S2, xor in S1+S2] (the ``closing``
    variable once S1+S2 holds all of c's units) joins them — the
    reference's chain of rank-3 parity tensors (tndecoder.jl:221-238)
    arranged as a tree instead of a path.  Every subtree's tensor then has
    one label per check it cuts.

    Returns ``(graph, tree)``: the FactorGraph (evidence leaves last, as the
    reference networks) and the ContractionTree over its non-evidence
    tensors."""
    from .planner import ContractionTree, dissect, recompute_masks
    var_unit = {v: u for u, vs in enumerate(units) for v in vs}
    members = [list(m) for m, _ in checks]
    closing = [c for _, c in checks]
    out = set(iy)
    cu = [sorted(set(var_unit[v] for v in m)) for m in members]
    total = [len(x) for x in cu]
    checks_of_unit: List[List[int]] = [[] for _ in units]
    for c, us in enumerate(cu):
        for u in us:
            checks_of_unit[u].append(c)
    dchildren = dissect(len(units), cu, [closing[c] in out for c in range(len(checks))], seed=seed)
    ixs: List[List[int]] = [list(vs) for vs in units]
    tensors: List[np.ndarray] = [np.asarray(p, dtype=np.float64) for p in priors]
    steps: List[tuple] = []          # (a, b) over ('t', tensor) / ('n', step) references
    nv = [nvars]

    def tensor(labels, data):
        ixs.append(list(labels))
        tensors.append(data)
        return ("t", len(ixs) - 1)

    def step(a, b):
        steps.append((a, b))
        return ("n", len(steps) - 1)

    def new_label():
        nv[0] += 1
        return nv[0] - 1

    state = {}   # dissection node -> (tree node, {check: partial label}, {check: units inside})
    U = len(units)
    for u in range(U):
        cur = ("t", u)
        part, cnt = {}, {}
        for c in checks_of_unit[u]:
            vs = [v for v in members[c] if var_unit[v] == u]
            if len(vs) == 1:
                lab = vs[0]
            else:
                lab = new_label()
                cur = step(cur, tensor(vs + [lab], parity_check_tensor(len(vs))))
            if total[c] == 1:
                cur = step(cur, tensor([lab, closing[c]], parity_check_tensor(1)))
            else:
                part[c] = lab
                cnt[c] = 1
        state[u] = (cur, part, cnt)
    for i, (a, b) in enumerate(dchildren):
        na, pa, ca = state.pop(a)
        nb, pb, cb = state.pop(b)
        common = sorted(set(pa) & set(pb))
        # join the checks on the side with fewer open labels, then contract
        if len(pa) <= len(pb):
            x, px, y = na, pa, nb
        else:
            x, px, y = nb, pb, na
        part = {c: l for c, l in list(pa.items()) + list(pb.items()) if c not in common}
        cnt = {c: ca.get(c, 0) + cb.get(c, 0) for c in set(ca) | set(cb)}
        for c in common:
            done = cnt[c] == total[c]
            lab = closing[c] if done else new_label()
            x = step(x, tensor([pa[c], pb[c], lab], parity_check_tensor(2)))
            if done:
                del cnt[c]
            else:
                part[c] = lab
        state[U + i] = (step(y, x), part, cnt)
    (root, _, _), = state.values() if state else [(("t", 0), {}, {})]
    for c in range(len(checks)):
        if total[c] == 0:
            root = step(root, tensor([closing[c]], parity_check_tensor(0)))
    L = len(ixs)
    children = [tuple(r[1] if r[0] == "t" else L + r[1] for r in s) for s in steps]
    tree = ContractionTree(L, children, [0] * (L + len(children)))
    ev_leaves = []
    for v in ev_labels:
        ev_leaves.append(len(ixs))
        ixs.append([v])
        tensors.append(np.array([1.0, 0.0]))
    g = FactorGraph(nv[0], ixs, tensors, list(iy), list(ev_labels), ev_leaves, dict(meta))
    ev = set(ev_labels)
    tree.masks = recompute_masks(tree, [[l for l in ix if l not in ev] for ix in ixs[:L]], iy)
    return g, tree


def dissected_dem_network(dem, seed: int = 0):
    """:func:`dissected_parity_network` of a detector error model: variables
    numbered as in tndecoder.jl:186-219 (mechanisms, detectors, logicals,
    then auxiliary labels), one unit per mechanism with prior ``[1-p, p]``,
    one check per detector (closed by its syndrome bit) and per logical
    (closed by its output label)."""
    tanner = dem.tanner()
    nq, ns = tanner.nq, tanner.ns
    l2q = dem.l2q()
    nl = len(l2q)
    checks = [(list(c), nq + i) for i, c in enumerate(tanner.s2q)] + \
             [(list(c), nq + ns + i) for i, c in enumerate(l2q)]
    priors = [np.array([1.0 - p, p]) for p in dem.error_rates]
    return dissected_parity_network(nq + ns + nl, [[m] for m in range(nq)], priors, checks,
                                    list(range(nq + ns, nq + ns + nl)), list(range(nq, nq + ns)),
                                    {"kind": "dem", "nq": nq, "ns": ns, "nl": nl, "l2q": l2q, "factorize": True,
                                     "dissected": True}, seed=seed)


def dissected_css_network(tanner, em, lxlz=None, seed: int = 0):
    """:func:`dissected_parity_network` of a CSS code under independent
    depolarizing noise: variables numbered as in tndecoder.jl:94-96, one unit
    per qubit (its x- and z-error under the prior ``[1-px-py-pz pz; px
    py]``), the X checks on z-errors and Z checks on x-errors closed by their
    syndrome bits, and the lz / lx parities closed by the output labels."""
    n = tanner.nq
    lx, lz = logical_operator(tanner) if lxlz is None else lxlz
    k = lx.shape[0]
    nsx, nsz = tanner.stgx.ns, tanner.stgz.ns
    ns = nsx + nsz
    checks = [([q + n for q in c], 2 * n + i) for i, c in enumerate(tanner.stgx.s2q)] + \
             [(list(c), 2 * n + nsx + i) for i, c in enumerate(tanner.stgz.s2q)] + \
             [([int(q) for q in np.flatnonzero(lz[i])], 2 * n + ns + k + i) for i in range(k)] + \
             [([int(q) + n for q in np.flatnonzero(lx[i])], 2 * n + ns + i) for i in range(k)]
    priors = [qubit_prior(em.px[q], em.py[q], em.pz[q]) for q in range(n)]
    return dissected_parity_network(2 * n + ns + 2 * k, [[q, q + n] for q in range(n)], priors, checks,
                                    list(range(2 * n + ns, 2 * n + ns + 2 * k)), list(range(2 * n, 2 * n + ns)),
                                    {"kind": "css", "n": n, "k": k, "nsx": nsx, "nsz": nsz, "lx": lx, "lz": lz,
                                     "factorize": True, "dissected": True}, seed=seed)


def tnmmap_css_network(tanner, em, factorize: bool = False, lxlz=None) -> FactorGraph:
    """TNMMAP network for a CSS code under independent depolarizing noise.

    Variable order (0-based): x-errors [0, n), z-errors [n, 2n), X-syndromes,
    Z-syndromes, logical-X parities, logical-Z parities (tndecoder.jl:94-96).
    Tensors in reference order: X checks on z-error vars, Z checks on x-error
    vars, qubit priors, lz parities on x-errors, lx parities on z-errors,
    evidence vectors (tndecoder.jl:106-139).  ``factorize=False`` reproduces
    the reference network exactly; ``factorize=True`` applies the rank-3
    chain factorization the reference uses for DEMs to every check and
    logical parity (an exact rewriting of the same distribution).
    """
    n = tanner.nq
    lx, lz = logical_operator(tanner) if lxlz is None else lxlz
    k = lx.shape[0]
    nsx, nsz = tanner.stgx.ns, tanner.stgz.ns
    ns = nsx + nsz
    nvars = 2 * n + ns + 2 * k
    ixs: List[List[int]] = []
    tensors: List[np.ndarray] = []
    for i, c in enumerate(tanner.stgx.s2q):
        nvars = _push_check(ixs, tensors, [q + n for q in c], i + 2 * n, nvars, factorize)
    for i, c in enumerate(tanner.stgz.s2q):
        nvars = _push_check(ixs, tensors, list(c), i + 2 * n + nsx, nvars, factorize)
    for i in range(n):
        ixs.append([i, i + n])
        tensors.append(qubit_prior(em.px[i], em.py[i], em.pz[i]))
    for i in range(k):
        sup = [int(q) for q in np.flatnonzero(lz[i])]
        nvars = _push_check(ixs, tensors, sup, i + 2 * n + ns + k, nvars, factorize)
    for i in range(k):
        sup = [int(q) + n for q in np.flatnonzero(lx[i])]
        nvars = _push_check(ixs, tensors, sup, i + 2 * n + ns, nvars, factorize)
    ev_labels = list(range(2 * n, 2 * n + ns))
    ev_leaves = []
    for v in ev_labels:
        ev_leaves.append(len(ixs))
        ixs.append([v])
        tensors.append(np.array([1.0, 0.0]))
    iy = list(range(2 * n + ns, 2 * n + ns + 2 * k))
    return FactorGraph(nvars, ixs, tensors, iy, ev_labels, ev_leaves,
                       {"kind": "css", "n": n, "k": k, "nsx": nsx, "nsz": nsz, "lx": lx, "lz": lz,
                        "factorize": factorize})


def tnmmap_dem_network(dem, factorize: bool = True) -> FactorGraph:
    """TNMMAP network for a detector error model (tndecoder.jl:186-219).

    Variable order: mechanisms, detectors, logicals, then auxiliary chain
    variables.  Checks (detectors, then logicals) come first, then the
    ``[1-p, p]`` priors, then the evidence vectors.
    """
    tanner = dem.tanner()
    nq, ns = tanner.nq, tanner.ns
    nl = len(dem.logical_list)
    l2q = dem.l2q()
    nvars = nq + ns + nl
    iy = list(range(nq + ns, nvars))
    ixs: List[List[int]] = []
    tensors: List[np.ndarray] = []
    for i, c in enumerate(tanner.s2q):
        if factorize and len(c) < 2:
            raise ValueError(f"detector {i + 1} has {len(c)} member(s); the rank-3 chain needs at least 2 "
                             "(the same restriction as tndecoder.jl:227)")
        nvars = _push_check(ixs, tensors, list(c), i + nq, nvars, factorize)
    for i, c in enumerate(l2q):
        if factorize and len(c) < 2:
            raise ValueError(f"logical {i + 1} has {len(c)} member(s); the rank-3 chain needs at least 2")
        nvars = _push_check(ixs, tensors, list(c), i + nq + ns, nvars, factorize)
    for i, p in enumerate(dem.error_rates):
        ixs.append([i])
        tensors.append(np.array([1.0 - p, p]))
    ev_labels = list(range(nq, nq + ns))
    ev_leaves = []
    for v in ev_labels:
        ev_leaves.append(len(ixs))
        ixs.append([v])
        tensors.append(np.array([1.0, 0.0]))
    return FactorGraph(nvars, ixs, tensors, iy, ev_labels, ev_leaves,
                       {"kind": "dem", "nq": nq, "ns": ns, "nl": nl, "l2q": l2q, "factorize": factorize})
