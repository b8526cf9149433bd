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
