"""CPU oracle for the TNMMAP decoding path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module, and only as the checker.  The product (the HIP library and its
host package) never imports it.

It restates, in FP64 numpy, what the reference computes for one syndrome:

* the TNMMAP factor graph of a CSS code — src/decoding/tndecoder.jl:97-139
  (variable order :94-96, parity tensor :42-44, qubit prior :118-121,
  evidence vectors :136-139);
* ``update_syndrome!`` — tndecoder.jl:148-158 (one-hot ``[1,0]``/``[0,1]``);
* the contraction ``ct.code(ct.tensors...)`` — tndecoder.jl:162, here with
  the oracle's own greedy pairwise order (independent of the product's
  planner), batched over syndromes by a batch label on the evidence tensors;
* ``findmax`` — tndecoder.jl:163 (first maximum, column-major);
* the logical fix of ``error_pattern`` — tndecoder.jl:167-174;
* the DEM network — tndecoder.jl:186-238 and the sampler semantics of
  src/decoding/error_model.jl:97-117.

and an exact brute-force enumerator of P(syndrome, logical class), which
pins the network and contraction without relying on either.

Pinning (see tests/test_oracle.py): the reference's marginal golden
(test/decoding/tndecoder.jl:109), brute-force enumeration over every error
pattern for d=3 surface, Steane, Shor and ToricCode(2,2), total probability
one, structural zeros, and agreement between two contraction orders.
"""
from __future__ import annotations

import itertools
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_LETTERS = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"


def parity_tensor(w: int) -> np.ndarray:
    """tndecoder.jl:42-44: shape (2,)*(w+1), entry 1 iff the index sum is even."""
    idx = np.indices((2,) * (w + 1)).sum(axis=0)
    return (1 - (idx % 2)).astype(np.float64)


def css_network(Hx, Hz, lx, lz, px, py, pz):
    """TNMMAP CSS network of tndecoder.jl:97-139 (0-based variables).

    Returns (ixs, tensors, iy, syndrome_vars); tensors are numpy arrays with
    one axis per label, evidence tensors start as [1, 0].
    """
    Hx, Hz, lx, lz = (np.asarray(a, dtype=np.uint8) for a in (Hx, Hz, lx, lz))
    n = Hx.shape[1]
    nsx, nsz, k = Hx.shape[0], Hz.shape[0], lx.shape[0]
    ns = nsx + nsz
    ixs, ts = [], []
    for i in range(nsx):
        c = [int(q) for q in np.flatnonzero(Hx[i])]
        ixs.append([q + n for q in c] + [2 * n + i])
        ts.append(parity_tensor(len(c)))
    for i in range(nsz):
        c = [int(q) for q in np.flatnonzero(Hz[i])]
        ixs.append(c + [2 * n + nsx + i])
        ts.append(parity_tensor(len(c)))
    for i in range(n):
        ixs.append([i, i + n])
        ts.append(np.array([[1 - px[i] - py[i] - pz[i], pz[i]], [px[i], py[i]]], dtype=np.float64))
    for i in range(k):
        c = [int(q) for q in np.flatnonzero(lz[i])]
        ixs.append(c + [2 * n + ns + k + i])
        ts.append(parity_tensor(len(c)))
    for i in range(k):
        c = [int(q) + n for q in np.flatnonzero(lx[i])]
        ixs.append(c + [2 * n + ns + i])
        ts.append(parity_tensor(len(c)))
    svars = list(range(2 * n, 2 * n + ns))
    for v in svars:
        ixs.append([v])
        ts.append(np.array([1.0, 0.0]))
    iy = list(range(2 * n + ns, 2 * n + ns + 2 * k))
    return ixs, ts, iy, svars


def _pair(a, la, b, lb, keep):
    """einsum of two tensors; keep = output label list."""
    labs = {}
    for l in list(la) + list(lb) + list(keep):
        if l not in labs:
            labs[l] = _LETTERS[len(labs)]
    sa = "".join(labs[l] for l in la)
    sb = "".join(labs[l] for l in lb)
    sc = "".join(labs[l] for l in keep)
    return np.einsum(f"{sa},{sb}->{sc}", a, b)


def contract(ixs: Sequence[Sequence], tensors: Sequence[np.ndarray], iy: Sequence,
             order: str = "greedy") -> np.ndarray:
    """Contract a network pairwise.  ``order="greedy"``: repeatedly merge the
    connected pair with the smallest result (first pair on ties);
    ``order="greedy_last"``: the same with the last pair on ties, which gives
    a different tree; ``order="sequential"``: left to right (small networks
    only — disjoint neighbours become outer products)."""
    items = [(list(l), t) for l, t in zip(ixs, tensors)]
    iyl = list(iy)

    def needed(l, skip):
        if l in iyl:
            return True
        return any(l in items[j][0] for j in range(len(items)) if j not in skip)

    while len(items) > 1:
        if order == "sequential":
            i, j = 0, 1
        else:
            best = None
            for i0 in range(len(items)):
                for j0 in range(i0 + 1, len(items)):
                    la, lb = items[i0][0], items[j0][0]
                    if not set(la) & set(lb):
                        continue
                    out = [l for l in dict.fromkeys(la + lb) if needed(l, (i0, j0))]
                    size = len(out)
                    if best is None or size < best[0] or (order == "greedy_last" and size == best[0]):
                        best = (size, i0, j0)
            if best is None:
                i, j = 0, 1
            else:
                _, i, j = best
        la, a = items[i]
        lb, b = items[j]
        out = [l for l in dict.fromkeys(la + lb) if needed(l, (i, j))]
        c = _pair(a, la, b, lb, out)
        items = [items[q] for q in range(len(items)) if q not in (i, j)] + [(out, c)]
    l, t = items[0]
    if set(l) != set(iyl):
        t = _pair(t, l, np.ones(()), [], [x for x in iyl])
        l = iyl
    perm = [l.index(x) for x in iyl]
    return np.transpose(t, perm)


def tnmmap_marginals(Hx, Hz, lx, lz, px, py, pz, syndromes: np.ndarray, order: str = "greedy") -> np.ndarray:
    """P(syndrome, class) for a batch, shape (B, 2^(2k)), classes in
    column-major order of iy (the flat index of a Julia array).

    The batch is a label ``"b"`` on every evidence tensor (tndecoder.jl:148-158
    applied to B syndromes at once)."""
    ixs, ts, iy, svars = css_network(Hx, Hz, lx, lz, px, py, pz)
    S = np.atleast_2d(np.asarray(syndromes, dtype=np.uint8))
    B = S.shape[0]
    ne = len(svars)
    for i in range(ne):
        e = np.zeros((B, 2))
        e[np.arange(B), S[:, i]] = 1.0
        ts[len(ts) - ne + i] = e
        ixs[len(ixs) - ne + i] = ["b", svars[i]]
    t = contract(ixs, ts, ["b"] + iy, order=order)
    # column-major flattening of the class axes
    nout = len(iy)
    return np.transpose(t, [0] + list(range(nout, 0, -1))).reshape(B, 1 << nout)


def findmax(mar: np.ndarray) -> np.ndarray:
    """Index of the first maximum of each row (Julia ``findmax`` on the
    column-major flattened marginal)."""
    return np.argmax(mar, axis=1).astype(np.int32)


def tie_flags(mar: np.ndarray, tol: float) -> np.ndarray:
    s = np.sort(mar, axis=1)
    if mar.shape[1] < 2:
        return np.zeros(mar.shape[0], dtype=np.uint8)
    return ((s[:, -1] - s[:, -2]) <= tol * s[:, -1]).astype(np.uint8)


def brute_force_css(Hx, Hz, lx, lz, px, py, pz) -> Dict[Tuple[int, ...], np.ndarray]:
    """Exact P(syndrome, class) by enumerating all 4^n Pauli errors.

    The class of an error (ex, ez) is the column-major index of
    (lx_1.ez, ..., lx_k.ez, lz_1.ex, ..., lz_k.ex) — the iy axes of the
    network.  Returns {syndrome tuple: vector over classes}."""
    Hx, Hz, lx, lz = (np.asarray(a, dtype=np.int64) for a in (Hx, Hz, lx, lz))
    n = Hx.shape[1]
    k = lx.shape[0]
    if n > 10:
        raise ValueError("brute force limited to n <= 10")
    paulis = np.array(list(itertools.product(range(4), repeat=n)), dtype=np.int64)  # 0=I,1=X,2=Y,3=Z
    ex = ((paulis == 1) | (paulis == 2)).astype(np.int64)
    ez = ((paulis == 2) | (paulis == 3)).astype(np.int64)
    probs = np.ones(len(paulis))
    P = np.stack([1 - np.asarray(px) - np.asarray(py) - np.asarray(pz), px, py, pz], axis=1)
    for q in range(n):
        probs *= P[q][paulis[:, q]]
    sx = (ez @ Hx.T) & 1
    sz = (ex @ Hz.T) & 1
    cls = np.zeros(len(paulis), dtype=np.int64)
    for i in range(k):
        cls |= ((ez @ lx[i]) & 1) << i
        cls |= ((ex @ lz[i]) & 1) << (k + i)
    out: Dict[Tuple[int, ...], np.ndarray] = {}
    keys = np.concatenate([sx, sz], axis=1)
    for key, c, p in zip(map(tuple, keys), cls, probs):
        if key not in out:
            out[key] = np.zeros(1 << (2 * k))
        out[key][c] += p
    return out


def logical_fix(ex, ez, lx, lz, cls: int):
    """tndecoder.jl:169-172 for one syndrome: make the error pattern's
    logical parities equal the decided class."""
    ex = np.array(ex, dtype=np.uint8)
    ez = np.array(ez, dtype=np.uint8)
    k = lx.shape[0]
    for i in range(k):
        if (int(lz[i].astype(np.int64) @ ex) & 1) != ((cls >> (i + k)) & 1):
            ex ^= lx[i].astype(np.uint8)
        if (int(lx[i].astype(np.int64) @ ez) & 1) != ((cls >> i) & 1):
            ez ^= lz[i].astype(np.uint8)
    return ex, ez


# ---------------------------------------------------------------------------
# Detector error models (tndecoder.jl:186-238)
# ---------------------------------------------------------------------------

def parse_dem_text(text: str):
    """stim_parser.jl:342-375: error rates, 1-based detector lists,
    logicals offset past the largest detector."""
    rates, dets, logs = [], [], []
    for line in text.split("\n"):
        if not line:
            continue
        s = line.split()
        if s[0].startswith("error("):
            rates.append(float(s[0][6:-1]))
            d, l = [], []
            for tok in s[1:]:
                if tok.startswith("D"):
                    d.append(int(tok[1:]) + 1)
                if tok.startswith("L"):
                    l.append(int(tok[1:]) + 1)
            dets.append(d)
            logs.append(l)
        elif s[0].startswith("repeat"):
            raise ValueError("repeat blocks are not supported")
    largest = max(max(d) for d in dets if d)
    logs = [[x + largest for x in l] for l in logs]
    largest_l = max([max(l) for l in logs if l] + [0])
    flipped = [sorted(set(d) | set(l)) for d, l in zip(dets, logs)]
    return rates, flipped, list(range(1, largest + 1)), list(range(largest + 1, largest_l + 1))


def brute_force_dem(rates, flipped, detector_list, logical_list) -> Dict[Tuple[int, ...], np.ndarray]:
    """Exact P(detectors, logical class) over every mechanism subset."""
    m = len(rates)
    if m > 22:
        raise ValueError("brute force limited to 22 mechanisms")
    nd, nl = len(detector_list), len(logical_list)
    Dm = np.zeros((m, nd), dtype=np.int64)
    Lm = np.zeros((m, nl), dtype=np.int64)
    for i, f in enumerate(flipped):
        for x in f:
            if x in detector_list:
                Dm[i, detector_list.index(x)] = 1
            if x in logical_list:
                Lm[i, logical_list.index(x)] = 1
    out: Dict[Tuple[int, ...], np.ndarray] = {}
    E = ((np.arange(1 << m)[:, None] >> np.arange(m)[None, :]) & 1).astype(np.int64)
    r = np.asarray(rates)
    pr = np.prod(np.where(E == 1, r[None, :], 1 - r[None, :]), axis=1)
    syn = (E @ Dm) & 1
    lg = (E @ Lm) & 1
    cls = (lg * (1 << np.arange(nl))[None, :]).sum(axis=1)
    for key, c, p in zip(map(tuple, syn), cls, pr):
        if key not in out:
            out[key] = np.zeros(1 << nl)
        out[key][c] += p
    return out


def dem_network(rates, flipped, detector_list, logical_list, factorize: bool):
    """tndecoder.jl:186-238 with dem2tanner (dem.jl:150-160)."""
    nq = len(rates)
    ns = max(detector_list)
    q2s = [[x for x in f if x not in logical_list] for f in flipped]
    s2q = [[j for j in range(nq) if (i + 1) in q2s[j]] for i in range(ns)]
    l2q = [[j for j in range(nq) if L in flipped[j]] for L in logical_list]
    nvars = nq + ns + len(logical_list)
    ixs, ts = [], []

    def push(c, node, nvars):
        w = len(c)
        if not factorize or w == 2:
            ixs.append(list(c) + [node])
            ts.append(parity_tensor(w))
            return nvars
        ixs.append([c[0], c[1], nvars])
        ts.append(parity_tensor(2))
        for j in range(2, w - 1):
            ixs.append([nvars + j - 1, c[j], nvars + j - 2])
            ts.append(parity_tensor(2))
        ixs.append([node, c[w - 1], nvars + w - 3])
        ts.append(parity_tensor(2))
        return nvars + w - 2

    for i, c in enumerate(s2q):
        nvars = push(c, nq + i, nvars)
    for i, c in enumerate(l2q):
        nvars = push(c, nq + ns + i, nvars)
    for i, p in enumerate(rates):
        ixs.append([i])
        ts.append(np.array([1 - p, p]))
    svars = list(range(nq, nq + ns))
    for v in svars:
        ixs.append([v])
        ts.append(np.array([1.0, 0.0]))
    iy = list(range(nq + ns, nq + ns + len(logical_list)))
    return ixs, ts, iy, svars


def dem_marginals(rates, flipped, detector_list, logical_list, syndromes, factorize=True):
    ixs, ts, iy, svars = dem_network(rates, flipped, detector_list, logical_list, factorize)
    S = np.atleast_2d(np.asarray(syndromes, dtype=np.uint8))
    B = S.shape[0]
    ne = len(svars)
    for i in range(ne):
        e = np.zeros((B, 2))
        e[np.arange(B), S[:, i]] = 1.0
        ts[len(ts) - ne + i] = e
        ixs[len(ixs) - ne + i] = ["b", svars[i]]
    t = contract(ixs, ts, ["b"] + iy)
    nout = len(iy)
    return np.transpose(t, [0] + list(range(nout, 0, -1))).reshape(B, 1 << nout)


def brute_force_map_css(Hx, Hz, px, py, pz):
    """Exact maximum of P(ex, ez) over error patterns with a given syndrome
    (the quantity TNMAP maximises, tndecoder.jl:53-57 on the network of
    stg2uaimodel, :33-40): {syndrome tuple: max probability}."""
    Hx, Hz = np.asarray(Hx, dtype=np.int64), np.asarray(Hz, dtype=np.int64)
    n = Hx.shape[1]
    if n > 10:
        raise ValueError("brute force limited to n <= 10")
    paulis = np.array(list(itertools.product(range(4), repeat=n)), dtype=np.int64)
    ex = ((paulis == 1) | (paulis == 2)).astype(np.int64)
    ez = ((paulis == 2) | (paulis == 3)).astype(np.int64)
    P = np.stack([1 - np.asarray(px) - np.asarray(py) - np.asarray(pz), px, py, pz], axis=1)
    probs = np.ones(len(paulis))
    for q in range(n):
        probs *= P[q][paulis[:, q]]
    keys = np.concatenate([(ez @ Hx.T) & 1, (ex @ Hz.T) & 1], axis=1)
    out = {}
    for key, pr in zip(map(tuple, keys), probs):
        if pr > out.get(key, -1.0):
            out[key] = pr
    return out


def css_config_probability(ex, ez, px, py, pz):
    """P(ex, ez) under independent depolarizing noise (rows are patterns)."""
    ex = np.atleast_2d(ex).astype(bool)
    ez = np.atleast_2d(ez).astype(bool)
    px, py, pz = (np.asarray(a, dtype=np.float64) for a in (px, py, pz))
    p_i = 1 - px - py - pz
    val = np.where(ex & ez, py, np.where(ex, px, np.where(ez, pz, p_i)))
    return np.prod(val, axis=1)
