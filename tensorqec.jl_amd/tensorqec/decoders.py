"""TNMMAP decoding through the HIP library, behind the reference's
``decode(compile(decoder, problem), syndrome)`` interface.

Reference: src/decoding/interfaces.jl:1-161 (problem types, compile/decode
dispatch, DecodingResult), src/decoding/tndecoder.jl:68-174 (TNMMAP for CSS
codes), :176-271 (TNMMAP for detector error models).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional, Tuple

import numpy as np

from . import lib as _lib
from .codes import CSSTannerGraph, SimpleTannerGraph
from .errors import (CSSErrorPattern, IndependentDepolarizingError, IndependentFlipError, css_syndrome_rows,
                     iid_error)
from .gf2 import gf2_right_inverse, logical_operator
from .network import FactorGraph, tnmmap_css_network, tnmmap_dem_network
from .planner import Plan, anneal_tree, check_tree, flatten, search_tree, tree_cost


# ----------------------------------------------------------------------------
# Contraction-order optimizers (the reference's CodeOptimizer argument)
# ----------------------------------------------------------------------------

@dataclass(frozen=True)
class TreeSA:
    """Seeded tree search: best of greedy / min-fill candidates, refined by
    simulated annealing over tree rotations.  Deterministic for a seed."""
    seed: int = 1
    ntrials: int = 2
    sweeps: int = 300
    restarts: int = 4
    sc_target: Optional[int] = None
    primary: Optional[tuple] = None   # variables of a time-ordered sweep candidate


@dataclass(frozen=True)
class GreedyMethod:
    seed: int = 1


@dataclass(frozen=True)
class Dissection:
    """Nested dissection of the error units (network.dissected_parity_network):
    the parity checks rewritten as trees of rank-3 parity tensors along a
    recursive bisection of the mechanisms or qubits, contracted in that
    order, then refined by ``sweeps`` of tree rotations; the best of
    ``seeds`` dissections.  Narrow where greedy and elimination orders run
    wide (circuit-level detector error models, codes on a torus).  Works on
    its own network, so only with evidence given as slices.  ``narrow``
    adds a second annealing pass with a penalty on the widest steps, kept
    when it is cheaper (multiply-adds, then width): on circuit-level
    detector error models it finds trees both narrower and cheaper."""
    seeds: int = 2
    sweeps: int = 200
    narrow: bool = True


def dissection_plan(build, opt: "Dissection"):
    """The cheapest (multiply-adds) of ``opt.seeds`` dissected networks
    ``build(seed) -> (graph, tree)``, annealed: ``(graph, plan, cost)``."""
    from .planner import anneal_tree
    best = None
    for seed in range(max(1, opt.seeds)):
        g, tree = build(seed)
        if opt.sweeps > 0:
            ref = anneal_tree(tree, g.iy, seed=seed, sweeps=opt.sweeps)
            if tree_cost(ref)[0] < tree_cost(tree)[0]:
                tree = ref
            if opt.narrow:
                c0 = tree_cost(tree)
                nar = anneal_tree(tree, g.iy, seed=seed * 31 + 1, sweeps=opt.sweeps, sc_target=c0[2] - 1)
                c1 = tree_cost(nar)
                if (c1[0], c1[2]) < (c0[0], c0[2]):
                    tree = nar
        c = tree_cost(tree)
        if best is None or (c[0], c[2]) < (best[2][0], best[2][2]):
            best = (g, tree, c)
    g, tree, c = best
    return g, flatten(g.without_evidence_leaves(), tree, True), c


@dataclass(frozen=True)
class NoOptimizer:
    """Contract in leaf order, ``((t1 t2) t3) ...`` (tndecoder.jl:64)."""


def optimize_tree(ixs, iy, optimizer):
    from .planner import ContractionTree, greedy_tree, recompute_masks
    if isinstance(optimizer, NoOptimizer):
        L = len(ixs)
        children = [(0, 1)] + [(L + i, i + 2) for i in range(L - 2)]
        tr = ContractionTree(L, children, [0] * (2 * L - 1))
        tr.masks = recompute_masks(tr, ixs, iy)
        return tr
    if isinstance(optimizer, GreedyMethod):
        return greedy_tree(ixs, iy, seed=optimizer.seed)
    opt = optimizer if isinstance(optimizer, TreeSA) else TreeSA()
    base = search_tree(ixs, iy, seed=opt.seed, ntrials=opt.ntrials, sc_target=opt.sc_target,
                       primary=opt.primary)
    best, best_key = base, tree_cost(base)[0]
    for r in range(opt.restarts):
        tr = anneal_tree(base, iy, seed=opt.seed * 7919 + r, sweeps=opt.sweeps, sc_target=opt.sc_target)
        c = tree_cost(tr)
        pen = 0 if opt.sc_target is None or c[2] <= opt.sc_target else c[2] - opt.sc_target
        key = (pen, c[0])
        bp = 0 if opt.sc_target is None or tree_cost(best)[2] <= opt.sc_target else 1
        if key < (bp, best_key):
            best, best_key = tr, c[0]
    check_tree(best, ixs, iy)
    return best


# ----------------------------------------------------------------------------
# Decoder types, problems and results
# ----------------------------------------------------------------------------

@dataclass
class TNMMAP:
    """Marginal-MAP tensor-network decoder (tndecoder.jl:68-80).

    ``factorize`` rewrites every parity tensor of three or more members into
    a chain of rank-3 parity tensors (tndecoder.jl:221-238; the reference
    applies it to DEM networks only, here it applies to CSS networks too —
    the joint distribution is the same).  ``mode`` picks the executor:
    "plain_f64", "fused_f64" or "fused_f32".  ``evidence`` is "slices"
    (syndrome variables fixed by slicing the tensors that hold them) or
    "leaves" (one-hot evidence vectors, as in the reference network).
    """
    optimizer: Any = field(default_factory=TreeSA)
    factorize: bool = True
    mode: str = "fused_f32"
    evidence: str = "slices"
    tabulate_bits: Optional[int] = None   # None: 0 for plain_f64, 4 otherwise


@dataclass
class DecodingResult:
    """interfaces.jl:116-119."""
    success_tag: bool
    error_pattern: Any


@dataclass
class CSSSyndrome:
    sx: np.ndarray
    sz: np.ndarray


@dataclass
class SimpleSyndrome:
    s: np.ndarray


@dataclass
class IndependentDepolarizingDecodingProblem:
    tanner: CSSTannerGraph
    pvec: IndependentDepolarizingError


@dataclass
class BatchResult:
    """Per-syndrome outputs of :func:`decode_batch`."""
    marginals: Optional[np.ndarray]   # (B, 2^n_out): P(syndrome, class), column-major class order
    classes: np.ndarray                # (B,) int32, Julia findmax position - 1
    ties: np.ndarray                   # (B,) uint8
    error_patterns: Any = None         # CSSErrorPattern of (B, n) arrays, or (B, nq) for DEMs


@dataclass
class CompiledTNMMAP:
    """tndecoder.jl:85-92 plus the HIP plan."""
    tanner: CSSTannerGraph
    lx: np.ndarray
    lz: np.ndarray
    graph: FactorGraph
    plan: Plan
    hip: Any
    mode: str
    syndrome_indices: list
    pure_x: np.ndarray = None   # maps sz -> an ex with Hz ex = sz
    pure_z: np.ndarray = None   # maps sx -> an ez with Hx ez = sx

    def contraction_complexity(self):
        """(tc, sc, rw) in log2 like ``contraction_complexity(ct)``
        (tndecoder.jl:277-279)."""
        s = self.plan.stats()
        return {"tc": np.log2(max(s["tc"], 1)), "sc": s["sc"], "rwc": np.log2(max(s["rw"], 1))}


def build_plan(graph: FactorGraph, decoder: TNMMAP) -> Plan:
    if decoder.evidence == "slices":
        g = graph.without_evidence_leaves()
        ev = set(g.evidence_labels)
        ixs = [[l for l in ix if l not in ev] for ix in g.ixs]
        tree = optimize_tree(ixs, g.iy, decoder.optimizer)
        return flatten(g, tree, True)
    if decoder.evidence == "leaves":
        tree = optimize_tree(graph.ixs, graph.iy, decoder.optimizer)
        return flatten(graph, tree, False)
    raise ValueError(f"unknown evidence handling {decoder.evidence!r}")


def compile(decoder, tanner, em=None):
    """``compile(decoder, tanner, em)`` (interfaces.jl:74-79).

    As in the reference, ``em`` defaults to ``iid_error(0.05, tanner)``,
    i.e. px = py = pz = 0.05 each.
    """
    if isinstance(tanner, IndependentDepolarizingDecodingProblem):
        tanner, em = tanner.tanner, tanner.pvec
    if hasattr(tanner, "error_rates") and em is None:
        return compile_dem(decoder, tanner)
    if not isinstance(decoder, TNMMAP):
        from .tnmap import TNMAP, compile_tnmap
        if isinstance(decoder, TNMAP):
            return compile_tnmap(decoder, tanner, em)
        raise TypeError(f"unsupported decoder {decoder!r}")
    if em is None:
        em = iid_error(0.05, 0.05, 0.05, tanner.nq)
    lx, lz = logical_operator(tanner)
    graph = tnmmap_css_network(tanner, em, factorize=decoder.factorize, lxlz=(lx, lz))
    opt = decoder.optimizer
    if isinstance(opt, TreeSA) and opt.primary is None:
        # extra candidate: sweep the qubits in index order, x and z error of a
        # qubit together (a row-by-row sweep of a planar layout)
        from dataclasses import replace
        n = tanner.nq
        decoder = replace(decoder, optimizer=replace(opt, primary=tuple(v for q in range(n) for v in (q, q + n))))
    if isinstance(opt, Dissection):
        if decoder.evidence != "slices":
            raise ValueError("the dissection optimizer builds its own network: evidence must be 'slices'")
        from .network import dissected_css_network
        graph, plan, _ = dissection_plan(lambda s: dissected_css_network(tanner, em, lxlz=(lx, lz), seed=s), opt)
    else:
        plan = build_plan(graph, decoder)
    tab = decoder.tabulate_bits
    if tab is None:
        tab = 0 if decoder.mode == "plain_f64" else 4
    hip = _lib.HipPlan(plan, tabulate_bits=tab)
    n = tanner.nq
    ct = CompiledTNMMAP(tanner, lx, lz, graph, plan, hip, decoder.mode,
                        list(range(2 * n, 2 * n + tanner.ns)))
    ct.pure_x = gf2_right_inverse(tanner.stgz.H)
    ct.pure_z = gf2_right_inverse(tanner.stgx.H)
    return ct


def _syndrome_rows(ct, syndromes) -> np.ndarray:
    if isinstance(syndromes, CSSSyndrome):
        return np.concatenate([np.asarray(syndromes.sx, np.uint8), np.asarray(syndromes.sz, np.uint8)])[None, :]
    if isinstance(syndromes, SimpleSyndrome):
        return np.asarray(syndromes.s, np.uint8)[None, :]
    if isinstance(syndromes, tuple) and len(syndromes) == 2:
        sx, sz = np.atleast_2d(syndromes[0]), np.atleast_2d(syndromes[1])
        return np.concatenate([sx, sz], axis=1).astype(np.uint8)
    return np.atleast_2d(np.asarray(syndromes, dtype=np.uint8))


def css_error_patterns(ct: CompiledTNMMAP, rows: np.ndarray, classes: np.ndarray) -> CSSErrorPattern:
    """Pure error plus logical fix for a batch (tndecoder.jl:167-174).

    The pure error comes from a precomputed GF(2) right inverse instead of
    the MIP of ipdecoder.jl:150-169 (any solution of H e = s is valid there
    too); the logical fix is the reference's loop: add lx[i] to ex when
    lz[i].ex disagrees with the class bit of axis k+i, and lz[i] to ez when
    lx[i].ez disagrees with the bit of axis i.
    """
    nsx = ct.tanner.stgx.ns
    sx = rows[:, :nsx].astype(np.int64)
    sz = rows[:, nsx:].astype(np.int64)
    ex = ((sz @ ct.pure_x.T.astype(np.int64)) & 1).astype(np.uint8)
    ez = ((sx @ ct.pure_z.T.astype(np.int64)) & 1).astype(np.uint8)
    k = ct.lx.shape[0]
    cls = classes.astype(np.int64)
    for i in range(k):
        want_x = (cls >> (i + k)) & 1
        have_x = ((ex.astype(np.int64) @ ct.lz[i].astype(np.int64)) & 1)
        ex ^= ((have_x != want_x)[:, None] * ct.lx[i][None, :]).astype(np.uint8)
        want_z = (cls >> i) & 1
        have_z = ((ez.astype(np.int64) @ ct.lx[i].astype(np.int64)) & 1)
        ez ^= ((have_z != want_z)[:, None] * ct.lz[i][None, :]).astype(np.uint8)
    return CSSErrorPattern(ex, ez)


def decode_batch(ct, syndromes, mode: Optional[str] = None, want_marginals: bool = True,
                 error_patterns: bool = True) -> BatchResult:
    """Decode a batch of syndromes (rows ``[sx | sz]``) in one library call."""
    if not isinstance(ct, CompiledTNMMAP):
        from .dem import CompiledDEMTNMMAP, decode_batch_dem
        if isinstance(ct, CompiledDEMTNMMAP):
            return decode_batch_dem(ct, syndromes, mode, want_marginals, error_patterns)
        raise TypeError(f"unsupported compiled decoder {type(ct)}")
    rows = _syndrome_rows(ct, syndromes)
    mar, cls, tie = ct.hip.decode(rows, mode or ct.mode, want_marginals)
    ep = css_error_patterns(ct, rows, cls) if error_patterns else None
    return BatchResult(mar, cls, tie, ep)


def decode(ct, syndrome) -> DecodingResult:
    """``decode(ct, syndrome)`` (tndecoder.jl:160-165): one syndrome."""
    if not isinstance(ct, CompiledTNMMAP):
        from .dem import CompiledDEMTNMMAP, decode_dem
        if isinstance(ct, CompiledDEMTNMMAP):
            return decode_dem(ct, syndrome)
        from .tnmap import CompiledTNMAP, decode_tnmap
        if isinstance(ct, CompiledTNMAP):
            return decode_tnmap(ct, syndrome)
        raise TypeError(f"unsupported compiled decoder {type(ct)}")
    r = decode_batch(ct, syndrome)
    ep = CSSErrorPattern(r.error_patterns.xerror[0], r.error_patterns.zerror[0])
    return DecodingResult(True, ep)


def compile_dem(decoder, dem):
    from .dem import compile_dem as _c
    return _c(decoder, dem)
