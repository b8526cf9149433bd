"""TNMAP: most probable error configuration by max-product contraction.

Reference: src/decoding/tndecoder.jl:1-57 (``TNMAP``, ``stg2uaimodel``,
``compile``/``decode`` through TensorInference's ``most_probable_config``),
src/decoding/interfaces.jl:129-142 and src/decoding/general_decoding.jl:1-32
(the CSS and classical reductions to a general decoding problem).

The network is the reference's UAI model: one even-parity tensor per check
over the check's error variables plus its syndrome variable, and the prior
tensors of the problem; the syndrome variables are the evidence (fixed by
slicing).  The HIP library contracts it in the max-product semiring and
backtracks the arg-max per syndrome.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from . import lib as _lib
from .codes import CSSTannerGraph, SimpleTannerGraph
from .errors import CSSErrorPattern, IndependentDepolarizingError, IndependentFlipError
from .network import FactorGraph, _push_check, parity_check_tensor, qubit_prior


@dataclass
class TNMAP:
    """``TNMAP(; optimizer=TreeSA())`` (tndecoder.jl:9-11).  ``factorize``
    (default False, the reference network) splits checks of 3+ members into
    rank-3 parity chains, which narrows wide networks."""
    optimizer: Any = None
    factorize: bool = False


@dataclass
class SimpleTensorNetwork:
    """(ixs, tensors) prior network over error variables 0..nq-1
    (src/nonclifford/tensornetwork.jl:134-202)."""
    ixs: List[List[int]]
    tensors: List[np.ndarray]


@dataclass
class GeneralDecodingProblem:
    tanner: SimpleTannerGraph
    ptn: SimpleTensorNetwork


def reduce2general(tanner: CSSTannerGraph, em) -> GeneralDecodingProblem:
    """general_decoding.jl:24-32: error variables x 0..n-1 and z n..2n-1;
    checks = X checks on z variables, then Z checks on x variables; priors
    ``[1-px-py-pz pz; px py]`` on (i, i+n)."""
    n = tanner.nq
    checks = [[q + n for q in c] for c in tanner.stgx.s2q] + [list(c) for c in tanner.stgz.s2q]
    if isinstance(em, SimpleTensorNetwork):
        tn = em
    else:
        tn = SimpleTensorNetwork([[i, i + n] for i in range(n)],
                                 [qubit_prior(em.px[i], em.py[i], em.pz[i]) for i in range(n)])
    return GeneralDecodingProblem(SimpleTannerGraph.from_checks(2 * n, checks), tn)


def classical_problem(tanner: SimpleTannerGraph, em: IndependentFlipError) -> GeneralDecodingProblem:
    """interfaces.jl:139-142: priors ``[1-p, p]`` per bit."""
    return GeneralDecodingProblem(tanner, SimpleTensorNetwork([[i] for i in range(tanner.nq)],
                                                              [np.array([1 - p, p]) for p in em.p]))


def uai_network(problem: GeneralDecodingProblem, factorize: bool = False) -> FactorGraph:
    """stg2uaimodel (tndecoder.jl:33-40): checks on (s2q[s], nq + s), then
    the prior tensors; syndrome variables nq..nq+ns-1 are the evidence."""
    t = problem.tanner
    nq, ns = t.nq, t.ns
    nvars = nq + ns
    ixs, tensors = [], []
    for s, c in enumerate(t.s2q):
        nvars = _push_check(ixs, tensors, list(c), nq + s, nvars, factorize and len(c) >= 2)
    for ix, ten in zip(problem.ptn.ixs, problem.ptn.tensors):
        ixs.append(list(ix))
        tensors.append(np.asarray(ten, dtype=np.float64).reshape(-1, order="F")
                       if np.asarray(ten).ndim > 1 else np.asarray(ten, dtype=np.float64))
    ev = list(range(nq, nq + ns))
    return FactorGraph(nvars, ixs, tensors, [], ev, [-1] * ns, {"kind": "uai", "nq": nq, "ns": ns})


@dataclass
class CompiledTNMAP:
    graph: FactorGraph
    plan: Any
    hip: Any
    qubit_num: int            # error variables reported
    css_n: Optional[int] = None   # set for CSS problems: split into (ex, ez)


def compile_tnmap(decoder: TNMAP, tanner, em=None) -> CompiledTNMAP:
    """``compile(TNMAP(), problem)`` (tndecoder.jl:46-50 via interfaces.jl:129-142)."""
    from .decoders import TreeSA, optimize_tree
    from .planner import flatten
    css_n = None
    if isinstance(tanner, GeneralDecodingProblem):
        prob = tanner
    elif isinstance(tanner, CSSTannerGraph):
        if em is None:
            from .errors import iid_error
            em = iid_error(0.05, 0.05, 0.05, tanner.nq)
        prob = reduce2general(tanner, em)
        css_n = tanner.nq
    else:
        if em is None:
            em = IndependentFlipError(np.full(tanner.nq, 0.05))
        prob = classical_problem(tanner, em)
    g = uai_network(prob, decoder.factorize)
    ev = set(g.evidence_labels)
    ixs = [[l for l in ix if l not in ev] for ix in g.ixs]
    tree = optimize_tree(ixs, g.iy, decoder.optimizer or TreeSA())
    plan = flatten(g, tree, True)
    hip = _lib.HipPlan(plan, tabulate_bits=0, semiring=1)
    return CompiledTNMAP(g, plan, hip, prob.tanner.nq, css_n)


def decode_tnmap_batch(ct: CompiledTNMAP, rows: np.ndarray):
    """Configurations (B, qubit_num) and log of the maximum (B,)."""
    rows = np.atleast_2d(np.asarray(rows, dtype=np.uint8))
    return ct.hip.map_decode(rows, ct.qubit_num)


def decode_tnmap(ct: CompiledTNMAP, syndrome):
    """``decode(ct, syndrome)`` (tndecoder.jl:53-57; CSS: interfaces.jl:135-137
    and extract_decoding, general_decoding.jl:6-9)."""
    from .decoders import CSSSyndrome, DecodingResult, SimpleSyndrome
    if isinstance(syndrome, CSSSyndrome):
        row = np.concatenate([np.asarray(syndrome.sx, np.uint8), np.asarray(syndrome.sz, np.uint8)])
    elif isinstance(syndrome, SimpleSyndrome):
        row = np.asarray(syndrome.s, np.uint8)
    else:
        row = np.asarray(syndrome, np.uint8)
    cfg, _ = decode_tnmap_batch(ct, row[None, :])
    cfg = cfg[0]
    if ct.css_n is not None:
        n = ct.css_n
        return DecodingResult(True, CSSErrorPattern(cfg[:n].copy(), cfg[n:2 * n].copy()))
    return DecodingResult(True, cfg.copy())
