"""Detector error models: container, ``.dem`` parser and TNMMAP decoding.

Reference: src/decoding/dem.jl:1-6 (``DetectorErrorModel``), :150-164
(``dem2tanner``, ``random_error_pattern``), src/stim_parser/stim_parser.jl:
337-375 (``parse_dem_file`` / ``parse_dem_string``),
src/decoding/tndecoder.jl:176-271 (compile / decode / error_pattern).
Detector and logical ids are 1-based as in the reference; logicals are
numbered after the largest detector.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import lib as _lib
from .codes import SimpleTannerGraph
from .errors import IndependentFlipError, sample_flips, syndrome_extraction_H
from .gf2 import gf2_right_inverse
from .network import tnmmap_dem_network


@dataclass
class DetectorErrorModel:
    error_rates: np.ndarray
    flipped_detectors: List[List[int]]   # per mechanism: detector and logical ids (1-based)
    detector_list: List[int]
    logical_list: List[int]

    def tanner(self) -> SimpleTannerGraph:
        """dem.jl:150-160: mechanisms are the bits, detectors the checks
        (``ns = maximum(detector_list)``)."""
        nq = len(self.error_rates)
        logs = set(self.logical_list)
        q2s = [[x for x in v if x not in logs] for v in self.flipped_detectors]
        ns = max(self.detector_list)
        checks = [[j for j in range(nq) if (i + 1) in q2s[j]] for i in range(ns)]
        return SimpleTannerGraph.from_checks(nq, checks)

    def l2q(self) -> List[List[int]]:
        """Mechanisms (0-based) touching each logical (tndecoder.jl:188)."""
        return [[j for j, f in enumerate(self.flipped_detectors) if L in f] for L in self.logical_list]

    def detector_matrix(self) -> np.ndarray:
        return self.tanner().H

    def logical_matrix(self) -> np.ndarray:
        M = np.zeros((len(self.logical_list), len(self.error_rates)), dtype=np.uint8)
        for i, ms in enumerate(self.l2q()):
            M[i, ms] = 1
        return M


def parse_dem_string(content: str) -> DetectorErrorModel:
    """stim_parser.jl:342-375: only ``error(p) D.. L..`` lines are read;
    ``repeat`` blocks are rejected; detector/annotation lines are ignored."""
    rates, dets, logs = [], [], []
    for line in content.split("\n"):
        if not line:
            continue
        s = line.split()
        if not s:
            continue
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
            raise ValueError("Repeat is not supported, use `circuit.detector_error_model(flatten_loops=True)` "
                             "to flatten the loops in stim")
    largest = max(max(d) for d in dets if d)
    logs = [[x + largest for x in l] for l in logs]
    largest_l = max([max(l) for l in logs if l] + [0])
    flipped = [sorted(set(d) | set(l)) for d, l in zip(dets, logs)]
    return DetectorErrorModel(np.asarray(rates, dtype=np.float64), flipped, list(range(1, largest + 1)),
                              list(range(largest + 1, largest_l + 1)))


def parse_dem_file(path: str) -> DetectorErrorModel:
    with open(path) as f:
        return parse_dem_string(f.read())


def random_error_pattern(dem: DetectorErrorModel, seed: int, count: int, first: int = 0) -> np.ndarray:
    """dem.jl:162-164 with the seeded sampler (rows are samples)."""
    return sample_flips(IndependentFlipError(dem.error_rates), seed, count, first)


@dataclass
class CompiledDEMTNMMAP:
    """tndecoder.jl:176-182 plus the HIP plan."""
    tanner: SimpleTannerGraph
    l2q: List[List[int]]
    graph: object
    plan: object
    hip: object
    mode: str
    syndrome_indices: list
    pure: np.ndarray = None        # mechanisms x detectors: e = pure s satisfies H e = s
    l2q_mask: np.ndarray = None    # logicals x mechanisms


# Reference networks with more tensors than this are not searched by the
# greedy / elimination candidates of TreeSA (the dissection is used alone).
_TREESA_TENSOR_LIMIT = 1200


def compile_dem(decoder, dem: DetectorErrorModel) -> CompiledDEMTNMMAP:
    """``compile(TNMMAP(), dem)`` (tndecoder.jl:186-219).

    With the default optimizer two networks compete: the reference's
    (chain-factorized checks, tree from the seeded TreeSA search with a
    time-sweep candidate) and the nested-dissection rewrite
    (``decoders.Dissection``); the plan with fewer multiply-adds per
    syndrome is kept.  Both give the reference's marginals (the same joint
    distribution; only auxiliary variables differ)."""
    from dataclasses import replace
    from .decoders import Dissection, TreeSA, build_plan, dissection_plan
    from .network import dissected_dem_network
    g = tnmmap_dem_network(dem, factorize=decoder.factorize)
    opt = decoder.optimizer
    plan = None
    if isinstance(opt, Dissection) or (isinstance(opt, TreeSA) and decoder.evidence == "slices"):
        gd, pd, cd = dissection_plan(lambda s: dissected_dem_network(dem, seed=s),
                                     opt if isinstance(opt, Dissection) else Dissection())
        if isinstance(opt, Dissection) or len(g.ixs) > _TREESA_TENSOR_LIMIT:
            g, plan = gd, pd
    if plan is None:
        if isinstance(opt, TreeSA) and opt.primary is None:
            # mechanisms in detector (time) order seed a sweep candidate
            opt = replace(opt, primary=tuple(range(len(dem.error_rates))))
        plan = build_plan(g, replace(decoder, optimizer=opt))
        st = plan.stats()
        if isinstance(decoder.optimizer, TreeSA) and decoder.evidence == "slices" and \
                (cd[0], cd[2]) < (st["tc"], st["sc"]):
            g, plan = gd, pd
    tab = decoder.tabulate_bits
    if tab is None:
        tab = 0 if decoder.mode == "plain_f64" else 4
    hip = _lib.HipPlan(plan, tabulate_bits=tab)
    tanner = g.meta.get("tanner") or dem.tanner()
    ct = CompiledDEMTNMMAP(tanner, dem.l2q(), g, plan, hip, decoder.mode,
                           list(range(tanner.nq, tanner.nq + tanner.ns)))
    ct.pure = gf2_right_inverse(tanner.H)
    ct.l2q_mask = dem.logical_matrix()
    return ct


def dem_error_patterns(ct: CompiledDEMTNMMAP, rows: np.ndarray, classes: np.ndarray) -> np.ndarray:
    """tndecoder.jl:255-271 for a batch: a pure error, then — for each
    logical whose mechanism parity disagrees with the decided class bit —
    every mechanism touching that logical is flipped (the reference's rule)."""
    ep = ((rows.astype(np.int64) @ ct.pure.T.astype(np.int64)) & 1).astype(np.uint8)
    for i, mask in enumerate(ct.l2q_mask):
        par = (ep.astype(np.int64) @ mask.astype(np.int64)) & 1
        want = (classes.astype(np.int64) >> i) & 1
        ep ^= ((par != want)[:, None] * mask[None, :]).astype(np.uint8)
    return ep


def decode_batch_dem(ct: CompiledDEMTNMMAP, syndromes, mode: Optional[str] = None, want_marginals: bool = True,
                     error_patterns: bool = True):
    from .decoders import BatchResult, SimpleSyndrome
    if isinstance(syndromes, SimpleSyndrome):
        rows = np.asarray(syndromes.s, np.uint8)[None, :]
    else:
        rows = np.atleast_2d(np.asarray(syndromes, dtype=np.uint8))
    mar, cls, tie = ct.hip.decode(rows, mode or ct.mode, want_marginals)
    ep = dem_error_patterns(ct, rows, cls) if error_patterns else None
    return BatchResult(mar, cls, tie, ep)


def decode_dem(ct: CompiledDEMTNMMAP, syndrome):
    """``decode(ct::CompiledDEMTNMMAP, syndrome)`` (tndecoder.jl:248-253)."""
    from .decoders import DecodingResult
    r = decode_batch_dem(ct, syndrome)
    return DecodingResult(True, r.error_patterns[0])


def syndrome_extraction(ep: np.ndarray, tanner: SimpleTannerGraph) -> np.ndarray:
    return syndrome_extraction_H(ep, tanner.H)
