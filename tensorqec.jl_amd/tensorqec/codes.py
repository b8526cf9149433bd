"""Quantum code constructors and Tanner graphs.

Stabilizers are kept in the reference's order, because that order fixes the
syndrome bit order on the decoder boundary.

Reference: src/codes/codes.jl:37-66 (rotated surface), :73-110 (Shor,
Steane), :184-244 (bivariate bicycle / toric), src/codes/ldpc.jl:27-115
(Tanner graphs), src/codes/code_distance.jl:300-304 (dependency removal).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from .gf2 import Bimatrix, css_gaussian_elimination

# Pauli ids follow the reference: 0=I, 1=X, 2=Y, 3=Z.
I_, X_, Y_, Z_ = 0, 1, 2, 3


def _pauli(nq: int, support: Sequence[int], p: int) -> np.ndarray:
    s = np.zeros(nq, dtype=np.int8)
    for q in support:
        s[q - 1] = p
    return s


@dataclass(frozen=True)
class SurfaceCode:
    """Rotated surface code with ``m`` rows and ``n`` columns."""
    m: int
    n: int


@dataclass(frozen=True)
class ToricCode:
    """Toric code on an ``m`` x ``n`` torus (a bivariate bicycle code,
    src/codes/codes.jl:242)."""
    m: int
    n: int


@dataclass(frozen=True)
class SteaneCode:
    pass


@dataclass(frozen=True)
class ShorCode:
    pass


def stabilizers(code) -> List[np.ndarray]:
    """Stabilizer generators as Pauli id vectors (length nq)."""
    if isinstance(code, SurfaceCode):
        return _surface(code.m, code.n)
    if isinstance(code, ToricCode):
        return _toric(code.m, code.n)
    if isinstance(code, SteaneCode):
        nq = 7
        sup = [(1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)]
        return [_pauli(nq, s, X_) for s in sup] + [_pauli(nq, s, Z_) for s in sup]
    if isinstance(code, ShorCode):
        nq = 9
        xs = [(1, 2, 3, 4, 5, 6), (1, 2, 3, 7, 8, 9)]
        zs = [(1, 2), (1, 3), (4, 5), (4, 6), (7, 8), (7, 9)]
        return [_pauli(nq, s, X_) for s in xs] + [_pauli(nq, s, Z_) for s in zs]
    raise TypeError(f"unsupported code {code!r}")


def _surface(m: int, n: int) -> List[np.ndarray]:
    """src/codes/codes.jl:37-66; qubit (i, j) (1-based) is (i-1)*n + j."""
    nq = m * n

    def q(i, j):
        return (i - 1) * n + j

    out = []
    for i in range(1, m):
        for j in range(1, n):
            if (i + j) % 2 == 0:
                out.append(_pauli(nq, (q(i, j), q(i + 1, j), q(i, j + 1), q(i + 1, j + 1)), X_))
    for i in range(1, m // 2 + 1):
        if 2 * i + (n + 1) % 2 <= m:
            out.append(_pauli(nq, (q(2 * i - 1 + (n + 1) % 2, n), q(2 * i + (n + 1) % 2, n)), X_))
        if 2 * i + 1 <= m:
            out.append(_pauli(nq, (q(2 * i, 1), q(2 * i + 1, 1)), X_))
    for i in range(1, m):
        for j in range(1, n):
            if (i + j) % 2 == 1:
                out.append(_pauli(nq, (q(i, j), q(i + 1, j), q(i, j + 1), q(i + 1, j + 1)), Z_))
    for j in range(1, n // 2 + 1):
        out.append(_pauli(nq, (q(1, 2 * j - 1), q(1, 2 * j)), Z_))
        if 2 * j + m % 2 <= n:
            out.append(_pauli(nq, (q(m, 2 * j - 1 + m % 2), q(m, 2 * j + m % 2)), Z_))
    return out


def _toric(m: int, n: int) -> List[np.ndarray]:
    """Bivariate bicycle code with vc=((1,0),(0,0)), hd=((0,1),(0,0)) and
    linear dependencies removed (src/codes/codes.jl:205-231, 242;
    src/codes/code_distance.jl:300-304)."""
    nsite = m * n
    nq = 2 * nsite
    vc = ((1, 0), (0, 0))
    hd = ((0, 1), (0, 0))

    def mod1(a, b):
        return (a - 1) % b + 1

    def ve(a, b):
        return (b - 1) * m + a

    def he(a, b):
        return nsite + (b - 1) * m + a

    out = []
    for j in range(1, n + 1):
        for i in range(1, m + 1):
            st = [ve(mod1(i + s, m), mod1(j + t, n)) for (s, t) in hd]
            st += [he(mod1(i + s, m), mod1(j + t, n)) for (s, t) in vc]
            out.append(_pauli(nq, st, X_))
    for j in range(1, n + 1):
        for i in range(1, m + 1):
            st = [ve(mod1(i - s, m), mod1(j - t, n)) for (s, t) in vc]
            st += [he(mod1(i - s, m), mod1(j - t, n)) for (s, t) in hd]
            out.append(_pauli(nq, st, Z_))
    return remove_linear_dependency(out)


def stabilizers2bimatrix(sts: List[np.ndarray]) -> Bimatrix:
    """src/codes/encoder.jl:40-47: X-type rows then Z-type rows, [X | Z]."""
    nq = len(sts[0])
    xs = [s for s in sts if np.all((s == I_) | (s == X_))]
    zs = [s for s in sts if np.all((s == I_) | (s == Z_))]
    if len(xs) + len(zs) != len(sts):
        raise ValueError("Invalid PauliString")
    A = np.array([(s == X_).astype(np.uint8) for s in xs]).reshape(len(xs), nq)
    B = np.array([(s == Z_).astype(np.uint8) for s in zs]).reshape(len(zs), nq)
    M = np.zeros((len(xs) + len(zs), 2 * nq), dtype=np.uint8)
    M[:len(xs), :nq] = A
    M[len(xs):, nq:] = B
    return Bimatrix(M, list(range(nq)), None, len(xs))


def remove_linear_dependency(sts: List[np.ndarray]) -> List[np.ndarray]:
    """Keep the stabilizers whose row survives CSS elimination
    (src/codes/code_distance.jl:300-304).

    ``stabilizers2bimatrix`` groups X rows before Z rows; the reference then
    indexes the *original* list with the row mask, which is the same thing
    when the list is already X-then-Z (true for every code built here).
    """
    b = stabilizers2bimatrix(sts)
    css_gaussian_elimination(b)
    keep = b.matrix.any(axis=1)
    return [s for s, k in zip(sts, keep) if k]


@dataclass
class SimpleTannerGraph:
    """Tanner graph of a classical linear code (src/codes/ldpc.jl:27-58).

    ``s2q[i]`` / ``q2s[j]`` are 0-based index lists; ``H`` is ns x nq uint8.
    """
    nq: int
    ns: int
    q2s: List[List[int]]
    s2q: List[List[int]]
    H: np.ndarray

    @classmethod
    def from_checks(cls, nq: int, sts: List[List[int]]) -> "SimpleTannerGraph":
        ns = len(sts)
        H = np.zeros((ns, nq), dtype=np.uint8)
        for i, c in enumerate(sts):
            for j in c:
                H[i, j] = 1
        q2s = [[i for i in range(ns) if H[i, j]] for j in range(nq)]
        s2q = [sorted(set(c)) for c in sts]
        return cls(nq, ns, q2s, s2q, H)

    @classmethod
    def from_H(cls, H) -> "SimpleTannerGraph":
        H = np.asarray(H, dtype=np.uint8) & 1
        return cls.from_checks(H.shape[1], [list(np.flatnonzero(r)) for r in H])


@dataclass
class CSSTannerGraph:
    """X-check and Z-check Tanner graphs of a CSS code
    (src/codes/ldpc.jl:72-115)."""
    stgx: SimpleTannerGraph
    stgz: SimpleTannerGraph

    @property
    def nq(self) -> int:
        return self.stgx.nq

    @property
    def ns(self) -> int:
        return self.stgx.ns + self.stgz.ns

    @classmethod
    def from_code(cls, code) -> "CSSTannerGraph":
        return cls.from_stabilizers(stabilizers(code))

    @classmethod
    def from_stabilizers(cls, sts: List[np.ndarray]) -> "CSSTannerGraph":
        """Split by the type of the first non-identity Pauli
        (src/codes/ldpc.jl:105-111)."""
        nq = len(sts[0])
        xs, zs = [], []
        for s in sts:
            first = s[np.flatnonzero(s)[0]]
            sup = [int(j) for j in np.flatnonzero(s)]
            if first == X_:
                xs.append(sup)
            elif first == Z_:
                zs.append(sup)
        return cls(SimpleTannerGraph.from_checks(nq, xs), SimpleTannerGraph.from_checks(nq, zs))


def CSSTannerGraph_of(code) -> CSSTannerGraph:
    return CSSTannerGraph.from_code(code)
