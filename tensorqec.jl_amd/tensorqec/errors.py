"""Error models, the seeded counter-based sampler, syndrome extraction and the
logical-error check.

The sampler replaces Julia's global ``rand()`` with Philox4x32-10 (Salmon,
Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), so
that the host, the CPU oracle and the GPU sampler produce bit-identical
syndrome batches from (seed, sample index, qubit index).

Reference semantics: src/decoding/error_model.jl:13-117 (models and the
u < py / < px+py / < px+py+pz rule), :131-146 (syndromes), :161-181
(logical check).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

PHILOX_M0 = np.uint64(0xD2511F53)
PHILOX_M1 = np.uint64(0xCD9E8D57)
PHILOX_W0 = np.uint32(0x9E3779B9)
PHILOX_W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)

# Stream tags keep the draws of different purposes independent.
STREAM_DEPOLARIZING = 0x51D0
STREAM_FLIP = 0xF11B


def philox4x32(c0, c1, c2, c3, k0, k1, rounds: int = 10):
    """Vectorised Philox4x32 on uint32 arrays; returns four uint32 arrays."""
    c0 = np.asarray(c0, dtype=np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, dtype=np.uint32).astype(np.uint64)
    c2 = np.asarray(c2, dtype=np.uint32).astype(np.uint64)
    c3 = np.asarray(c3, dtype=np.uint32).astype(np.uint64)
    k0 = np.uint64(np.uint32(k0))
    k1 = np.uint64(np.uint32(k1))
    for _ in range(rounds):
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK32, lo1, (hi0 ^ c3 ^ k1) & _MASK32, lo0
        k0 = (k0 + np.uint64(PHILOX_W0)) & _MASK32
        k1 = (k1 + np.uint64(PHILOX_W1)) & _MASK32
    return (c0.astype(np.uint32), c1.astype(np.uint32), c2.astype(np.uint32), c3.astype(np.uint32))


def uniform53(x0, x1) -> np.ndarray:
    """Two uint32 draws to a double in [0, 1) with 53 random bits."""
    a = (np.asarray(x0, dtype=np.uint64) >> np.uint64(5)).astype(np.float64)
    b = (np.asarray(x1, dtype=np.uint64) >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0)


def uniforms(seed: int, stream: int, first: int, count: int, width: int) -> np.ndarray:
    """``count x width`` uniforms; entry (s, j) uses counter
    (j, s mod 2^32, s >> 32, stream) and key (seed mod 2^32, seed >> 32)."""
    s = np.arange(first, first + count, dtype=np.uint64)
    j = np.arange(width, dtype=np.uint64)
    S, J = np.meshgrid(s, j, indexing="ij")
    c0 = J.astype(np.uint32)
    c1 = (S & _MASK32).astype(np.uint32)
    c2 = (S >> np.uint64(32)).astype(np.uint32)
    c3 = np.full_like(c0, stream & 0xFFFFFFFF)
    x0, x1, _, _ = philox4x32(c0, c1, c2, c3, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    return uniform53(x0, x1)


@dataclass
class IndependentFlipError:
    """Independent bit flips with per-bit probability (error_model.jl:13-15)."""
    p: np.ndarray

    def __post_init__(self):
        self.p = np.asarray(self.p, dtype=np.float64)


@dataclass
class IndependentDepolarizingError:
    """Per-qubit X/Y/Z probabilities (error_model.jl:41-45)."""
    px: np.ndarray
    py: np.ndarray
    pz: np.ndarray

    def __post_init__(self):
        self.px = np.asarray(self.px, dtype=np.float64)
        self.py = np.asarray(self.py, dtype=np.float64)
        self.pz = np.asarray(self.pz, dtype=np.float64)

    @property
    def n(self) -> int:
        return self.px.shape[0]


def iid_error(*args):
    """``iid_error(p, n)`` -> flips; ``iid_error(px, py, pz, n)`` -> depolarizing
    (error_model.jl:28, 61).  Note: ``iid_error(p, tanner)`` in the reference
    means px = py = pz = p; pass p/3 each for total depolarizing rate p."""
    if len(args) == 2:
        p, n = args
        return IndependentFlipError(np.full(int(n), float(p)))
    px, py, pz, n = args
    n = int(n)
    return IndependentDepolarizingError(np.full(n, float(px)), np.full(n, float(py)), np.full(n, float(pz)))


@dataclass
class CSSErrorPattern:
    """X and Z error bits (error_model.jl:81-84); arrays are uint8 0/1, and
    may carry a leading batch axis."""
    xerror: np.ndarray
    zerror: np.ndarray


def sample_depolarizing(em: IndependentDepolarizingError, seed: int, count: int, first: int = 0):
    """``count`` error patterns, rows are samples (error_model.jl:97-117).

    For qubit q of sample s with u = U(s, q): Y if u < py, X if u < px+py,
    Z if u < px+py+pz, else I.
    """
    u = uniforms(seed, STREAM_DEPOLARIZING, first, count, em.n)
    t1 = em.py[None, :]
    t2 = (em.px + em.py)[None, :]
    t3 = (em.px + em.py + em.pz)[None, :]
    y = u < t1
    x = (~y) & (u < t2)
    z = (~y) & (~x) & (u < t3)
    ex = (y | x).astype(np.uint8)
    ez = (y | z).astype(np.uint8)
    return CSSErrorPattern(ex, ez)


def sample_flips(em: IndependentFlipError, seed: int, count: int, first: int = 0) -> np.ndarray:
    """``count`` flip patterns: bit i set iff U(s, i) < p[i] (error_model.jl:69-71)."""
    u = uniforms(seed, STREAM_FLIP, first, count, em.p.shape[0])
    return (u < em.p[None, :]).astype(np.uint8)


def syndrome_extraction_H(errors: np.ndarray, H: np.ndarray) -> np.ndarray:
    """``H * e`` over GF(2) for one pattern or a batch (rows)
    (error_model.jl:131-136)."""
    e = np.asarray(errors, dtype=np.uint8)
    H = np.asarray(H, dtype=np.uint8)
    if e.ndim == 1:
        return ((H.astype(np.int64) @ e.astype(np.int64)) & 1).astype(np.uint8)
    return ((e.astype(np.int64) @ H.T.astype(np.int64)) & 1).astype(np.uint8)


def syndrome_extraction(ep: CSSErrorPattern, tanner):
    """``(sx, sz)`` with sx = Hx * ez and sz = Hz * ex (error_model.jl:144-146)."""
    sx = syndrome_extraction_H(ep.zerror, tanner.stgx.H)
    sz = syndrome_extraction_H(ep.xerror, tanner.stgz.H)
    return sx, sz


def css_syndrome_rows(ep: CSSErrorPattern, tanner) -> np.ndarray:
    """Batch of syndromes as uint8 rows ``[sx | sz]`` — the boundary layout."""
    sx, sz = syndrome_extraction(ep, tanner)
    return np.ascontiguousarray(np.concatenate([np.atleast_2d(sx), np.atleast_2d(sz)], axis=1))


def check_logical_error_H(e1, e2, l) -> np.ndarray:
    """Any odd overlap of ``e1 - e2`` with a row of ``l`` (error_model.jl:161-163)."""
    d = (np.asarray(e1, dtype=np.uint8) ^ np.asarray(e2, dtype=np.uint8)).astype(np.int64)
    return (((d @ np.asarray(l, dtype=np.int64).T) & 1) != 0).any(axis=-1)


def check_logical_error(ep1: CSSErrorPattern, ep2: CSSErrorPattern, lx, lz) -> np.ndarray:
    """error_model.jl:179-181."""
    return check_logical_error_H(ep1.zerror, ep2.zerror, lx) | check_logical_error_H(ep1.xerror, ep2.xerror, lz)
