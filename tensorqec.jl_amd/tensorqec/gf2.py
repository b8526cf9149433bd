"""GF(2) linear algebra used at compile time: Gaussian elimination with
optional column (qubit) swaps, null spaces and logical operators.

All matrices are numpy ``uint8`` arrays holding 0/1.  Indices are 0-based;
the algorithms follow the reference step by step so that the logical
operators (and therefore the meaning of every marginal axis) are the same
rows the reference produces.

Reference: src/codes/gaussian_elimination.jl:1-102,
src/codes/code_distance.jl:24-100, src/codes/encoder.jl:14-47.
"""
from __future__ import annotations

import numpy as np


class Bimatrix:
    """Binary matrix plus row-operation record ``Q`` and qubit ``ordering``.

    Mirrors ``SimpleBimatrix`` / ``CSSBimatrix`` (src/codes/encoder.jl:14-33).
    ``xcodenum`` is only meaningful for the CSS form.
    """

    def __init__(self, matrix, ordering=None, Q=None, xcodenum=None):
        self.matrix = np.array(matrix, dtype=np.uint8) & 1
        m, n = self.matrix.shape
        self.Q = np.eye(m, dtype=np.uint8) if Q is None else np.array(Q, dtype=np.uint8)
        self.ordering = list(range(n)) if ordering is None else list(ordering)
        self.xcodenum = xcodenum
        self.css = xcodenum is not None

    def copy(self):
        b = Bimatrix(self.matrix.copy(), list(self.ordering), self.Q.copy(), self.xcodenum)
        return b

    def switch_qubits(self, i: int, j: int) -> None:
        """Swap qubits i and j (src/codes/gaussian_elimination.jl:1-15)."""
        self.ordering[i], self.ordering[j] = self.ordering[j], self.ordering[i]
        M = self.matrix
        M[:, [i, j]] = M[:, [j, i]]
        if self.css:
            nq = M.shape[1] // 2
            M[:, [nq + i, nq + j]] = M[:, [nq + j, nq + i]]


def gaussian_elimination(bimat: Bimatrix, rows: range, col_offset: int, qubit_offset: int,
                         allow_col_operation: bool = True) -> Bimatrix:
    """In-place GF(2) elimination over ``rows``.

    Restates src/codes/gaussian_elimination.jl:32-96 with 0-based indices:
    with column operations the pivot of the ``offset``-th non-zero row is
    swapped into column ``qubit_offset + offset``; without them, rows are
    swapped to find a pivot in the next non-zero column.
    """
    M = bimat.matrix
    ncols = M.shape[1]
    start_col = col_offset + qubit_offset
    zero_row = 0
    zero_col = 0
    r0 = rows.start
    for i in rows:
        if allow_col_operation:
            offset = i - r0 - zero_row
            nz = np.flatnonzero(M[i, start_col:])
            if nz.size == 0:
                zero_row += 1
                continue
            j = int(nz[0])  # relative to start_col
            bimat.switch_qubits(qubit_offset + offset, j + qubit_offset)
        else:
            if i + zero_col >= ncols:
                return bimat
            col = i + zero_col
            nz = np.flatnonzero(M[i:rows.stop, col])
            while nz.size == 0:
                zero_col += 1
                if i + zero_col >= ncols:
                    return bimat
                col = i + zero_col
                nz = np.flatnonzero(M[i:rows.stop, col])
            j = i + int(nz[0])
            if j != i:
                bimat.Q[[i, j], :] = bimat.Q[[j, i], :]
                M[[i, j], :] = M[[j, i], :]
            offset = i - r0 + zero_col
        pc = offset + start_col
        for k in rows:
            if k != i and M[k, pc]:
                M[k, :] ^= M[i, :]
                bimat.Q[k, :] ^= bimat.Q[i, :]
    return bimat


def css_gaussian_elimination(bimat: Bimatrix) -> Bimatrix:
    """Elimination of a CSS bimatrix: X block then Z block
    (src/codes/gaussian_elimination.jl:97-102)."""
    nq = bimat.matrix.shape[1] // 2
    gaussian_elimination(bimat, range(0, bimat.xcodenum), 0, 0)
    gaussian_elimination(bimat, range(bimat.xcodenum, bimat.matrix.shape[0]), nq, bimat.xcodenum)
    return bimat


def row_echelon_form(H) -> np.ndarray:
    """Row-only elimination (src/codes/code_distance.jl:24-28)."""
    b = Bimatrix(H)
    gaussian_elimination(b, range(0, b.matrix.shape[0]), 0, 0, allow_col_operation=False)
    return b.matrix


def null_space(H) -> np.ndarray:
    """Basis of ker(H) over GF(2), one row per free column
    (src/codes/code_distance.jl:30-51)."""
    H = np.asarray(H, dtype=np.uint8)
    reH = row_echelon_form(H)
    m, n = H.shape
    pivots = []
    for i in range(m):
        nz = np.flatnonzero(reH[i, :])
        if nz.size:
            pivots.append(int(nz[0]))
    rank = len(pivots)
    ns = np.zeros((n - rank, n), dtype=np.uint8)
    row = 0
    pivset = set(pivots)
    for j in range(n):
        if j not in pivset:
            ns[row, j] = 1
            for i in range(len(pivots)):
                if reH[i, j]:
                    ns[row, pivots[i]] = 1
            row += 1
    return ns


def _logical_operator_half(Hx, Hz) -> np.ndarray:
    """Rows of ker(Hx) outside span(Hz) after elimination with qubit swaps,
    un-permuted (src/codes/code_distance.jl:53-64)."""
    Hx = np.asarray(Hx, dtype=np.uint8)
    Hz = np.asarray(Hz, dtype=np.uint8)
    kerHx = null_space(Hx)
    H = np.vstack([Hz, kerHx])
    b = Bimatrix(H)
    gaussian_elimination(b, range(0, H.shape[0]), 0, 0)
    l = b.matrix[Hz.shape[0]:, :]
    l = l[l.any(axis=1), :]
    out = np.zeros_like(l)
    out[:, b.ordering] = l
    return out


def same_qubit_order(lx: np.ndarray, lz: np.ndarray):
    """Pair lx rows with lz rows (src/codes/code_distance.jl:84-100)."""
    lx_new = np.zeros_like(lx)
    lxc = lx.copy()
    for j in range(lz.shape[0]):
        hits = [i for i in range(lxc.shape[0]) if int(lxc[i] @ lz[j]) & 1]
        if not hits:
            raise ValueError("The logical operator is linearly dependent!")
        lx_new[j] = lxc[hits[0]]
        for i in hits[1:]:
            lxc[i] ^= lxc[hits[0]]
    return lx_new, lz


def logical_operator(tanner):
    """``(lx, lz)`` of a CSS Tanner graph (src/codes/code_distance.jl:78-82).

    ``lz`` spans ker(Hx) modulo the Z stabilizers; ``lx`` spans ker(Hz)
    modulo the X stabilizers; rows are paired by :func:`same_qubit_order`.
    """
    Hx = tanner.stgx.H
    Hz = tanner.stgz.H
    lz = _logical_operator_half(Hx, Hz)
    lx = _logical_operator_half(Hz, Hx)
    return same_qubit_order(lx, lz)


def gf2_rank(H) -> int:
    b = Bimatrix(np.asarray(H, dtype=np.uint8))
    gaussian_elimination(b, range(0, b.matrix.shape[0]), 0, 0)
    return int(b.matrix.any(axis=1).sum())


def gf2_right_inverse(H) -> np.ndarray:
    """Matrix ``R`` (n x m) with ``H @ R @ s == s (mod 2)`` for every ``s`` in
    the column space of ``H``.

    Used as the pure-error map that replaces the per-syndrome MIP feasibility
    solve of src/decoding/ipdecoder.jl:150-169: for a consistent syndrome
    ``s``, ``e = R s`` satisfies ``H e = s``.
    """
    H = np.asarray(H, dtype=np.uint8) & 1
    m, n = H.shape
    # Row-reduce [H | I_m]; pivot columns give a solution basis.
    A = np.concatenate([H.copy(), np.eye(m, dtype=np.uint8)], axis=1)
    piv_cols = []
    r = 0
    for c in range(n):
        rows = np.flatnonzero(A[r:, c]) if r < m else np.array([], dtype=int)
        if rows.size == 0:
            continue
        p = r + int(rows[0])
        if p != r:
            A[[r, p]] = A[[p, r]]
        for k in range(m):
            if k != r and A[k, c]:
                A[k] ^= A[r]
        piv_cols.append(c)
        r += 1
        if r == m:
            break
    # After reduction: row i (< r) reads  e[piv_cols[i]] = (T s)_i  where T = A[:, n:].
    T = A[:, n:]
    R = np.zeros((n, m), dtype=np.uint8)
    for i, c in enumerate(piv_cols):
        R[c, :] = T[i, :]
    return R
