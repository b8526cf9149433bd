"""Stim-format circuits and detector-error-model extraction.

* :func:`parse_stim` reads the stim text subset used by the configs (gates,
  resets, measurements, noise channels, DETECTOR / OBSERVABLE_INCLUDE,
  REPEAT blocks, coordinates ignored) into a flat instruction list with
  absolute measurement indices.
* :func:`detector_error_model` propagates every fault mechanism as a Pauli
  frame through the Clifford circuit (all frames at once, one row per fault)
  and merges mechanisms with identical symptoms, as
  src/decoding/dem.jl:22-148 does one fault at a time:
    - X_ERROR / Y_ERROR / Z_ERROR(p): one mechanism of probability p;
    - DEPOLARIZE1(p): X, Y, Z as independent mechanisms of probability
      q = (1 - sqrt(1 - 4p/3)) / 2 each (dem.jl:63-66); when one of them has
      no symptom the other two coincide and merge to 2p/3 (dem.jl:53-62);
    - DEPOLARIZE2(p): the 15 two-qubit Paulis as independent mechanisms of
      probability q = (1 - (1 - 16p/15)^(1/8)) / 2 each (the same
      decomposition for 4^2 - 1 Paulis; the reference rejects this channel,
      dem.jl:43);
    - identical symptom sets merge as p <- p (1 - q) + q (1 - p)
      (dem.jl:137-148); mechanisms without symptoms are dropped.
  Detectors are numbered 1..nd in order; observables follow as nd+1.. (the
  numbering of parse_dem_string).
* :func:`rotated_memory_z` writes the stim text of a rotated surface code
  memory-Z experiment with circuit-level noise.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

import numpy as np

from .codes import CSSTannerGraph, SurfaceCode
from .dem import DetectorErrorModel
from .gf2 import logical_operator

_GATES_1Q = {"H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "C_XYZ", "C_ZYX", "X", "Y", "Z", "I", "H_YZ", "H_XY"}
_GATES_2Q = {"CX", "CNOT", "ZCX", "CZ", "ZCZ", "CY", "ZCY", "SWAP"}
_MEAS = {"M": "Z", "MZ": "Z", "MX": "X", "MY": "Y", "MR": "Z", "MRZ": "Z", "MRX": "X", "MRY": "Y"}
_RESETS = {"R", "RZ", "RX", "RY"}
_NOISE = {"X_ERROR", "Y_ERROR", "Z_ERROR", "DEPOLARIZE1", "DEPOLARIZE2"}
_IGNORED = {"TICK", "QUBIT_COORDS", "SHIFT_COORDS"}


def _tokens(line: str):
    line = line.split("#", 1)[0].strip()
    if not line:
        return None
    head, _, rest = line.partition(" ")
    name, args = head, []
    if "(" in head:
        name = head[:head.index("(")]
        inner = head[head.index("(") + 1:]
        # arguments may contain spaces: re-join until ')'
        while ")" not in inner and rest:
            a, _, rest = rest.partition(" ")
            inner += a
        args = [float(x) for x in inner[:inner.index(")")].split(",") if x.strip()]
    return name.upper(), args, rest.split()


def parse_stim(text: str) -> List[Tuple[str, List[float], List]]:
    """Flat instruction list: (name, args, targets).  Measurement records of
    DETECTOR / OBSERVABLE_INCLUDE become absolute measurement indices."""
    lines = text.split("\n")
    out: List[Tuple[str, List[float], List]] = []
    nmeas = [0]

    def block(i: int, reps: int) -> int:
        start = i
        end = i
        for _ in range(reps):
            i = start
            while i < len(lines):
                raw = lines[i].strip()
                if raw.startswith("}"):
                    end = i + 1
                    break
                tok = _tokens(lines[i])
                i += 1
                if tok is None:
                    continue
                name, args, tg = tok
                if name == "REPEAT":
                    n = int(tg[0])
                    i = block(i, n)
                    continue
                if name in _IGNORED:
                    continue
                if name in ("DETECTOR", "OBSERVABLE_INCLUDE"):
                    recs = []
                    for t in tg:
                        if not t.startswith("rec[-"):
                            raise ValueError(f"unsupported target {t}")
                        recs.append(nmeas[0] - int(t[5:-1]))
                    out.append((name, args, recs))
                    continue
                qs = [int(t) for t in tg]
                if name in _MEAS:
                    for q in qs:
                        out.append((name, args, [q]))
                        nmeas[0] += 1
                    continue
                if name in _GATES_1Q or name in _GATES_2Q or name in _RESETS or name in _NOISE:
                    out.append((name, args, qs))
                    continue
                raise ValueError(f"unsupported stim instruction {name}")
            else:
                end = i
        return end

    block(0, 1)
    return out


def _pauli_components(name: str, p: float, qs: List[int]):
    """(probability, [(qubit, x, z), ...]) components of a noise channel."""
    if name == "X_ERROR":
        return [(p, [(q, 1, 0)]) for q in qs]
    if name == "Y_ERROR":
        return [(p, [(q, 1, 1)]) for q in qs]
    if name == "Z_ERROR":
        return [(p, [(q, 0, 1)]) for q in qs]
    if name == "DEPOLARIZE1":
        if p > 0.75:
            raise ValueError("single-qubit depolarizing probability above 3/4 cannot be decomposed")
        q1 = (1 - math.sqrt(1 - 4 * p / 3)) / 2
        comps = []
        for q in qs:
            for x, z in ((1, 0), (1, 1), (0, 1)):
                comps.append((q1, [(q, x, z)]))
        return comps
    if name == "DEPOLARIZE2":
        if p > 15 / 16:
            raise ValueError("two-qubit depolarizing probability above 15/16 cannot be decomposed")
        q2 = (1 - (1 - 16 * p / 15) ** (1 / 8)) / 2
        comps = []
        for a, b in zip(qs[0::2], qs[1::2]):
            for pa in range(4):
                for pb in range(4):
                    if pa == 0 and pb == 0:
                        continue
                    xa, za = pa & 1, pa >> 1
                    xb, zb = pb & 1, pb >> 1
                    comps.append((q2, [(a, xa, za), (b, xb, zb)]))
        return comps
    raise ValueError(name)


def detector_error_model(instrs) -> DetectorErrorModel:
    """Forward Pauli-frame analysis of every fault (see the module notes)."""
    nq = 1 + max([t for name, _, tg in instrs if name not in ("DETECTOR", "OBSERVABLE_INCLUDE") for t in tg]
                 + [0])
    # pass 1: count faults
    nf = sum(len(_pauli_components(name, args[0], tg)) for name, args, tg in instrs if name in _NOISE)
    X = np.zeros((nf, nq), dtype=bool)
    Z = np.zeros((nf, nq), dtype=bool)
    probs = np.zeros(nf)
    meas: List[np.ndarray] = []
    dets: List[np.ndarray] = []
    obs: Dict[int, np.ndarray] = {}
    f = 0
    for name, args, tg in instrs:
        if name in _NOISE:
            for pr, paulis in _pauli_components(name, args[0], tg):
                for q, x, z in paulis:
                    X[f, q] ^= bool(x)
                    Z[f, q] ^= bool(z)
                probs[f] = pr
                f += 1
        elif name in ("H",):
            for q in tg:
                X[:, q], Z[:, q] = Z[:, q].copy(), X[:, q].copy()
        elif name in ("S", "S_DAG"):
            for q in tg:
                Z[:, q] ^= X[:, q]
        elif name in ("SQRT_X", "SQRT_X_DAG"):
            for q in tg:
                X[:, q] ^= Z[:, q]
        elif name == "H_YZ":          # Y <-> Z, X -> -X
            for q in tg:
                X[:, q] ^= Z[:, q]
        elif name == "H_XY":          # X <-> Y, Z -> -Z
            for q in tg:
                Z[:, q] ^= X[:, q]
        elif name == "C_XYZ":         # X -> Y -> Z -> X
            for q in tg:
                x, z = X[:, q].copy(), Z[:, q].copy()
                X[:, q], Z[:, q] = x ^ z, x
        elif name == "C_ZYX":         # Z -> Y -> X -> Z
            for q in tg:
                x, z = X[:, q].copy(), Z[:, q].copy()
                X[:, q], Z[:, q] = z, x ^ z
        elif name in ("X", "Y", "Z", "I"):
            pass
        elif name in ("CX", "CNOT", "ZCX"):
            for c, t in zip(tg[0::2], tg[1::2]):
                X[:, t] ^= X[:, c]
                Z[:, c] ^= Z[:, t]
        elif name in ("CZ", "ZCZ"):
            for a, b in zip(tg[0::2], tg[1::2]):
                Z[:, a] ^= X[:, b]
                Z[:, b] ^= X[:, a]
        elif name in ("CY", "ZCY"):   # CY = S_t CX S_t^dag
            for c, t in zip(tg[0::2], tg[1::2]):
                Z[:, t] ^= X[:, t]
                X[:, t] ^= X[:, c]
                Z[:, c] ^= Z[:, t]
                Z[:, t] ^= X[:, t]
        elif name == "SWAP":
            for a, b in zip(tg[0::2], tg[1::2]):
                X[:, [a, b]] = X[:, [b, a]]
                Z[:, [a, b]] = Z[:, [b, a]]
        elif name in _MEAS:
            q = tg[0]
            basis = _MEAS[name]
            flip = X[:, q].copy() if basis == "Z" else (Z[:, q].copy() if basis == "X" else X[:, q] ^ Z[:, q])
            meas.append(flip)
            if name.startswith("MR"):
                X[:, q] = False
                Z[:, q] = False
        elif name in _RESETS:
            for q in tg:
                X[:, q] = False
                Z[:, q] = False
        elif name == "DETECTOR":
            v = np.zeros(nf, dtype=bool)
            for m in tg:
                v ^= meas[m]
            dets.append(v)
        elif name == "OBSERVABLE_INCLUDE":
            k = int(args[0]) if args else 0
            v = obs.get(k, np.zeros(nf, dtype=bool))
            for m in tg:
                v = v ^ meas[m]
            obs[k] = v
        else:
            raise ValueError(name)
    nd = len(dets)
    nobs = (max(obs) + 1) if obs else 0
    S = np.zeros((nf, nd + nobs), dtype=bool)
    for i, v in enumerate(dets):
        S[:, i] = v
    for k, v in obs.items():
        S[:, nd + k] = v
    merged: Dict[bytes, float] = {}
    keys: Dict[bytes, Tuple[int, ...]] = {}
    for r in range(nf):
        row = S[r]
        if not row.any():
            continue
        key = np.packbits(row).tobytes()
        if key in merged:
            a = merged[key]
            merged[key] = (1 - a) * probs[r] + a * (1 - probs[r])
        else:
            merged[key] = probs[r]
            keys[key] = tuple(int(i) + 1 for i in np.flatnonzero(row))
    items = sorted(((keys[k], merged[k]) for k in merged), key=lambda kv: kv[0])
    rates = np.array([p for _, p in items])
    flipped = [list(s) for s, _ in items]
    return DetectorErrorModel(rates, flipped, list(range(1, nd + 1)), list(range(nd + 1, nd + nobs + 1)))


def detector_error_model_from_stim(text: str) -> DetectorErrorModel:
    return detector_error_model(parse_stim(text))


def to_dem_text(dem: DetectorErrorModel) -> str:
    """stim ``.dem`` text of a DEM (0-based D/L ids), readable by parse_dem_string."""
    nd = len(dem.detector_list)
    lines = []
    for p, f in zip(dem.error_rates, dem.flipped_detectors):
        toks = [f"D{x - 1}" for x in f if x <= nd] + [f"L{x - nd - 1}" for x in f if x > nd]
        lines.append(f"error({float(p)!r}) " + " ".join(toks))
    return "\n".join(lines) + "\n"


def _stabilizer_schedule(d: int):
    """For each stabilizer of SurfaceCode(d, d) (reference order): its type
    and its data qubits in the order of the four CNOT layers (None where the
    weight-2 boundary check has no qubit)."""
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))

    def rc(q):
        return divmod(q, d)  # (row, col), 0-based

    sched = []
    for typ, checks in (("X", t.stgx.s2q), ("Z", t.stgz.s2q)):
        for c in checks:
            rows = [rc(q)[0] for q in c]
            cols = [rc(q)[1] for q in c]
            if len(c) == 4:
                i, j = min(rows), min(cols)
            else:
                # virtual plaquette of a boundary check
                if len(set(cols)) == 1:          # vertical pair on the left or right edge
                    i = min(rows)
                    j = cols[0] if cols[0] == d - 1 else cols[0] - 1
                else:                            # horizontal pair on the top or bottom edge
                    j = min(cols)
                    i = rows[0] - 1 if rows[0] == 0 else rows[0]
            pos = {}
            for q in c:
                r_, c_ = rc(q)
                pos[(r_ - i, c_ - j)] = q
            nw, ne, sw, se = pos.get((0, 0)), pos.get((0, 1)), pos.get((1, 0)), pos.get((1, 1))
            order = [nw, ne, sw, se] if typ == "X" else [nw, sw, ne, se]
            sched.append((typ, order))
    return t, sched


def rotated_memory_z(d: int, rounds: int, p: float) -> str:
    """Stim text of a rotated surface code memory-Z experiment.

    Data qubits 0..d^2-1 (row-major), one ancilla per stabilizer in the
    reference's stabilizer order.  Noise: DEPOLARIZE1(p) after every H,
    DEPOLARIZE2(p) after every CNOT layer, X_ERROR(p) after every reset
    (including the reset of MR) and before every measurement.  Detectors:
    Z checks in round 1, every check from round 2 on (this round xor the
    previous), and each Z check against the final data measurements; the
    observable is a Z logical of the data measurements."""
    t, sched = _stabilizer_schedule(d)
    lx, lz = logical_operator(t)
    nd = d * d
    anc = [nd + i for i in range(len(sched))]
    xanc = [a for a, (typ, _) in zip(anc, sched) if typ == "X"]
    L: List[str] = []
    data = list(range(nd))
    L.append("R " + " ".join(map(str, data + anc)))
    L.append(f"X_ERROR({p}) " + " ".join(map(str, data + anc)))
    L.append("TICK")
    nanc = len(anc)
    for r in range(rounds):
        L.append("H " + " ".join(map(str, xanc)))
        L.append(f"DEPOLARIZE1({p}) " + " ".join(map(str, xanc)))
        L.append("TICK")
        for layer in range(4):
            pairs = []
            used = set()
            for a, (typ, order) in zip(anc, sched):
                q = order[layer]
                if q is None:
                    continue
                pair = (a, q) if typ == "X" else (q, a)
                if a in used or q in used:
                    raise AssertionError("CNOT schedule collision")
                used.update(pair)
                pairs.extend(pair)
            L.append("CX " + " ".join(map(str, pairs)))
            L.append(f"DEPOLARIZE2({p}) " + " ".join(map(str, pairs)))
            L.append("TICK")
        L.append("H " + " ".join(map(str, xanc)))
        L.append(f"DEPOLARIZE1({p}) " + " ".join(map(str, xanc)))
        L.append("TICK")
        L.append(f"X_ERROR({p}) " + " ".join(map(str, anc)))
        L.append("MR " + " ".join(map(str, anc)))
        L.append(f"X_ERROR({p}) " + " ".join(map(str, anc)))
        for k, (a, (typ, _)) in enumerate(zip(anc, sched)):
            back = nanc - k
            if r == 0:
                if typ == "Z":
                    L.append(f"DETECTOR rec[-{back}]")
            else:
                L.append(f"DETECTOR rec[-{back}] rec[-{back + nanc}]")
        L.append("TICK")
    L.append(f"X_ERROR({p}) " + " ".join(map(str, data)))
    L.append("M " + " ".join(map(str, data)))
    # final Z-check detectors against the last ancilla round
    for k, (a, (typ, order)) in enumerate(zip(anc, sched)):
        if typ != "Z":
            continue
        recs = [nd - q for q in order if q is not None]
        back = nd + (nanc - k)
        L.append("DETECTOR " + " ".join(f"rec[-{x}]" for x in recs) + f" rec[-{back}]")
    L.append("OBSERVABLE_INCLUDE(0) " + " ".join(f"rec[-{nd - int(q)}]" for q in np.flatnonzero(lz[0])))
    return "\n".join(L) + "\n"
