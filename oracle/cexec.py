"""ctypes wrapper of oracle/cexec.c — TEST INFRASTRUCTURE ONLY (tests/,
smoke() and bench.py's CPU baseline).

``OraclePlan(plan)`` accepts any object with the attributes of a flattened
plan (``leaf_labels``, ``leaf_data``, ``leaf_slot``, ``lhs``, ``rhs``,
``out_labels``, ``out_labels_final``, ``n_syndrome_bits``,
``evidence_label_bits``); it does not import the product package.
"""
import ctypes
import os

import numpy as np

from .build import OUT, build

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(OUT):
            build()
        lib = ctypes.CDLL(OUT)
        P = ctypes.POINTER
        i32p, f64p, u8p = P(ctypes.c_int32), P(ctypes.c_double), P(ctypes.c_uint8)
        lib.ox_plan_new.restype = ctypes.c_void_p
        lib.ox_plan_new.argtypes = [ctypes.c_int, i32p, i32p, f64p, i32p, ctypes.c_int, i32p, i32p, i32p, i32p,
                                    ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int, i32p, i32p]
        lib.ox_plan_free.argtypes = [ctypes.c_void_p]
        lib.ox_plan_free.restype = None
        lib.ox_plan_macs.argtypes = [ctypes.c_void_p]
        lib.ox_plan_macs.restype = ctypes.c_double
        lib.ox_decode.argtypes = [ctypes.c_void_p, u8p, ctypes.c_int64, f64p, i32p, u8p, ctypes.c_double,
                                  ctypes.c_int]
        lib.ox_decode.restype = ctypes.c_int
        _lib = lib
    return _lib


def _i32(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
    return a if a.size else np.zeros(1, np.int32)


class OraclePlan:
    def __init__(self, plan):
        lib = load()
        P = ctypes.POINTER
        i32p, f64p = P(ctypes.c_int32), P(ctypes.c_double)
        L = len(plan.leaf_labels)
        ranks = _i32([len(l) for l in plan.leaf_labels])
        labels = _i32([x for l in plan.leaf_labels for x in l])
        data = np.ascontiguousarray(np.concatenate(
            [np.zeros(1 << len(l)) if d is None else np.asarray(d, np.float64)
             for l, d in zip(plan.leaf_labels, plan.leaf_data)]))
        ev = _i32(plan.leaf_slot)
        lhs, rhs = _i32(plan.lhs), _i32(plan.rhs)
        orank = _i32([len(o) for o in plan.out_labels])
        olab = _i32([x for o in plan.out_labels for x in o])
        fin = _i32(plan.out_labels_final)
        sl = sorted(plan.evidence_label_bits.items())
        sll, slb = _i32([a for a, _ in sl]), _i32([b for _, b in sl])
        self._keep = (ranks, labels, data, ev, lhs, rhs, orank, olab, fin, sll, slb)
        c = lambda a, t: a.ctypes.data_as(t)
        self.h = lib.ox_plan_new(L, c(ranks, i32p), c(labels, i32p), c(data, f64p), c(ev, i32p), len(plan.lhs),
                                 c(lhs, i32p), c(rhs, i32p), c(orank, i32p), c(olab, i32p),
                                 len(plan.out_labels_final), c(fin, i32p), plan.n_syndrome_bits, len(sl),
                                 c(sll, i32p), c(slb, i32p))
        if not self.h:
            raise MemoryError("oracle plan allocation failed")
        self.n_out = len(plan.out_labels_final)
        self.n_syn = plan.n_syndrome_bits

    def __del__(self):
        if getattr(self, "h", None):
            load().ox_plan_free(self.h)
            self.h = None

    def macs(self) -> float:
        return load().ox_plan_macs(self.h)

    def decode(self, syndromes, tie_tol=1e-12, nthreads=0):
        lib = load()
        P = ctypes.POINTER
        s = np.ascontiguousarray(np.atleast_2d(np.asarray(syndromes, dtype=np.uint8)))
        B = s.shape[0]
        mar = np.empty((B, 1 << self.n_out))
        cls = np.empty(B, np.int32)
        tie = np.empty(B, np.uint8)
        lib.ox_decode(self.h, s.ctypes.data_as(P(ctypes.c_uint8)), B, mar.ctypes.data_as(P(ctypes.c_double)),
                      cls.ctypes.data_as(P(ctypes.c_int32)), tie.ctypes.data_as(P(ctypes.c_uint8)),
                      float(tie_tol), int(nthreads))
        return mar, cls, tie
