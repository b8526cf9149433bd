"""ctypes binding of ``libtqec_hip.so`` (declared in include/tqec_hip.h).

The product path always goes through this library; there is no CPU
fallback.  Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TQEC_LIB names another build of the same library (the sanitized CPU-test
# build); by default the in-tree libtqec_hip.so.
LIB_PATH = os.environ.get("TQEC_LIB") or os.path.join(_HERE, "libtqec_hip.so")

TQEC_OK = 0
TQEC_EINVAL = -1
TQEC_ENOMEM = -2
TQEC_EDEVICE = -3
TQEC_EUNSUPPORTED = -4
TQEC_ENULL = -5

MODE_PLAIN_F64 = 0
MODE_FUSED_F64 = 1
MODE_FUSED_F32 = 2
MODE_WIDE_F32 = 3
MODE_SPEC_F32 = 4
MODE_SPEC_F64 = 5
MODE_SPEC_BF16 = 6
MODES = {"plain_f64": MODE_PLAIN_F64, "fused_f64": MODE_FUSED_F64, "fused_f32": MODE_FUSED_F32,
         "wide_f32": MODE_WIDE_F32, "spec_f32": MODE_SPEC_F32, "spec_f64": MODE_SPEC_F64,
         "spec_bf16": MODE_SPEC_BF16}
NOALLOC_MAX_ROWS = 1 << 20   # TQEC_NOALLOC_MAX_ROWS

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("n_vars", ctypes.c_int32),
        ("n_leaves", ctypes.c_int32),
        ("leaf_rank", _i32p),
        ("leaf_labels", _i32p),
        ("leaf_data", _f64p),
        ("leaf_evidence", _i32p),
        ("n_steps", ctypes.c_int32),
        ("step_lhs", _i32p),
        ("step_rhs", _i32p),
        ("step_out_rank", _i32p),
        ("step_out_labels", _i32p),
        ("n_out", ctypes.c_int32),
        ("out_labels", _i32p),
        ("n_syndrome_bits", ctypes.c_int32),
        ("n_sliced", ctypes.c_int32),
        ("sliced_label", _i32p),
        ("sliced_bit", _i32p),
        ("tabulate_bits", ctypes.c_int32),
        ("semiring", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("tc", ctypes.c_double),
        ("rw", ctypes.c_double),
        ("sc", ctypes.c_int32),
        ("n_steps", ctypes.c_int32),
        ("n_dep_steps", ctypes.c_int32),
        ("dep_tc", ctypes.c_double),
        ("dep_rw", ctypes.c_double),
        ("plain_ws_entries", ctypes.c_int64),
        ("fused_lds_entries", ctypes.c_int64),
        ("n_out", ctypes.c_int32),
        ("fused_tile_f64", ctypes.c_int32),
        ("fused_tile_f32", ctypes.c_int32),
        ("const_entries", ctypes.c_int64),
        ("table_entries", ctypes.c_int64),
        ("tabulate_bits", ctypes.c_int32),
    ]


class MixedBatch(ctypes.Structure):
    _fields_ = [
        ("plan", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("syndromes", ctypes.c_void_p),
        ("marginals", ctypes.c_void_p),
        ("classes", ctypes.c_void_p),
        ("ties", ctypes.c_void_p),
    ]


class TqecError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tqec error {code}: {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes) of every exported symbol.
EXPORTS = [
    ("tqec_plan_create", ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_plan_destroy", None, [ctypes.c_void_p]),
    ("tqec_plan_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("tqec_decode_batch", ctypes.c_int,
     [ctypes.c_void_p, _u8p, ctypes.c_int64, ctypes.c_int32, _f64p, _i32p, _u8p, ctypes.c_void_p]),
    ("tqec_decode_batch_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("tqec_decode_batch_device_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("tqec_plan_set_tolerances", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    ("tqec_map_decode_batch", ctypes.c_int,
     [ctypes.c_void_p, _u8p, ctypes.c_int64, ctypes.c_int32, _u8p, _f64p, ctypes.c_void_p]),
    ("tqec_map_decode_batch_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("tqec_last_recheck_count", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    ("tqec_plan_autotune", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, _i32p, _i32p]),
    ("tqec_plan_set_tile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("tqec_plan_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("tqec_kernel_time_totals", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_int64)]),
    ("tqec_sampler_create_css", ctypes.c_int,
     [ctypes.c_int32, _f64p, _f64p, _f64p, ctypes.c_int32, _u8p, ctypes.c_int32, _u8p, ctypes.c_int32, _u8p, _u8p,
      ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_sampler_create_flips", ctypes.c_int,
     [ctypes.c_int32, _f64p, ctypes.c_int32, _u8p, ctypes.c_int32, _u8p, ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_sample", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("tqec_sampler_destroy", None, [ctypes.c_void_p]),
    ("tqec_recovery_create_css", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _u8p, _u8p, ctypes.c_int32, _u8p, _u8p,
      ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_recovery_create_dem", ctypes.c_int,
     [ctypes.c_int32, ctypes.c_int32, _u8p, ctypes.c_int32, _u8p, ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_recover", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("tqec_recovery_destroy", None, [ctypes.c_void_p]),
    ("tqec_decode_list_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("tqec_mlp_create", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_mlp_num_params", ctypes.c_int64, [ctypes.c_void_p]),
    ("tqec_mlp_set_params", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("tqec_mlp_get_params", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("tqec_mlp_forward", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    ("tqec_mlp_gradients", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                           ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]),
    ("tqec_mlp_train_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                            ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]),
    ("tqec_mlp_gate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_float,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    ("tqec_mlp_set_inference", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("tqec_mlp_inference", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    ("tqec_mlp_destroy", None, [ctypes.c_void_p]),
    ("tqec_posteriors_from_marginals", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                                       ctypes.c_void_p, ctypes.c_void_p]),
    ("tqec_mixed_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_mixed_set_graph", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("tqec_mixed_set_order", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("tqec_mixed_set_kernel", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("tqec_mixed_spec_source", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("tqec_mixed_plan_config", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _i32p, _i32p]),
    ("tqec_mixed_decode_device", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MixedBatch), ctypes.c_int32,
                                                 ctypes.c_void_p]),
    ("tqec_mixed_last_call", ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_int32,
                                             ctypes.POINTER(ctypes.c_int64)]),
    ("tqec_mixed_destroy", None, [ctypes.c_void_p]),
    ("tqec_plan_spec_source", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("tqec_plan_set_spec", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32]),
    ("tqec_plan_spec_config", ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _i32p, _i32p]),
    ("tqec_plan_set_map_kernel", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("tqec_plan_map_kernel", ctypes.c_int, [ctypes.c_void_p, _i32p, ctypes.c_char_p, ctypes.c_size_t]),
    ("tqec_ler_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    ("tqec_ler_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]),
    ("tqec_ler_destroy", None, [ctypes.c_void_p]),
    ("tqec_last_error", ctypes.c_char_p, []),
]


def load(path: Optional[str] = None):
    """Load the library (once).  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library not built: {p} (run tensorqec.jl_amd/build.py)")
    lib = ctypes.CDLL(p)
    for name, res, args in EXPORTS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    msg = load().tqec_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    if rc != TQEC_OK:
        raise TqecError(rc, last_error())


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class HipPlan:
    """Owner of a ``tqec_plan*`` built from a :class:`planner.Plan`."""

    def __init__(self, plan, tabulate_bits: int = 4, semiring: int = 0):
        lib = load()
        L = plan.n_leaves
        ranks = _i32([len(l) for l in plan.leaf_labels])
        labels = _i32([x for l in plan.leaf_labels for x in l]) if ranks.sum() else np.zeros(1, np.int32)
        data = []
        for t in range(L):
            d = plan.leaf_data[t]
            data.append(np.zeros(1 << len(plan.leaf_labels[t])) if d is None else np.asarray(d, dtype=np.float64))
        data = np.ascontiguousarray(np.concatenate(data))
        ev = _i32(plan.leaf_slot)
        lhs, rhs = _i32(plan.lhs), _i32(plan.rhs)
        orank = _i32([len(o) for o in plan.out_labels])
        olab = _i32([x for o in plan.out_labels for x in o]) if orank.sum() else np.zeros(1, np.int32)
        out = _i32(plan.out_labels_final) if plan.out_labels_final else np.zeros(1, np.int32)
        sl = sorted(plan.evidence_label_bits.items())
        sl_l = _i32([a for a, _ in sl]) if sl else np.zeros(1, np.int32)
        sl_b = _i32([b for _, b in sl]) if sl else np.zeros(1, np.int32)
        nvars = 1 + max([x for l in plan.leaf_labels for x in l] + [x for o in plan.out_labels for x in o] + [0])
        self._keep = (ranks, labels, data, ev, lhs, rhs, orank, olab, out, sl_l, sl_b)
        desc = PlanDesc(nvars, L, _ptr(ranks, _i32p), _ptr(labels, _i32p), _ptr(data, _f64p), _ptr(ev, _i32p),
                        plan.n_steps, _ptr(lhs, _i32p), _ptr(rhs, _i32p), _ptr(orank, _i32p), _ptr(olab, _i32p),
                        len(plan.out_labels_final), _ptr(out, _i32p), plan.n_syndrome_bits, len(sl),
                        _ptr(sl_l, _i32p), _ptr(sl_b, _i32p), int(tabulate_bits), int(semiring))
        h = ctypes.c_void_p()
        check(lib.tqec_plan_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h
        self.n_out = len(plan.out_labels_final)
        self.n_syn = plan.n_syndrome_bits

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().tqec_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def stats(self) -> dict:
        s = Stats()
        check(load().tqec_plan_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def map_decode(self, syndromes: np.ndarray, n_config: int):
        """Max-product plan: (configurations (B, n_config) uint8, log max (B,))."""
        s = np.ascontiguousarray(np.atleast_2d(np.asarray(syndromes, dtype=np.uint8)))
        if s.shape[1] != self.n_syn:
            raise ValueError(f"syndromes must be (B, {self.n_syn}); got {s.shape}")
        B = s.shape[0]
        cfg = np.empty((B, n_config), dtype=np.uint8)
        lp = np.empty(B, dtype=np.float64)
        check(load().tqec_map_decode_batch(self._h, _ptr(s, _u8p), B, int(n_config),
                                           _ptr(cfg, _u8p) if cfg.size else None, _ptr(lp, _f64p), None))
        return cfg, lp

    def map_decode_device(self, d_syn: int, n: int, n_config: int, d_config: int, d_logp: int, stream: int = 0) -> None:
        """Device-pointer max-product decode; enqueues only."""
        check(load().tqec_map_decode_batch_device(self._h, ctypes.c_void_p(d_syn), int(n), int(n_config),
                                                  ctypes.c_void_p(d_config) if d_config else None,
                                                  ctypes.c_void_p(d_logp) if d_logp else None,
                                                  ctypes.c_void_p(stream) if stream else None))

    def set_tolerances(self, tie_tol: float, recheck_tol: float) -> None:
        check(load().tqec_plan_set_tolerances(self._h, float(tie_tol), float(recheck_tol)))

    def decode(self, syndromes: np.ndarray, mode="fused_f32", want_marginals=True) -> Tuple:
        """Host-buffer batch decode: returns (marginals or None, classes, ties)."""
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        s = np.ascontiguousarray(np.asarray(syndromes, dtype=np.uint8))
        if s.ndim == 1:
            s = s[None, :]
        if s.ndim != 2 or s.shape[1] != self.n_syn:
            raise ValueError(f"syndromes must be (B, {self.n_syn}); got {s.shape}")
        B = s.shape[0]
        nout = 1 << self.n_out
        mar = np.empty((B, nout), dtype=np.float64) if want_marginals else None
        cls = np.empty(B, dtype=np.int32)
        tie = np.empty(B, dtype=np.uint8)
        check(load().tqec_decode_batch(self._h, _ptr(s, _u8p), B, m,
                                       _ptr(mar, _f64p) if mar is not None else None,
                                       _ptr(cls, _i32p), _ptr(tie, _u8p), None))
        return mar, cls, tie

    def decode_device(self, d_syn: int, n: int, mode, d_mar: int, d_cls: int, d_tie: int, stream: int = 0) -> None:
        """Device-pointer decode (integers are device addresses); enqueues only."""
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        check(load().tqec_decode_batch_device(self._h, ctypes.c_void_p(d_syn), int(n), m,
                                              ctypes.c_void_p(d_mar) if d_mar else None,
                                              ctypes.c_void_p(d_cls) if d_cls else None,
                                              ctypes.c_void_p(d_tie) if d_tie else None,
                                              ctypes.c_void_p(stream) if stream else None))

    def decode_device_noalloc(self, d_syn: int, n: int, mode, d_mar: int, d_cls: int, d_tie: int,
                              stream: int = 0) -> None:
        """As decode_device, with no device allocation (specialised modes, at
        most NOALLOC_MAX_ROWS rows, the plan prepared on the device by an
        earlier decode_device call in the same mode); enqueues only."""
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        check(load().tqec_decode_batch_device_noalloc(self._h, ctypes.c_void_p(d_syn), int(n), m,
                                                      ctypes.c_void_p(d_mar) if d_mar else None,
                                                      ctypes.c_void_p(d_cls) if d_cls else None,
                                                      ctypes.c_void_p(d_tie) if d_tie else None,
                                                      ctypes.c_void_p(stream) if stream else None))

    def autotune(self, mode, d_syn: int, n: int, stream: int = 0):
        """Pick the fastest fused tile configuration on a device-resident
        trial batch; returns (syndromes per workgroup, threads)."""
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        tile, thr = ctypes.c_int32(), ctypes.c_int32()
        check(load().tqec_plan_autotune(self._h, m, ctypes.c_void_p(d_syn), int(n),
                                        ctypes.c_void_p(stream) if stream else None,
                                        ctypes.byref(tile), ctypes.byref(thr)))
        return tile.value, thr.value

    def set_tile(self, mode, tile: int, threads: int) -> None:
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        check(load().tqec_plan_set_tile(self._h, m, int(tile), int(threads)))

    def set_spec(self, tile: int, threads: int, tables_in_lds: bool, waves: int = 0, slab: bool = False,
                 unroll: int = 1, reg: bool = False, dyn: bool = False, mfma32: bool = False) -> None:
        """Pin the SPEC_F32 configuration; `dyn` (register-resident only):
        per-step scaling, bit-identical to the generic FP32 kernel, instead
        of plan-time static exponents; `mfma32` (tiles of one syndrome):
        GEMM-shaped steps on FP32 MFMA."""
        flags = (int(bool(tables_in_lds)) | (4 if slab else 0) | (16 if reg else 0) | (32 if dyn else 0) |
                 (64 if mfma32 else 0))
        check(load().tqec_plan_set_spec(self._h, int(tile), int(threads), flags, int(waves) | (int(unroll) << 8)))

    def set_map_kernel(self, specialised: bool) -> None:
        """Max-product plans: plan-specialised (default) or generic kernel."""
        check(load().tqec_plan_set_map_kernel(self._h, 1 if specialised else 0))

    def map_kernel(self):
        """(True if the plan-specialised max-product kernel is in use, reason
        it was given up or '')."""
        v = ctypes.c_int32()
        buf = ctypes.create_string_buffer(512)
        check(load().tqec_plan_map_kernel(self._h, ctypes.byref(v), buf, 512))
        return bool(v.value), buf.value.decode()

    def spec_config(self) -> dict:
        v = [ctypes.c_int32() for _ in range(4)]
        check(load().tqec_plan_spec_config(self._h, *[ctypes.byref(x) for x in v]))
        return {"tile": v[0].value, "threads": v[1].value, "tables_in_lds": bool(v[2].value & 1),
                "slab": bool(v[2].value & 4), "reg": bool(v[2].value & 16), "dyn": bool(v[2].value & 32),
                "mfma32": bool(v[2].value & 64),
                "waves": v[3].value & 0xff,
                "unroll": max(1, v[3].value >> 8)}

    def spec_source(self, tile: int, threads: int, tables_in_lds: bool = True, fp64: bool = False,
                    slab: bool = False, unroll: int = 1, mfma: bool = False, reg: bool = False,
                    dyn: bool = False, mfma32: bool = False) -> str:
        """HIP source of the plan-specialised kernel for a tile choice."""
        flags = (int(bool(tables_in_lds)) | (2 if fp64 else 0) | (4 if slab else 0) | (8 if mfma else 0)
                 | (16 if reg else 0) | (32 if dyn else 0) | (64 if mfma32 else 0) | (int(unroll) << 8))
        n = ctypes.c_size_t()
        check(load().tqec_plan_spec_source(self._h, int(tile), int(threads), flags, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(load().tqec_plan_spec_source(self._h, int(tile), int(threads), flags, buf, n.value + 1,
                                           ctypes.byref(n)))
        return buf.value.decode()

    def set_timing(self, enable: bool) -> None:
        check(load().tqec_plan_set_timing(self._h, 1 if enable else 0))

    def kernel_time_totals(self, stream: int = 0):
        """(main kernel ms, FP64 re-check ms, calls) summed over the calls on
        ``stream`` since the previous query (waits for their events)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        check(load().tqec_kernel_time_totals(self._h, ctypes.c_void_p(stream) if stream else None,
                                             ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def decode_list_device(self, d_syn: int, d_list: int, d_count: int, n_max: int, mode, d_mar: int,
                           d_cls: int, d_tie: int, stream: int = 0) -> None:
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        check(load().tqec_decode_list_device(self._h, ctypes.c_void_p(d_syn), ctypes.c_void_p(d_list),
                                             ctypes.c_void_p(d_count), int(n_max), m,
                                             ctypes.c_void_p(d_mar) if d_mar else None,
                                             ctypes.c_void_p(d_cls) if d_cls else None,
                                             ctypes.c_void_p(d_tie) if d_tie else None,
                                             ctypes.c_void_p(stream) if stream else None))

    def last_recheck_count(self, stream: int = 0) -> int:
        c = ctypes.c_int64()
        check(load().tqec_last_recheck_count(self._h, ctypes.c_void_p(stream) if stream else None, ctypes.byref(c)))
        return c.value


def _u8(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint8))
    return a if a.size else np.zeros(1, np.uint8)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _vp(x):
    return ctypes.c_void_p(x) if x else None


class HipSampler:
    """Device sampler (``tqec_sampler``): bit-identical to
    :func:`tensorqec.errors.sample_depolarizing` / :func:`sample_flips`."""

    def __init__(self, handle, nsyn, nvar, kind):
        self._h, self.nsyn, self.nvar, self.kind = handle, nsyn, nvar, kind

    @classmethod
    def css(cls, tanner, em, lx, lz):
        lib = load()
        h = ctypes.c_void_p()
        px, py, pz = _f64(em.px), _f64(em.py), _f64(em.pz)
        Hx, Hz, LX, LZ = _u8(tanner.stgx.H), _u8(tanner.stgz.H), _u8(lx), _u8(lz)
        check(lib.tqec_sampler_create_css(tanner.nq, _ptr(px, _f64p), _ptr(py, _f64p), _ptr(pz, _f64p),
                                          tanner.stgx.ns, _ptr(Hx, _u8p), tanner.stgz.ns, _ptr(Hz, _u8p),
                                          int(lx.shape[0]), _ptr(LX, _u8p), _ptr(LZ, _u8p), ctypes.byref(h)))
        return cls(h, tanner.ns, tanner.nq, "css")

    @classmethod
    def flips(cls, p, D, Lm):
        lib = load()
        h = ctypes.c_void_p()
        p, D, Lm = _f64(p), _u8(D), _u8(Lm)
        nd = D.shape[0] if D.ndim == 2 else 0
        nl = Lm.shape[0] if Lm.ndim == 2 else 0
        check(lib.tqec_sampler_create_flips(len(p), _ptr(p, _f64p), nd, _ptr(D, _u8p), nl, _ptr(Lm, _u8p),
                                            ctypes.byref(h)))
        return cls(h, nd, len(p), "flips")

    def sample(self, seed: int, first: int, n: int, d_syn: int, d_cls: int = 0, d_err: int = 0, stream: int = 0):
        check(load().tqec_sample(self._h, int(seed), int(first), int(n), _vp(d_syn), _vp(d_cls), _vp(d_err),
                                 _vp(stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().tqec_sampler_destroy(h)
            except Exception:
                pass
            self._h = None


class HipRecovery:
    """Device error-pattern recovery (``tqec_recovery``)."""

    def __init__(self, handle, nvar, kind):
        self._h, self.nvar, self.kind = handle, nvar, kind

    @classmethod
    def css(cls, tanner, Rx, Rz, lx, lz):
        lib = load()
        h = ctypes.c_void_p()
        Rx, Rz, LX, LZ = _u8(Rx), _u8(Rz), _u8(lx), _u8(lz)
        check(lib.tqec_recovery_create_css(tanner.nq, tanner.stgx.ns, tanner.stgz.ns, _ptr(Rx, _u8p),
                                           _ptr(Rz, _u8p), int(lx.shape[0]), _ptr(LX, _u8p), _ptr(LZ, _u8p),
                                           ctypes.byref(h)))
        return cls(h, tanner.nq, "css")

    @classmethod
    def dem(cls, R, l2q_mask):
        lib = load()
        h = ctypes.c_void_p()
        R, M = _u8(R), _u8(l2q_mask)
        nl = M.shape[0] if M.ndim == 2 else 0
        check(lib.tqec_recovery_create_dem(R.shape[0], R.shape[1], _ptr(R, _u8p), nl, _ptr(M, _u8p),
                                           ctypes.byref(h)))
        return cls(h, R.shape[0], "dem")

    def recover(self, d_syn: int, d_cls: int, n: int, d_a: int, d_b: int = 0, stream: int = 0):
        check(load().tqec_recover(self._h, _vp(d_syn), _vp(d_cls), int(n), _vp(d_a), _vp(d_b), _vp(stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().tqec_recovery_destroy(h)
            except Exception:
                pass
            self._h = None


class HipMixed:
    """Owner of a ``tqec_mixed*``: batches of several compiled plans decoded
    by one persistent launch (FP32 with FP64 re-check), optionally replayed
    from a captured HIP graph."""

    def __init__(self, plans, max_batches: int, max_rows: int, graph: bool = True):
        lib = load()
        self._plans = list(plans)                 # keep the plans alive
        arr = (ctypes.c_void_p * len(self._plans))(*[p.handle.value for p in self._plans])
        h = ctypes.c_void_p()
        check(lib.tqec_mixed_create(arr, len(self._plans), int(max_batches), int(max_rows), ctypes.byref(h)))
        self._h = h
        self.max_batches = int(max_batches)
        self.n_last = 0
        self.set_graph(graph)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().tqec_mixed_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_graph(self, enable: bool) -> None:
        check(load().tqec_mixed_set_graph(self._h, 1 if enable else 0))

    def set_kernel(self, specialised: bool) -> None:
        """Generic interpreter kernel, or one kernel generated from the plans'
        specialised code (compiled here, seconds)."""
        check(load().tqec_mixed_set_kernel(self._h, 1 if specialised else 0))

    def set_order(self, cheapest_first: bool) -> None:
        check(load().tqec_mixed_set_order(self._h, 1 if cheapest_first else 0))

    def plan_config(self, i: int):
        """(tile, slab flags) of plan i: bit 0 FP32 slab, bit 1 FP64 slab."""
        t, sl = ctypes.c_int32(), ctypes.c_int32()
        check(load().tqec_mixed_plan_config(self._h, int(i), ctypes.byref(t), ctypes.byref(sl)))
        return t.value, sl.value

    def decode_device(self, batches, stream: int = 0) -> None:
        """``batches``: sequence of (plan index, n, d_syn, d_mar, d_cls, d_tie)
        with device addresses (0 for an absent output).  Enqueues only."""
        tab = (MixedBatch * max(1, len(batches)))()
        for i, (pi, n, syn, mar, cls, tie) in enumerate(batches):
            tab[i] = MixedBatch(int(pi), 0, int(n), syn or None, mar or None, cls or None, tie or None)
        check(load().tqec_mixed_decode_device(self._h, tab, len(batches), ctypes.c_void_p(stream) if stream else None))
        self.n_last = len(batches)

    def last_call(self):
        """(per-batch latency in microseconds, FP64 re-check count) of the
        last call; waits for it."""
        us = np.zeros(max(1, self.n_last), dtype=np.float64)
        rc = ctypes.c_int64()
        check(load().tqec_mixed_last_call(self._h, _ptr(us, _f64p), self.n_last, ctypes.byref(rc)))
        return us[:self.n_last], rc.value


def mixed_spec_source(plans, threads: int = 128) -> str:
    """Source of the mixed-code kernel generated for these HipPlans."""
    arr = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
    n = ctypes.c_size_t()
    check(load().tqec_mixed_spec_source(arr, len(plans), int(threads), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(load().tqec_mixed_spec_source(arr, len(plans), int(threads), buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


class HipLER:
    """Logical-error-rate loop on the device (``tqec_ler``): sample, decode,
    recover, check against the logical operators, count."""

    def __init__(self, plan: "HipPlan", sampler: "HipSampler", recovery: "HipRecovery", batch: int = 65536):
        self._keep = (plan, sampler, recovery)
        h = ctypes.c_void_p()
        check(load().tqec_ler_create(plan.handle, sampler._h, recovery._h, int(batch), ctypes.byref(h)))
        self._h = h

    def run(self, mode, seed: int, first: int, n: int, stream: int = 0) -> dict:
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        c = (ctypes.c_int64 * 4)()
        check(load().tqec_ler_run(self._h, m, int(seed), int(first), int(n), c, _vp(stream)))
        return {"samples": int(n), "logical_errors": c[0], "x_half": c[1], "z_half": c[2], "inconsistent": c[3]}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().tqec_ler_destroy(h)
            except Exception:
                pass
            self._h = None
