"""The C-ABI library without a GPU: it loads, exports every symbol the
header declares, and rejects malformed input with an error code (never an
abort).  No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.errors import iid_error
from tensorqec import decoders as D


def _compiler_env():
    """Environment for hipcc child processes: without a preloaded sanitizer
    runtime (the sanitized-library run preloads one into this process)."""
    return {k: v for k, v in os.environ.items() if k != "LD_PRELOAD"}

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(REPO, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            out |= set(re.findall(r"\b(tqec_[a-z0-9_]+)\s*\(", txt))
    return out


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    syms = declared_symbols()
    assert "tqec_plan_create" in syms and "tqec_decode_batch" in syms
    for s in sorted(syms):
        assert hasattr(lib, s), s
    # and the Python binding declares a signature for each of them
    bound = {name for name, _, _ in L.EXPORTS}
    assert syms <= bound, syms - bound


def small_plan():
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    p = 0.05 / 3
    ct = D.compile(D.TNMMAP(factorize=False, evidence="leaves", optimizer=D.GreedyMethod()), t,
                   iid_error(p, p, p, 9))
    return ct


def desc_from(plan, **over):
    """Build a PlanDesc (plus keep-alive arrays) from a Plan with overrides."""
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32)) if len(a) else np.zeros(1, np.int32)
    f = dict(
        ranks=i32([len(l) for l in plan.leaf_labels]),
        labels=i32([x for l in plan.leaf_labels for x in l]),
        data=np.ascontiguousarray(np.concatenate([np.zeros(1 << len(l)) if d is None else d
                                                  for l, d in zip(plan.leaf_labels, plan.leaf_data)])),
        ev=i32(plan.leaf_slot), lhs=i32(plan.lhs), rhs=i32(plan.rhs),
        orank=i32([len(o) for o in plan.out_labels]), olab=i32([x for o in plan.out_labels for x in o]),
        out=i32(plan.out_labels_final), n_steps=plan.n_steps, n_leaves=plan.n_leaves,
        n_out=len(plan.out_labels_final), n_syn=plan.n_syndrome_bits, n_vars=200)
    f.update(over)
    P = lambda a, t: a.ctypes.data_as(t)
    d = L.PlanDesc(f["n_vars"], f["n_leaves"], P(f["ranks"], L._i32p), P(f["labels"], L._i32p),
                   P(f["data"], L._f64p), P(f["ev"], L._i32p), f["n_steps"], P(f["lhs"], L._i32p),
                   P(f["rhs"], L._i32p), P(f["orank"], L._i32p), P(f["olab"], L._i32p), f["n_out"],
                   P(f["out"], L._i32p), f["n_syn"], 0, None, None, f.get("tab", 0), f.get("semiring", 0))
    return d, f


def create(d):
    h = ctypes.c_void_p()
    rc = L.load().tqec_plan_create(ctypes.byref(d), ctypes.byref(h))
    if rc == 0:
        L.load().tqec_plan_destroy(h)
    return rc


def test_valid_descriptor_accepted():
    ct = small_plan()
    d, _ = desc_from(ct.plan)
    assert create(d) == 0


def test_malformed_descriptors_rejected():
    ct = small_plan()
    plan = ct.plan
    lib = L.load()
    h = ctypes.c_void_p()
    assert lib.tqec_plan_create(None, ctypes.byref(h)) == L.TQEC_ENULL
    d, _ = desc_from(plan, n_steps=plan.n_steps - 1)
    assert create(d) == L.TQEC_EINVAL
    bad = np.array([x for l in plan.leaf_labels for x in l], np.int32)
    bad[0] = 10_000
    d, _ = desc_from(plan, labels=bad)
    assert create(d) == L.TQEC_EINVAL
    assert "out of range" in L.last_error()
    data = desc_from(plan)[1]["data"].copy()
    data[3] = np.nan
    d, _ = desc_from(plan, data=data)
    assert create(d) == L.TQEC_EINVAL
    data[3] = -0.5
    d, _ = desc_from(plan, data=data)
    assert create(d) == L.TQEC_EINVAL
    assert "non-negative" in L.last_error()
    lhs = np.array(plan.lhs, np.int32)
    lhs[-1] = lhs[0]          # a node consumed twice
    d, _ = desc_from(plan, lhs=lhs)
    assert create(d) == L.TQEC_EINVAL
    olab = [list(o) for o in plan.out_labels]
    # drop a label that is still needed later: remove the first kept label of the first step with >1 label
    for i, o in enumerate(olab[:-1]):
        if len(o) > 1:
            olab[i] = o[1:]
            break
    orank = np.array([len(o) for o in olab], np.int32)
    d, _ = desc_from(plan, olab=np.array([x for o in olab for x in o], np.int32), orank=orank)
    assert create(d) == L.TQEC_EINVAL
    out = np.array(plan.out_labels_final[::-1][:1], np.int32)
    d, _ = desc_from(plan, out=out, n_out=1)
    assert create(d) == L.TQEC_EINVAL
    ev = np.array(plan.leaf_slot, np.int32)
    first_ev = int(np.flatnonzero(ev >= 0)[0])
    ev[first_ev] = 99          # syndrome bit out of range
    d, _ = desc_from(plan, ev=ev)
    assert create(d) == L.TQEC_EINVAL


def test_null_and_argument_errors():
    lib = L.load()
    assert lib.tqec_plan_stats(None, None) == L.TQEC_ENULL
    assert lib.tqec_decode_batch(None, None, 1, 0, None, None, None, None) == L.TQEC_ENULL
    assert lib.tqec_decode_batch_device(None, None, 1, 0, None, None, None, None) == L.TQEC_ENULL
    ct = small_plan()
    h = ct.hip.handle
    assert lib.tqec_decode_batch(h, None, -1, 0, None, None, None, None) == L.TQEC_EINVAL
    assert lib.tqec_decode_batch(h, None, 4, 0, None, None, None, None) == L.TQEC_ENULL
    buf = np.zeros((1, 8), np.uint8)
    assert lib.tqec_decode_batch(h, buf.ctypes.data_as(L._u8p), 1, 7, None, None, None, None) == L.TQEC_EINVAL
    assert lib.tqec_plan_set_tolerances(h, -1.0, 1e-3) == L.TQEC_EINVAL
    assert lib.tqec_plan_set_tolerances(h, float("nan"), 1e-3) == L.TQEC_EINVAL
    lib.tqec_plan_destroy(None)  # no-op
    with pytest.raises(ValueError):
        ct.hip.decode(np.zeros((2, 7), np.uint8))   # wrong syndrome length


def test_stats_mirror_plan():
    ct = small_plan()
    s = ct.hip.stats()
    ps = ct.plan.stats()
    assert s["tc"] == ps["tc"] and 0 < s["dep_tc"] <= ps["tc_dep"] and s["sc"] == ps["sc"]
    assert s["n_out"] == 2 and s["n_steps"] == ct.plan.n_steps
    assert s["plain_ws_entries"] > 0


@pytest.mark.parametrize("tile,threads,tab,fp64,slab,unroll", [(1, 64, True, False, False, 1),
                                                               (4, 128, False, False, False, 4),
                                                               (16, 256, True, False, True, 2),
                                                               (1, 64, False, True, False, 2),
                                                               (8, 128, True, True, True, 1)])
def test_specialised_kernel_source_compiles_for_gfx950(tmp_path, tile, threads, tab, fp64, slab, unroll):
    """The plan-specialised FP32 kernel (TQEC_MODE_SPEC_F32) is generated on
    the host and compiled for gfx950 at first use; check here, without a
    GPU, that the generated source compiles and uses no scratch memory."""
    import shutil
    import subprocess
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    t = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    p = 0.05 / 3
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, iid_error(p, p, p, t.nq))
    src = ct.hip.spec_source(tile, threads, tab, fp64, slab, unroll)
    assert ("tqec_spec_f64" if fp64 else "tqec_spec_f32") in src and f"constexpr int G = {tile}," in src
    assert ("using T = double;" in src) == fp64
    f = tmp_path / "spec.hip"
    f.write_text("#include <hip/hip_runtime.h>\n" + src)
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-c",
                        str(f), "-o", str(tmp_path / "spec.o"), "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=300, env=_compiler_env())
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)
    assert m and int(m.group(1)) == 0
    with pytest.raises(L.TqecError):
        ct.hip.spec_source(3, 64, True)          # not a power of two


def test_fp32_mfma_kernel_source_compiles_for_gfx950(tmp_path):
    """TQEC_SPEC_MFMA32: on a circuit-level DEM plan (d=3, 3 rounds) the
    GEMM-shaped steps are emitted as v_mfma_f32_32x32x2f32 tiles with FP32
    operands (no BF16 rounding), the other steps as scalar code; the kernel
    compiles for gfx950 without scratch; the flag is refused with BF16 MFMA
    steps, on tiles of several syndromes and on the register-resident
    kernel."""
    import shutil
    import subprocess
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    ct = D.compile(D.TNMMAP(mode="spec_f32"), detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001)))
    src = ct.hip.spec_source(1, 256, False, False, False, 4, mfma32=True)
    plain = ct.hip.spec_source(1, 256, False, False, False, 4)
    assert src.count("__builtin_amdgcn_mfma_f32_32x32x2f32") > 0 and "(short)to_bf16(" not in src
    assert "mfma" not in plain
    f = tmp_path / "spec.hip"
    f.write_text("#include <hip/hip_runtime.h>\n" + src)
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-c",
                        str(f), "-o", str(tmp_path / "spec.o"), "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=300, env=_compiler_env())
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)
    assert m and int(m.group(1)) == 0
    with pytest.raises(L.TqecError):
        ct.hip.spec_source(1, 256, False, mfma=True, mfma32=True)
    with pytest.raises(L.TqecError):
        ct.hip.set_spec(4, 256, False, mfma32=True)
    with pytest.raises(L.TqecError):
        ct.hip.set_spec(64, 256, False, reg=True, mfma32=True)


@pytest.mark.parametrize("threads,tab", [(64, False), (256, True)])
def test_register_resident_kernel_source_compiles_for_gfx950(tmp_path, threads, tab):
    """The register-resident variant (one syndrome per quad of lanes, every
    tensor in registers, DPP quad exchanges) compiles for gfx950 without
    scratch on the d=5 plan, with static exponents (no per-step maxima or
    rescales) and with per-step scaling; configurations it cannot take are
    refused."""
    import shutil
    import subprocess
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    t = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    p = 0.05 / 3
    ct = D.compile(D.TNMMAP(mode="fused_f32"), t, iid_error(p, p, p, t.nq))
    for dyn in (False, True):
        src = ct.hip.spec_source(threads // 4, threads, tab, reg=True, dyn=dyn)
        assert "tqec_spec_f32" in src and f"constexpr int G = {threads // 4}," in src
        assert "__builtin_amdgcn_mov_dpp" in src
        body = src[src.index('extern "C" __global__'):]
        steps = body[body.index("{ // step 0"):body.index("float rv[")]
        assert "__syncthreads" not in steps and "__shared__" not in steps   # the step program never synchronises
        # flagged rows are re-contracted in FP64 inside the kernel, by a call of its own
        assert "noinline)) void recheck64(" in src and "recheck64(syn, nf, flagged" in body
        # static exponents: no per-tensor maxima, exponents or rescales between steps
        assert ("lmax" in steps and "ilogbf" in steps) == dyn
        assert ("static exponents" in src) != dyn
        f = tmp_path / "reg.hip"
        f.write_text("#include <hip/hip_runtime.h>\n" + src)
        r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only",
                            "-c", str(f), "-o", str(tmp_path / "reg.o"), "-Xclang", "-target-feature", "-Xclang",
                            "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"],
                           capture_output=True, text=True, timeout=300, env=_compiler_env())
        assert r.returncode == 0, r.stderr[-2000:]
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)
        assert m and int(m.group(1)) == 0
    with pytest.raises(L.TqecError):
        ct.hip.spec_source(threads // 2, threads, tab, reg=True)   # tile must be threads / 4
    # the FP64 form (SPEC_F64): doubles in register pairs, no scaling, no re-check list
    src64 = ct.hip.spec_source(threads // 4, threads, tab, fp64=True, reg=True)
    assert "tqec_spec_f64" in src64 and "ldexpf" not in src64 and "atomicAdd" not in src64
    f64 = tmp_path / "reg64.hip"
    f64.write_text("#include <hip/hip_runtime.h>\n" + src64)
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-c",
                        str(f64), "-o", str(tmp_path / "reg64.o"), "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=300, env=_compiler_env())
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)
    assert m and int(m.group(1)) == 0
    # a plan whose live tensors exceed the register file is refused
    t3 = CSSTannerGraph.from_code(ToricCode(3, 3))
    ct3 = D.compile(D.TNMMAP(mode="fused_f32"), t3, iid_error(p, p, p, t3.nq))
    with pytest.raises(L.TqecError) as e:
        ct3.hip.spec_source(16, 64, True, reg=True)
    assert e.value.code == L.TQEC_EUNSUPPORTED


def test_mixed_kernel_source_compiles_for_gfx950(tmp_path):
    """The mixed-code kernel generated from several plans' specialised code
    (one LDS plan, one slab plan) compiles for gfx950 without scratch."""
    import shutil
    import subprocess
    from tensorqec.codes import SteaneCode, ToricCode
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    plans = []
    for code in (SurfaceCode(3, 3), SteaneCode(), ToricCode(3, 3)):
        t = CSSTannerGraph.from_code(code)
        p = 0.05 / 3
        plans.append(D.compile(D.TNMMAP(mode="spec_f32"), t, iid_error(p, p, p, t.nq)).hip)
    src = L.mixed_spec_source(plans, 128)
    assert "tqec_mixed_spec" in src and "p2_f64" in src and "slabw" in src
    f = tmp_path / "mixed.hip"
    f.write_text("#include <hip/hip_runtime.h>\n" + src)
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-c",
                        str(f), "-o", str(tmp_path / "mixed.o"), "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=600, env=_compiler_env())
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)
    assert m and int(m.group(1)) == 0
