"""Throughput of TNMMAP decoding on one MI355X.

Workload: rotated surface code d=7 (49 data qubits, 48 stabilizers), code-
capacity depolarizing noise p=0.05 (X, Y, Z each p/3), 65,536 Philox-seeded
syndromes per step, already resident in HBM.  One step decodes one batch:
marginals P(syndrome, class) for the 4 logical classes, the arg-max class and
the tie flag per syndrome.

The reported value is the chosen executor (FP32 with per-tensor scaling and
FP64 re-check of low-margin syndromes, on the chain-factorized network; the
plan-specialised kernel, register-resident where the plan fits) over K
timed steps, each on its own batch, after W warm-up steps cycling over 20
other batches.  The kernel duration behind `roofline` comes from a second,
instrumented pass over the same K batches (per-launch HIP events on the
launch stream).  The plain FP64 executor (one kernel per step, reference
network with one-hot evidence) is timed on the same batches; both logical
error rates are reported with 95% Wilson intervals.  The CPU baseline is the
oracle's C FP64 per-syndrome executor on the reference network, timed on a
bounded sample on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import gc
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "tensorqec.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "syndromes/s, d=7 surface code (chosen vs plain FP64 TN, same GPU, same LER)"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def wilson(k, n, z=1.96):
    if n == 0:
        return (0.0, 0.0)
    p = k / n
    den = 1 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return (max(0.0, c - h), min(1.0, c + h))


def true_classes(ep, lx, lz):
    import numpy as np
    k = lx.shape[0]
    cls = np.zeros(ep.xerror.shape[0], dtype=np.int64)
    for i in range(k):
        cls |= ((ep.zerror.astype(np.int64) @ lx[i].astype(np.int64)) & 1) << i
        cls |= ((ep.xerror.astype(np.int64) @ lz[i].astype(np.int64)) & 1) << (k + i)
    return cls


STREAM_CODES = (("surface3", ("SurfaceCode", 3)), ("surface5", ("SurfaceCode", 5)), ("surface7", ("SurfaceCode", 7)),
                ("surface9", ("SurfaceCode", 9)), ("toric3", ("ToricCode", 3)), ("toric4", ("ToricCode", 4)),
                ("steane", ("SteaneCode", 0)), ("shor", ("ShorCode", 0)))


def measured_traffic(mode, plan_cfg, batch):
    """Calibrated HBM bytes per launch of the chosen kernel, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when that summary
    holds the same plan (tree seed, width target, tabulation depth) with the
    same executor and table placement at this batch; otherwise None.  The
    traffic of these kernels is their syndromes, outputs and tables: it does
    not depend on the tile, threads, register cap or unroll."""
    import re
    if batch != 65536 or mode != "spec_f32" or plan_cfg.get("sc_target") is not None:
        return None
    if plan_cfg.get("reg"):
        path = os.path.join(REPO, "profiles", "spec_f32_reg_d7_p005_counters.md")
        if plan_cfg.get("tables_in_lds") or not os.path.exists(path):
            return None
        m = re.search(r"seed=%d tab=%d: fetch ([0-9.]+) MB \+ write ([0-9.]+) MB"
                      % (plan_cfg.get("seed", -1), plan_cfg.get("tabulate_bits", -1)), open(path).read())
    else:
        path = os.path.join(REPO, "profiles", "spec_f32_d7_p005_counters.md")
        if (plan_cfg.get("seed"), plan_cfg.get("tabulate_bits"), plan_cfg.get("slab")) != (3, 3, False) or \
                not os.path.exists(path):
            return None
        m = re.search(r"fetch ([0-9.]+) MB \+ write ([0-9.]+) MB", open(path).read())
    return (float(m.group(1)) + float(m.group(2))) * 1e6 if m else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def stream_leg(torch, dev, seed, calls=6, warmup=1, window=48, pool=8192, codes=STREAM_CODES, compare=True):
    """Uneven-size batched stream: rotated surface codes d=3/5/7/9, toric codes
    L=3/4, Steane and Shor codes, each under code-capacity depolarizing noise
    with p drawn uniformly from [0.01, 0.1]; every call decodes a window of
    batches (code drawn uniformly, size log-uniform in [1, 4096]) in one
    persistent launch of the mixed-code executor (kernel generated from every
    plan's autotuned specialised code), replayed from a HIP graph.  The
    first timed window is decoded again, batch by batch, by the exact FP64
    executor, and its marginals and classes are compared.  With `codes`
    including toric L=5 (`--only stream_full`) its batches join the stream
    too (1.3 G multiply-adds per syndrome, so they dominate the windows;
    `compare` off skips the generic-kernel and one-launch-per-batch passes,
    which would take minutes on them); L=6 needs ~2^39 per syndrome."""
    import numpy as np
    from tensorqec import codes as C
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.codes import CSSTannerGraph
    from tensorqec.errors import iid_error
    from tensorqec.gf2 import logical_operator

    rng = np.random.default_rng(seed)
    cuda_stream = torch.cuda.Stream()
    sp = cuda_stream.cuda_stream
    t0 = time.time()
    plans, info = [], []
    for name, (cls_name, size) in codes:
        code = getattr(C, cls_name)(size, size) if size else getattr(C, cls_name)()
        t = CSSTannerGraph.from_code(code)
        p = float(rng.uniform(0.01, 0.1))
        em = iid_error(p / 3, p / 3, p / 3, t.nq)
        if name == "toric5":
            # nested dissection: the narrowest exact contraction found (width 2^21)
            ct = D.compile(D.TNMMAP(mode="fused_f32", optimizer=D.Dissection(), tabulate_bits=0), t, em)
        else:
            opt = D.TreeSA(seed=1, ntrials=8, sweeps=1000, restarts=8) if name == "toric4" else D.TreeSA(seed=1)
            ct = D.compile(D.TNMMAP(mode="fused_f32", optimizer=opt), t, em)
        lx, lz = logical_operator(t)
        smp = L.HipSampler.css(t, em, lx, lz)
        S = torch.empty((pool, t.ns), dtype=torch.uint8, device=dev)
        truth = torch.empty(pool, dtype=torch.int32, device=dev)
        smp.sample(seed + len(plans), 0, pool, S.data_ptr(), truth.data_ptr())
        st = ct.hip.stats()
        if st["sc"] <= 20:
            try:
                ct.hip.autotune("fused_f32", S.data_ptr(), pool)
            except L.TqecError:
                pass                              # slab plan: no LDS tile to tune
            ct.hip.autotune("spec_f32", S.data_ptr(), min(pool, 2048 if st["sc"] > 16 else pool))
        # wider plans (toric L=5) keep the default slab configuration with
        # FP32-MFMA steps: each candidate kernel takes tens of seconds to build
        plans.append(ct)
        info.append(dict(name=name, p=p, ns=t.ns, nout=1 << ct.hip.n_out, S=S, truth=truth,
                         macs=st["dep_tc"], sc=st["sc"], spec=ct.hip.spec_config()))
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    nb_total = (warmup + calls) * window
    code_of = rng.integers(0, len(plans), nb_total)
    size_of = np.minimum(4096, np.floor(np.exp(rng.uniform(0, np.log(4097), nb_total)))).astype(np.int64)
    size_of = np.maximum(size_of, 1)
    off_of = np.array([rng.integers(0, pool - n + 1) for n in size_of])
    max_rows = int(max(size_of[i * window:(i + 1) * window].sum() for i in range(warmup + calls)))
    mx = L.HipMixed([ct.hip for ct in plans], max_batches=window, max_rows=max_rows)
    mx.set_kernel(True)                           # one kernel generated from every plan's specialised code
    out_mar = [torch.empty(max_rows * 16, dtype=torch.float64, device=dev) for _ in range(2)]
    out_cls = [torch.empty(max_rows, dtype=torch.int32, device=dev) for _ in range(2)]
    out_tie = [torch.empty(max_rows, dtype=torch.uint8, device=dev) for _ in range(2)]

    def table(ci):
        tab, row, mrow = [], 0, 0
        k = ci % 2
        for b in range(ci * window, (ci + 1) * window):
            j, n, o = int(code_of[b]), int(size_of[b]), int(off_of[b])
            it = info[j]
            tab.append((j, n, it["S"][o:].data_ptr(), out_mar[k][mrow:].data_ptr(), out_cls[k][row:].data_ptr(),
                        out_tie[k][row:].data_ptr()))
            row += n
            mrow += n * it["nout"]
        return tab, row

    tabs = [table(ci) for ci in range(warmup + calls)]
    rows_timed = sum(r for _, r in tabs[warmup:])
    for ci in range(warmup):
        mx.decode_device(tabs[ci][0], stream=sp)
    cuda_stream.synchronize()
    # throughput: calls back to back (two in flight)
    t1 = time.perf_counter()
    for ci in range(warmup, warmup + calls):
        mx.decode_device(tabs[ci][0], stream=sp)
    cuda_stream.synchronize()
    dt = time.perf_counter() - t1
    # latency: one call at a time, per-batch device-clock latencies
    lat, call_ms, rechecks, errors, decided = [], [], 0, 0, 0
    f64 = (torch.empty(4096 * 16, dtype=torch.float64, device=dev), torch.empty(4096, dtype=torch.int32, device=dev),
           torch.empty(4096, dtype=torch.uint8, device=dev))
    chk = {"reference": "exact FP64 single-plan executor (spec_f64; plain_f64 where a plan has no FP64 specialised kernel) on the first timed window's batches",
           "syndromes": 0, "max_rel_marginal_err_vs_fp64": 0.0, "non_tie": 0, "class_agree_non_tie": 0}
    for ci in range(warmup, warmup + calls):
        h0 = time.perf_counter()
        mx.decode_device(tabs[ci][0], stream=sp)
        us, nre = mx.last_call()
        call_ms.append((time.perf_counter() - h0) * 1e3)
        lat.extend(us.tolist())
        rechecks += nre
        k = ci % 2
        cls_np = out_cls[k].cpu().numpy()
        mar_np = out_mar[k].cpu().numpy() if ci == warmup else None
        row, mrow = 0, 0
        for b in range(ci * window, (ci + 1) * window):
            j, n, o = int(code_of[b]), int(size_of[b]), int(off_of[b])
            tr = info[j]["truth"][o:o + n].cpu().numpy()
            errors += int((cls_np[row:row + n] != tr).sum())
            decided += n
            if mar_np is not None:
                no = info[j]["nout"]
                try:
                    plans[j].hip.decode_device(info[j]["S"][o:].data_ptr(), n, "spec_f64", f64[0].data_ptr(),
                                               f64[1].data_ptr(), f64[2].data_ptr())
                except L.TqecError:
                    plans[j].hip.decode_device(info[j]["S"][o:].data_ptr(), n, "plain_f64", f64[0].data_ptr(),
                                               f64[1].data_ptr(), f64[2].data_ptr())
                torch.cuda.synchronize()
                m64 = f64[0][:n * no].cpu().numpy().reshape(n, no)
                c64, t64 = f64[1][:n].cpu().numpy(), f64[2][:n].cpu().numpy()
                m32 = mar_np[mrow:mrow + n * no].reshape(n, no)
                valid = m64.max(axis=1) > 0
                if valid.any():
                    e = np.abs(m32[valid] - m64[valid]) / m64[valid].max(axis=1, keepdims=True)
                    chk["max_rel_marginal_err_vs_fp64"] = max(chk["max_rel_marginal_err_vs_fp64"], float(e.max()))
                ok = t64 == 0
                chk["syndromes"] += n
                chk["non_tie"] += int(ok.sum())
                chk["class_agree_non_tie"] += int((cls_np[row:row + n][ok] == c64[ok]).sum())
            row += n
            mrow += n * info[j]["nout"]
    if not compare:
        return stream_result(info, mx, codes, window, calls, rows_timed, dt, lat, call_ms, rechecks, errors, decided,
                             chk, setup_s)
    # the same batches through the generic (descriptor-interpreting) mixed kernel
    mx.set_kernel(False)
    for ci in range(warmup):
        mx.decode_device(tabs[ci][0], stream=sp)
    cuda_stream.synchronize()
    t3 = time.perf_counter()
    for ci in range(warmup, warmup + calls):
        mx.decode_device(tabs[ci][0], stream=sp)
    cuda_stream.synchronize()
    dt_gen = time.perf_counter() - t3
    # the same batches as one launch sequence per batch (single-plan executors)
    dev_outs = (torch.empty(4096 * 16, dtype=torch.float64, device=dev), torch.empty(4096, dtype=torch.int32, device=dev),
                torch.empty(4096, dtype=torch.uint8, device=dev))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for ci in range(warmup, warmup + calls):
        for b in range(ci * window, (ci + 1) * window):
            j, n, o = int(code_of[b]), int(size_of[b]), int(off_of[b])
            mode = "spec_f32"
            plans[j].hip.decode_device(info[j]["S"][o:].data_ptr(), n, mode, dev_outs[0].data_ptr(),
                                       dev_outs[1].data_ptr(), dev_outs[2].data_ptr())
    torch.cuda.synchronize()
    dt_sep = time.perf_counter() - t2
    r = stream_result(info, mx, codes, window, calls, rows_timed, dt, lat, call_ms, rechecks, errors, decided, chk,
                      setup_s)
    r["generic_mixed_kernel_syndromes_per_s"] = rows_timed / dt_gen
    r["separate_launches_syndromes_per_s"] = rows_timed / dt_sep
    r["separate_launches"] = "one specialised single-plan launch (plus its FP64 re-check) per batch"
    return r


def stream_result(info, mx, codes, window, calls, rows_timed, dt, lat, call_ms, rechecks, errors, decided, chk,
                  setup_s):
    names = {n for n, _ in codes}
    toric = "3/4/5" if "toric5" in names else "3/4"
    return {
        "workload": f"uneven stream: surface d=3/5/7/9, toric L={toric}, Steane, Shor; p ~ U[0.01, 0.1] per code; "
                    f"{window} batches per launch, sizes log-uniform in [1, 4096]",
        "codes": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in it.items()
                   if k in ("name", "p", "ns", "macs", "sc", "spec")} | {"tile_slab": mx.plan_config(i)}
                  for i, it in enumerate(info)],
        "syndromes_per_s": rows_timed / dt,
        "calls": calls, "batches_per_call": window, "syndromes_timed": rows_timed,
        "ms_per_call": dt / calls * 1e3,
        "batch_latency_us_p50": pct(lat, 50), "batch_latency_us_p99": pct(lat, 99),
        "call_latency_ms_p50": pct(call_ms, 50), "call_latency_ms_p99": pct(call_ms, 99),
        "fp64_rechecks": rechecks, "logical_error_rate_all_codes": errors / max(decided, 1),
        "fp64_check": {"reference": chk["reference"], "syndromes": chk["syndromes"],
                       "max_rel_marginal_err_vs_fp64": chk["max_rel_marginal_err_vs_fp64"],
                       "class_agreement_non_tie": chk["class_agree_non_tie"] / max(chk["non_tie"], 1)},
        "setup_s": setup_s,
    }


def learned_leg(torch, dev, seed, ct=None, mode="fused_f32", d=7, p=0.05, n_train=10_000_000, train_seconds=90.0,
                batch=65536, eval_batches=8, ref_mode="plain_f64"):
    """Learned front-end at d=7 code capacity p=0.05: a 48-512-512-512-4 MLP
    (BF16 weights, FP32 accumulation) trained by the library's own kernels on
    TN-labelled syndromes, then gated: confident syndromes keep the MLP's
    class, the rest go to the TN decoder on the device.  The gate threshold
    is the smallest of a fixed ladder whose decisions differ from the TN's on
    at most 1e-4 of a calibration batch; the evaluation batches are held
    out.  Reported against FP64 TN decisions on the same syndromes: the plain
    executor, or (``ref_mode="spec_f64"``, for codes where the plain one is
    too slow) the exact FP64 specialised kernel on the same plan."""
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.codes import CSSTannerGraph, SurfaceCode
    from tensorqec.errors import iid_error
    from tensorqec.gf2 import logical_operator

    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    if ct is None:
        ct = D.compile(D.TNMMAP(mode=mode), t, em)
    plain = None
    if ref_mode == "plain_f64":
        plain = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="plain_f64"), t, em)
    lx, lz = logical_operator(t)
    smp = L.HipSampler.css(t, em, lx, lz)
    out = learned_core(torch, dev, seed, ct, smp, t.ns, 4, mode, n_train, train_seconds, batch, eval_batches,
                       ref_mode, plain)
    return dict({"workload": f"learned front-end, rotated surface code d={d}, code-capacity p={p}: MLP "
                             f"{t.ns}-512-512-512-4 (BF16 weights, FP32 accumulation), trained on TN-labelled "
                             f"syndromes, confidence-gated, TN ({mode}) for the rest"}, **out)


def learned_dem_leg(torch, dev, seed, d=5, rounds=2, p=0.001, n_train=2_000_000, train_seconds=60.0):
    """The learned front-end on a circuit-level detector error model (memory-Z,
    rotated surface code, depolarizing p after every gate, reset and
    measurement flips p; d=5 with 2 rounds, the largest with an exact
    contraction in reach: 48 detectors): one input per detector, 2 classes
    (logical flip or not), trained on syndromes labelled by the TN decoder
    of the DEM network, gated, judged against the exact FP64 specialised
    kernel on the same plan."""
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    dem = detector_error_model_from_stim(rotated_memory_z(d, rounds, p))
    ct = D.compile(D.TNMMAP(mode="spec_f32"), dem)
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    S = torch.empty((16384, smp.nsyn), dtype=torch.uint8, device=dev)
    tr = torch.empty(16384, dtype=torch.int32, device=dev)
    smp.sample(seed + 4, 0, 16384, S.data_ptr(), tr.data_ptr())
    ct.hip.autotune("spec_f32", S.data_ptr(), 16384 if ct.hip.stats()["sc"] <= 16 else 1024)
    out = learned_core(torch, dev, seed, ct, smp, smp.nsyn, 2, "spec_f32", n_train, train_seconds, 65536, 4,
                       "spec_f64", None)
    return dict({"workload": f"learned front-end, circuit-level memory-Z, rotated surface code d={d}, {rounds} "
                             f"rounds, p={p}: MLP {smp.nsyn}-512-512-512-2 over the detectors, trained on "
                             "TN-labelled syndromes, confidence-gated, TN (spec_f32) for the rest",
                 "tn_macs_per_syndrome": ct.hip.stats()["dep_tc"], "tn_config": ct.hip.spec_config()}, **out)


def learned_core(torch, dev, seed, ct, smp, n_in, n_classes, mode, n_train, train_seconds, batch, eval_batches,
                 ref_mode, plain):
    """Train, calibrate the gate, evaluate on held-out batches (see learned_leg)."""
    from tensorqec.learned import GatedDecoder, HipMLP, train_on_tn
    mlp = HipMLP(n_in, n_classes, seed=seed % 1000)
    rep = train_on_tn(mlp, ct, smp, n_samples=n_train, batch=8192, lr=2e-3, lr_final=1e-4, seed=seed + 5,
                      torch=torch, mode=mode, max_seconds=train_seconds)
    stream = torch.cuda.current_stream().cuda_stream
    S = torch.empty((batch, n_in), dtype=torch.uint8, device=dev)
    truth = torch.empty(batch, dtype=torch.int32, device=dev)
    mar = torch.empty((batch, n_classes), dtype=torch.float64, device=dev)
    cls_tn = torch.empty(batch, dtype=torch.int32, device=dev)
    cls_g = torch.empty(batch, dtype=torch.int32, device=dev)
    conf = torch.empty(batch, dtype=torch.float32, device=dev)
    tie = torch.empty(batch, dtype=torch.uint8, device=dev)
    # calibration batch: pick the threshold
    smp.sample(seed + 77, 0, batch, S.data_ptr(), truth.data_ptr())
    ct.hip.decode_device(S.data_ptr(), batch, mode, mar.data_ptr(), cls_tn.data_ptr(), tie.data_ptr(), stream)
    # gate target: the MLP may disagree with the TN on at most 1e-4 of the
    # calibration syndromes and at most a tenth of the TN's own logical error
    # rate there, so the disagreements cannot move the logical error rate
    # out of the FP64 decoder's confidence interval at low noise
    torch.cuda.synchronize()
    tn_err = float((cls_tn != truth).float().mean().item())
    target = min(1e-4, 0.1 * tn_err)
    ladder = [0.9, 0.95, 0.98, 0.99, 0.995, 0.998, 0.999, 0.9995, 0.9999, 0.99999]
    chosen = ladder[-1]
    for th in ladder:
        gd = GatedDecoder(mlp, ct, th, batch, torch=torch, mode=mode)
        gd.decode_device(S.data_ptr(), batch, cls_g.data_ptr(), conf.data_ptr(), mar.data_ptr(), tie.data_ptr(), stream)
        torch.cuda.synchronize()
        if float((cls_g != cls_tn).float().mean().item()) <= target:
            chosen = th
            break
    gd = GatedDecoder(mlp, ct, chosen, batch, torch=torch, mode=mode)
    # held-out evaluation
    res = dict(fallback=0, disagree_tn=0, disagree_fp64=0, err_g=0, err_fp64=0, err_tn=0, n=0)
    tg, ttn = 0.0, 0.0
    for k in range(eval_batches):
        smp.sample(seed + 1000, k * batch, batch, S.data_ptr(), truth.data_ptr())
        torch.cuda.synchronize()
        h0 = time.perf_counter()
        ct.hip.decode_device(S.data_ptr(), batch, mode, mar.data_ptr(), cls_tn.data_ptr(), tie.data_ptr(), stream)
        torch.cuda.synchronize()
        h1 = time.perf_counter()
        gd.decode_device(S.data_ptr(), batch, cls_g.data_ptr(), conf.data_ptr(), mar.data_ptr(), tie.data_ptr(), stream)
        torch.cuda.synchronize()
        h2 = time.perf_counter()
        if k > 0:                       # first batch warms up
            ttn += h1 - h0
            tg += h2 - h1
        if plain is not None:
            m64, c64, t64 = plain.hip.decode(S.cpu().numpy(), "plain_f64")
        else:
            ct.hip.decode_device(S.data_ptr(), batch, ref_mode, mar.data_ptr(), cls_tn.data_ptr(), tie.data_ptr(),
                                 stream)
            c64 = cls_tn.cpu().numpy()
            ct.hip.decode_device(S.data_ptr(), batch, mode, mar.data_ptr(), cls_tn.data_ptr(), tie.data_ptr(), stream)
        tr = truth.cpu().numpy()
        a, b = cls_tn.cpu().numpy(), cls_g.cpu().numpy()
        res["fallback"] += gd.fallbacks()
        res["disagree_tn"] += int((a != b).sum())
        res["disagree_fp64"] += int((b != c64).sum())
        res["err_g"] += int((b != tr).sum())
        res["err_tn"] += int((a != tr).sum())
        res["err_fp64"] += int((c64 != tr).sum())
        res["n"] += batch
    n = res["n"]
    ler64 = res["err_fp64"] / n
    ci = wilson(res["err_fp64"], n)
    ler_g = res["err_g"] / n
    return {
        "train_samples": rep.samples, "train_steps": rep.steps, "train_s": rep.seconds,
        "train_samples_per_s": rep.samples / max(rep.seconds, 1e-9),
        "loss_first_last": [rep.losses[0][1], rep.losses[-1][1]] if rep.losses else None,
        "gate_threshold": chosen, "gate_target_disagreement": target, "eval_syndromes": n,
        "fallback_fraction": res["fallback"] / n,
        "disagreement_vs_tn": res["disagree_tn"] / n, "disagreement_vs_fp64": res["disagree_fp64"] / n,
        "ler_gated": ler_g, "ler_tn": res["err_tn"] / n, "ler_fp64": ler64, "ler_fp64_ci95": ci,
        "ler_gated_within_ci": ci[0] <= ler_g <= ci[1],
        "gated_syndromes_per_s": batch * (eval_batches - 1) / max(tg, 1e-9),
        "tn_only_syndromes_per_s": batch * (eval_batches - 1) / max(ttn, 1e-9),
    }


def learned_d9_leg(torch, dev, seed, p=0.05):
    """The learned front-end where the exact contraction is expensive: d=9
    (80 syndrome bits, 400 K multiply-adds per syndrome, LDS-tile specialised
    kernel, autotuned), MLP trained on 40 M TN-labelled syndromes (at most
    90 s), FP64 reference from the exact specialised kernel."""
    from tensorqec import decoders as D
    from tensorqec.codes import CSSTannerGraph, SurfaceCode
    from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing
    d = 9
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
    S = torch.from_numpy(css_syndrome_rows(sample_depolarizing(em, seed + 9, 65536), t)).to(dev)
    ct.hip.autotune("spec_f32", S.data_ptr(), 65536)
    r = learned_leg(torch, dev, seed, ct=ct, mode="spec_f32", d=d, p=p, n_train=40_000_000, train_seconds=90.0,
                    eval_batches=4, ref_mode="spec_f64")
    r["tn_config"] = ct.hip.spec_config()
    r["tn_macs_per_syndrome"] = ct.hip.stats()["dep_tc"]
    return r


def pipeline_leg(torch, dev, seed, ct=None, mode="spec_f32", d=7, p=0.05, batch=65536, steps=10):
    """The whole decode loop on the device at d=7, p=0.05: Philox sampling of
    errors and syndromes, TN decoding, recovery of the error pattern (GF(2)
    pure error + logical fix, the DecodingResult payload), and the logical
    error count against the sampled class; nothing crosses to the host but
    the final counts."""
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.codes import CSSTannerGraph, SurfaceCode
    from tensorqec.errors import iid_error
    from tensorqec.gf2 import logical_operator
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    if ct is None:
        ct = D.compile(D.TNMMAP(mode=mode), t, em)
    lx, lz = logical_operator(t)
    smp = L.HipSampler.css(t, em, lx, lz)
    rec = L.HipRecovery.css(t, ct.pure_x, ct.pure_z, ct.lx, ct.lz)
    S = torch.empty((batch, t.ns), dtype=torch.uint8, device=dev)
    truth = torch.empty(batch, dtype=torch.int32, device=dev)
    mar = torch.empty((batch, 4), dtype=torch.float64, device=dev)
    cls = torch.empty(batch, dtype=torch.int32, device=dev)
    tie = torch.empty(batch, dtype=torch.uint8, device=dev)
    ex = torch.empty((batch, t.nq), dtype=torch.uint8, device=dev)
    ez = torch.empty((batch, t.nq), dtype=torch.uint8, device=dev)
    errors = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def one(k):
        smp.sample(seed + 11, k * batch, batch, S.data_ptr(), truth.data_ptr(), stream=stream)
        ct.hip.decode_device(S.data_ptr(), batch, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), stream)
        rec.recover(S.data_ptr(), cls.data_ptr(), batch, ex.data_ptr(), ez.data_ptr(), stream=stream)
        errors.add_((cls != truth).sum())

    one(0)
    torch.cuda.synchronize()
    errors.zero_()
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        one(k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n = steps * batch
    e = int(errors.item())
    return {"workload": f"device loop at d={d}, p={p}: sample, decode ({mode}), recover error patterns, count "
                        f"logical errors; {steps} x {batch} syndromes",
            "syndromes_per_s": n / dt, "ms_per_batch": dt / steps * 1e3, "logical_errors": e,
            "ler": e / n, "ler_ci95": wilson(e, n)}


def tnmap_leg(torch, dev, seed, d=5, p=0.05, batch=16384, reps=3):
    """TNMAP (most probable error configuration, max-product contraction with
    an arg-max backtrace) at code capacity: throughput and LER against the
    sampled class, with the TNMMAP (marginal) decision on the same syndromes
    for comparison."""
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.codes import CSSTannerGraph, SurfaceCode
    from tensorqec.errors import iid_error
    from tensorqec.gf2 import logical_operator
    from tensorqec.tnmap import TNMAP, compile_tnmap
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    t0 = time.time()
    cm = compile_tnmap(TNMAP(), t, em)
    compile_s = time.time() - t0
    lx, lz = logical_operator(t)
    smp = L.HipSampler.css(t, em, lx, lz)
    S = torch.empty((batch, t.ns), dtype=torch.uint8, device=dev)
    truth = torch.empty(batch, dtype=torch.int32, device=dev)
    smp.sample(seed + 21, 0, batch, S.data_ptr(), truth.data_ptr())
    nconf = cm.qubit_num
    cfg = torch.empty((batch, nconf), dtype=torch.uint8, device=dev)
    lp = torch.empty(batch, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    cm.hip.map_decode_device(S.data_ptr(), batch, nconf, cfg.data_ptr(), lp.data_ptr(), stream)
    torch.cuda.synchronize()
    h0 = time.perf_counter()
    for _ in range(reps):
        cm.hip.map_decode_device(S.data_ptr(), batch, nconf, cfg.data_ptr(), lp.data_ptr(), stream)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - h0) / reps
    # class of the most probable configuration: logical parities of (ex, ez)
    import numpy as np
    c = cfg.cpu().numpy().astype(np.int64)
    n = t.nq
    exx, ezz = c[:, :n], c[:, n:2 * n]
    k = lx.shape[0]
    cls = np.zeros(batch, np.int64)
    for i in range(k):
        cls |= ((ezz @ lx[i].astype(np.int64)) & 1) << i
        cls |= ((exx @ lz[i].astype(np.int64)) & 1) << (k + i)
    tr = truth.cpu().numpy()
    ct = D.compile(D.TNMMAP(mode="spec_f32"), t, em)
    r = D.decode_batch(ct, S.cpu().numpy(), error_patterns=False)
    return {"workload": f"TNMAP, rotated surface code d={d}, code-capacity p={p}, {batch} syndromes",
            "compile_s": compile_s, "syndromes_per_s": batch / dt, "ler_tnmap": float((cls != tr).mean()),
            "ler_tnmmap": float((r.classes != tr).mean()),
            "class_agreement_with_tnmmap": float((cls == r.classes).mean())}


def mlp_leg(torch, dev, seed, n_in=48, batch=65536, reps=20):
    """Inference of the learned front-end alone: the 48-512-512-512-4 MLP
    (random weights) on a batch of syndromes, with the fused kernel (all
    layers per 128-row tile, activations in LDS) and with the
    layer-by-layer kernels training uses.  Algorithmic flops count the
    padded input width the GEMMs run (kin) and the 4 real output columns."""
    from tensorqec.learned import HipMLP
    mlp = HipMLP(n_in, 4, seed=seed % 1000)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    S = (torch.rand((batch, n_in), device=dev, generator=g) < 0.1).to(torch.uint8)
    z = torch.empty((batch, 4), dtype=torch.float32, device=dev)
    flops = 2.0 * batch * (mlp.kin * 512 + 2 * 512 * 512 + 512 * 4)
    out = {"workload": f"MLP {n_in}-512-512-512-4 forward, BF16 weights, FP32 accumulation, {batch} syndromes",
           "gflop_per_batch": flops / 1e9}
    for name, kern in (("fused", 1), ("layer_by_layer", 0)):
        mlp.set_inference(kern)
        for _ in range(3):
            mlp.forward(S.data_ptr(), batch, z.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            mlp.forward(S.data_ptr(), batch, z.data_ptr())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"ms_per_batch": dt * 1e3, "syndromes_per_s": batch / dt, "tflops": flops / dt / 1e12,
                     "frac_of_bf16_peak": flops / dt / 2.5e15}
    mlp.set_inference(1)
    return out


def dem_leg(torch, dev, seed, d=5, rounds=2, p=0.001, n=16384, reps=3):
    """Circuit-level memory-Z experiment (rotated surface code, depolarizing
    p after every 1- and 2-qubit gate, reset and measurement flips p),
    compiled to a detector error model by the library's own generator and
    decoded by TNMMAP over the DEM network (the cheaper of the reference's
    chain-factorized network and its nested-dissection rewrite); FP32 and
    BF16-MFMA executors against plain FP64 on the same seeded syndromes.
    d=5 with 2 rounds is the largest round count with an exact contraction
    in reach (21 M multiply-adds per syndrome; 3 rounds need ~2^39)."""
    import numpy as np
    from tensorqec import decoders as D
    from tensorqec import lib as L
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z

    t0 = time.time()
    dem = detector_error_model_from_stim(rotated_memory_z(d, rounds, p))
    ct = D.compile(D.TNMMAP(mode="wide_f32"), dem)
    st = ct.hip.stats()
    compile_s = time.time() - t0
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    ns = smp.nsyn
    S = torch.empty((n, ns), dtype=torch.uint8, device=dev)
    truth = torch.empty(n, dtype=torch.int32, device=dev)
    smp.sample(seed + 3, 0, n, S.data_ptr(), truth.data_ptr())
    tr = truth.cpu().numpy()
    nout = 1 << ct.hip.n_out
    stream = torch.cuda.current_stream().cuda_stream
    out = {"workload": f"circuit-level memory-Z, rotated surface code d={d}, {rounds} rounds, p={p}: "
                       f"{ns} detectors, {len(dem.error_rates)} mechanisms, {n} syndromes",
           "plan": {k: st[k] for k in ("sc", "dep_tc", "dep_rw", "plain_ws_entries", "n_dep_steps")},
           "compile_s": compile_s, "modes": {}}
    t1 = time.time()
    for m in ("spec_f32", "spec_bf16"):
        try:
            ct.hip.autotune(m, S.data_ptr(), min(n, 4096 if st["sc"] <= 16 else 1024), stream)
        except L.TqecError:
            pass
    try:
        out["spec_config"] = ct.hip.spec_config()
    except L.TqecError as e:
        out["spec_config"] = str(e)
    out["spec_tune_s"] = time.time() - t1
    ref = None
    for mode in ("plain_f64", "wide_f32", "spec_f32", "spec_bf16"):
        mar = torch.empty((n, nout), dtype=torch.float64, device=dev)
        cls = torch.empty(n, dtype=torch.int32, device=dev)
        tie = torch.empty(n, dtype=torch.uint8, device=dev)
        try:
            ct.hip.decode_device(S.data_ptr(), n, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), stream)
        except L.TqecError as e:
            out["modes"][mode] = {"unsupported": str(e)}
            continue
        torch.cuda.synchronize()
        h0 = time.perf_counter()
        for _ in range(reps):
            ct.hip.decode_device(S.data_ptr(), n, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - h0) / reps
        m, c, ti = mar.cpu().numpy(), cls.cpu().numpy(), tie.cpu().numpy()
        r = {"syndromes_per_s": n / dt, "ler": float((c != tr).mean())}
        if mode.startswith("spec"):
            r["fp64_rechecks"] = ct.hip.last_recheck_count(stream)
        if ref is None:
            ref = (m, c, ti)
        else:
            valid = ref[0].max(axis=1) > 0
            r["max_rel_marginal_err_vs_fp64"] = float(np.max(np.abs(m[valid] - ref[0][valid]) /
                                                             ref[0][valid].max(axis=1, keepdims=True)))
            ok = ref[2] == 0
            r["class_agreement_non_tie"] = float((c[ok] == ref[1][ok]).mean())
        out["modes"][mode] = r
    return out


def toric5_leg(torch, dev, seed, L=5, p=0.03, n=2048, reps=1):
    """Toric code L=5 (50 qubits, 2 logical qubits, 16 classes) under
    code-capacity depolarizing noise px = py = pz = p, decoded exactly: the
    nested-dissection network (width 2^21, 1.33 G multiply-adds per
    syndrome) in FP64 (plain executor) and FP32 (wide and specialised
    executors, FP64 re-check of low margins), on the same device-sampled
    syndromes; logical error rate against the sampled class."""
    import numpy as np
    from tensorqec import decoders as D
    from tensorqec import lib as L_
    from tensorqec.codes import CSSTannerGraph, ToricCode
    from tensorqec.errors import iid_error
    from tensorqec.gf2 import logical_operator
    t0 = time.time()
    t = CSSTannerGraph.from_code(ToricCode(L, L))
    em = iid_error(p, p, p, t.nq)
    ct = D.compile(D.TNMMAP(mode="plain_f64", optimizer=D.Dissection(), tabulate_bits=0), t, em)
    st = ct.hip.stats()
    compile_s = time.time() - t0
    lx, lz = logical_operator(t)
    smp = L_.HipSampler.css(t, em, lx, lz)
    S = torch.empty((n, t.ns), dtype=torch.uint8, device=dev)
    truth = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    smp.sample(seed + 17, 0, n, S.data_ptr(), truth.data_ptr(), stream=stream)
    tr = truth.cpu().numpy()
    nout = 1 << ct.hip.n_out
    out = {"workload": f"toric code L={L}, code-capacity depolarizing px=py=pz={p}, {n} syndromes, "
                       f"nested-dissection network",
           "plan": {k: st[k] for k in ("sc", "dep_tc", "dep_rw", "plain_ws_entries", "n_dep_steps")},
           "compile_s": compile_s, "spec_config": ct.hip.spec_config(), "modes": {}}
    ref = None
    for mode in ("plain_f64", "wide_f32", "spec_f32"):
        mar = torch.empty((n, nout), dtype=torch.float64, device=dev)
        cls = torch.empty(n, dtype=torch.int32, device=dev)
        tie = torch.empty(n, dtype=torch.uint8, device=dev)
        try:
            ct.hip.decode_device(S.data_ptr(), n, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), stream)
        except L_.TqecError as e:
            out["modes"][mode] = {"unsupported": str(e)}
            continue
        torch.cuda.synchronize()
        h0 = time.perf_counter()
        for _ in range(reps):
            ct.hip.decode_device(S.data_ptr(), n, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(), stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - h0) / reps
        m, c, ti = mar.cpu().numpy(), cls.cpu().numpy(), tie.cpu().numpy()
        r = {"syndromes_per_s": n / dt, "tflops": 2.0 * st["dep_tc"] * n / dt / 1e12, "ler": float((c != tr).mean())}
        if ref is None:
            ref = (m, c, ti)
        else:
            valid = ref[0].max(axis=1) > 0
            r["max_rel_marginal_err_vs_fp64"] = float(np.max(np.abs(m[valid] - ref[0][valid]) /
                                                             ref[0][valid].max(axis=1, keepdims=True)))
            ok = ref[2] == 0
            r["class_agreement_non_tie"] = float((c[ok] == ref[1][ok]).mean())
        out["modes"][mode] = r
    return out


def sliced_run(torch, plan, labels, S, truth, n, modes=("spec_f32", "spec_f64"), tabulate_bits=0):
    """Decode the ``n`` device syndromes ``S`` through a sliced plan (each
    syndrome as 2^k rows, slices summed in FP64) in each mode: rates, FP32
    against FP64, logical errors against the sampled classes."""
    import numpy as np
    from tensorqec.slicing import SlicedDecoder
    dec = SlicedDecoder(plan, labels, tabulate_bits=tabulate_bits)
    k = len(labels)
    st0, st1 = plan.stats(), dec.plan.stats()
    st1["tc_dep"] = dec.hip.stats()["dep_tc"]        # executed per row (tabulated subtrees excluded)
    out = {"plan": {"sliced_labels": k, "width": st1["sc_dep"], "macs_per_slice": st1["tc_dep"],
                    "tabulate_bits": tabulate_bits,
                    "macs_per_syndrome": st1["tc_dep"] * (1 << k),
                    "same_tree_unsliced": {"macs_per_syndrome": st0["tc_dep"], "width": st0["sc_dep"]}},
           "modes": {}}
    res = {}
    for mode in modes:
        h0 = time.time()
        dec.decode_device(torch, S, n, mode)          # compiles the kernels
        torch.cuda.synchronize()
        first = time.time() - h0
        h0 = time.perf_counter()
        mar, cls, tie = dec.decode_device(torch, S, n, mode)
        torch.cuda.synchronize()
        dt = time.perf_counter() - h0
        res[mode] = (mar.cpu().numpy(), cls.cpu().numpy(), tie.cpu().numpy())
        out["modes"][mode] = {"syndromes_per_s": n / dt, "first_call_s": first,
                              "tflops": 2.0 * st1["tc_dep"] * (n << k) / dt / 1e12}
        if mode == "spec_f32":
            out["modes"][mode]["fp64_rechecked_syndromes"] = dec.last_rechecked
        print(f"sliced decode {mode}: {n / dt:.3g} syndromes/s", file=sys.stderr, flush=True)
    m32, c32, _ = res["spec_f32"]
    m64, c64, t64 = res["spec_f64"]
    valid = m64.max(axis=1) > 0
    out["modes"]["spec_f32"]["max_rel_marginal_err_vs_fp64"] = float(
        np.max(np.abs(m32[valid] - m64[valid]) / m64[valid].max(axis=1, keepdims=True)))
    ok = t64 == 0
    out["modes"]["spec_f32"]["class_agreement_non_tie"] = float((c32[ok] == c64[ok]).mean())
    out["total_probability_max"] = float(m64.sum(axis=1).max())
    out["logical_errors_fp64"] = int((c64 != truth.cpu().numpy()).sum())
    return out


def toric6_leg(torch, dev, seed, L=6, p=0.03, n=64, target_rank=21, tabulate_bits=4):
    """Toric code L=6 (72 qubits, 16 classes), decoded exactly with label
    slicing: the nested-dissection network, its tree annealed together with
    the labels to slice for a widest per-row tensor of 2^21 entries
    (2^36.3 multiply-adds per syndrome over the slices), each syndrome run
    as 2^k rows and the rows' marginals summed in FP64; subtrees that depend
    on at most 4 syndrome bits tabulated (59 row-dependent steps instead of
    220).  FP32 (slab kernels, FP32-MFMA steps) against FP64 on the same
    syndromes."""
    from tensorqec import lib as L_
    from tensorqec.codes import CSSTannerGraph, ToricCode
    from tensorqec.errors import iid_error
    from tensorqec.gf2 import logical_operator
    from tensorqec.network import dissected_css_network
    from tensorqec.slicing import sliced_dissection_plan
    t0 = time.time()
    t = CSSTannerGraph.from_code(ToricCode(L, L))
    em = iid_error(p, p, p, t.nq)
    lx, lz = logical_operator(t)
    plan, labels, _, _ = sliced_dissection_plan(lambda s: dissected_css_network(t, em, lxlz=(lx, lz), seed=s),
                                                target_rank)
    plan_s = time.time() - t0
    print(f"toric L={L}: sliced plan in {plan_s:.0f} s", file=sys.stderr, flush=True)
    smp = L_.HipSampler.css(t, em, lx, lz)
    S = torch.empty((n, t.ns), dtype=torch.uint8, device=dev)
    truth = torch.empty(n, dtype=torch.int32, device=dev)
    smp.sample(seed + 23, 0, n, S.data_ptr(), truth.data_ptr())
    out = {"workload": f"toric code L={L}, code-capacity depolarizing px=py=pz={p}, {n} syndromes, "
                       f"nested-dissection network sliced on {len(labels)} labels",
           "planning_s": plan_s}
    out.update(sliced_run(torch, plan, labels, S, truth, n, tabulate_bits=tabulate_bits))
    return out


def dem_sliced_leg(torch, dev, seed, d=5, rounds=3, p=0.001, n=32, target_rank=21):
    """Circuit-level memory-Z experiment at a round count whose exact
    contraction is wider than one workgroup's slab (d=5, 3 rounds: 72
    detectors, 821 mechanisms, 2^39.3 multiply-adds per syndrome with a
    widest tensor of 2^28 entries unsliced), decoded with label slicing: the
    nested-dissection DEM network, tree and sliced labels annealed together
    for a widest per-row tensor of 2^21 entries.  FP32 against FP64 on the
    same syndromes, logical errors against the sampled classes.  Subtrees
    that depend on at most 4 syndrome bits are tabulated (as the DEM
    compile does), which keeps the generated kernel source small."""
    from tensorqec import lib as L
    from tensorqec.circuits import detector_error_model_from_stim, rotated_memory_z
    from tensorqec.network import dissected_dem_network
    from tensorqec.slicing import sliced_dissection_plan
    t0 = time.time()
    dem = detector_error_model_from_stim(rotated_memory_z(d, rounds, p))
    plan, labels, _, _ = sliced_dissection_plan(lambda s: dissected_dem_network(dem, seed=s), target_rank)
    plan_s = time.time() - t0
    print(f"d={d}, {rounds} rounds: sliced plan in {plan_s:.0f} s", file=sys.stderr, flush=True)
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    S = torch.empty((n, smp.nsyn), dtype=torch.uint8, device=dev)
    truth = torch.empty(n, dtype=torch.int32, device=dev)
    smp.sample(seed + 29, 0, n, S.data_ptr(), truth.data_ptr())
    out = {"workload": f"circuit-level memory-Z, rotated surface code d={d}, {rounds} rounds, p={p}: "
                       f"{smp.nsyn} detectors, {len(dem.error_rates)} mechanisms, {n} syndromes, "
                       f"nested-dissection network sliced on {len(labels)} labels",
           "planning_s": plan_s}
    out.update(sliced_run(torch, plan, labels, S, truth, n, tabulate_bits=4))
    return out


def copy_calibration(torch, dev, mib):
    """Kernels of known traffic: a fill writes and a sum reads `mib` MiB
    once; their WRITE_SIZE / FETCH_SIZE counters calibrate the HBM byte
    counters of the decode kernels in the same profiled run."""
    if mib <= 0:
        return
    src = torch.empty(mib * (1 << 20) // 4, dtype=torch.float32, device=dev)
    src.fill_(1.0)
    float(src.sum().item())
    torch.cuda.synchronize()
    del src


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--d", type=int, default=7)
    ap.add_argument("--p", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--config", default=None,
                    help="JSON line of an earlier run: reuse its plan and tile choice (no autotuning)")
    ap.add_argument("--plan", default=None,
                    help="SEED,SC_TARGET,MODE,TILE,THREADS[,TABLES_IN_LDS,WAVES[,UNROLL[,TABULATE_BITS[,REG]]]]: pin the "
                         "plan, executor and tile (no autotuning); SC_TARGET may be none; MODE fused_f32 or spec_f32; "
                         "REG 1 = register-resident specialised kernel (TILE = THREADS / 4)")
    ap.add_argument("--no-spec", action="store_true", help="do not consider the plan-specialised kernel")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the secondary workloads (uneven stream, learned front-ends, circuit-level DEM, "
                         "toric L=5 and L=6, MLP, device loop, TNMAP)")
    ap.add_argument("--only", default=None,
                    choices=["stream", "stream_full", "learned", "learned_d9", "learned_dem", "dem", "mlp", "pipeline",
                             "tnmap", "toric5", "toric6", "dem_sliced"],
                    help="run one secondary workload alone and print its JSON")
    ap.add_argument("--dem-d", type=int, default=5, help="distance of the circuit-level DEM workload")
    ap.add_argument("--dem-rounds", type=int, default=2, help="syndrome rounds of the circuit-level DEM workload")
    ap.add_argument("--plan-out", default=None,
                    help="write the chosen plan (tree, leaf tensors, per-syndrome work) to this JSON file")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied once by a plain device copy before the timed steps (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (syndromes are independent: "
              "a batch splits into per-GPU shards with no exchange, so N GPUs give N times the one-GPU rate).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.only:
        copy_calibration(torch, torch.device("cuda:0"), args.copy_calibration)
        kw = {"dem": {"d": args.dem_d, "rounds": args.dem_rounds},
              "dem_sliced": {"d": args.dem_d, "rounds": max(args.dem_rounds, 3)},
              "stream_full": {"codes": STREAM_CODES + (("toric5", ("ToricCode", 5)),), "compare": False}}.get(args.only, {})
        fn = {"stream": stream_leg, "stream_full": stream_leg, "learned": learned_leg, "learned_d9": learned_d9_leg, "learned_dem": learned_dem_leg,
              "dem": dem_leg,
              "mlp": mlp_leg, "pipeline": pipeline_leg, "tnmap": tnmap_leg, "toric5": toric5_leg, "toric6": toric6_leg,
              "dem_sliced": dem_sliced_leg}[args.only]
        print(json.dumps(fn(torch, torch.device("cuda:0"), args.seed, **kw)))
        return

    from tensorqec import decoders as D
    from tensorqec.codes import CSSTannerGraph, SurfaceCode
    from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing

    dev = torch.device("cuda:0")
    d, p, B = args.d, args.p, args.batch
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    # warm-up steps cycle over the first `nwarm` batches; every timed step has its own batch
    nwarm = max(1, min(args.warmup, 20))
    nb = nwarm + args.steps
    eps = [sample_depolarizing(em, args.seed, B, first=i * B) for i in range(nb)]
    S_host = [css_syndrome_rows(e, t) for e in eps]
    S_dev = [torch.from_numpy(s).to(dev) for s in S_host]
    stream = torch.cuda.current_stream().cuda_stream

    t0 = time.time()
    plain = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="plain_f64"), t, em)
    # Candidate contraction trees for the fused executor (seed, width target);
    # each is autotuned on the first batch and the fastest is kept.
    pinned = None
    if args.config:
        pinned = json.loads(open(args.config).read().strip().splitlines()[-1])["config"]["plan"]
    elif args.plan:
        f = args.plan.split(",")
        pinned = {"seed": int(f[0]), "sc_target": None if f[1].lower() == "none" else int(f[1]), "mode": f[2],
                  "tile": int(f[3]), "threads": int(f[4])}
        if len(f) > 5:
            pinned["tables_in_lds"] = bool(int(f[5]))
            pinned["waves"] = int(f[6])
        if len(f) > 7:
            pinned["unroll"] = int(f[7])
        if len(f) > 8:
            pinned["tabulate_bits"] = int(f[8])
        if len(f) > 9:
            pinned["reg"] = bool(int(f[9]))
        if len(f) > 10:
            pinned["dyn"] = bool(int(f[10]))
    if pinned:
        cands = [(pinned["seed"], pinned["sc_target"], pinned.get("tabulate_bits", 4))]
    else:
        cands = [(14, None, 7), (11, None, 8), (14, None, 6), (24, None, 6), (30, None, 6), (11, None, 7), (19, None, 6),
                 (31, None, 6), (13, None, 6), (24, None, 5), (14, None, 5), (24, None, 4)]
    outs = [torch.empty((B, 4), dtype=torch.float64, device=dev), torch.empty(B, dtype=torch.int32, device=dev),
            torch.empty(B, dtype=torch.uint8, device=dev)]

    def time_mode(ct, mode):
        # one untimed call first: it compiles a specialised kernel on first use
        ct.hip.decode_device(S_dev[0].data_ptr(), B, mode, outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(),
                             stream)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(3):
            ct.hip.decode_device(S_dev[0].data_ptr(), B, mode, outs[0].data_ptr(), outs[1].data_ptr(),
                                 outs[2].data_ptr(), stream)
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / 3

    # Candidate contraction trees (seed, width target, tabulation depth): each
    # is timed with the plan-specialised kernel in its default configuration
    # (the register-resident kernel where the plan's live tensors fit), the
    # three fastest are autotuned, and the fastest (plan, configuration) is
    # kept.  Without the specialised kernel the generic fused kernel is
    # autotuned instead.
    tried = []
    for seed, sct, tab in cands:
        ct = D.compile(D.TNMMAP(mode="fused_f32", optimizer=D.TreeSA(seed=seed, sc_target=sct), tabulate_bits=tab), t,
                       em)
        st = ct.hip.stats()
        if pinned:
            cfg = {"seed": seed, "sc_target": sct, "tabulate_bits": tab, "mode": pinned["mode"], "tile": pinned["tile"],
                   "threads": pinned["threads"]}
            if pinned["mode"] == "spec_f32":
                ct.hip.set_spec(pinned["tile"], pinned["threads"], pinned.get("tables_in_lds", False),
                                pinned.get("waves", 0), pinned.get("slab", False), pinned.get("unroll", 1),
                                pinned.get("reg", False), pinned.get("dyn", False))
                cfg.update(ct.hip.spec_config())
            else:
                ct.hip.set_tile("fused_f32", pinned["tile"], pinned["threads"])
            tried.append((0.0, ct, cfg))
            continue
        if args.no_spec:
            tile, thr = ct.hip.autotune("fused_f32", S_dev[0].data_ptr(), B, stream)
            ms = time_mode(ct, "fused_f32")
            cfg = {"seed": seed, "sc_target": sct, "tabulate_bits": tab, "mode": "fused_f32", "tile": tile,
                   "threads": thr}
        else:
            ms = time_mode(ct, "spec_f32")
            cfg = dict({"seed": seed, "sc_target": sct, "tabulate_bits": tab, "mode": "spec_f32"},
                       **ct.hip.spec_config())
        print(f"[bench] candidate seed={seed} sc_target={sct} tabulate_bits={tab}: steps={st['n_dep_steps']} "
              f"macs={st['dep_tc']:.0f} lds/syn={st['fused_lds_entries']} tables={st['table_entries']} "
              f"{cfg} {ms:.3f} ms/batch", file=sys.stderr, flush=True)
        tried.append((ms, ct, cfg))
    tried.sort(key=lambda x: x[0])
    best = tried[0]
    if not pinned and not args.no_spec:
        for ms_f, ct, cfg in tried[:3]:
            t1 = time.time()
            ct.hip.autotune("spec_f32", S_dev[0].data_ptr(), B, stream)
            sc = ct.hip.spec_config()
            ms = time_mode(ct, "spec_f32")
            print(f"[bench] candidate seed={cfg['seed']} sc_target={cfg['sc_target']} tabulate_bits={cfg['tabulate_bits']}: "
                  f"spec_f32 {sc} {ms:.3f} ms/batch "
                  f"(tuned in {time.time() - t1:.1f}s)", file=sys.stderr, flush=True)
            if ms < best[0]:
                best = (ms, ct, dict(cfg, mode="spec_f32", **sc))
    chosen, plan_cfg = best[1], best[2]
    if args.plan_out:
        from tensorqec.planio import save_plan
        save_plan(chosen.plan, args.plan_out, tabulate_bits=[plan_cfg["tabulate_bits"]],
                  meta={"code": f"rotated surface code d={d}", "noise": f"code-capacity depolarizing, p={p} (p/3 each)",
                        "tree": f"TreeSA(seed={plan_cfg['seed']}, sc_target={plan_cfg['sc_target']})",
                        "tabulate_bits": plan_cfg["tabulate_bits"], "mode": plan_cfg["mode"]})
    fused_ms = None
    if plan_cfg["mode"] == "spec_f32":
        plan_cfg = dict(plan_cfg, **chosen.hip.spec_config())   # the configuration the plan now holds
        if not pinned:
            # the generic (descriptor-interpreting) FP32 kernel on the chosen plan, for comparison
            chosen.hip.autotune("fused_f32", S_dev[0].data_ptr(), B, stream)
            fused_ms = time_mode(chosen, "fused_f32")
    print(f"[bench] plans compiled and tuned in {time.time() - t0:.1f}s: {plan_cfg}", file=sys.stderr, flush=True)
    truth = [true_classes(e, chosen.lx, chosen.lz) for e in eps]
    nout = 1 << chosen.hip.n_out
    mar = torch.empty((B, nout), dtype=torch.float64, device=dev)
    cls = torch.empty(B, dtype=torch.int32, device=dev)
    tie = torch.empty(B, dtype=torch.uint8, device=dev)
    print(f"[bench] {nb} batches of {B} syndromes resident", file=sys.stderr, flush=True)
    copy_calibration(torch, dev, args.copy_calibration)

    def run(ct, mode, label, warmup=None, instrument=True):
        """Warm-up, then the timed region (K steps, no instrumentation:
        `value`), then, when `instrument`, the same K steps again with the
        library's per-launch HIP events on the launch stream (kernel and
        re-check durations for the roofline; the events add a few µs per
        step, so they stay out of the timed region)."""
        # per-step output buffers, written once before the timed region so
        # that it does not pay for the first touch of fresh device memory
        cls_all = torch.zeros((args.steps, B), dtype=torch.int32, device=dev)
        tie_all = torch.zeros((args.steps, B), dtype=torch.uint8, device=dev)
        mar_all = torch.zeros((args.steps, B, nout), dtype=torch.float64, device=dev)
        for i in range(args.warmup if warmup is None else warmup):
            ct.hip.decode_device(S_dev[i % nwarm].data_ptr(), B, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(),
                                 stream)
        torch.cuda.synchronize()

        # device addresses of every step's batch and outputs, taken before the
        # timed region: the loop itself only enqueues through the C ABI (no
        # tensor views, no allocations, and the cyclic collector paused, so a
        # host stall cannot drain the launch queue and idle the GPU)
        ptrs = [(S_dev[nwarm + k].data_ptr(), mar_all[k].data_ptr(), cls_all[k].data_ptr(), tie_all[k].data_ptr())
                for k in range(args.steps)]
        decode = ct.hip.decode_device

        def steps():
            for s_, m_, c_, t_ in ptrs:
                decode(s_, B, mode, m_, c_, t_, stream)

        torch.cuda.synchronize()
        gc.disable()
        t0 = time.perf_counter()
        steps()
        t_enq = time.perf_counter() - t0
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        gc.enable()
        kmain = krec = 0.0
        calls = 0
        if instrument:
            ct.hip.kernel_time_totals(stream)       # drop earlier events
            ct.hip.set_timing(True)
            steps()
            torch.cuda.synchronize()
            ct.hip.set_timing(False)
            kmain, krec, calls = ct.hip.kernel_time_totals(stream)
        print(f"[bench] {label}: {args.steps} steps in {dt * 1e3:.2f} ms (enqueued in {t_enq * 1e3:.2f} ms); "
              f"instrumented pass: kernels {kmain:.2f} ms + re-check {krec:.2f} ms", file=sys.stderr, flush=True)
        return dt, kmain, krec, calls, cls_all.cpu().numpy(), tie_all.cpu().numpy(), mar_all.cpu().numpy()

    mode_c = plan_cfg["mode"]
    dt_c, km_c, kr_c, calls_c, cls_c, tie_c, mar_c = run(chosen, mode_c, f"chosen {mode_c}")
    thr_c = args.steps * B / dt_c
    res = {}
    truth_t = np.stack(truth[nwarm:])
    n_total = truth_t.size
    err_c = int((cls_c != truth_t).sum())
    res["ler_chosen"] = err_c / n_total
    res["ler_chosen_ci95"] = wilson(err_c, n_total)
    if not args.no_plain:
        dt_p, km_p, kr_p, calls_p, cls_p, tie_p, mar_p = run(plain, "plain_f64", "plain plain_f64",
                                                             warmup=min(args.warmup, 3), instrument=False)
        thr_p = args.steps * B / dt_p
        err_p = int((cls_p != truth_t).sum())
        res["plain_f64_syndromes_per_s"] = thr_p
        res["speedup_vs_plain_f64"] = thr_c / thr_p
        res["ler_plain_f64"] = err_p / n_total
        res["ler_plain_f64_ci95"] = wilson(err_p, n_total)
        ok = tie_p == 0
        res["class_agreement_non_tie"] = float((cls_c[ok] == cls_p[ok]).mean())
        res["ties_fp64"] = int((~ok).sum())
        res["max_rel_marginal_err_fp32_vs_fp64"] = float(
            np.max(np.abs(mar_c - mar_p) / mar_p.max(axis=2, keepdims=True)))
        res["plain_ms_per_step"] = dt_p / args.steps * 1e3
        if mode_c == "spec_f32":
            # the exact FP64 specialised kernel on the chosen plan (register-resident when it fits)
            dt_6, _, _, _, cls_6, tie_6, mar_6 = run(chosen, "spec_f64", "chosen spec_f64", warmup=20, instrument=False)
            ok6 = (tie_p == 0) & (tie_6 == 0)
            res["spec_f64_syndromes_per_s"] = args.steps * B / dt_6
            res["speedup_vs_spec_f64"] = thr_c / res["spec_f64_syndromes_per_s"]
            res["spec_f64_class_agreement_with_plain_f64"] = float((cls_6[ok6] == cls_p[ok6]).mean())
            res["spec_f64_max_rel_marginal_err_vs_plain_f64"] = float(
                np.max(np.abs(mar_6 - mar_p) / mar_p.max(axis=2, keepdims=True)))
        # the plan as the Julia binding hands it over: the reference network
        # (no chain factorization), one-hot evidence leaves, default tree
        # search and tabulation depth, default specialised configuration
        dropin = D.compile(D.TNMMAP(factorize=False, evidence="leaves", mode="spec_f32", tabulate_bits=4), t, em)
        dj = {"plan": "reference network, one-hot evidence leaves, tabulate_bits 4, default spec_f32 configuration",
              "config": dropin.hip.spec_config(), "dep_tc": dropin.hip.stats()["dep_tc"]}
        for mode_j in ("spec_f32", "spec_f64"):
            dt_j, _, _, _, cls_j, tie_j, mar_j = run(dropin, mode_j, f"drop-in plan {mode_j}", warmup=20,
                                                     instrument=False)
            okj = (tie_p == 0) & (tie_j == 0)
            dj[mode_j] = {"syndromes_per_s": args.steps * B / dt_j,
                          "class_agreement_with_plain_f64": float((cls_j[okj] == cls_p[okj]).mean()),
                          "max_rel_marginal_err_vs_plain_f64": float(
                              np.max(np.abs(mar_j - mar_p) / mar_p.max(axis=2, keepdims=True)))}
        res["dropin_plan"] = dj
    st = chosen.hip.stats()
    traffic = measured_traffic(mode_c, plan_cfg, B)
    flops_per_syn = 2.0 * st["dep_tc"]
    avg_ms = km_c / max(calls_c, 1)
    achieved = flops_per_syn * B / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    res["recheck_ms_per_step"] = kr_c / max(calls_c, 1)
    if fused_ms:
        res["generic_fused_f32_syndromes_per_s"] = B / (fused_ms * 1e-3)
    res["recheck_fraction"] = None

    cpu = None
    if not args.no_cpu:
        from oracle.cexec import OraclePlan
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        op = OraclePlan(plain.plan)
        s_all = np.concatenate(S_host)

        def cpu_rate(nthreads, seconds):
            t0 = time.perf_counter()
            op.decode(s_all[:256], nthreads=nthreads)
            rate0 = 256 / max(time.perf_counter() - t0, 1e-6)
            ncpu = int(min(len(s_all), max(256, rate0 * seconds)))
            t0 = time.perf_counter()
            op.decode(s_all[:ncpu], nthreads=nthreads)
            return ncpu, time.perf_counter() - t0

        ncpu, dt = cpu_rate(cores, args.cpu_seconds)
        n1, dt1 = cpu_rate(1, min(6.0, args.cpu_seconds))
        cpu = {"value": ncpu / dt, "unit": "syndromes/s", "cores": cores, "kind": "port",
               "sample": f"first {ncpu} syndromes of the same d=7 p=0.05 batches, FP64, reference network "
                         f"(one-hot evidence, every step per syndrome), OpenMP over syndromes on every core of "
                         f"the process's affinity set, {dt:.1f}s wall; SCIP's pure error not timed",
               "value_1core": n1 / dt1, "sample_1core": f"first {n1} syndromes, one thread, {dt1:.1f}s wall",
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        print(f"[bench] cpu baseline {ncpu / dt:.0f} syn/s on {cores} threads, {n1 / dt1:.0f} syn/s on 1 "
              f"({cpu['cpu_model']})", file=sys.stderr, flush=True)

    if not args.no_extras:
        t0 = time.time()
        res["stream"] = stream_leg(torch, dev, args.seed)
        print(f"[bench] stream leg {time.time() - t0:.1f}s: {res['stream']['syndromes_per_s']:.0f} syn/s", file=sys.stderr,
              flush=True)
        t0 = time.time()
        res["learned"] = learned_leg(torch, dev, args.seed, ct=chosen, mode=mode_c, d=d, p=p)
        print(f"[bench] learned leg {time.time() - t0:.1f}s: {res['learned']}", file=sys.stderr, flush=True)
        t0 = time.time()
        res["learned_d9"] = learned_d9_leg(torch, dev, args.seed)
        print(f"[bench] learned d=9 leg {time.time() - t0:.1f}s: gated {res['learned_d9']['gated_syndromes_per_s']:.0f} "
              f"vs TN {res['learned_d9']['tn_only_syndromes_per_s']:.0f} syn/s", file=sys.stderr, flush=True)
        t0 = time.time()
        res["learned_dem"] = learned_dem_leg(torch, dev, args.seed)
        print(f"[bench] learned DEM leg {time.time() - t0:.1f}s: fallback {res['learned_dem']['fallback_fraction']:.4f}, "
              f"gated {res['learned_dem']['gated_syndromes_per_s']:.0f} vs TN "
              f"{res['learned_dem']['tn_only_syndromes_per_s']:.0f} syn/s", file=sys.stderr, flush=True)
        t0 = time.time()
        res["pipeline"] = pipeline_leg(torch, dev, args.seed, ct=chosen, mode=mode_c, d=d, p=p)
        print(f"[bench] pipeline leg {time.time() - t0:.1f}s: {res['pipeline']['syndromes_per_s']:.0f} syn/s",
              file=sys.stderr, flush=True)
        t0 = time.time()
        res["mlp"] = mlp_leg(torch, dev, args.seed)
        print(f"[bench] mlp leg {time.time() - t0:.1f}s: fused {res['mlp']['fused']['tflops']:.0f} TFLOP/s, "
              f"layer by layer {res['mlp']['layer_by_layer']['tflops']:.0f} TFLOP/s", file=sys.stderr, flush=True)
        t0 = time.time()
        res["tnmap"] = tnmap_leg(torch, dev, args.seed)
        print(f"[bench] tnmap leg {time.time() - t0:.1f}s: {res['tnmap']}", file=sys.stderr, flush=True)
        t0 = time.time()
        res["dem"] = dem_leg(torch, dev, args.seed)
        print(f"[bench] dem leg {time.time() - t0:.1f}s: {res['dem']['modes']}", file=sys.stderr, flush=True)
        t0 = time.time()
        res["toric5"] = toric5_leg(torch, dev, args.seed)
        print(f"[bench] toric5 leg {time.time() - t0:.1f}s: {res['toric5']['modes']}", file=sys.stderr, flush=True)
        t0 = time.time()
        try:
            res["toric6"] = toric6_leg(torch, dev, args.seed)
        except Exception as e:      # recorded in the JSON line, which still prints
            res["toric6"] = {"error": f"{type(e).__name__}: {e}"}
        print(f"[bench] toric6 leg {time.time() - t0:.1f}s: {res['toric6'].get('modes', res['toric6'])}",
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": thr_c,
        "unit": "syndromes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_c / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox4x32-10 seeded depolarizing samples)",
        "config": {"workload": f"rotated surface code d={d}, code-capacity depolarizing p={p} (p/3 each), "
                               f"{B} syndromes per step, TNMMAP marginals over 4 logical classes",
                   "batch": B, "d": d, "p": p, "seed": args.seed,
                   "chosen": f"{mode_c} (chain-factorized network, evidence as slices, subtrees on <= "
                             f"{plan_cfg.get('tabulate_bits', 4)} syndrome bits tabulated, FP32 with per-tensor "
                             "power-of-two scaling, FP64 re-check"
                             + (", plan-specialised kernel compiled with hiprtc" if mode_c == "spec_f32" else "")
                             + (", one syndrome per quad of lanes with every tensor in registers"
                                if plan_cfg.get("reg") else "")
                             + ")",
                   "plan": plan_cfg,
                   "plan_stats": {k: st[k] for k in ("n_dep_steps", "dep_tc", "dep_rw", "fused_lds_entries",
                                                     "table_entries", "const_entries")},
                   "plain": "plain_f64 (reference network, one-hot evidence, one kernel per step)"},
        "roofline": {"bound": "fp32_valu" if mode_c in ("spec_f32", "fused_f32") else "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "tqec_spec_f32" if mode_c == "spec_f32" else "fused_kernel<float, true>",
                     "kernel_ms_avg": avg_ms,
                     "flops_per_syndrome": flops_per_syn,
                     "compute": ("FP32 vector ALU, tensors in registers, DPP quad exchanges, tables in "
                                 + ("LDS" if plan_cfg.get("tables_in_lds") else "global memory")
                                 + " (no MFMA on this path)") if plan_cfg.get("reg") else
                                "FP32 vector ALU with LDS-resident operands (no MFMA on this path)"},
        "cpu_baseline": cpu,
        "extra": res,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
