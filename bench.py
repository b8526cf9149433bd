"""Throughput of TNMMAP decoding on one MI355X.

Workload: rotated surface code d=7 (49 data qubits, 48 stabilizers), code-
capacity depolarizing noise p=0.05 (X, Y, Z each p/3), 65,536 Philox-seeded
syndromes per step, already resident in HBM.  One step decodes one batch:
marginals P(syndrome, class) for the 4 logical classes, the arg-max class and
the tie flag per syndrome.

The reported value is the chosen executor (fused FP32 with per-tensor
scaling and FP64 re-check of low-margin syndromes, on the chain-factorized
network).  The plain FP64 executor (one kernel per step, reference network
with one-hot evidence) is timed on the same batches; both logical error
rates are reported with 95% Wilson intervals.  The CPU baseline is the
oracle's C FP64 per-syndrome executor on the reference network, timed on a
bounded sample on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
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
                    help="SEED,SC_TARGET,MODE,TILE,THREADS[,TABLES_IN_LDS,WAVES]: pin the plan, executor and tile "
                         "(no autotuning); SC_TARGET may be none; MODE fused_f32 or spec_f32")
    ap.add_argument("--no-spec", action="store_true", help="do not consider the plan-specialised kernel")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied once by a plain device copy before the timed steps (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (syndrome batches shard "
              "across GPUs with no exchange, see the design notes).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from tensorqec import decoders as D
    from tensorqec.codes import CSSTannerGraph, SurfaceCode
    from tensorqec.errors import css_syndrome_rows, iid_error, sample_depolarizing

    dev = torch.device("cuda:0")
    d, p, B = args.d, args.p, args.batch
    t = CSSTannerGraph.from_code(SurfaceCode(d, d))
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    nb = args.warmup + args.steps
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
    if pinned:
        cands = [(pinned["seed"], pinned["sc_target"])]
    else:
        cands = [(1, 10), (2, 10), (3, 10), (1, 9), (3, None), (2, 9), (1, None)]
    outs = [torch.empty((B, 4), dtype=torch.float64, device=dev), torch.empty(B, dtype=torch.int32, device=dev),
            torch.empty(B, dtype=torch.uint8, device=dev)]

    def time_mode(ct, mode):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(3):
            ct.hip.decode_device(S_dev[0].data_ptr(), B, mode, outs[0].data_ptr(), outs[1].data_ptr(),
                                 outs[2].data_ptr(), stream)
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / 3

    # Candidate contraction trees (seed, width target): each is autotuned with
    # the generic fused kernel; the two fastest are also autotuned with the
    # plan-specialised kernel; the fastest (plan, executor, tile) is kept.
    tried = []
    for seed, sct in cands:
        ct = D.compile(D.TNMMAP(mode="fused_f32", optimizer=D.TreeSA(seed=seed, sc_target=sct)), t, em)
        st = ct.hip.stats()
        if pinned:
            cfg = {"seed": seed, "sc_target": sct, "mode": pinned["mode"], "tile": pinned["tile"],
                   "threads": pinned["threads"]}
            if pinned["mode"] == "spec_f32":
                ct.hip.set_spec(pinned["tile"], pinned["threads"], pinned.get("tables_in_lds", False),
                                pinned.get("waves", 0))
                cfg.update(ct.hip.spec_config())
            else:
                ct.hip.set_tile("fused_f32", pinned["tile"], pinned["threads"])
            tried.append((0.0, ct, cfg))
            continue
        tile, thr = ct.hip.autotune("fused_f32", S_dev[0].data_ptr(), B, stream)
        ms = time_mode(ct, "fused_f32")
        print(f"[bench] candidate seed={seed} sc_target={sct}: steps={st['n_dep_steps']} macs={st['dep_tc']:.0f} "
              f"lds/syn={st['fused_lds_entries']} fused_f32 tile={tile} threads={thr} {ms:.3f} ms/batch",
              file=sys.stderr, flush=True)
        tried.append((ms, ct, {"seed": seed, "sc_target": sct, "mode": "fused_f32", "tile": tile, "threads": thr}))
    tried.sort(key=lambda x: x[0])
    best = tried[0]
    if not pinned and not args.no_spec:
        for ms_f, ct, cfg in tried[:2]:
            t1 = time.time()
            ct.hip.autotune("spec_f32", S_dev[0].data_ptr(), B, stream)
            sc = ct.hip.spec_config()
            ms = time_mode(ct, "spec_f32")
            print(f"[bench] candidate seed={cfg['seed']} sc_target={cfg['sc_target']}: spec_f32 {sc} {ms:.3f} ms/batch "
                  f"(tuned in {time.time() - t1:.1f}s)", file=sys.stderr, flush=True)
            if ms < best[0]:
                best = (ms, ct, dict(cfg, mode="spec_f32", **sc))
    chosen, plan_cfg = best[1], best[2]
    print(f"[bench] plans compiled and tuned in {time.time() - t0:.1f}s: {plan_cfg}", file=sys.stderr, flush=True)
    truth = [true_classes(e, chosen.lx, chosen.lz) for e in eps]
    nout = 1 << chosen.hip.n_out
    mar = torch.empty((B, nout), dtype=torch.float64, device=dev)
    cls = torch.empty(B, dtype=torch.int32, device=dev)
    tie = torch.empty(B, dtype=torch.uint8, device=dev)
    print(f"[bench] {nb} batches of {B} syndromes resident", file=sys.stderr, flush=True)
    if args.copy_calibration > 0:
        # one plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the HBM byte counters of the decode kernels
        n_el = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n_el, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def run(ct, mode, label):
        outs = []
        for i in range(args.warmup):
            ct.hip.decode_device(S_dev[i].data_ptr(), B, mode, mar.data_ptr(), cls.data_ptr(), tie.data_ptr(),
                                 stream)
        torch.cuda.synchronize()
        ct.hip.kernel_time_totals(stream)       # drop warm-up events
        ct.hip.set_timing(True)
        cls_all = torch.empty((args.steps, B), dtype=torch.int32, device=dev)
        tie_all = torch.empty((args.steps, B), dtype=torch.uint8, device=dev)
        mar_all = torch.empty((args.steps, B, nout), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            i = args.warmup + k
            ct.hip.decode_device(S_dev[i].data_ptr(), B, mode, mar_all[k].data_ptr(), cls_all[k].data_ptr(),
                                 tie_all[k].data_ptr(), stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ct.hip.set_timing(False)
        kmain, krec, calls = ct.hip.kernel_time_totals(stream)
        print(f"[bench] {label}: {args.steps} steps in {dt * 1e3:.2f} ms", file=sys.stderr, flush=True)
        return dt, kmain, krec, calls, cls_all.cpu().numpy(), tie_all.cpu().numpy(), mar_all.cpu().numpy()

    mode_c = plan_cfg["mode"]
    dt_c, km_c, kr_c, calls_c, cls_c, tie_c, mar_c = run(chosen, mode_c, f"chosen {mode_c}")
    thr_c = args.steps * B / dt_c
    res = {}
    truth_t = np.stack(truth[args.warmup:])
    n_total = truth_t.size
    err_c = int((cls_c != truth_t).sum())
    res["ler_chosen"] = err_c / n_total
    res["ler_chosen_ci95"] = wilson(err_c, n_total)
    if not args.no_plain:
        dt_p, km_p, kr_p, calls_p, cls_p, tie_p, mar_p = run(plain, "plain_f64", "plain plain_f64")
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
        res["plain_kernel_ms_per_step"] = km_p / max(calls_p, 1)
    st = chosen.hip.stats()
    flops_per_syn = 2.0 * st["dep_tc"]
    avg_ms = km_c / max(calls_c, 1)
    achieved = flops_per_syn * B / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    res["recheck_ms_per_step"] = kr_c / max(calls_c, 1)
    res["recheck_fraction"] = None

    cpu = None
    if not args.no_cpu:
        from oracle.cexec import OraclePlan
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        cores = max(1, min(cores, 16))
        op = OraclePlan(plain.plan)
        s_all = np.concatenate(S_host)
        t0 = time.perf_counter()
        op.decode(s_all[:512], nthreads=cores)
        rate0 = 512 / max(time.perf_counter() - t0, 1e-6)
        ncpu = int(min(len(s_all), max(1024, rate0 * args.cpu_seconds)))
        t0 = time.perf_counter()
        _, ccls, _ = op.decode(s_all[:ncpu], nthreads=cores)
        dt = time.perf_counter() - t0
        cpu = {"value": ncpu / dt, "unit": "syndromes/s", "cores": cores, "kind": "port",
               "sample": f"first {ncpu} syndromes of the same d=7 p=0.05 batches, FP64, reference network "
                         f"(one-hot evidence, every step per syndrome), OpenMP over syndromes, "
                         f"{dt:.1f}s wall"}
        print(f"[bench] cpu baseline {ncpu / dt:.0f} syn/s on {cores} threads", file=sys.stderr, flush=True)

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
                   "chosen": f"{mode_c} (chain-factorized network, evidence as slices, subtrees on <= 4 "
                             "syndrome bits tabulated, FP32 with per-tensor power-of-two scaling, FP64 re-check"
                             + (", plan-specialised kernel compiled with hiprtc" if mode_c == "spec_f32" else "")
                             + ")",
                   "plan": plan_cfg,
                   "plan_stats": {k: st[k] for k in ("n_dep_steps", "dep_tc", "dep_rw", "fused_lds_entries",
                                                     "table_entries", "const_entries")},
                   "plain": "plain_f64 (reference network, one-hot evidence, one kernel per step)"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "tqec_spec" if mode_c == "spec_f32" else "fused_kernel<float, true>",
                     "kernel_ms_avg": avg_ms,
                     "flops_per_syndrome": flops_per_syn,
                     "compute": "FP32 vector ALU with LDS-resident operands (no MFMA on this path)"},
        "cpu_baseline": cpu,
        "extra": res,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
