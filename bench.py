#!/usr/bin/env python3
"""Benchmark of the rcsr device arm on one MI355X.

Workload (one "step"): one batched LM fit of 1024 candidates with
k ~ U{1..20} constants over N = 262144 rows of 3 inputs, 30 LM iterations
plus the exact f64 re-baseline and 4-step polish, through the C ABI
(rc_fit_batch) with every input already resident in HBM.  Metric: LM-fitted
candidates per second of the mixed-precision arm (FP32 interpreter and
adjoint, FP32-MFMA split Gram, FP64 solve and acceptance).

Beside it, in the same run:
  - the plain FP64 device arm on the same batch (the speed-up denominator);
  - the fitted-loss relative error of the mixed arm against the FP64 device
    arm (all candidates) and against the CPU f64 oracle (a bounded sample);
  - the CPU f64 oracle (a C++ restatement of the reference's fit_lm) timed on
    the host cores on a bounded sample of the same batch;
  - the assembly kernel's roofline: algorithmic flops 2 N (k^2 + k) per
    candidate per assembly over the measured assembly time, against the FP32
    MFMA peak.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "the-rule-and-compass_amd"))

METRIC = "LM-fitted candidates/s at N=262144 (mixed-prec+adjoint vs plain f64); row-evals/s"
FP32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_SUMMARY = os.path.join("profiles", "vm_asm_c3_traffic.md")


def committed_traffic():
    """HBM bytes per assembly launch at this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy of
    known size); None when no summary is present."""
    import re
    try:
        with open(os.path.join(ROOT, TRAFFIC_SUMMARY)) as fh:
            text = fh.read()
    except OSError:
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*(\d+)", text)
    return int(m.group(1)) if m else None


def copy_probe():
    """A plain 256 MiB device-to-device copy, three times: under rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE it gives the counters' reading for a known
    number of HBM bytes (256 MiB read and 256 MiB written per dispatch)."""
    import torch
    nbytes = 256 << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                      "dispatches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--cands", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=30)
    ap.add_argument("--fp64-cands", type=int, default=-1,
                    help="candidates in the FP64-arm timing (-1: all)")
    ap.add_argument("--cpu-sample", type=int, default=20,
                    help="CPU baseline sample: one candidate per constant count, at most this many")
    ap.add_argument("--verify-per-k", type=int, default=3,
                    help="C3 convergence / CPU-verification sample: candidates per constant count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-eval", action="store_true",
                    help="skip the batched-evaluation (row-evals/s) part")
    ap.add_argument("--eval-only", action="store_true",
                    help="only the batched-evaluation part (one JSON object)")
    ap.add_argument("--eval-n", type=int, default=1048576)
    ap.add_argument("--eval-progs", type=int, default=4096)
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the uneven-batch (autotuned vs fixed shapes) part")
    ap.add_argument("--no-arms", action="store_true",
                    help="skip the logistic / correlation fit arm lines")
    ap.add_argument("--no-order", action="store_true",
                    help="skip the exact order / rank / maximum stages part")
    ap.add_argument("--order-only", action="store_true",
                    help="only the exact order / rank / maximum stages (one JSON object)")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the dispatch device-rate probe")
    ap.add_argument("--no-implicit", action="store_true",
                    help="skip the implicit-relation fit part")
    ap.add_argument("--implicit-cands", type=int, default=2048)
    ap.add_argument("--implicit-cpu", type=int, default=7,
                    help="implicit CPU-oracle sample: one candidate per constant count 2..8")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a device copy of 256 MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one MI355X; "
              f"--gpus {args.gpus} is not supported (candidate sharding needs no "
              f"communication, see the design notes).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from rcsr_hip import host, lib, workloads

    if args.eval_only:
        print(json.dumps(eval_part(args, host, workloads)), flush=True)
        return
    if args.order_only:
        print(json.dumps(order_part(host, workloads)), flush=True)
        return

    t0 = time.time()
    wl = workloads.config3(n=args.n, ncand=args.cands, iters=args.iters)
    plans = [host.gpu_fit_plan(t) for t in wl.trees]
    assert all(p is not None for p in plans)
    ks = np.asarray([p.nconst for p in plans])
    log(f"workload ready in {time.time() - t0:.1f}s: {len(plans)} candidates, N={wl.n}, "
        f"k mean {ks.mean():.2f}, nodes mean {np.mean([p.nodes for p in plans]):.1f}")

    ev = host.GpuEvaluator(wl.cols32(), wl.n, wl.v, wl.y.astype(np.float32))
    assert ev.set_fit_data(wl.cols64(), wl.y)
    L = lib.load()

    def fit():
        outs, ms = ev.fit_batch(plans, args.iters, 0, kmax=20)
        if ms < 0:
            raise RuntimeError("rc_fit_batch failed")
        return outs, ms

    for w in range(args.warmup):
        ts = time.perf_counter()
        outs, ms = fit()
        log(f"warmup {w}: {time.perf_counter() - ts:.3f}s (device {ms:.1f} ms)")
    times = []
    for s in range(args.steps):
        ts = time.perf_counter()
        outs, ms = fit()
        times.append(time.perf_counter() - ts)
        log(f"step {s}: {times[-1]:.3f}s (device {ms:.1f} ms)")
    step_s = float(np.mean(times))
    value = len(plans) / step_s
    statuses = np.asarray([o.status for o in outs])
    mixed_loss = np.asarray([o.sse for o in outs])

    # Per-pass timing of one extra (untimed) fit for the roofline.
    import ctypes as C
    L.rc_fit_profile(ev.ctx, 1)
    fit()
    rep = lib.RcFitReport()
    per = np.zeros(5 * len(plans), np.int32)
    L.rc_fit_report(ev.ctx, C.byref(rep), lib.ptr(per, lib.i32p), len(plans))
    L.rc_fit_profile(ev.ctx, 0)
    per = per.reshape(-1, 5)
    flops = float(np.sum(per[:, 0] * 2.0 * wl.n * (ks.astype(np.float64) ** 2 + ks)))
    asm_tflops = flops / (rep.asm_ms * 1e-3) / 1e12 if rep.asm_ms > 0 else 0.0
    # Exact-loss pass bytes: cols64 (3 x 8 B) + y64 (8 B) per row per candidate pass.
    log(f"profile: asm {rep.asm_ms:.1f} ms / {rep.asm_launches} passes, trial {rep.trial_ms:.1f} ms "
        f"/ {rep.trial_launches}, exact {rep.exact_ms:.1f} ms / {rep.exact_launches}, ctrl "
        f"{rep.ctrl_ms:.1f} ms / {rep.ctrl_launches}; assemblies/cand {per[:, 0].mean():.1f}")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 rows + f64 solve/acceptance",
        "data": "synthetic (Philox-seeded C3 shape)",
        "config": {
            "workload": wl.name,
            "candidates": len(plans),
            "n_rows": wl.n,
            "inputs": wl.v,
            "lm_iterations": args.iters,
            "k_mean": float(ks.mean()),
            "metric": "squared",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "vm_asm_kernel (fit VM forward + adjoint + FP32-MFMA Gram)",
            "achieved": asm_tflops,
            "peak": FP32_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": asm_tflops / FP32_MFMA_PEAK_TFLOPS,
            "traffic": committed_traffic(),
            "traffic_source": TRAFFIC_SUMMARY if committed_traffic() is not None else None,
            "asm_ms_per_fit": rep.asm_ms,
            "asm_timing": "HIP events on the fit stream around each assembly pass, whose class "
                          "kernels run on two streams (the union of their intervals)",
            "asm_launches": rep.asm_launches,
            "trial_ms_per_fit": rep.trial_ms,
            "exact_ms_per_fit": rep.exact_ms,
            "ctrl_ms_per_fit": rep.ctrl_ms,
        },
        "status_counts": {str(int(s)): int(np.sum(statuses == s)) for s in np.unique(statuses)},
    }

    # Per-candidate LM accept/reject record of the last (profiled) mixed fit.
    mixed_bits, mixed_nd = per[:, 3].copy(), per[:, 4].copy()

    # Candidates stratified by constant count: up to `per_k` of each k, in
    # batch order (the CPU legs run on such samples, so their k mix matches
    # the batch's instead of its first rows').
    def stratified(per_k):
        pick = []
        for kk in range(1, 21):
            pick += [i for i in range(len(plans)) if ks[i] == kk][:per_k]
        return sorted(pick)

    # The plain FP64 device arm on the same batch.
    if not args.no_fp64:
        nsub = len(plans) if args.fp64_cands < 0 else min(args.fp64_cands, len(plans))
        ev.set_fit_mode(host.FIT_MODE_FP64)
        ts = time.perf_counter()
        f64outs, ms64 = ev.fit_batch(plans[:nsub], args.iters, 0, kmax=20)
        t64 = time.perf_counter() - ts
        rep64 = lib.RcFitReport()
        per64 = np.zeros(5 * nsub, np.int32)
        L.rc_fit_report(ev.ctx, C.byref(rep64), lib.ptr(per64, lib.i32p), nsub)
        per64 = per64.reshape(-1, 5)
        ev.set_fit_mode(host.FIT_MODE_MIXED)
        if ms64 < 0:
            raise RuntimeError("FP64 arm failed")
        v64 = nsub / t64
        f64loss = np.asarray([o.sse for o in f64outs])
        okm = np.array([host.fit_status_usable(outs[i].status) and outs[i].status != 7 and
                        host.fit_status_usable(f64outs[i].status) and f64outs[i].status != 7
                        and np.isfinite(f64loss[i]) and f64loss[i] > 0 for i in range(nsub)])
        signed = (mixed_loss[:nsub][okm] - f64loss[okm]) / f64loss[okm]
        rel = np.abs(signed)
        st_m = statuses[:nsub]
        st_f = np.asarray([o.status for o in f64outs])
        log(f"fp64 arm: {nsub} candidates in {t64:.2f}s -> {v64:.2f} cand/s; mixed vs fp64 "
            f"loss rel err median {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e}, "
            f"max {np.max(rel):.2e}; statuses agree {np.mean(st_m == st_f):.3f}")
        # The first LM decision where the two arms part (their accept/reject
        # records), for the candidates outside 1e-6.
        idx_ok = np.nonzero(okm)[0]
        diverge = []
        for i in idx_ok[rel > 1e-6]:
            a, b = int(mixed_bits[i]), int(per64[i, 3])
            nd = min(int(mixed_nd[i]), int(per64[i, 4]), 31)
            x = (a ^ b) & ((1 << nd) - 1)
            diverge.append((x & -x).bit_length() - 1 if x else nd)
        result["fp64_device"] = {
            "value": v64, "unit": "candidates/s", "candidates": nsub, "seconds": t64,
            "design": "same decomposition as the mixed arm (work items, hipGraph, control "
                      "kernel); f64 rows, f64 value and adjoint sweeps, FP64-MFMA Gram, "
                      "f64 trial losses"}
        result["speedup_vs_fp64"] = value / v64
        result["loss_rel_err_vs_fp64_device"] = {
            "median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
            "max": float(np.max(rel)), "frac_le_1e-6": float(np.mean(rel <= 1e-6)),
            "compared": int(okm.sum()),
            "signed_quantiles": {q: float(np.quantile(signed, float(q)))
                                 for q in ("0.01", "0.05", "0.5", "0.95", "0.99")},
            "signed_min": float(signed.min()), "signed_max": float(signed.max()),
            "frac_mixed_worse_gt_1e-6": float(np.mean(signed > 1e-6)),
            "frac_mixed_better_gt_1e-6": float(np.mean(signed < -1e-6)),
            "status_agreement": float(np.mean(st_m == st_f)),
            "status_pairs": {f"{a}->{b}": int(np.sum((st_m == a) & (st_f == b)))
                             for a in np.unique(st_m) for b in np.unique(st_f)
                             if np.sum((st_m == a) & (st_f == b)) > 0},
            "first_lm_decision_that_differs_outside_1e-6": {
                str(d): int(diverge.count(d)) for d in sorted(set(diverge))}}
        if not args.no_cpu and args.verify_per_k > 0:
            result["c3_convergence_and_verification"] = verify_part(
                args, wl, plans, ks, outs, f64outs, ev, stratified(args.verify_per_k))

    # The CPU f64 oracle on a bounded sample of the same batch: one candidate
    # of each constant count 1..20.
    if not args.no_cpu and args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = min(os.cpu_count() or 1, 16)
        sample = stratified(1)[:args.cpu_sample]
        ns = len(sample)
        trees_s = [wl.trees[i] for i in sample]
        ts = time.perf_counter()
        cl, _ = oracle.fit_batch(trees_s, wl.x, wl.y, None, args.iters, threads=threads, kmax=20)
        tc = time.perf_counter() - ts
        ok = np.isfinite(cl) & (cl > 0) & np.array([outs[i].status != 7 for i in sample])
        rel = np.abs(mixed_loss[sample][ok] - cl[ok]) / cl[ok]
        log(f"cpu oracle: {ns} candidates on {threads} threads in {tc:.1f}s; mixed vs oracle "
            f"median {np.median(rel):.2e} max {np.max(rel):.2e}")
        # The reference's CPU fallback runs min(cores, 8) workers.
        w8 = min(os.cpu_count() or 1, 8)
        ts = time.perf_counter()
        oracle.fit_batch(trees_s, wl.x, wl.y, None, args.iters, threads=w8, kmax=20)
        t8 = time.perf_counter() - ts
        log(f"cpu oracle at {w8} workers: {ns / t8:.2f} candidates/s")
        kmix = ks[sample]
        result["cpu_baseline"] = {
            "value": ns / tc, "unit": "candidates/s", "cores": threads, "kind": "port",
            "sample": f"{ns} of the {len(plans)} candidates, one per constant count "
                      f"k = {int(kmix.min())}..{int(kmix.max())} (sample mean k "
                      f"{kmix.mean():.2f}, batch mean {ks.mean():.2f}), full N={wl.n}, "
                      f"{args.iters} LM iterations + 4 polish, one candidate per thread",
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
            "value_at_min_cores_8": ns / t8, "workers_at_min_cores_8": w8}
        result["loss_rel_err_vs_cpu_oracle"] = {
            "median": float(np.median(rel)), "max": float(np.max(rel)),
            "compared": int(ok.sum())}
    ev.close()

    # Batched evaluation: 4096 programs (3-60 nodes) x N = 1048576 rows of 5
    # f32 columns, inputs resident; the metric's row-evals/s.
    if not args.no_eval:
        result["eval"] = eval_part(args, host, workloads)
    # Uneven batch (8 islands x 512, N log-uniform in [512, 4194304], mixed
    # squared / L1): the built-in launch shapes against the autotuned ones.
    if not args.no_c4:
        result["c4_shapes"] = c4_part(host, workloads)
        result["small_batch_shapes"] = small_batch_part(host, workloads)
    # Exact order / rank / maximum stages on the C3 rows.
    if not args.no_order:
        result["order_stages"] = order_part(host, workloads)
    # The logistic and correlation fit arms on a sample of the C3 batch.
    if not args.no_arms:
        result["other_fit_arms"] = arms_part(host, workloads)
    # The dispatch cost model's device rates, probed on this device at the
    # reference's calibration shape, and the fit route they give this batch.
    if not args.no_probe:
        result["device_rates"] = probe_part(host, result)
    # Implicit relations: 2048 candidates f(x1, x2, y) = 0 with up to 8
    # constants over N = 1048576 rows, 8 LM iterations, LM control on the
    # device; a CPU-oracle sample beside it.
    if not args.no_implicit:
        result["implicit"] = implicit_part(args, host, workloads)
    print(json.dumps(result), flush=True)


def verify_part(args, wl, plans, ks, outs, f64outs, ev, sample):
    """On a k-stratified sample of the C3 batch: which candidates the f64 LM
    has converged by the bench's iteration count (the FP64 arm at 120
    iterations agrees to 1e-9), the mixed arm's loss error against the FP64
    arm on the converged and unconverged ones, and the CPU verification of
    both arms (refit_finalists_exact, fit/gpu.rs:666-687: the exact f64 LM
    warm-started from each arm's constants) with the two verified Pareto
    fronts over (loss, node count)."""
    import math
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from rcsr_hip import host
    sub = [plans[i] for i in sample]
    ev.set_fit_mode(host.FIT_MODE_FP64)
    long_outs, ms = ev.fit_batch(sub, 120, 0, kmax=20)
    ev.set_fit_mode(host.FIT_MODE_MIXED)
    if ms < 0:
        raise RuntimeError("FP64 arm (120 iterations) failed")
    rows = []
    for j, i in enumerate(sample):
        m, f, fl = outs[i], f64outs[i], long_outs[j]
        usable = all(host.fit_status_usable(o.status) and o.status != 7 for o in (m, f, fl))
        if not usable or not (f.sse > 0 and fl.sse > 0):
            continue
        conv = abs(f.sse - fl.sse) <= 1e-9 * fl.sse
        rows.append((i, conv, (m.sse - f.sse) / f.sse))
    conv = np.array([r[1] for r in rows])
    err = np.array([r[2] for r in rows])
    # CPU verification of both arms from their constants.
    threads = min(os.cpu_count() or 1, 16)
    def verified(res):
        ts = [wl.trees[i].with_constants(res[i].c) if host.fit_status_usable(res[i].status)
              else wl.trees[i] for i in sample]
        l, _ = oracle.fit_batch(ts, wl.x, wl.y, None, args.iters, threads=threads, kmax=20)
        return np.asarray(l)
    t0 = time.perf_counter()
    vm, vf = verified(outs), verified(f64outs)
    tv = time.perf_counter() - t0
    sizes = [plans[i].nodes for i in sample]
    def front(scores):
        idx = [j for j in range(len(sample)) if math.isfinite(scores[j])]
        idx.sort(key=lambda j: (sizes[j], scores[j]))
        out, best = [], math.inf
        for j in idx:
            if scores[j] < best * (1 - 1e-9):
                out.append((int(sample[j]), int(sizes[j]), float(scores[j])))
                best = scores[j]
        return out
    fm, ff = front(vm), front(vf)
    # Identical fronts: the same (node count, verified loss) points to 1e-6;
    # two candidates that tie on both may swap places.
    same = len(fm) == len(ff) and all(a[1] == b[1] and abs(a[2] - b[2]) <= 1e-6 * abs(b[2])
                                      for a, b in zip(fm, ff))
    same_members = same and all(a[0] == b[0] for a, b in zip(fm, ff))
    # The CPU oracle's own 30-iteration fit on the sample: how far each
    # device arm's unconverged trajectories are from an independent f64 LM.
    cpu30, _ = oracle.fit_batch([wl.trees[i] for i in sample], wl.x, wl.y, None, args.iters,
                                threads=threads, kmax=20)
    cpu30 = np.asarray(cpu30)
    def vs_cpu(res):
        rel = np.array([abs(res[i].sse - cpu30[j]) / cpu30[j] if cpu30[j] > 0 else np.nan
                        for j, i in enumerate(sample)])
        cv = np.array([r[1] for r in rows]) if rows else np.array([], bool)
        idx = {r[0]: q for q, r in enumerate(rows)}
        conv_j = np.array([idx.get(i) is not None and cv[idx[i]] for i in sample])
        ok = np.isfinite(rel)
        out = {"median": float(np.median(rel[ok])), "max": float(np.max(rel[ok])),
               "frac_le_1e-6": float(np.mean(rel[ok] <= 1e-6))}
        if (ok & conv_j).any():
            out["converged_max"] = float(np.max(rel[ok & conv_j]))
        if (ok & ~conv_j).any():
            out["unconverged_median"] = float(np.median(rel[ok & ~conv_j]))
            out["unconverged_frac_le_1e-6"] = float(np.mean(rel[ok & ~conv_j] <= 1e-6))
        return out
    mixed_cpu, fp64_cpu = vs_cpu(outs), vs_cpu(f64outs)
    fin = np.isfinite(vm) & np.isfinite(vf) & (vf > 0)
    vrel = (vm[fin] - vf[fin]) / vf[fin]
    log(f"C3 verification sample: {len(sample)} candidates, {int(conv.sum())} converged by "
        f"{args.iters} iterations; mixed vs fp64 on converged max |rel| "
        f"{np.max(np.abs(err[conv])) if conv.any() else 0:.2e}, unconverged median |rel| "
        f"{np.median(np.abs(err[~conv])) if (~conv).any() else 0:.2e}; verified fronts "
        f"identical: {same} ({len(fm)} points); CPU refit {tv:.1f}s")
    return {
        "sample": f"{len(sample)} candidates, up to {args.verify_per_k} per constant count",
        "compared": len(rows),
        "converged": int(conv.sum()),
        "converged_max_abs_rel_err": float(np.max(np.abs(err[conv]))) if conv.any() else None,
        "converged_frac_le_1e-6": float(np.mean(np.abs(err[conv]) <= 1e-6)) if conv.any() else None,
        "unconverged_median_abs_rel_err": float(np.median(np.abs(err[~conv]))) if (~conv).any() else None,
        "unconverged_signed_quantiles": {
            q: float(np.quantile(err[~conv], float(q))) for q in ("0.05", "0.5", "0.95")}
            if (~conv).any() else None,
        "verified_rel_mixed_vs_fp64": {
            "median": float(np.median(vrel)), "max": float(np.max(vrel)),
            "min": float(np.min(vrel)), "frac_mixed_worse_gt_1e-6": float(np.mean(vrel > 1e-6))},
        "verified_front_mixed": fm, "verified_front_fp64": ff,
        "verified_fronts_identical": bool(same),
        "verified_fronts_same_members": bool(same_members),
        "mixed_vs_cpu_oracle_30": mixed_cpu, "fp64_device_vs_cpu_oracle_30": fp64_cpu,
        "cpu_refit_seconds": tv,
    }


def arms_part(host, workloads, ncand=256):
    """The logistic (metric 5, labels y > midpoint of the C3 target) and
    correlation (metric 2) arms: the first 256 C3 candidates, N = 262144,
    30 LM iterations, inputs resident; candidates/s over the best of 2."""
    import numpy as np
    wl = workloads.config3(n=262144, ncand=ncand, iters=30)
    plans = [host.gpu_fit_plan(t) for t in wl.trees]
    ev = host.GpuEvaluator(wl.cols32(), wl.n, wl.v, wl.y.astype(np.float32))
    assert ev.set_fit_data(wl.cols64(), wl.y)
    mid = 0.5 * (wl.y.min() + wl.y.max())
    assert ev.set_fit_sign(np.where(wl.y > mid, 1.0, -1.0))
    out = {"workload": f"first {ncand} C3 candidates, N={wl.n}, 30 LM iterations"}
    for name, metric in (("logistic", 5), ("correlation", 2)):
        outs, ms = ev.fit_batch(plans, 30, metric, kmax=20)
        walls = []
        for _ in range(2):
            t0 = time.perf_counter()
            outs, ms = ev.fit_batch(plans, 30, metric, kmax=20)
            walls.append(time.perf_counter() - t0)
        if ms < 0:
            raise RuntimeError(f"{name} arm failed")
        st = {}
        for o in outs:
            st[str(o.status)] = st.get(str(o.status), 0) + 1
        out[name] = {"candidates_per_s": ncand / min(walls), "ms_wall": 1e3 * min(walls),
                     "status_counts": st}
        log(f"{name} arm: {ncand / min(walls):.0f} candidates/s ({1e3 * min(walls):.1f} ms)")
    ev.close()
    return out


def probe_part(host, result):
    """host.probe_device_rates (row-evals/s over 2,456,306 rows, fits/s at
    4 constants x 8 LM iterations on 2,456,306 rows, upload bytes/s, launch
    seconds) beside the reference's anchors, and the route the cost model
    then picks for this batch's fit stage with the measured CPU rate."""
    from rcsr_hip import dispatch as dsp
    t0 = time.perf_counter()
    rates = host.probe_device_rates()
    probe_s = time.perf_counter() - t0
    anchors = dsp.default_device_rates()
    out = {"probe_seconds": probe_s,
           "measured": {"r_eval": rates.r_eval, "r_fit": rates.r_fit, "bw": rates.bw,
                        "launch_s": rates.launch_s},
           "reference_anchors": {"r_eval": anchors.r_eval, "r_fit": anchors.r_fit,
                                 "bw": anchors.bw, "launch_s": anchors.launch_s},
           "fit_ref": {"rows": dsp.FIT_REF_ROWS, "constants": dsp.FIT_REF_CONSTANTS,
                       "sweeps": dsp.FIT_REF_SWEEPS}}
    log(f"device rates: r_eval {rates.r_eval:.3e}/s, r_fit {rates.r_fit:.2f}/s, bw "
        f"{rates.bw:.3e} B/s, launch {rates.launch_s * 1e6:.1f} us ({probe_s:.1f}s)")
    cfg = result.get("config", {})
    cpu = result.get("cpu_baseline")
    if cpu and cfg:
        rows, k, sweeps = cfg["n_rows"], cfg["k_mean"], cfg["lm_iterations"]
        # The CPU rate at the reference shape, from the measured per-candidate
        # rate at this batch's shape (the model's inverse scaling).
        cpu_fit_ref = cpu["value"] * (rows / dsp.FIT_REF_ROWS) * (k / dsp.FIT_REF_CONSTANTS) * \
            (sweeps / dsp.FIT_REF_SWEEPS)
        cal = dsp.with_device_rates(dsp.Calibration(
            r_eval=anchors.r_eval, r_fit=anchors.r_fit, bw=anchors.bw, launch_s=anchors.launch_s,
            cpu_eval=1.5e8, cpu_fit=cpu_fit_ref), rates)
        d = dsp.Dispatcher(cal).plan(dsp.Stage.FIT, dsp.Workload.fit(
            cfg["candidates"], rows, max(1, int(round(k))), sweeps))
        out["fit_route_for_this_batch"] = {
            "route": d.route.value, "reason": d.reason.value,
            "predicted_device_s": d.cost.fit_s + d.cost.transfer_s + d.cost.launch_s,
            "predicted_cpu_s": d.cost.cpu_s, "cpu_fit_at_ref_shape": cpu_fit_ref}
    return out


def implicit_part(args, host, workloads):
    """The implicit-relation fit at its full shape: config5's candidates
    (2-8 constants) over N = 1048576 rows of (x1, x2, y) on the noisy unit
    circle, 8 LM iterations of the geometric residual f / |grad_x f| with a
    central-difference Jacobian; f64 rows resident.  Candidates/s over the
    whole batch, and the oracle on one candidate per constant count over a
    row sample."""
    import os
    import numpy as np
    import oracle
    wl = workloads.config5(n=1048576, ncand=args.implicit_cands, iters=8)
    plans = [host.gpu_fit_plan(t) for t in wl.trees]
    X = wl.x
    n, v = X.shape
    ev = host.GpuEvaluator(np.ascontiguousarray(X.T, np.float32).reshape(-1), n, v,
                           np.zeros(n, np.float32))
    assert ev.set_fit_data(np.ascontiguousarray(X.T).reshape(-1), np.zeros(n))
    outs, ms = ev.implicit_fit_batch(plans[:32], 8)
    if ms < 0:
        raise RuntimeError("implicit fit failed")
    t0 = time.perf_counter()
    outs, ms = ev.implicit_fit_batch(plans, 8)
    wall = time.perf_counter() - t0
    if ms < 0:
        raise RuntimeError("implicit fit failed")
    ev.close()
    st = {}
    for o in outs:
        st[str(o.status)] = st.get(str(o.status), 0) + 1
    ncand = len(plans)
    log(f"implicit: {ncand} candidates x N={n} in {wall:.2f}s (device {ms:.1f} ms) -> "
        f"{ncand / wall:.1f} candidates/s")
    res = {"workload": f"implicit: {ncand} candidates (2-8 constants) x N={n} rows (x1, x2, y), "
                       "8 LM iterations, geometric residual, central-difference Jacobian",
           "candidates_per_s": ncand / wall, "seconds": wall, "device_ms": ms,
           "status_counts": st, "rows": "f64 (the Jacobian's steps of 1e-6 (1 + |c|) "
           "need f64 arithmetic)"}
    if args.no_cpu or args.implicit_cpu <= 0:
        return res
    # CPU baseline on a bounded sample: one candidate per constant count on
    # the first N / 16 rows (the fit's cost is linear in the rows, so the
    # rate is scaled by 1 / 16 to the full N), one candidate per thread.  The
    # device fits the same sample on the same rows for the parity figures.
    by_k = {}
    for i, p in enumerate(plans):
        by_k.setdefault(p.nconst, i)
    sample = [by_k[k] for k in sorted(by_k)][:args.implicit_cpu]
    ns = n // 16
    Xs = np.ascontiguousarray(X[:ns])
    threads = min(len(sample), max(1, min(16, os.cpu_count() or 1)))
    t0 = time.perf_counter()
    cl, _ = oracle.implicit_fit_batch([wl.trees[i] for i in sample], Xs, None, 8,
                                      threads=threads, kmax=8)
    cpu_s = time.perf_counter() - t0
    evs = host.GpuEvaluator(np.ascontiguousarray(Xs.T, np.float32).reshape(-1), ns, v,
                            np.zeros(ns, np.float32))
    assert evs.set_fit_data(np.ascontiguousarray(Xs.T).reshape(-1), np.zeros(ns))
    souts, sms = evs.implicit_fit_batch([plans[i] for i in sample], 8)
    evs.close()
    if sms < 0:
        raise RuntimeError("implicit sample fit failed")
    rel = []
    for j in range(len(sample)):
        if np.isfinite(cl[j]) and cl[j] > 0 and np.isfinite(souts[j].sse):
            rel.append((souts[j].sse - cl[j]) / cl[j])
    rel = np.asarray(rel)
    res["cpu_baseline"] = {
        "value": len(sample) / cpu_s * (ns / n), "unit": "candidates/s", "cores": threads,
        "kind": "port",
        "sample": f"{len(sample)} candidates, one per constant count "
                  f"{sorted(plans[i].nconst for i in sample)}, on the first {ns} of the {n} rows "
                  f"({cpu_s:.1f} s on {threads} threads, one candidate per thread), 8 LM "
                  f"iterations; rate scaled by {ns}/{n} to the full rows"}
    res["loss_rel_err_vs_cpu_oracle"] = {
        "rows": ns, "compared": int(len(rel)),
        "median_abs": float(np.median(np.abs(rel))) if len(rel) else None,
        "max_abs": float(np.max(np.abs(rel))) if len(rel) else None,
        "signed": [float(r) for r in rel]}
    log(f"implicit cpu oracle: {len(sample)} candidates x {ns} rows on {threads} threads in "
        f"{cpu_s:.1f}s -> {res['cpu_baseline']['value']:.4f} candidates/s at full N; device vs "
        f"oracle median |rel| {res['loss_rel_err_vs_cpu_oracle']['median_abs']}")
    return res


def order_part(host, workloads, ncand=64, ntrials=4):
    """The exact f64 stages over the C3 rows: 64 candidates x 4 constant
    vectors (256 segments of N = 262144 rows)."""
    import numpy as np
    ts = time.perf_counter()
    wl = workloads.config3(n=262144, ncand=ncand, iters=30)
    plans = [host.gpu_fit_plan(t) for t in wl.trees]
    kmax = max(p.nconst for p in plans)
    c = np.zeros((ncand * ntrials, kmax))
    for i, p in enumerate(plans):
        for t in range(ntrials):
            c[i * ntrials + t, :p.nconst] = np.asarray(p.c0) * (1.0 + 0.01 * t)
    ev = host.GpuEvaluator(wl.cols32(), wl.n, wl.v, wl.y.astype(np.float32))
    assert ev.set_fit_data(wl.cols64(), wl.y)
    log(f"order workload ready in {time.perf_counter() - ts:.1f}s: {ncand} x {ntrials} segments "
        f"of N={wl.n}")
    segrows = float(ncand * ntrials * wl.n)
    out = {"workload": f"{ncand} C3 candidates x {ntrials} constant vectors, N={wl.n}, f64",
           "segment_rows": segrows}
    first = ev.rank_score_batch(plans, c, ntrials)[3]
    runs = {
        "rank_sums": lambda: ev.rank_score_batch(plans, c, ntrials)[3],
        "median": lambda: ev.order_stat_batch(plans, 0, c, ntrials)[2],
        "interquartile_mean": lambda: ev.order_stat_batch(plans, 1, c, ntrials)[2],
        "max_scan": lambda: ev.max_scan_batch(plans, c, ntrials)[1],
        "max_scan_with_level": lambda: ev.max_scan_batch(
            plans, c, ntrials, coords=[0] * (ncand * ntrials), levels=[1.0] * (ncand * ntrials))[1],
    }
    for name, f in runs.items():
        f()
        ms = min(f() for _ in range(3))
        if ms < 0:
            raise RuntimeError(f"{name} failed")
        out[name] = {"ms": ms, "segment_rows_per_s": segrows / (ms * 1e-3)}
    out["rank_sums_first_call_ms"] = first
    ev.close()
    log("order stages: " + ", ".join(f"{k} {v['ms']:.1f} ms" for k, v in out.items()
                                     if isinstance(v, dict)))
    return out


def small_batch_part(host, workloads, ncand=48):
    """A batch too small to fill the GPU with the built-in work items: the
    first 48 C3 candidates (N = 262144, 30 LM iterations, 32 items of 8192
    rows each), built-in shapes against the autotuner's choice, timed on the
    whole batch (best of 2 fits each)."""
    import numpy as np
    from rcsr_hip import autotune
    wl = workloads.config3(n=262144, ncand=ncand, iters=30)
    plans = [host.gpu_fit_plan(t) for t in wl.trees]
    ev = host.GpuEvaluator(wl.cols32(), wl.n, wl.v, wl.y.astype(np.float32))
    assert ev.set_fit_data(wl.cols64(), wl.y)

    def fit_ms():
        best = float("inf")
        for _ in range(2):
            outs, ms = ev.fit_batch(plans, 30, 0, kmax=20)
            if ms < 0:
                raise RuntimeError("rc_fit_batch failed")
            best = min(best, ms)
        return best, outs
    fixed, o_fixed = fit_ms()
    ts = time.perf_counter()
    res = autotune.tune_fit(ev, plans, 30, rows=[wl.n] * len(plans), reps=1)
    tune_s = time.perf_counter() - ts
    tuned, o_tuned = fit_ms()
    ev.close()
    agree = float(np.mean([abs(a.sse - b.sse) <= 1e-6 * abs(b.sse) for a, b in zip(o_tuned, o_fixed)]))
    log(f"small batch ({ncand} C3 candidates): built-in shapes {fixed:.1f} ms, autotuned "
        f"{tuned:.1f} ms ({res.to_json()['best']}, search {tune_s:.1f}s)")
    return {
        "workload": f"first {ncand} C3 candidates, N={wl.n}, 30 LM iterations",
        "fixed_ms": fixed,
        "autotuned_ms": tuned,
        "speedup": fixed / tuned,
        "autotuned": res.to_json()["best"],
        "search_seconds": tune_s,
        "search_table_ms": res.to_json()["table"],
        "losses_within_1e-6_of_built_in": agree,
    }


def c4_part(host, workloads):
    import numpy as np
    from rcsr_hip import autotune
    ts = time.perf_counter()
    wl = workloads.config4()
    X, Y, sets = wl.x, wl.y, wl.rowsets
    plans, isl, met = [], [], []
    for t, i, m in zip(wl.trees, wl.island_of, wl.metrics):
        p = host.gpu_fit_plan(t)
        if p is None:
            continue
        plans.append(p)
        isl.append(int(i))
        met.append(int(m))
    x0, y0 = X[sets[0]], Y[sets[0]]
    ev = host.GpuEvaluator(np.ascontiguousarray(x0.T, np.float32).reshape(-1), len(x0), wl.v,
                           y0.astype(np.float32))
    assert ev.set_fit_data(np.ascontiguousarray(x0.T).reshape(-1), y0)
    ids = {0: 0}
    for s in range(1, len(sets)):
        xs, ys = X[sets[s]], Y[sets[s]]
        ids[s] = ev.add_rows(np.ascontiguousarray(xs.T, np.float32).reshape(-1), len(xs), wl.v,
                             ys.astype(np.float32), np.ascontiguousarray(xs.T).reshape(-1), ys,
                             None)
    rowset = [ids[i] for i in isl]
    rows = [len(sets[i]) for i in isl]
    log(f"c4 workload ready in {time.perf_counter() - ts:.1f}s: {len(plans)} candidates, "
        f"island rows {[len(r) for r in sets]}")

    def fit_ms():
        best = float("inf")
        for _ in range(2):
            outs, ms = ev.fit_batch_hetero(plans, rowset, met, wl.iters, kmax=20)
            if ms < 0:
                raise RuntimeError("rc_fit_batch_hetero failed")
            best = min(best, ms)
        return best
    fixed = fit_ms()
    ts = time.perf_counter()
    res = autotune.tune_fit(ev, plans, wl.iters, rows=rows, rowset=rowset, metrics=met,
                            sample=512, reps=1, confirm=True)
    tune_s = time.perf_counter() - ts
    tuned = fit_ms()
    ev.close()
    log(f"c4: built-in shapes {fixed:.1f} ms, autotuned {tuned:.1f} ms "
        f"({res.to_json()['best']}, search {tune_s:.1f}s)")
    return {
        "workload": wl.name,
        "candidates": len(plans),
        "lm_iterations": wl.iters,
        "fixed_ms": fixed,
        "autotuned_ms": tuned,
        "speedup": fixed / tuned,
        "autotuned": res.to_json()["best"],
        "search_seconds": tune_s,
        "search_table_ms_on_512_candidate_sample": res.to_json()["table"],
        "confirm_table_ms_on_whole_batch": res.to_json()["confirm_table"],
    }


def tree_vars(t):
    """Distinct input columns a program reads."""
    from rcsr_hip.expr import VAR
    seen = set()
    stack = [t]
    while stack:
        n = stack.pop()
        if n.op == VAR:
            seen.add(int(n.val))
        stack.extend(n.kids)
    return seen


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def eval_part(args, host, workloads):
    import numpy as np
    ts = time.perf_counter()
    ew = workloads.config2(n=args.eval_n, nprog=args.eval_progs)
    evq = host.GpuEvaluator(ew.cols32(), ew.n, ew.v, ew.y.astype(np.float32))
    pop = host.Population()
    for t in ew.trees:
        pop.add(t)
    prog, off = pop.arrays()
    log(f"eval workload ready in {time.perf_counter() - ts:.1f}s: {len(pop)} programs x N={ew.n}")
    # The first batch also computes the derived columns (unary operators of
    # input columns the batch shares), kept for the evaluator's lifetime.
    _, ms0 = evq.eval_words(prog, off, 0)
    if ms0 < 0:
        raise RuntimeError("rc_eval failed")
    times, walls = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        _, ms = evq.eval_words(prog, off, 0)
        walls.append((time.perf_counter() - t0) * 1e3)
        times.append(ms)
    # A batch of different programs on the same rows: new hoisted constants,
    # and derived columns only for pairs not seen before.
    rng = np.random.default_rng(12345)
    pop2 = host.Population()
    for _ in range(len(ew.trees)):
        pop2.add(workloads.sized_tree(rng, ew.v, int(rng.integers(3, 61))))
    prog2, off2 = pop2.arrays()
    _, ms_new = evq.eval_words(prog2, off2, 0)
    evq.close()
    if ms_new < 0 or min(times) < 0:
        raise RuntimeError("rc_eval failed")
    ms = float(np.median(times))
    nprog = len(pop)
    row_evals = float(ew.n) * nprog
    # Algorithmic bytes: every program streams the columns it reads and y (f32).
    alg = float(sum(ew.n * 4 * (len(tree_vars(t)) + 1) for t in ew.trees))
    gbps = alg / (ms * 1e-3) / 1e9
    log(f"eval: {ms:.2f} ms per batch -> {row_evals / (ms * 1e-3):.3e} row-evals/s, "
        f"algorithmic {gbps:.0f} GB/s")
    wall = float(np.median(walls))
    log(f"eval: first batch {ms0:.2f} ms (with derived columns), new programs {ms_new:.2f} ms, "
        f"full path {wall:.2f} ms wall")
    return {
        "workload": ew.name,
        "row_evals_per_s": row_evals / (ms * 1e-3),
        "ms_per_batch": ms,
        "ms_first_batch": ms0,
        "ms_batch_of_new_programs": ms_new,
        "full_path_ms_wall": wall,
        "row_evals_per_s_full_path": row_evals / (wall * 1e-3),
        "programs": nprog,
        "rows": ew.n,
        "algorithmic_bytes_per_batch": alg,
        "algorithmic_GBps": gbps,
        "hbm_peak_GBps": HBM_PEAK_GBS,
        "frac_of_hbm_peak": gbps / HBM_PEAK_GBS,
        "note": ("algorithmic bytes above 1x HBM: each program re-reads the row tiles, "
                 "which the L2 and the Infinity Cache serve") if gbps > HBM_PEAK_GBS else None,
    }


if __name__ == "__main__":
    main()
