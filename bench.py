#!/usr/bin/env python3
"""Pairwise interactions/s of the direct-sum gravity field on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (c2, seed 1002): 10,000 Plummer/Kroupa stars (20% hard-binary members) and
1,024,000 AMR cell centres (2,000 leaf blocks of 8^3).  One step is one bridge field
evaluation in both directions: stars -> cells (acceleration + potential at the cell
centres) and cells as point masses (m = rho dx^3, eps = dx) -> stars, each preceded by
loading its source set.  Inputs are resident in HBM (device tensors) before timing.
2 x 1.024e10 pair interactions per step.

The headline `value` is the mixed-precision rate; the plain-FP64 kernel is timed the same
way and reported beside it (`fp64_value`, `mixed_over_fp64`).  Mixed-precision error is
measured over all targets against the FP64 kernel, which is itself checked against the
CPU oracle on a seeded subsample (the checker leg).  Prints one JSON line on stdout;
progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_INTERACTION = 21          # acc + pot, FMA = 2
PEAK_FP32_TFLOPS = 157.3           # MI355X vector FP32 (packed) peak
PEAK_FP64_TFLOPS = 78.6            # MI355X vector FP64 peak
PEAK_FP32_UNPACKED_TFLOPS = 78.6   # one FP32 op per lane per cycle (packed FP32 disabled)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def metric_name():
    try:
        with open(os.path.join(ROOT, "BASELINE.json")) as fh:
            return json.load(fh)["metric"]
    except Exception:
        return "pairwise interactions/s, mixed vs FP64 GPU kernel, at <=1e-6 median rel accel error"


def rel_err(got, ref):
    ga = np.stack(got[:3], 1)
    ra = np.stack(ref[:3], 1)
    ea = np.linalg.norm(ga - ra, axis=1) / np.linalg.norm(ra, axis=1)
    ep = np.abs(got[3] - ref[3]) / np.abs(ref[3])
    return ea, ep


def run_mode(mode, dev, st, ce, steps, warmup):
    import torch

    from torch_amd import GasGravityField, StellarGravityField, autotune
    fwd = StellarGravityField(mode, validate=False)
    rev = GasGravityField(mode, validate=False)
    cfg = (autotune.apply_eval(fwd, len(st["x"]), len(ce["x"])),
           autotune.apply_eval(rev, len(ce["x"]), len(st["x"])))
    log(f"{mode}: tuned launch shapes fwd={cfg[0]} rev={cfg[1]}")
    S = [dev[k] for k in ("sx", "sy", "sz", "sm", "se")]
    C = [dev[k] for k in ("cx", "cy", "cz", "cm", "ce")]

    def step(ev=None):
        fwd.set_sources(*S)
        if ev is not None:
            ev[0].record()
        of = fwd.evaluate(None, C[0], C[1], C[2])
        if ev is not None:
            ev[1].record()
        rev.set_sources(*C)
        orv = rev.evaluate(None, S[0], S[1], S[2])
        return of, orv

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0.record()
    for i in range(steps):
        out = step(kev[i])
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1) / steps
    kms = float(np.mean([a.elapsed_time(b) for a, b in kev]))
    res = tuple(tuple(a.cpu().numpy() for a in o) for o in out)
    fwd.close()
    rev.close()
    return ms, kms, wall, res


def cpu_baseline(st, ce, G, budget_s):
    """The oracle (C port of the reference loop nest, -O2) on a seeded target subsample of the
    forward direction: one thread, then all host cores (<= 16)."""
    import oracle
    ns = len(st["x"])
    rng = np.random.Generator(np.random.PCG64(4242))
    probe = np.sort(rng.choice(len(ce["x"]), 200, replace=False))
    t = time.perf_counter()
    oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][probe], ce["y"][probe],
                        ce["z"][probe], None, G, mode=oracle.SEQ, threads=1)
    rate1 = 200 * ns / (time.perf_counter() - t)
    out = {}
    cores = max(1, min(16, os.cpu_count() or 1))
    for threads, share in ((1, 0.35), (cores, 0.65)):
        n = int(min(len(ce["x"]), max(256, budget_s * share * rate1 * (0.8 * threads) / ns)))
        idx = np.sort(rng.choice(len(ce["x"]), n, replace=False))
        t = time.perf_counter()
        oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][idx], ce["y"][idx],
                            ce["z"][idx], None, G, mode=oracle.SEQ, threads=threads)
        dt = time.perf_counter() - t
        out[threads] = dict(value=n * ns / dt, seconds=dt, targets=n)
        log(f"cpu oracle {threads} thread(s): {n} targets x {ns} stars in {dt:.2f} s "
            f"= {n * ns / dt:.3e} interactions/s")
    return out, cores


def oracle_check(st, ce, res64, G):
    """FP64 kernel vs the compensated oracle on seeded subsamples (the checker)."""
    import oracle
    rng = np.random.Generator(np.random.PCG64(77))
    i_f = np.sort(rng.choice(len(ce["x"]), 1024, replace=False))
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][i_f], ce["y"][i_f],
                              ce["z"][i_f], None, G, mode=oracle.NEUMAIER, threads=16)
    ea_f, ep_f = rel_err(tuple(a[i_f] for a in res64[0]), ref)
    i_r = np.sort(rng.choice(len(st["x"]), 96, replace=False))
    ref = oracle.eval_plummer(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"], st["x"][i_r], st["y"][i_r],
                              st["z"][i_r], None, G, mode=oracle.NEUMAIER, threads=16)
    ea_r, ep_r = rel_err(tuple(a[i_r] for a in res64[1]), ref)
    return dict(fwd_targets=len(i_f), rev_targets=len(i_r),
                max_rel_acc=float(max(ea_f.max(), ea_r.max())),
                max_rel_pot=float(max(ep_f.max(), ep_r.max())))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle check")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--modes", default="mixed,fp64")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the field evaluation is built and measured on a "
              "single MI355X (multi-GPU sharding is a design note, not part of this build).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from torch_amd import workloads as W
    G = W.G_PC_MSUN_KMS
    t = time.perf_counter()
    w = W.c2(1002)
    st, ce = w["stars"], w["cells"]
    ns, nc = len(st["x"]), len(ce["x"])
    log(f"c2 generated in {time.perf_counter() - t:.1f} s: {ns} stars, {nc} cells")
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    dev = dict(sx=d(st["x"]), sy=d(st["y"]), sz=d(st["z"]), sm=d(st["m"]), se=d(st["eps"]),
               cx=d(ce["x"]), cy=d(ce["y"]), cz=d(ce["z"]), cm=d(ce["m"]), ce=d(ce["eps"]))
    torch.cuda.synchronize()
    inter_step = 2.0 * ns * nc

    results = {}
    for mode in args.modes.split(","):
        ms, kms, wall, res = run_mode(mode, dev, st, ce, args.steps, args.warmup)
        results[mode] = dict(ms=ms, kms=kms, wall=wall, res=res)
        log(f"{mode}: {ms:.3f} ms/step ({inter_step / ms * 1e3:.3e} interactions/s), "
            f"forward kernel {kms:.3f} ms, wall {wall:.2f} s")

    head = "mixed" if "mixed" in results else list(results)[0]
    r = results[head]
    value = inter_step / (r["ms"] * 1e-3)
    kern_s = r["kms"] * 1e-3
    achieved = FLOP_PER_INTERACTION * ns * nc / kern_s / 1e12
    peak = PEAK_FP32_TFLOPS if head == "mixed" else PEAK_FP64_TFLOPS
    line = {
        "metric": metric_name(),
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64" if head == "mixed" else "f64",
        "data": "synthetic (seeded c2 workload)",
        "config": {"workload": "c2: 10,000 stars (20% binary members) <-> 1,024,000 AMR cell "
                               "centres (2,000 leaf blocks of 8^3); forward + reverse, acc + pot",
                   "seed": 1002, "mode": head, "n_stars": ns, "n_cells": nc,
                   "interactions_per_step": inter_step},
        "roofline": {"bound": "valu", "kernel": "k_eval_mixed (forward, stars -> cells)"
                     if head == "mixed" else "k_eval_fp64 (forward)",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "flop_per_interaction": FLOP_PER_INTERACTION,
                     "kernel_ms": r["kms"], "peak_unpacked_fp32": PEAK_FP32_UNPACKED_TFLOPS},
    }
    if "fp64" in results:
        r64 = results["fp64"]
        line["fp64_value"] = inter_step / (r64["ms"] * 1e-3)
        line["fp64_kernel_ms"] = r64["kms"]
        line["mixed_over_fp64"] = line["value"] / line["fp64_value"] if head == "mixed" else 1.0
        line["fp64_roofline_frac"] = FLOP_PER_INTERACTION * ns * nc / (r64["kms"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
        if "mixed" in results:
            errs = {}
            for k, name in ((0, "fwd"), (1, "rev")):
                ea, ep = rel_err(results["mixed"]["res"][k], r64["res"][k])
                errs[name] = dict(acc_median=float(np.median(ea)), acc_max=float(ea.max()),
                                  pot_median=float(np.median(ep)), pot_max=float(ep.max()))
                if name == "rev":
                    b = st["binary"] >= 0
                    errs["rev_binary_members"] = dict(acc_median=float(np.median(ea[b])),
                                                      acc_max=float(ea[b].max()))
            line["mixed_error_vs_fp64"] = errs
            line["error_gate_met"] = all(e["acc_median"] <= 1e-6 and e["acc_max"] <= 1e-4
                                         for e in errs.values())
    if not args.no_cpu:
        if "fp64" in results:
            line["fp64_vs_oracle"] = oracle_check(st, ce, results["fp64"]["res"], G)
            log("fp64 vs oracle:", line["fp64_vs_oracle"])
        cb, cores = cpu_baseline(st, ce, G, args.cpu_seconds)
        line["cpu_baseline"] = {
            "value": cb[cores]["value"], "unit": "interactions/s", "cores": cores, "kind": "port",
            "sample": f"forward c2 direction, {cb[cores]['targets']} seeded cell targets x {ns} stars "
                      f"(acc + pot, FP64, OpenMP over targets); one thread: "
                      f"{cb[1]['value']:.3e} interactions/s on {cb[1]['targets']} targets",
            "scalar_value": cb[1]["value"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
