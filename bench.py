#!/usr/bin/env python3
"""Pairwise interactions/s of the direct-sum gravity field on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4|c5|c2|c3|c2fused|c2flash|c1periodic|c2pm]
                    [--no-cpu] [--no-c5]

Default workload c4 (seed 1004), the configuration the metric's target is quoted on: 64 independent
problems (N_s log-uniform in [100, 50,000], Plummer/Kroupa stars or gas cells as sources, per-source
eps in [0.005, 0.1] pc; N_t log-uniform in [1e3, 1e6] as 8^3 leaf blocks), acceleration + potential of
all of them in one persistent launch replayed from a hipGraph.  One step is one replay.  The plan
takes the autotuner's launch shape (torch_amd/tune_cache.json).
  The same problems with the targets as explicit cell centres are timed beside it, and a c5 sub-record
  (`c5`) runs C5 as one graph-replayed plan: 100,000 stars (30% binary members) -> 4,194,304 cells
  (8,192 leaf blocks) and the cells as point masses -> the stars.  The CPU leg checks both kernels
  against the oracle on 256 seeded targets of every c4 problem, and on 65,536 seeded cells (forward)
  and 65,536 seeded stars plus every binary member (reverse) of c5.
  c5: that C5 record as the line.
  c2: 10,000 stars <-> 1,024,000 cells (2,000 leaf blocks of 8^3), forward + reverse, each preceded by
      loading its source set (eager calls).
  c3: the c1 inputs (1,024 stars <-> 4,096 cells of a 16^3 grid, explicit arrays) in both
      directions; both kernels against the oracle over every target, with the condition number
      kappa = sum |terms| / |sum| of each target's acceleration.
  c2fused: c2 with symmetric softening (eps2 = eps_cell^2 + eps_star^2) evaluated by the fused
      bidirectional pass (one pass gives both directions) on the cells as 8^3 blocks; the fused
      pass on explicit arrays and two directed passes in each format are timed beside it.
  c2flash: the reverse direction of c2 (1,024,000 cells as 8^3 blocks -> 10,000 stars) with
      FLASH's own spline softening (global radius 0.078125 pc = 2.5 finest cells, the reference's
      potential), FP64.
  c1periodic: c1 reverse (4,096 cells -> 1,024 stars, 4 pc box) with FLASH's spline softening
      and its periodic image sum (nrep = 6: images within 24 pc), FP64; the rate counts pairs,
      `image_interactions_per_step` the images summed.  The pairs are grouped by separation
      (sort, grouped walk, per-star sums, all in the step); `per_target_walk` times each star's
      image walk on its own lane beside it.
  c2pm: the particle-mesh pieces beside the kick on c2's 2,000 leaf blocks with 4 guard layers
      (8,192,000 cells): CIC deposit of the 9,174 stars inside the box, finite-difference
      acceleration fused with the grid kick, CIC interpolation back to the stars (the Poisson
      solve between them is not part of this library; a Plummer potential stands in).  FP64;
      the rate is guarded cells per second of the whole step.
Inputs are resident in HBM (device tensors) before timing.

The headline `value` is the mixed-precision rate; the plain-FP64 kernel is timed the same
way and reported beside it (`fp64_value`, `mixed_over_fp64`).  The roofline uses the dominant
evaluation kernel alone, timed with HIP events on its own stream through the library's
instrumentation hook (eager launches of the same plan).  Prints one JSON line on stdout; progress
on stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_INTERACTION = 21          # acc + pot: 3 sub, 1 add, 3 FMA, rsqrt, 2 mul, 1 mul, 1 add, 3 FMA
PEAK_FP32_TFLOPS = 157.3           # MI355X vector FP32 peak (packed FP32 instructions)
PEAK_FP64_TFLOPS = 78.6            # MI355X vector FP64 peak
PEAK_FP32_UNPACKED_TFLOPS = 78.6   # one FP32 FMA per lane per clock (packed FP32 disabled here)
BYTES_PER_TARGET = 64              # x, y, z, eps in + ax, ay, az, phi out (FP64)
BYTES_PER_SOURCE = 40              # x, y, z, m, eps (FP64), read once per launch

# HBM bytes per launch of the persistent or block kernels, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes over `python -m torch_amd.profile_driver` running the same launch (KiB per
# dispatch, means over the dispatches), FETCH_SIZE divided by 0.5 and WRITE_SIZE by 1.0: the
# factors those counters read for a plain 512 MiB device copy in the same runs.  Keyed by the launch
# shape measured (c2: mode, source split, waves per workgroup; the plans: shape_key); reported only
# when the bench runs that shape, null otherwise.
C2_FWD_TRAFFIC = {("mixed", 6, 4): (3602.66 / 0.5 + 193676 / 1.0) * 1024,       # k_eval_grid
                  ("fp64", 2, 4): (2494.88 / 0.5 + 64014.9 / 1.0) * 1024,
                  ("mixed", 2, 4): (2511.5 / 0.5 + 64283.9 / 1.0) * 1024,
                  ("fp64", 6, 4): (3591.22 / 0.5 + 192287 / 1.0) * 1024}


def shape_key(mode, r):
    """(mode, item size, column / block-source item scales, waves) of a plan's tuned launch."""
    c = r.get("cfg") or {}
    return (mode, float(c.get("item", 0.0)), float(c.get("col_scale", 1.0)), float(c.get("gs_scale", 1.0)),
            r.get("waves"))


C4_PLAN_TRAFFIC = {("mixed", 8e6, 2.5, 1.0, 4): (61747.7 / 0.5 + 292839 / 1.0) * 1024,   # k_batch_eval_grid
                   ("fp64", 0.0, 1.0, 1.0, 4): (59845.0 / 0.5 + 339885 / 1.0) * 1024}
C5_PLAN_TRAFFIC = {("mixed", 1e7, 2.5, 1.5, 4): {"fwd": (16814.2 / 0.5 + 394011 / 1.0) * 1024,
                                                 "rev": (91200.4 / 0.5 + 113708 / 1.0) * 1024},
                   ("fp64", 2.5e7, 1.0, 1.0, 4): {"fwd": (19378.3 / 0.5 + 394009 / 1.0) * 1024,
                                                  "rev": (89607.5 / 0.5 + 135066 / 1.0) * 1024}}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def metric_name():
    return "pairwise interactions/s, mixed vs FP64 GPU kernel, at \u22641e-6 median rel accel error"


def rel_err(got, ref):
    ga = np.stack(got[:3], 1)
    ra = np.stack(ref[:3], 1)
    ea = np.linalg.norm(ga - ra, axis=1) / np.linalg.norm(ra, axis=1)
    ep = np.abs(got[3] - ref[3]) / np.abs(ref[3])
    return ea, ep


def err_stats(ea, ep=None):
    d = dict(acc_median=float(np.median(ea)), acc_max=float(ea.max()))
    if ep is not None:
        d.update(pot_median=float(np.median(ep)), pot_max=float(ep.max()))
    return d


# ----------------------------------------------------------------- c2 / c5 ----
def run_pair(mode, dev, ns, nc, steps, warmup, blocks=True):
    """Forward (stars -> cells) and reverse (cells -> stars) per step.  blocks=True hands the
    cells over as 8^3 leaf blocks (origin, width, per-cell masses), the format FLASH holds
    them in; blocks=False as explicit per-cell arrays."""
    import torch

    from torch_amd import GasGravityField, StellarGravityField, autotune
    if blocks:
        fwd = StellarGravityField(mode, validate=False)
        rev = GasGravityField(mode, validate=False)
        nblk = nc // 512
        cfg = (autotune.apply_blocks(fwd, "blockfwd", ns, nblk), autotune.apply_blocks(rev, "blocksrc", ns, nblk))
        waves = fwd.column_waves()
    else:
        # The library applies the tuned shapes itself: explicit arrays recognised as 8^3 blocks
        # take the block kernels' entries, others the generic kernels' (reported here).
        fwd = StellarGravityField(mode, validate=False, tune_cache=autotune.CACHE)
        rev = GasGravityField(mode, validate=False, tune_cache=autotune.CACHE)
        cache = autotune.load_cache()
        m = "mixed" if mode == "mixed" else "fp64"
        cfg = (cache.get(autotune.eval_key(m, ns, nc)), cache.get(autotune.eval_key(m, nc, ns)))
        waves = fwd.column_waves()
    log(f"{mode}: {'block' if blocks else 'explicit'} format, tuned launch shapes fwd={cfg[0]} rev={cfg[1]}")
    S = [dev[k] for k in ("sx", "sy", "sz", "sm", "se")]
    C = [dev[k] for k in ("cx", "cy", "cz", "cm", "ce")]
    B = (dev.get("bo"), dev.get("bdx"))

    def step():
        fwd.set_sources(*S)
        if blocks:
            of = fwd.evaluate_blocks(B[0], B[1])
            rev.set_blocks(B[0], B[1], C[3])
        else:
            of = fwd.evaluate(None, C[0], C[1], C[2])
            rev.set_sources(*C)
        orv = rev.evaluate(None, S[0], S[1], S[2])
        return of, orv

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    fwd.set_timing(True)
    rev.set_timing(True)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record()
    for _ in range(steps):
        out = step()
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    kf = [r[0] for r in fwd.kernel_times()]
    kr = [r[0] for r in rev.kernel_times()]
    res = tuple(tuple(a.cpu().numpy() for a in o) for o in out)
    fwd.close()
    rev.close()
    return dict(ms=t0.elapsed_time(t1) / steps, fwd_kernel_ms=float(np.mean(kf)),
                rev_kernel_ms=float(np.mean(kr)), wall=wall, res=res, cfg=cfg, waves=waves)


# ---------------------------------------------------------------------- c4 ----
def run_batch(mode, probs_dev, steps, warmup, eager_timed=3):
    """Graph replays of one batch plan: ms per replay, and each persistent kernel's average
    duration over `eager_timed` instrumented eager runs (HIP events on the plan's stream)."""
    import torch

    from torch_amd import GravityField, autotune
    from torch_amd.batch import BatchPlan
    f = GravityField(mode, validate=False)
    cfg = autotune.apply_batch(f, [{k: {q: v for q, v in p[k].items()} for k in ("sources", "targets")}
                                   for p in probs_dev])
    log(f"{mode}: tuned batch shape {cfg}")
    plan = BatchPlan(f, probs_dev)
    for _ in range(warmup):
        plan.run(graph=True)
    torch.cuda.synchronize()
    # the persistent kernels alone, eager launches, instrumented: one record per persistent kernel
    # (generic, column and block-source work lists, those the plan has) and run
    f.set_timing(True)
    recs = []
    for _ in range(eager_timed):
        plan.run(graph=False)
        torch.cuda.synchronize()
        recs.append([r[0] for r in f.kernel_times()])
    f.set_timing(False)
    per_kernel = [float(np.mean([r[q] for r in recs])) for q in range(len(recs[0]))] if recs and recs[0] else []
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record()
    for _ in range(steps):
        out = plan.run(graph=True)
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    res = [tuple(a.cpu().numpy() for a in o) for o in out]
    info = dict(ms=t0.elapsed_time(t1) / steps, kernel_ms=float(np.sum(per_kernel)) if per_kernel else None,
                kernels_ms=per_kernel, wall=wall, res=res, cfg=cfg, items=plan.n_items, waves=f.column_waves(),
                interactions=plan.interactions)
    plan.close()
    f.close()
    return info


def c5_plan_problems(dev):
    """C5 as one plan: stars -> the 8,192 leaf blocks (column work list) and the blocks' cells as
    point masses (m = rho dx^3, eps = dx) -> the stars (block-source work list)."""
    S = dict(x=dev["sx"], y=dev["sy"], z=dev["sz"], m=dev["sm"], eps=dev["se"])
    return [dict(sources=S, targets=dict(block_origin=dev["bo"], block_dx=dev["bdx"])),
            dict(sources=dict(block_origin=dev["bo"], block_dx=dev["bdx"], block_eps=dev["bdx"], m=dev["cm"]),
                 targets=dict(x=dev["sx"], y=dev["sy"], z=dev["sz"]))]


# ---------------------------------------------------------------- c2fused ----
def run_fused(mode, dev, ns, nc, steps, warmup):
    """One fused pass per step (cells <-> stars, eps2 = eps_cell^2 + eps_star^2), with the cells
    as 8^3 leaf blocks (the headline) and as explicit arrays; the same fields from two directed
    evaluations per format are timed beside them."""
    import torch

    from torch_amd import GravityField
    f = GravityField(mode, validate=False)
    S = [dev[k] for k in ("sx", "sy", "sz", "sm", "se")]
    C = [dev[k] for k in ("cx", "cy", "cz", "cm", "ce")]
    BO, BD = dev["bo"], dev["bdx"]

    def fused_blocks():
        return f.bidirectional_blocks(BO, BD, C[3], BD, *S)

    def fused():
        return f.bidirectional(*C, *S)

    def directed_blocks():
        f.set_sources(*S)
        a = f.evaluate_blocks(BO, BD, eps=BD)
        f.set_source_blocks(BO, BD, C[3], eps=BD)
        b = f.evaluate(S[4], S[0], S[1], S[2])
        return a, b

    def directed():
        f.set_sources(*S)
        a = f.evaluate(C[4], C[0], C[1], C[2])
        f.set_sources(*C)
        b = f.evaluate(S[4], S[0], S[1], S[2])
        return a, b

    res = {}
    for name, fn in (("fused_blocks", fused_blocks), ("fused", fused), ("directed_blocks", directed_blocks),
                     ("directed", directed)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        f.set_timing(True)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps):
            out = fn()
        t1.record()
        torch.cuda.synchronize()
        ks = f.kernel_times()
        f.set_timing(False)
        res[name] = dict(ms=t0.elapsed_time(t1) / steps, kernel_ms=float(np.mean([k[0] for k in ks])),
                         res=tuple(tuple(a.cpu().numpy() for a in o) for o in out))
    f.close()
    r = res["fused_blocks"]
    return dict(ms=r["ms"], fwd_kernel_ms=r["kernel_ms"], res=r["res"],
                others={k: res[k]["ms"] for k in ("fused", "directed_blocks", "directed")},
                explicit_res=res["fused"]["res"], cfg=None)


# ------------------------------------------------------- c2flash / c1periodic ----
def run_flash(dev, steps, warmup, rsoft, box=None, nrep=6, per_target=False):
    """Cells -> stars with FLASH's spline kernel (reference potential), FP64; cells as 8^3
    leaf blocks when the workload has them, else as explicit arrays.  Periodic: the pairs
    grouped by separation, or (per_target) each star's image walk on its own lane."""
    import torch

    from torch_amd import GasGravityField
    rev = GasGravityField("fp64", validate=False)
    rev.set_softening("spline", rsoft, quirks=True)
    if box is not None:
        rev.set_periodic(box, nrep, per_target=per_target)
    S = [dev[k] for k in ("sx", "sy", "sz", "sm", "se")]
    C = [dev[k] for k in ("cx", "cy", "cz", "cm", "ce")]

    def step():
        if "bo" in dev:
            rev.set_blocks(dev["bo"], dev["bdx"], C[3])
        else:
            rev.set_sources(*C)
        return rev.evaluate(None, S[0], S[1], S[2])

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    rev.set_timing(True)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        out = step()
    t1.record()
    torch.cuda.synchronize()
    kr = [r[0] for r in rev.kernel_times()]
    res = tuple(a.cpu().numpy() for a in out)
    rev.close()
    return dict(ms=t0.elapsed_time(t1) / steps, kernel_ms=float(np.mean(kr)), res=res)


def run_pm(B, P, phi, steps, warmup):
    """One PM step: deposit, gradient + grid kick, interpolation (device arrays)."""
    import torch

    from torch_amd import ParticleMesh
    pm = ParticleMesh(validate=False)
    rho = torch.zeros(B.ncell, dtype=torch.float64, device="cuda")
    vel = [torch.zeros(B.ncell, dtype=torch.float64, device="cuda") for _ in range(3)]
    g = [torch.empty(B.ncell, dtype=torch.float64, device="cuda") for _ in range(3)]

    def step():
        pm.update_grid_var(P["x"], P["y"], P["z"], P["m"], P["blk"], B, rho=rho)
        gg = pm.accel_list_of_blocks(phi, B, dt=1e-3, vel=vel)
        return pm.map_mesh_to_particles(P["x"], P["y"], P["z"], P["blk"], B, *gg)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    pm.set_timing(True)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        out = step()
    t1.record()
    torch.cuda.synchronize()
    ks = pm.kernel_times()         # per step: deposit, gradient, interpolation
    names = ("deposit", "gradient_kick", "interpolate")
    kms = {nm: float(np.mean([k[0] for k in ks[i::3]])) for i, nm in enumerate(names)}
    res = dict(rho=rho.cpu().numpy(), a=tuple(a.cpu().numpy() for a in out))
    pm.close()
    del g
    return dict(ms=t0.elapsed_time(t1) / steps, kernel_ms=kms, res=res)


def mean_images_in_cutoff(S, T, box, nrep, npairs=2000, seed=99):
    """Mean number of images per (cell, point) pair inside nrep * min(box), over seeded pairs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    i = rng.integers(0, len(T["x"]), npairs)
    j = rng.integers(0, len(S["x"]), npairs)
    d = np.stack([T[k][i] - S[k][j] for k in ("x", "y", "z")], 1)
    r = np.arange(-nrep, nrep + 1)
    n = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * np.asarray(box)
    cnt = 0
    for q in range(0, npairs, 100):
        rr = np.sqrt(((d[q:q + 100, None, :] + n[None, :, :]) ** 2).sum(-1))
        cnt += int((rr < nrep * min(box)).sum())
    return cnt / npairs


# ----------------------------------------------------------- checker legs ----
def host_threads():
    """Threads of the CPU legs: the host's CPU share when OMP_NUM_THREADS sets it, else every core
    this process may run on."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    if v.isdigit() and int(v) > 0:
        return int(v)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    """The host CPU as /proc/cpuinfo names it (what lscpu prints as the model name), with its
    family / model / stepping numbers, which identify the part when the name is generic."""
    info = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if ":" in line:
                    k, v = (x.strip() for x in line.split(":", 1))
                    info.setdefault(k, v)
                elif info:
                    break                       # first processor only
    except OSError:
        pass
    if "model name" in info:
        ids = [f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info]
        return info["model name"] + (f" ({', '.join(ids)})" if ids else "")
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(S, T, G, budget_s, label, flash=None, reverse=None):
    """The oracle (C port of the reference loop nest, -O2) on seeded target subsamples:
    one thread, then the host's threads (host_threads()).  flash = dict(rsoft, box, nrep) times the FLASH
    kernel instead of the Plummer one.  reverse = (S2, T2): the other direction too, on all
    cores, whole when it fits in half the budget (c2's 1.024e10 interactions do)."""
    import oracle
    ns = len(S["x"])
    rng = np.random.Generator(np.random.PCG64(4242))

    def run(idx, threads):
        if flash is None:
            oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], T["x"][idx], T["y"][idx],
                                T["z"][idx], None, G, mode=oracle.SEQ, threads=threads)
        else:
            oracle.eval_flash(S["x"], S["y"], S["z"], S["m"], T["x"][idx], T["y"][idx], T["z"][idx], G,
                              flash["rsoft"], quirks=True, threads=threads, box=flash.get("box"),
                              nrep=flash.get("nrep", 6))

    work = 1.0 if flash is None or flash.get("box") is None else flash["images"]
    nprobe = min(len(T["x"]), max(8, int(2e6 / (ns * work))))
    probe = np.sort(rng.choice(len(T["x"]), nprobe, replace=False))
    t = time.perf_counter()
    run(probe, 1)
    rate1 = nprobe * ns / (time.perf_counter() - t)
    out = {}
    cores = host_threads()
    for threads, share in ((1, 0.35), (cores, 0.65)):
        n = int(min(len(T["x"]), max(16, budget_s * share * rate1 * (0.8 * threads) / ns)))
        idx = np.sort(rng.choice(len(T["x"]), n, replace=False))
        t = time.perf_counter()
        run(idx, threads)
        dt = time.perf_counter() - t
        out[threads] = dict(value=n * ns / dt, seconds=dt, targets=n)
        log(f"cpu oracle {threads} thread(s): {n} targets x {ns} sources in {dt:.2f} s "
            f"= {n * ns / dt:.3e} interactions/s")
    rev = None
    if reverse is not None:
        S2, T2 = reverse
        ns2 = len(S2["x"])
        n = int(min(len(T2["x"]), max(16, budget_s * 0.5 * rate1 * (0.8 * cores) / ns2)))
        idx = np.sort(rng.choice(len(T2["x"]), n, replace=False))
        t = time.perf_counter()
        oracle.eval_plummer(S2["x"], S2["y"], S2["z"], S2["m"], S2["eps"], T2["x"][idx], T2["y"][idx],
                            T2["z"][idx], None, G, mode=oracle.SEQ, threads=cores)
        dt = time.perf_counter() - t
        rev = dict(value=n * ns2 / dt, seconds=dt, targets=n, sources=ns2, whole=n == len(T2["x"]))
        log(f"cpu oracle reverse, {cores} thread(s): {n} targets x {ns2} sources in {dt:.2f} s "
            f"= {n * ns2 / dt:.3e} interactions/s")
    cb = {"value": out[cores]["value"], "unit": "interactions/s", "cores": cores, "kind": "port",
          "sample": f"{label}: {out[cores]['targets']} seeded targets x {ns} sources "
                    f"(acc + pot, FP64, {'Plummer' if flash is None else 'FLASH spline'}"
                    f"{'' if flash is None or flash.get('box') is None else ', periodic images'}, "
                    f"OpenMP over targets); one thread: "
                    f"{out[1]['value']:.3e} interactions/s on {out[1]['targets']} targets",
          "scalar_value": out[1]["value"], "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
          "whole_direction": out[cores]["targets"] == len(T["x"])}
    if rev is not None:
        cb["reverse"] = rev
    return cb


def acc_condition(S, T, G):
    """kappa = sum_j |G m_j d_ij / r_ij^3| / |a_i| per target (numpy; small inputs)."""
    s = np.stack([S["x"], S["y"], S["z"]], 1)
    t = np.stack([T["x"], T["y"], T["z"]], 1)
    d = s[None, :, :] - t[:, None, :]
    r2 = (d * d).sum(-1) + (S["eps"] ** 2)[None, :] + (T["eps_t"] ** 2 if "eps_t" in T else 0.0)
    ok = r2 > 0
    w = np.where(ok, G * S["m"][None, :] / np.where(ok, r2, 1.0) ** 1.5, 0.0)
    a = np.einsum("ij,ijk->ik", w, d)
    return (w * np.sqrt((d * d).sum(-1))).sum(1) / np.linalg.norm(a, axis=1)


def oracle_errors(results, k, S, T, idx, G, label, extra=None):
    """Each mode's result (direction k) against the compensated oracle on targets idx;
    extra = {name: boolean mask over idx} adds stats for subsets."""
    import oracle
    t0 = time.perf_counter()
    et = T.get("eps_t")
    ref = [np.empty(len(idx)) for _ in range(4)]
    step = max(1, int(4e10 // max(1, len(S["x"]))))     # ~10 s of oracle work between progress lines
    for a in range(0, len(idx), step):
        j = idx[a:a + step]
        part = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S.get("eps"), T["x"][j], T["y"][j],
                                   T["z"][j], None if et is None else et[j], G, mode=oracle.NEUMAIER,
                                   threads=host_threads())
        for q in range(4):
            ref[q][a:a + len(j)] = part[q]
        if len(idx) > step:
            log(f"oracle {label}: {a + len(j)} / {len(idx)} targets")
    ref = tuple(ref)
    log(f"oracle {label}: {len(idx)} targets x {len(S['x'])} sources in {time.perf_counter() - t0:.1f} s")
    out = {}
    for mode, r in results.items():
        ea, ep = rel_err(tuple(a[idx] for a in r["res"][k]), ref)
        out[mode] = err_stats(ea, ep)
        for name, mask in (extra or {}).items():
            out[mode][name] = err_stats(ea[mask], ep[mask])
    return out


def oracle_check(pairs, G):
    """FP64 GPU results vs the compensated oracle on seeded subsamples.
    pairs: [(S, T, gpu_result, n_sub)]."""
    import oracle
    rng = np.random.Generator(np.random.PCG64(77))
    worst_a = worst_p = 0.0
    n = 0
    for S, T, res, k in pairs:
        idx = np.sort(rng.choice(len(T["x"]), min(k, len(T["x"])), replace=False))
        et = T.get("eps_t")
        ref = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S.get("eps"), T["x"][idx], T["y"][idx],
                                  T["z"][idx], None if et is None else et[idx], G, mode=oracle.NEUMAIER,
                                  threads=host_threads())
        ea, ep = rel_err(tuple(a[idx] for a in res), ref)
        worst_a, worst_p, n = max(worst_a, ea.max()), max(worst_p, ep.max()), n + len(idx)
    return dict(targets_checked=n, max_rel_acc=float(worst_a), max_rel_pot=float(worst_p))


# ------------------------------------------------------------------ c4 / c5 ----
# Flops the column kernels execute per interaction (the algorithmic count is 21): per cell the r^2
# FMA, rsqrt, m y, the potential add, (m y) y y and three FMAs = 13; per source the 8 dz, formed once
# for the workgroup's 64 columns (1/64 per cell); per source and column of 8 cells dx, dy and two FMAs
# for dx^2 + dy^2 + eps^2 = 6, i.e. 0.75 per cell.  FP64 adds the refinement of the FP32 seed (r2 y,
# two FMAs, y e, one FMA = 8).
EXECUTED_FLOP_COLUMN = {"mixed": 13.77, "fp64": 21.77}


def roofline_of(mode, interactions, kernel_ms, kernel, traffic=None, algorithmic_bytes=None):
    achieved = FLOP_PER_INTERACTION * interactions / (kernel_ms * 1e-3) / 1e12
    peak = PEAK_FP32_TFLOPS if mode == "mixed" else PEAK_FP64_TFLOPS
    ex = EXECUTED_FLOP_COLUMN[mode] * interactions / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "kernel_ms": kernel_ms,
            "flop_per_interaction": FLOP_PER_INTERACTION, "interactions_per_launch": interactions,
            "algorithmic_bytes_per_launch": algorithmic_bytes,
            "executed": {"flop_per_interaction": EXECUTED_FLOP_COLUMN[mode], "achieved": ex,
                         "peak_unpacked": PEAK_FP32_UNPACKED_TFLOPS if mode == "mixed" else PEAK_FP64_TFLOPS,
                         "frac_of_unpacked": ex / (PEAK_FP32_UNPACKED_TFLOPS if mode == "mixed"
                                                   else PEAK_FP64_TFLOPS)}}


def c5_record(args, modes, d, steps, warmup):
    """C5 as one graph-replayed plan (forward: 100,000 stars -> 8,192 leaf blocks; reverse: their
    4,194,304 cells -> the stars): rates per mode, each direction's persistent kernel, mixed vs FP64,
    and (CPU leg) both kernels against the oracle on 65,536 seeded cells (forward) and 65,536 seeded
    stars plus every binary member (reverse)."""
    import torch

    from torch_amd import workloads as W
    G = W.G_PC_MSUN_KMS
    t = time.perf_counter()
    w = W.c5(1005)
    st, ce = w["stars"], w["cells"]
    ns, nc = len(st["x"]), len(ce["x"])
    dev = dict(sx=d(st["x"]), sy=d(st["y"]), sz=d(st["z"]), sm=d(st["m"]), se=d(st["eps"]),
               cm=d(ce["m"]), bo=d(ce["block_origin"]), bdx=d(ce["block_size"] / 8))
    torch.cuda.synchronize()
    log(f"c5 generated in {time.perf_counter() - t:.1f} s: {ns} stars, {nc} cells")
    probs = c5_plan_problems(dev)
    inter = 2.0 * ns * nc
    results = {}
    for mode in modes:
        r = run_batch(mode, probs, steps, warmup, eager_timed=2)
        results[mode] = r
        log(f"c5 {mode}: {r['ms']:.2f} ms/replay ({inter / r['ms'] * 1e3:.3e} interactions/s); persistent "
            f"kernels {['%.2f' % k for k in r['kernels_ms']]} ms, {r['items']} work items")
    rec = {"workload": "c5: 100,000 stars (30% binary members) <-> 4,194,304 AMR cell centres (8,192 leaf "
                       "blocks of 8^3, lognormal sigma = 2, 10 pc box); forward + reverse, acc + pot, one "
                       "graph-replayed plan", "seed": w["seed"], "n_stars": ns, "n_cells": nc,
           "interactions_per_step": inter, "steps": steps, "warmup": warmup}
    for mode, r in results.items():
        kf, kr = (r["kernels_ms"] + [None, None])[:2]
        peak = PEAK_FP32_TFLOPS if mode == "mixed" else PEAK_FP64_TFLOPS
        frac = lambda k: None if k is None else FLOP_PER_INTERACTION * ns * nc / (k * 1e-3) / 1e12 / peak  # noqa
        tr = C5_PLAN_TRAFFIC.get(shape_key(mode, r), {})
        rec[mode] = {"value": inter / (r["ms"] * 1e-3), "ms_per_step": r["ms"], "fwd_kernel_ms": kf,
                     "rev_kernel_ms": kr, "fwd_roofline_frac": frac(kf), "rev_roofline_frac": frac(kr),
                     "peak": peak, "launch_shape": r["cfg"], "work_items": r["items"], "waves": r["waves"],
                     "fwd_traffic": tr.get("fwd"), "rev_traffic": tr.get("rev"),
                     "fwd_algorithmic_bytes": BYTES_PER_TARGET * nc + BYTES_PER_SOURCE * ns,
                     "rev_algorithmic_bytes": BYTES_PER_TARGET * ns + BYTES_PER_SOURCE * nc}
    if "mixed" in results and "fp64" in results:
        rec["mixed_over_fp64"] = rec["mixed"]["value"] / rec["fp64"]["value"]
        errs = {}
        for k, nm in ((0, "fwd"), (1, "rev")):
            ea, ep = rel_err(results["mixed"]["res"][k], results["fp64"]["res"][k])
            errs[nm] = err_stats(ea, ep)
            if nm == "rev":
                errs["rev_binary_members"] = err_stats(ea[st["binary"] >= 0])
        rec["mixed_error_vs_fp64"] = errs
    if not args.no_cpu and "fp64" in results:
        S = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=st["eps"])
        C = dict(x=ce["x"], y=ce["y"], z=ce["z"], m=ce["m"], eps=ce["eps"])
        rng = np.random.Generator(np.random.PCG64(55))
        idx_f = np.sort(rng.choice(nc, 65536, replace=False))
        pick = np.zeros(ns, bool)
        pick[rng.choice(ns, 65536, replace=False)] = True
        pick |= st["binary"] >= 0
        idx_r = np.nonzero(pick)[0]
        rep = {"fwd": oracle_errors(results, 0, S, C, idx_f, G, "c5 fwd"),
               "rev": oracle_errors(results, 1, S=C, T=S, idx=idx_r, G=G, label="c5 rev",
                                    extra={"binary_members": st["binary"][idx_r] >= 0})}
        rep["targets"] = {"fwd": len(idx_f), "rev": len(idx_r), "rev_binary_members": int((st["binary"] >= 0).sum())}
        rec["error_vs_oracle"] = rep
        rec["fp64_vs_oracle"] = dict(targets_checked=len(idx_f) + len(idx_r),
                                     max_rel_acc=max(rep[k]["fp64"]["acc_max"] for k in ("fwd", "rev")),
                                     max_rel_pot=max(rep[k]["fp64"]["pot_max"] for k in ("fwd", "rev")))
        if "mixed" in results:
            rec["error_gate_met"] = all(rep[k]["mixed"]["acc_median"] <= 1e-6 and rep[k]["mixed"]["acc_max"] <= 1e-4
                                        for k in ("fwd", "rev"))
        log("c5 vs oracle:", json.dumps(rep))
    return rec


def bench_c4_c5(args, line, modes, head, d):
    """The metric's configuration: C4, 64 uneven problems in one graph-replayed plan (the headline
    `value`), with C5 as a sub-record; `--workload c5` makes C5 the line."""
    import torch

    from torch_amd import workloads as W
    G = W.G_PC_MSUN_KMS
    steps5 = args.c5_steps if args.c5_steps > 0 else max(2, args.steps // 4)
    warm5 = max(1, min(args.warmup, 2))
    if args.workload == "c5":
        rec = c5_record(args, modes, d, args.steps, args.warmup)
        r = rec[head]
        line.update({"value": r["value"], "ms_per_step": r["ms_per_step"],
                     "config": {k: rec[k] for k in ("workload", "seed", "n_stars", "n_cells", "interactions_per_step")}})
        line["config"]["mode"] = head
        line["roofline"] = roofline_of(head, float(rec["n_stars"]) * rec["n_cells"], r["fwd_kernel_ms"],
                                       f"k_batch_eval_grid<{head}> (forward, stars -> 8^3 cell blocks)",
                                       traffic=r["fwd_traffic"], algorithmic_bytes=r["fwd_algorithmic_bytes"])
        for k in ("fp64", "mixed_over_fp64", "mixed_error_vs_fp64", "error_vs_oracle", "fp64_vs_oracle",
                  "error_gate_met"):
            if k in rec:
                line["fp64_value" if k == "fp64" else k] = rec[k]["value"] if k == "fp64" else rec[k]
        line["c5_detail"] = rec
        return
    t = time.perf_counter()
    probs = W.c4(1004)
    pdev = [{side: {k: d(v) for k, v in p[side].items()} for side in ("sources", "targets")} for p in probs]
    torch.cuda.synchronize()
    sizes = [(len(p["sources"]["x"]), len(p["targets"]["x"])) for p in probs]
    inter_step = float(sum(a * b for a, b in sizes))
    log(f"c4 generated in {time.perf_counter() - t:.1f} s: {len(probs)} problems, {inter_step:.3e} interactions")
    # the same problems with the targets as explicit cell centres (the blocks' C order)
    pdev_x = [{"sources": p["sources"], "targets": {k: p["targets"][k] for k in ("x", "y", "z", "eps")}}
              for p in pdev]
    results, explicit = {}, {}
    for mode in modes:
        r = run_batch(mode, pdev, args.steps, args.warmup)
        results[mode] = r
        log(f"c4 {mode}: {r['ms']:.3f} ms/replay ({inter_step / r['ms'] * 1e3:.3e} interactions/s), "
            f"persistent kernels {r['kernel_ms']:.3f} ms, {r['items']} work items")
        rx = run_batch(mode, pdev_x, args.steps, args.warmup)
        same = all(np.array_equal(a, b) for o, q in zip(rx["res"], r["res"]) for a, b in zip(o, q))
        explicit[mode] = {"value": inter_step / (rx["ms"] * 1e-3), "ms_per_step": rx["ms"],
                          "kernel_ms": rx["kernel_ms"], "bit_identical_to_block_form": same}
        log(f"c4 {mode} explicit targets: {rx['ms']:.3f} ms/replay, bit-identical to the block form: {same}")
    r = results[head]
    line["value"] = inter_step / (r["ms"] * 1e-3)
    line["ms_per_step"] = r["ms"]
    line["config"] = {"workload": "c4: 64 uneven problems (N_s log-uniform 100-50,000, per-source eps 0.005-0.1 pc; "
                                  "N_t log-uniform 1e3-1e6 as 8^3 leaf blocks) in one graph-replayed persistent "
                                  "launch, acc + pot", "seed": 1004, "mode": head, "problems": len(probs),
                      "interactions_per_step": inter_step,
                      "launch_shapes": {m: results[m]["cfg"] for m in results}}
    kern_bytes = sum(BYTES_PER_TARGET * b + BYTES_PER_SOURCE * a for a, b in sizes)
    line["roofline"] = roofline_of(head, inter_step, r["kernel_ms"],
                                   f"k_batch_eval_grid<{head}> (persistent column kernel over the 64 problems' "
                                   f"leaf blocks)", algorithmic_bytes=kern_bytes,
                                   traffic=C4_PLAN_TRAFFIC.get(shape_key(head, r)))
    line["explicit_targets"] = explicit
    if "fp64" in results and head == "mixed":
        r64 = results["fp64"]
        line["fp64_value"] = inter_step / (r64["ms"] * 1e-3)
        line["mixed_over_fp64"] = line["value"] / line["fp64_value"]
        line["fp64_kernel_ms"] = r64["kernel_ms"]
        line["fp64_roofline_frac"] = (FLOP_PER_INTERACTION * inter_step / (r64["kernel_ms"] * 1e-3) / 1e12 /
                                      PEAK_FP64_TFLOPS)
        pairs = [rel_err(m, q) for m, q in zip(results["mixed"]["res"], r64["res"]) if len(q[0])]
        line["mixed_error_vs_fp64"] = {"all_problems": err_stats(np.concatenate([p[0] for p in pairs]),
                                                                 np.concatenate([p[1] for p in pairs]))}
    if not args.no_cpu:
        # both kernels against the oracle on a seeded subsample of every problem (256 targets each)
        import oracle
        rng = np.random.Generator(np.random.PCG64(44))
        t0 = time.perf_counter()
        errs = {m: ([], []) for m in results}
        worst = {m: [] for m in results}
        for i, p in enumerate(probs):
            S, T = p["sources"], p["targets"]
            nt = len(T["x"])
            if nt == 0 or len(S["x"]) == 0:
                continue
            idx = np.sort(rng.choice(nt, min(nt, 256), replace=False))
            ref = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S.get("eps"), T["x"][idx], T["y"][idx],
                                      T["z"][idx], T["eps"][idx] if "eps" in T else None, G, mode=oracle.NEUMAIER,
                                      threads=host_threads())
            for m in results:
                ea, ep = rel_err(tuple(a[idx] for a in results[m]["res"][i]), ref)
                errs[m][0].append(ea)
                errs[m][1].append(ep)
                worst[m].append(float(ea.max()))
        rep = {m: err_stats(np.concatenate(errs[m][0]), np.concatenate(errs[m][1])) for m in results}
        for m in results:
            rep[m]["worst_problem_acc_max"] = max(worst[m])
        rep["targets_checked"] = int(sum(len(e) for e in errs[head][0]))
        rep["problems_checked"] = len(worst[head])
        line["error_vs_oracle"] = rep
        if "fp64" in rep:
            line["fp64_vs_oracle"] = dict(targets_checked=rep["targets_checked"], max_rel_acc=rep["fp64"]["acc_max"],
                                          max_rel_pot=rep["fp64"]["pot_max"])
        if "mixed" in rep:
            line["error_gate_met"] = rep["mixed"]["acc_median"] <= 1e-6 and rep["mixed"]["acc_max"] <= 1e-4
        log(f"c4 vs oracle ({time.perf_counter() - t0:.1f} s):", json.dumps(rep))
    if not args.no_c5:
        line["c5"] = c5_record(args, modes, d, steps5, warm5)
    if not args.no_cpu:
        big = max(range(len(probs)), key=lambda i: sizes[i][0] * sizes[i][1])
        line["cpu_baseline"] = cpu_baseline(probs[big]["sources"], probs[big]["targets"], G, args.cpu_seconds,
                                            f"c4 problem {big} ({sizes[big][0]} sources x {sizes[big][1]} targets)")


# -------------------------------------------------------------------- main ----
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c4", choices=["c4", "c5", "c2", "c3", "c2fused", "c2flash",
                                                          "c1periodic", "c2pm"])
    ap.add_argument("--no-c5", action="store_true", help="c4: leave out the c5 sub-record")
    ap.add_argument("--c5-steps", type=int, default=0, help="replays timed for c5 (0: max(2, steps / 4))")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle check")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--modes", default="mixed,fp64")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the field evaluation is built and measured on a "
              "single MI355X (multi-GPU sharding is a design note, not implemented here).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from torch_amd import workloads as W
    G = W.G_PC_MSUN_KMS
    modes = args.modes.split(",")
    head = "mixed" if "mixed" in modes else modes[0]
    t = time.perf_counter()
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    line = {"metric": metric_name(), "unit": "interactions/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32/f64" if head == "mixed" else "f64", "data": f"synthetic (seeded {args.workload})"}
    results = {}
    if args.workload == "c2pm":
        from torch_amd import LeafBlocks
        w = W.c2(1002)
        st, ce = w["stars"], w["cells"]
        blk = W.containing_block(st["x"], st["y"], st["z"], ce["block_origin"], ce["block_size"])
        keep = blk >= 0
        Bh = LeafBlocks(ce["block_origin"], ce["block_size"], nb=8, ng=4)
        x, y, z = Bh.centres()
        phi = -500.0 / np.sqrt(x * x + y * y + z * z + 0.36)
        Ph = dict(x=st["x"][keep], y=st["y"][keep], z=st["z"][keep], m=st["m"][keep], blk=blk[keep])
        B = LeafBlocks(d(Bh.origin), d(Bh.size), nb=8, ng=4)
        P = {k: d(v) for k, v in Ph.items()}
        phid = d(phi)
        torch.cuda.synchronize()
        r = run_pm(B, P, phid, args.steps, args.warmup)
        ncell, npart = Bh.ncell, len(Ph["x"])
        log(f"c2pm: {r['ms']:.3f} ms/step ({ncell / r['ms'] * 1e3:.3e} cells/s), kernels {r['kernel_ms']}")
        kg = r["kernel_ms"]["gradient_kick"]
        grad_bytes = 80.0 * ncell          # phi in; gx, gy, gz out; vx, vy, vz in and out (FP64)
        achieved = grad_bytes / (kg * 1e-3) / 1e9
        line.update({"metric": "particle-mesh bridge pieces: guarded cells per second (deposit + "
                               "finite-difference acceleration with grid kick + interpolation)",
                     "unit": "cells/s", "value": ncell / (r["ms"] * 1e-3), "ms_per_step": r["ms"], "dtype": "f64"})
        line["config"] = {"workload": f"c2pm: {npart} stars inside the 4 pc box -> 2,000 leaf blocks of 8^3 with "
                                      f"4 guard layers ({ncell} cells); Plummer potential stands in for the "
                                      f"Poisson solve", "seed": w["seed"], "n_particles": npart, "n_cells": ncell,
                          "kernel_ms": r["kernel_ms"]}
        line["roofline"] = {"bound": "hbm", "kernel": "k_pm_gradient (acceleration + grid kick)",
                            "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                            "traffic": None, "kernel_ms": kg, "algorithmic_bytes_per_launch": grad_bytes,
                            "bytes_per_cell": 80}
        if not args.no_cpu:
            import oracle
            ob = dict(origin=Bh.origin, size=Bh.size, nb=8, ng=4)
            t = time.perf_counter()
            rho = oracle.pm_deposit(Ph["x"], Ph["y"], Ph["z"], Ph["m"], Ph["blk"], ob)
            g = oracle.pm_gradient(phi, ob)
            a = oracle.pm_interpolate(Ph["x"], Ph["y"], Ph["z"], Ph["blk"], ob, *g)
            v = [1e-3 * q for q in g]
            dt_cpu = time.perf_counter() - t
            err_rho = float(np.max(np.abs(r["res"]["rho"] - rho) / np.maximum(np.abs(rho), 1e-300)))
            same_a = all(np.array_equal(r["res"]["a"][q], a[q]) for q in range(3))
            line["fp64_vs_oracle"] = {"deposit_max_rel": err_rho, "interpolation_bit_identical": same_a}
            log("vs oracle:", line["fp64_vs_oracle"])
            line["cpu_baseline"] = {"value": ncell / dt_cpu, "unit": "cells/s", "cores": 1, "kind": "port",
                                    "sample": f"the whole c2pm step once (deposit of {npart} stars, gradient + "
                                              f"kick over {ncell} cells, interpolation), scalar C oracle"}
            del v
        print(json.dumps(line), flush=True)
        return
    if args.workload in ("c2flash", "c1periodic"):
        periodic = args.workload == "c1periodic"
        w = W.c1(1001) if periodic else W.c2(1002)
        st, ce = w["stars"], w["cells"]
        ns, nc = len(st["x"]), len(ce["x"])
        rsoft, box, nrep = (0.25, (4.0, 4.0, 4.0), 6) if periodic else (2.5 * 0.03125, None, 6)
        dev = dict(sx=d(st["x"]), sy=d(st["y"]), sz=d(st["z"]), sm=d(st["m"]), se=d(st["eps"]),
                   cx=d(ce["x"]), cy=d(ce["y"]), cz=d(ce["z"]), cm=d(ce["m"]), ce=d(ce["eps"]))
        if not periodic:
            dev.update(bo=d(ce["block_origin"]), bdx=d(ce["block_size"] / 8))
        torch.cuda.synchronize()
        r = run_flash(dev, args.steps, args.warmup, rsoft, box, nrep)
        inter_step = float(ns) * nc
        images = mean_images_in_cutoff(ce, st, box, nrep) if periodic else 1.0
        log(f"{args.workload}: {r['ms']:.3f} ms/step ({inter_step / r['ms'] * 1e3:.3e} pair interactions/s), "
            f"kernel {r['kernel_ms']:.3f} ms, {images:.1f} images per pair")
        line["value"] = inter_step / (r["ms"] * 1e-3)
        line["ms_per_step"] = r["ms"]
        line["dtype"] = "f64"
        if not periodic:   # the same cells as explicit arrays (block C order, detected on the device)
            rx = run_flash({k: v for k, v in dev.items() if k not in ("bo", "bdx")}, args.steps, args.warmup,
                           rsoft, box, nrep)
            same = all(np.array_equal(p, q) for p, q in zip(rx["res"], r["res"]))
            log(f"c2flash explicit arrays: {rx['ms']:.3f} ms/step, kernel {rx['kernel_ms']:.3f} ms, "
                f"bit-identical to the block form: {same}")
            line["explicit_arrays"] = {"value": inter_step / (rx["ms"] * 1e-3), "ms_per_step": rx["ms"],
                                       "kernel_ms": rx["kernel_ms"], "bit_identical_to_block_form": same}
        line["config"] = {"workload": (f"c1periodic: 4,096 cells (16^3, 4 pc box) -> 1,024 stars, FLASH spline "
                                       f"softening (radius {rsoft} pc), periodic images n in [-6, 6]^3 within "
                                       f"24 pc, acc + pot" if periodic else
                                       f"c2flash: 1,024,000 cells (2,000 leaf blocks of 8^3) -> 10,000 stars, "
                                       f"FLASH spline softening (radius {rsoft} pc), isolated, acc + pot"),
                          "seed": w["seed"], "mode": "fp64", "n_stars": ns, "n_cells": nc,
                          "interactions_per_step": inter_step,
                          "image_interactions_per_step": inter_step * images}
        kern_inter = inter_step * images
        achieved = FLOP_PER_INTERACTION * kern_inter / (r["kernel_ms"] * 1e-3) / 1e12
        if periodic:   # the per-target image walk on the same inputs, for comparison
            rt = run_flash(dev, args.steps, args.warmup, rsoft, box, nrep, per_target=True)
            ach_t = FLOP_PER_INTERACTION * inter_step * images / (rt["kernel_ms"] * 1e-3) / 1e12
            ea_t, ep_t = rel_err(rt["res"], r["res"])
            log(f"c1periodic per-target walk: {rt['ms']:.3f} ms/step, kernel {rt['kernel_ms']:.3f} ms "
                f"({ach_t / PEAK_FP64_TFLOPS:.3f} of peak)")
            line["per_target_walk"] = {"kernel": "k_eval_flash<1,256,periodic>", "ms_per_step": rt["ms"],
                                       "kernel_ms": rt["kernel_ms"], "frac": ach_t / PEAK_FP64_TFLOPS,
                                       "max_rel_acc_vs_grouped": float(ea_t.max()),
                                       "max_rel_pot_vs_grouped": float(ep_t.max())}
        line["roofline"] = {"bound": "valu", "kernel": ("k_eval_flash_pairs (pairs grouped by separation)"
                                                          if periodic else "k_eval_flash_gridsrc<2,256>"),
                            "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                            "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "kernel_ms": r["kernel_ms"],
                            "flop_per_interaction": FLOP_PER_INTERACTION,
                            "interactions_per_launch": kern_inter}
        if not args.no_cpu:
            import oracle
            S = dict(x=ce["x"], y=ce["y"], z=ce["z"], m=ce["m"], eps=ce["eps"])
            T = dict(x=st["x"], y=st["y"], z=st["z"])
            rng = np.random.Generator(np.random.PCG64(78))
            idx = np.sort(rng.choice(ns, 24 if periodic else 64, replace=False))
            ref = oracle.eval_flash(S["x"], S["y"], S["z"], S["m"], T["x"][idx], T["y"][idx], T["z"][idx], G,
                                    rsoft, quirks=True, threads=host_threads(), box=box, nrep=nrep)
            ea, ep = rel_err(tuple(a[idx] for a in r["res"]), ref)
            line["fp64_vs_oracle"] = dict(targets_checked=len(idx), max_rel_acc=float(ea.max()),
                                          max_rel_pot=float(ep.max()))
            log("fp64 vs oracle:", line["fp64_vs_oracle"])
            line["cpu_baseline"] = cpu_baseline(S, T, G, args.cpu_seconds, f"{args.workload} reverse direction",
                                                flash=dict(rsoft=rsoft, box=box, nrep=nrep, images=images))
        print(json.dumps(line), flush=True)
        return
    if args.workload == "c3":
        w = W.c1(1001)
        st, ce = w["stars"], w["cells"]
        ns, nc = len(st["x"]), len(ce["x"])
        dev = dict(sx=d(st["x"]), sy=d(st["y"]), sz=d(st["z"]), sm=d(st["m"]), se=d(st["eps"]),
                   cx=d(ce["x"]), cy=d(ce["y"]), cz=d(ce["z"]), cm=d(ce["m"]), ce=d(ce["eps"]))
        torch.cuda.synchronize()
        inter_step = 2.0 * ns * nc
        for mode in modes:
            results[mode] = run_pair(mode, dev, ns, nc, args.steps, args.warmup, blocks=False)
            r = results[mode]
            log(f"{mode}: {r['ms']:.3f} ms/step ({inter_step / r['ms'] * 1e3:.3e} interactions/s); eval "
                f"kernels fwd {r['fwd_kernel_ms']:.3f} ms, rev {r['rev_kernel_ms']:.3f} ms")
        r = results[head]
        line["value"] = inter_step / (r["ms"] * 1e-3)
        line["ms_per_step"] = r["ms"]
        line["config"] = {"workload": "c3: the c1 inputs through both kernels -- 1,024 stars (eps 0.01 pc) -> "
                                      "4,096 cell centres of a 16^3 grid on [-2, 2]^3 pc, and the cells "
                                      "(m = rho dx^3, eps = dx) -> the stars; explicit arrays, acc + pot",
                          "seed": w["seed"], "mode": head, "n_stars": ns, "n_cells": nc,
                          "interactions_per_step": inter_step}
        kern_ms = r["fwd_kernel_ms"]
        achieved = FLOP_PER_INTERACTION * ns * nc / (kern_ms * 1e-3) / 1e12
        peak = PEAK_FP32_TFLOPS if head == "mixed" else PEAK_FP64_TFLOPS
        line["roofline"] = {"bound": "valu", "kernel": f"k_eval_{head} (forward, generic targets)",
                            "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                            "traffic": None, "kernel_ms": kern_ms, "flop_per_interaction": FLOP_PER_INTERACTION,
                            "interactions_per_launch": float(ns) * nc,
                            "note": "4.2e6 interactions per launch: launch-latency bound, an accuracy workload"}
        if not args.no_cpu:
            S = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=st["eps"])
            C = dict(x=ce["x"], y=ce["y"], z=ce["z"], m=ce["m"], eps=ce["eps"])
            acc = {}
            for k, (src, tgt, nm) in enumerate(((S, C, "fwd"), (C, S, "rev"))):
                kap = acc_condition(src, tgt, G)
                acc[nm] = oracle_errors(results, k, src, tgt, np.arange(len(tgt["x"])), G, f"c3 {nm}")
                acc[nm]["kappa"] = dict(median=float(np.median(kap)), max=float(kap.max()))
            line["error_vs_oracle"] = acc
            line["error_gate_met"] = all(acc[nm]["mixed"]["acc_median"] <= 1e-6 and acc[nm]["mixed"]["acc_max"] <= 1e-4
                                         for nm in acc) if "mixed" in results else None
            log("vs oracle:", json.dumps(acc))
            line["cpu_baseline"] = cpu_baseline(S, C, G, args.cpu_seconds, "forward c3 direction", reverse=(C, S))
        print(json.dumps(line), flush=True)
        return
    if args.workload in ("c4", "c5"):
        bench_c4_c5(args, line, modes, head, d)
        print(json.dumps(line), flush=True)
        return
    if args.workload in ("c2", "c2fused"):
        w = W.c2(1002)
        st, ce = w["stars"], w["cells"]
        ns, nc = len(st["x"]), len(ce["x"])
        log(f"{args.workload} generated in {time.perf_counter() - t:.1f} s: {ns} stars, {nc} cells")
        dev = dict(sx=d(st["x"]), sy=d(st["y"]), sz=d(st["z"]), sm=d(st["m"]), se=d(st["eps"]),
                   cx=d(ce["x"]), cy=d(ce["y"]), cz=d(ce["z"]), cm=d(ce["m"]), ce=d(ce["eps"]),
                   bo=d(ce["block_origin"]), bdx=d(ce["block_size"] / 8))
        torch.cuda.synchronize()
        inter_step = 2.0 * ns * nc
        for mode in modes:
            if args.workload == "c2fused":
                r = run_fused(mode, dev, ns, nc, args.steps, args.warmup)
                log(f"{mode}: fused (8^3 blocks) {r['ms']:.3f} ms/step ({inter_step / r['ms'] * 1e3:.3e} "
                    f"interactions/s), kernel {r['fwd_kernel_ms']:.3f} ms; fused explicit, directed blocks, "
                    f"directed explicit: {r['others']}")
            else:
                r = run_pair(mode, dev, ns, nc, args.steps, args.warmup, blocks=True)
                log(f"{mode}: {r['ms']:.3f} ms/step ({inter_step / r['ms'] * 1e3:.3e} interactions/s); "
                    f"eval kernels fwd {r['fwd_kernel_ms']:.3f} ms, rev {r['rev_kernel_ms']:.3f} ms")
                rx = run_pair(mode, dev, ns, nc, args.steps, args.warmup, blocks=False)
                log(f"{mode} explicit arrays: {rx['ms']:.3f} ms/step "
                    f"({inter_step / rx['ms'] * 1e3:.3e} interactions/s); eval kernels fwd "
                    f"{rx['fwd_kernel_ms']:.3f} ms, rev {rx['rev_kernel_ms']:.3f} ms")
                r["explicit"] = rx
            results[mode] = r
        r = results[head]
        kern_ms = r["fwd_kernel_ms"]
        kern_inter = float(ns) * nc * (2 if args.workload == "c2fused" else 1)
        kern_bytes = BYTES_PER_TARGET * nc + BYTES_PER_SOURCE * ns
        if args.workload == "c2fused":
            kern_bytes = (BYTES_PER_SOURCE + 32) * (nc + ns)
        name = ("10,000 stars (20% binary members) <-> 1,024,000 AMR cell centres (2,000 leaf "
                "blocks of 8^3)")
        if args.workload == "c2fused":
            name += ", symmetric softening eps2 = eps_cell^2 + eps_star^2, one fused pass per step"
        line["config"] = {"workload": f"{args.workload}: {name}; forward + reverse, acc + pot",
                          "seed": w["seed"], "mode": head, "n_stars": ns, "n_cells": nc,
                          "interactions_per_step": inter_step,
                          "format": "8^3 leaf blocks (origin, width, cell masses); explicit-array "
                                    "rates in explicit_arrays"}
        kernel_name = f"k_eval_grid<{'mixed' if head == 'mixed' else 'fp64'}> (forward, stars -> 8^3 cell blocks)"
        if "explicit" in results[head]:
            line["explicit_arrays"] = {m: {"value": inter_step / (results[m]["explicit"]["ms"] * 1e-3),
                                           "ms_per_step": results[m]["explicit"]["ms"],
                                           "fwd_kernel_ms": results[m]["explicit"]["fwd_kernel_ms"],
                                           "rev_kernel_ms": results[m]["explicit"]["rev_kernel_ms"],
                                           "generic_launch_shapes": results[m]["explicit"]["cfg"],
                                           "kernels": "the cells are in the blocks' C order, so the library "
                                                      "detects the 8^3 structure on the device and both "
                                                      "directions run the column kernels; the generic "
                                                      "shapes apply to unstructured points"}
                                       for m in results}
        if args.workload == "c2fused":
            kernel_name = f"k_bidir_grid<{'mixed' if head == 'mixed' else 'fp64'}> (both directions, 8^3 cell blocks)"
            line["other_forms_ms_per_step"] = {m: results[m]["others"] for m in results}
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    line["value"] = inter_step / (results[head]["ms"] * 1e-3)
    line["ms_per_step"] = results[head]["ms"]
    achieved = FLOP_PER_INTERACTION * kern_inter / (kern_ms * 1e-3) / 1e12
    peak = PEAK_FP32_TFLOPS if head == "mixed" else PEAK_FP64_TFLOPS
    traffic = None
    if args.workload == "c2" and results[head]["cfg"][0]:
        traffic = C2_FWD_TRAFFIC.get((head, results[head]["cfg"][0]["nsplit"], results[head]["waves"]))
    line["roofline"] = {"bound": "valu", "kernel": kernel_name, "achieved": achieved, "peak": peak,
                        "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                        "kernel_ms": kern_ms, "flop_per_interaction": FLOP_PER_INTERACTION,
                        "interactions_per_launch": kern_inter,
                        "algorithmic_bytes_per_launch": kern_bytes,
                        "peak_unpacked_fp32": PEAK_FP32_UNPACKED_TFLOPS,
                        "frac_of_unpacked_fp32": achieved / PEAK_FP32_UNPACKED_TFLOPS}
    if "fp64" in results and head == "mixed":
        r64 = results["fp64"]
        line["fp64_value"] = inter_step / (r64["ms"] * 1e-3)
        line["mixed_over_fp64"] = line["value"] / line["fp64_value"]
        k64 = r64["fwd_kernel_ms"] if "fwd_kernel_ms" in r64 else r64["kernel_ms"]
        line["fp64_kernel_ms"] = k64
        line["fp64_roofline_frac"] = FLOP_PER_INTERACTION * kern_inter / (k64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
        if args.workload in ("c2", "c2fused"):
            errs = {}
            for k, nm in ((0, "fwd"), (1, "rev")):
                ea, ep = rel_err(results["mixed"]["res"][k], r64["res"][k])
                errs[nm] = err_stats(ea, ep)
                if nm == "rev":
                    b = st["binary"] >= 0
                    errs["rev_binary_members"] = err_stats(ea[b])
        else:
            pairs = [rel_err(m, q) for m, q in zip(results["mixed"]["res"], r64["res"]) if len(q[0])]
            errs = {"all_problems": err_stats(np.concatenate([p[0] for p in pairs]),
                                              np.concatenate([p[1] for p in pairs]))}
        line["mixed_error_vs_fp64"] = errs
        line["error_gate_met"] = all(e["acc_median"] <= 1e-6 and e["acc_max"] <= 1e-4 for e in errs.values())
    if not args.no_cpu:
        if args.workload in ("c2", "c2fused"):
            S = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=st["eps"])
            C = dict(x=ce["x"], y=ce["y"], z=ce["z"], m=ce["m"], eps=ce["eps"])
            if args.workload == "c2fused":      # symmetric softening: targets carry their eps too
                S["eps_t"], C["eps_t"] = st["eps"], ce["eps"]
            if "fp64" in results:
                line["fp64_vs_oracle"] = oracle_check(
                    [(S, C, results["fp64"]["res"][0], 1024),
                     (C, S, results["fp64"]["res"][1], 96)], G)
                log("fp64 vs oracle:", line["fp64_vs_oracle"])
            line["cpu_baseline"] = cpu_baseline(S, C, G, args.cpu_seconds, f"forward {args.workload} direction",
                                                reverse=None if args.workload == "c2fused" else (C, S))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
