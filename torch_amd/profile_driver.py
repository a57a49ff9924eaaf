"""Short, fixed GPU workload for counter runs (rocprofv3 --pmc / --kernel-trace).

    python -m torch_amd.profile_driver --workload c2 --runs mixed:fwd fp64:fwd:1 --reps 3 --copy-mb 512
    python -m torch_amd.profile_driver --workload c4 --runs mixed:plan fp64:plan --reps 3

Each run spec is MODE:DIRECTION[:VARIANT[:NSPLIT]] (direction fwd = stars -> cells,
rev = cells -> stars, fwdb / revb = the same with the cells as 8^3 leaf blocks (column
kernels; VARIANT is ignored there); variant/nsplit default to the tuned shape of the run's
kernel (the block entries for fwdb / revb).  Workloads c4 and c5 take MODE:plan: the bench's
batch plan (c4: the 64 problems with block targets; c5: stars -> blocks and blocks -> stars),
run eagerly --reps times with the tuned shape.  --waves sets the waves per workgroup of the 8^3
block kernels (default: the library's); --alternate runs the point runs in turn (rep 1 of each run,
then rep 2, ...), as a bridge step alternates its two directions, instead of one run after the
other; --no-block-kernel keeps explicit arrays on the generic kernels.  After the evaluations it performs one plain
device-to-device copy of --copy-mb MiB (torch copy_), whose FETCH_SIZE / WRITE_SIZE readings
calibrate the traffic counters against a known byte count.  Everything runs on torch's current
stream; the program prints one line per step.
"""
import argparse
import time

import numpy as np


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def plan_problems(workload):
    """The bench's batch-plan problems of c4 or c5, as device tensors."""
    from . import workloads as W
    if workload == "c4":
        return [{side: {k: _dev(v) for k, v in p[side].items()} for side in ("sources", "targets")}
                for p in W.c4(1004)]
    w5 = W.c5(1005)
    st, ce = w5["stars"], w5["cells"]
    S = dict(x=_dev(st["x"]), y=_dev(st["y"]), z=_dev(st["z"]), m=_dev(st["m"]), eps=_dev(st["eps"]))
    bo, bdx = _dev(ce["block_origin"]), _dev(ce["block_size"] / 8)
    return [dict(sources=S, targets=dict(block_origin=bo, block_dx=bdx)),
            dict(sources=dict(block_origin=bo, block_dx=bdx, block_eps=bdx, m=_dev(ce["m"])),
                 targets=dict(x=S["x"], y=S["y"], z=S["z"]))]


def run_plans(args):
    import torch

    from . import autotune
    from .batch import BatchPlan
    from .field import GravityField
    probs = plan_problems(args.workload)
    torch.cuda.synchronize()
    for spec in args.runs:
        mode = spec.split(":")[0]
        with GravityField(mode, validate=False) as f:
            if args.waves:
                f.set_column_waves(args.waves)
            e = autotune.apply_batch(f, probs)
            plan = BatchPlan(f, probs)
            f.set_timing(True)
            for _ in range(args.reps):
                plan.run(graph=False)
            torch.cuda.synchronize()
            ks = f.kernel_times()
            print(f"{spec}: tuned={e} items={plan.n_items} kernel ms={[round(k[0], 3) for k in ks]}", flush=True)
            plan.close()


def run_points(args):
    import torch

    from . import autotune, workloads as W
    from .field import GravityField
    w = {"c1": W.c1, "c2": W.c2, "c5": W.c5}[args.workload]()
    st, ce = w["stars"], w["cells"]
    S = [_dev(st[k]) for k in ("x", "y", "z", "m", "eps")]
    C = [_dev(ce[k]) for k in ("x", "y", "z", "m", "eps")]
    torch.cuda.synchronize()
    BO = _dev(ce["block_origin"])
    BD = _dev(ce["block_size"] / 8)
    runs = []
    for spec in args.runs:
        parts = spec.split(":")
        mode, direction = parts[0], parts[1]
        src, tgt = (S, C) if direction.startswith("fwd") else (C, S)
        f = GravityField(mode, validate=False)
        if args.waves:
            f.set_column_waves(args.waves)
        if args.no_block_kernel:
            f.set_block_kernel(False)
        nblk = int(BD.numel())
        if direction == "fwdb":
            e = autotune.apply_blocks(f, "blockfwd", S[0].numel(), nblk)
        elif direction == "revb":
            e = autotune.apply_blocks(f, "blocksrc", S[0].numel(), nblk)
        else:
            e = autotune.apply_eval(f, src[0].numel(), tgt[0].numel())
        if len(parts) > 2:
            f.set_config(variant=int(parts[2]), nsplit=int(parts[3]) if len(parts) > 3 else 0)
        if direction == "revb":
            f.set_source_blocks(BO, BD, C[3])
        else:
            f.set_sources(*src)
        f.set_timing(True)
        if direction == "fwdb":
            call = (lambda f=f: f.evaluate_blocks(BO, BD))
        else:
            call = (lambda f=f, tgt=tgt: f.evaluate(None, tgt[0], tgt[1], tgt[2]))
        runs.append((spec, e, f, call))
    if args.alternate:                 # the runs in turn, as a bridge step alternates its directions
        for _ in range(args.reps):
            for _, _, _, call in runs:
                call()
    else:
        for _, _, _, call in runs:
            for _ in range(args.reps):
                call()
    torch.cuda.synchronize()
    for spec, e, f, _ in runs:
        ks = f.kernel_times()
        print(f"{spec}: tuned={e} kernel ms={[round(k[0], 3) for k in ks]}", flush=True)
        f.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c4", "c5"])
    ap.add_argument("--runs", nargs="+", default=["mixed:fwd", "fp64:fwd"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mb", type=int, default=512)
    ap.add_argument("--waves", type=int, default=0, choices=[0, 1, 2, 4])
    ap.add_argument("--alternate", action="store_true", help="run the point runs in turn, rep by rep")
    ap.add_argument("--no-block-kernel", action="store_true",
                    help="generic kernels for explicit arrays even when they are 8^3 blocks in C order")
    args = ap.parse_args()

    import torch

    if all(spec.endswith(":plan") for spec in args.runs):
        if args.workload not in ("c4", "c5"):
            raise SystemExit("MODE:plan runs need --workload c4 or c5")
        run_plans(args)
    else:
        run_points(args)
    if args.copy_mb > 0:
        n = args.copy_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        t = time.perf_counter()
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"copy {args.copy_mb} MiB read + {args.copy_mb} MiB write in "
              f"{(time.perf_counter() - t) * 1e3:.3f} ms", flush=True)


if __name__ == "__main__":
    main()
