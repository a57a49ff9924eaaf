"""Short, fixed GPU workload for counter runs (rocprofv3 --pmc / --kernel-trace).

    python -m torch_amd.profile_driver --workload c2 --runs mixed:fwd fp64:fwd:1 --reps 3 --copy-mb 512

Each run spec is MODE:DIRECTION[:VARIANT[:NSPLIT]] (direction fwd = stars -> cells,
rev = cells -> stars, fwdb / revb = the same with the cells as 8^3 leaf blocks (column
kernels; VARIANT is ignored there); variant/nsplit default to the tuned shape of the run's
kernel (the block entries for fwdb / revb).  After the evaluations
it performs one plain device-to-device copy of --copy-mb MiB (torch copy_), whose
FETCH_SIZE / WRITE_SIZE readings calibrate the traffic counters against a known byte count.
Everything runs on torch's current stream; the program prints one line per step.
"""
import argparse
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c5"])
    ap.add_argument("--runs", nargs="+", default=["mixed:fwd", "fp64:fwd"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mb", type=int, default=512)
    args = ap.parse_args()

    import torch

    from . import autotune, workloads as W
    from .field import GravityField
    w = {"c1": W.c1, "c2": W.c2, "c5": W.c5}[args.workload]()
    st, ce = w["stars"], w["cells"]
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    S = [d(st[k]) for k in ("x", "y", "z", "m", "eps")]
    C = [d(ce[k]) for k in ("x", "y", "z", "m", "eps")]
    torch.cuda.synchronize()
    BO = d(ce["block_origin"])
    BD = d(ce["block_size"] / 8)
    for spec in args.runs:
        parts = spec.split(":")
        mode, direction = parts[0], parts[1]
        src, tgt = (S, C) if direction.startswith("fwd") else (C, S)
        with GravityField(mode, validate=False) as f:
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
            for r in range(args.reps):
                if direction == "fwdb":
                    f.evaluate_blocks(BO, BD)
                else:
                    f.evaluate(None, tgt[0], tgt[1], tgt[2])
            torch.cuda.synchronize()
            ks = f.kernel_times()
            print(f"{spec}: tuned={e} kernel ms={[round(k[0], 3) for k in ks]}", flush=True)
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
