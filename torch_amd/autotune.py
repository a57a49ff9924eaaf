"""Autotuner for the launch shape of the direct-sum kernels, timed on the device.

Tunable knobs (all exposed by the C ABI):
  * kernel variant: targets per lane K and workgroup size NT (tgrav_variant_info lists them);
  * source split of a single evaluation (tgrav_set_config's nsplit; 0 = built-in heuristic);
  * batch plans: interactions per work item and persistent workgroups per CU
    (tgrav_set_batch_config).
Every candidate is run on synthetic inputs of the requested size and timed with HIP events
on the stream the kernels use; the fastest is kept.  Results are cached in a JSON file keyed
by device, mode and log2 size buckets, so a tuned shape is reused for every problem of
similar size.  ``apply_eval`` / ``apply_batch`` look the cache up and configure a field.

    python -m torch_amd.autotune --eval 10000 1024000 --eval 1024000 10000 --batch c4 --batch c5
"""
import argparse
import ctypes
import json
import math
import os
import time

import numpy as np

from . import _lib
from .field import MODES, GravityField

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.environ.get("TGRAV_TUNE_CACHE", os.path.join(HERE, "tune_cache.json"))


def _bucket(n):
    return int(round(math.log2(max(float(n), 1.0))))


def device_key():
    import torch
    p = torch.cuda.get_device_properties(0)
    return f"{getattr(p, 'gcnArchName', p.name).split(':')[0]}/{p.multi_processor_count}cu"


def load_cache(path=CACHE):
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def save_cache(cache, path=CACHE):
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def eval_key(mode, ns, nt):
    return f"{device_key()}|eval|{mode}|s{_bucket(ns)}|t{_bucket(nt)}"


def problem_sizes(p):
    """(sources, targets) of a batch problem: explicit arrays, or 8^3 leaf blocks (512 cells each)."""
    S, T = p["sources"], p["targets"]
    ns = len(S["m"]) if S.get("block_origin") is not None else len(S["x"])
    nt = 512 * len(T["block_dx"]) if T.get("block_origin") is not None else len(T["x"])
    return ns, nt


def waves_suffix(waves):
    """Key suffix of entries tuned with more than one wave per column workgroup (the library
    looks such entries up under the same suffix)."""
    return f"/w{int(waves)}" if int(waves) > 1 else ""


def batch_key(mode, problems, waves=1):
    ns = [problem_sizes(p)[0] for p in problems]
    nt = [problem_sizes(p)[1] for p in problems]
    inter = sum(a * b for a, b in zip(ns, nt))
    return (f"{device_key()}|batch{waves_suffix(waves)}|{mode}|p{len(problems)}|i{_bucket(inter)}"
            f"|s{_bucket(np.median(ns))}|t{_bucket(np.median(nt))}")


def variants(mode):
    lib = _lib.load()
    m = MODES[mode]
    out = []
    for v in range(lib.tgrav_num_variants(m)):
        k, nt, ts = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.tgrav_variant_info(m, v, ctypes.byref(k), ctypes.byref(nt), ctypes.byref(ts))
        out.append(dict(variant=v, k=k.value, threads=nt.value, tile_sources=ts.value))
    return out


_FLUSH = []


def _flush():
    """Other work between timed launches (a 256 MiB copy), as in real use: an evaluation is
    rarely launched right after itself with warm caches."""
    import torch
    if not _FLUSH:
        _FLUSH.append(torch.empty(1 << 25, dtype=torch.float64, device="cuda"))
        _FLUSH.append(torch.empty(1 << 25, dtype=torch.float64, device="cuda"))
    _FLUSH[1].copy_(_FLUSH[0])


def _time(fn, reps, before=None):
    """Median of `reps` timed calls of fn.  Before each: `before()` when given (the work that
    precedes the candidate where it is used, e.g. the other direction of a bridge step), else a
    256 MiB copy.  Two untimed calls first: the GPU clock ramps up over the first milliseconds
    of work after idling, so the first launches of a sequence run slower."""
    import torch
    fn()
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        if before is not None:
            before()
        else:
            _flush()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def _synthetic(ns, nt, seed=0):
    import torch
    from . import workloads as W
    rng = W.rng_for(seed)
    st = W.star_cluster(rng, ns)
    t = rng.uniform(-2.0, 2.0, (3, nt))
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    return [d(st[k]) for k in ("x", "y", "z", "m", "eps")], [d(t[q]) for q in range(3)]


def tune_eval(mode, ns, nt, reps=5, nsplits=(0, 1, 2, 4, 8, 16, 32, 64, 128, 256), log=print,
              in_sequence=True):
    """Best (variant, nsplit) for one evaluation of nt targets from ns sources.  in_sequence: each
    timed candidate follows the reverse evaluation (the nt targets as sources onto the ns sources,
    from a second context), as the bridge alternates the two directions, instead of a plain copy."""
    S, T = _synthetic(ns, nt)
    table = []
    partner = None
    if in_sequence:
        g = GravityField(mode, validate=False, tune_cache=None)
        import torch
        g.set_sources(T[0], T[1], T[2], torch.ones_like(T[0]))
        partner = (g, lambda: g.evaluate(None, S[0], S[1], S[2]))
    with GravityField(mode, validate=False, tune_cache=None) as f:
        f.set_sources(*S)
        for v in variants(mode):
            for sp in nsplits:
                tile = v["k"] * v["threads"]
                if sp > 1 and sp * v["tile_sources"] > ns:
                    continue
                if sp > 1 and -(-nt // tile) >= 4096:
                    continue      # plenty of target tiles already
                f.set_config(variant=v["variant"], nsplit=sp)
                ms = _time(lambda: f.evaluate(None, *T), reps, before=partner[1] if partner else None)
                table.append(dict(v, nsplit=sp, ms=ms, rate=ns * nt / ms * 1e3))
                log(f"  {mode} ns={ns} nt={nt} variant={v['variant']} (K={v['k']}, NT={v['threads']})"
                    f" nsplit={sp}: {ms:.3f} ms")
    if partner:
        partner[0].close()
    best = min(table, key=lambda r: r["ms"])
    return best, table


def tune_batch(mode, problems, reps=3, items=(0.0, 2e6, 8e6, 3.2e7), wg_per_cu=(0, 1, 2),
               log=print):
    """Best (variant, item size, workgroups per CU) for a batch plan of these problems (numpy or
    device arrays).  The variant and the workgroups per CU only shape the generic work list: a plan
    with none (targets or sources all 8^3 blocks) is tuned over the item size alone."""
    import torch
    from .batch import BatchPlan, problems_from_workload
    dev = problems if _on_device(problems) else problems_from_workload(problems)
    lib = _lib.load()
    table = []
    generic = any(p["targets"].get("block_origin") is None and p["sources"].get("block_origin") is None
                  for p in problems)
    with GravityField(mode, validate=False, tune_cache=None) as f:
        waves = f.column_waves()
        for v in (variants(mode) if generic else variants(mode)[:1]):
            for it in items:
                for w in (wg_per_cu if generic else (0,)):
                    f.set_config(variant=v["variant"])
                    lib.tgrav_set_batch_config(f.handle, float(it), int(w))
                    plan = BatchPlan(f, dev)
                    ms = _time(lambda: plan.run(graph=True), reps)
                    table.append(dict(v, item=it, wg_per_cu=w, ms=ms, items_n=plan.n_items,
                                      partial_bytes=plan.partial_bytes, col_scale=1.0, gs_scale=1.0,
                                      rate=plan.interactions / ms * 1e3))
                    log(f"  batch {mode} variant={v['variant']} (K={v['k']}, NT={v['threads']}) "
                        f"item={it:.1e} wg/cu={w}: {ms:.3f} ms, {plan.n_items} items")
                    plan.close()
                    torch.cuda.synchronize()
        lib.tgrav_set_batch_config(f.handle, 0.0, 0)
        blocks = [any(p[side].get("block_origin") is not None for p in problems) for side in ("targets", "sources")]
        if any(blocks):
            # Larger items for the block lists at the fastest shape's item size: each source chunk
            # writes 32 B of partial sums per target, and these lists have many items to balance.
            fast = min(table, key=lambda r: r["ms"])
            for cs, gs in LIST_SCALES:
                if (cs != 1.0 and not blocks[0]) or (gs != 1.0 and not blocks[1]):
                    continue
                f.set_config(variant=fast["variant"])
                lib.tgrav_set_batch_config(f.handle, float(fast["item"]), int(fast["wg_per_cu"]))
                lib.tgrav_set_batch_scales(f.handle, float(cs), float(gs))
                plan = BatchPlan(f, dev)
                ms = _time(lambda: plan.run(graph=True), reps)
                table.append(dict(fast, col_scale=cs, gs_scale=gs, ms=ms, items_n=plan.n_items,
                                  partial_bytes=plan.partial_bytes, rate=plan.interactions / ms * 1e3))
                log(f"  batch {mode} item={fast['item']:.1e} list items x{cs:g} / x{gs:g}: {ms:.3f} ms, "
                    f"{plan.partial_bytes / 1e6:.0f} MB of partials")
                plan.close()
                torch.cuda.synchronize()
            lib.tgrav_set_batch_scales(f.handle, 0.0, 0.0)
            lib.tgrav_set_batch_config(f.handle, 0.0, 0)
    best = pick(table, lambda r: (r["partial_bytes"], r["items_n"]))   # fewer partial sums
    return dict(best, waves=waves), table


def _on_device(problems):
    import torch
    return any(torch.is_tensor(v) for p in problems for side in ("sources", "targets") for v in p[side].values())


def c5_plan_problems():
    """C5 as the bench runs it: stars -> the 8,192 leaf blocks, and the blocks' cells -> the stars."""
    from .profile_driver import plan_problems
    return plan_problems("c5")


TIE = 0.01   # candidates within 1% of the fastest count as equally fast (repeated timings of one
             # shape on one box differ by about that much)
# Each source chunk of the block-target kernel writes 32 B of FP64 partial sums per cell, half the
# algorithmic bytes of the launch (64 B per cell), and the reduction reads them back: its split keeps
# the fewest chunks within 2% of the fastest.
TIE_BLOCK_TARGETS = 0.02


# (column, block-source) work lists' item sizes against the plan's, tried by tune_batch
LIST_SCALES = ((1.0, 1.0), (2.5, 1.0), (1.0, 1.5), (2.5, 1.5))


def pick(table, fewer_partials, tie=TIE):
    """The fastest row, except that among the rows within `tie` of it the one that writes the
    fewest partial sums to memory wins (fewer source chunks: less memory and less traffic for no
    measurable time).  fewer_partials(row) orders the rows by that (smaller = fewer)."""
    best = min(r["ms"] for r in table)
    near = [r for r in table if r["ms"] <= best * (1.0 + tie)]
    return min(near, key=lambda r: (fewer_partials(r), r["ms"]))


def _chunks(r):
    """Source chunks of a block-kernel row: nsplit, with 0 (the built-in heuristic, whose count
    is not known here) ranked last."""
    return r["nsplit"] if r["nsplit"] > 0 else float("inf")


def blocks_key(mode, direction, ns, nblk, waves=1):
    """direction "blockfwd" or "blocksrc", keyed by the kernels' waves per workgroup."""
    return f"{device_key()}|{direction}{waves_suffix(waves)}|{mode}|s{_bucket(ns)}|b{_bucket(nblk)}"


def tune_blocks(mode, workload="c2", reps=5, log=print):
    """Source split of the 8^3 column kernels on a workload's own leaf blocks:
    forward (stars -> block targets) and reverse (block sources -> stars)."""
    import torch
    from . import workloads as W
    w = {"c2": W.c2, "c5": W.c5}[workload]()
    st, ce = w["stars"], w["cells"]
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    S = [d(st[k]) for k in ("x", "y", "z", "m", "eps")]
    BO, BD, CM = d(ce["block_origin"]), d(ce["block_size"] / 8), d(ce["m"])
    ns, nblk = len(st["x"]), len(ce["block_size"])
    out = {}
    with GravityField(mode, validate=False, tune_cache=None) as f:
        f.set_sources(*S)
        table = []
        for sp in (0, 1, 2, 3, 4, 6, 8, 12, 16):
            f.set_config(nsplit=sp)
            ms = _time(lambda: f.evaluate_blocks(BO, BD), reps)
            table.append(dict(nsplit=sp, ms=ms))
            log(f"  {mode} block targets nsplit={sp}: {ms:.3f} ms")
        out[blocks_key(mode, "blockfwd", ns, nblk, f.column_waves())] = pick(table, _chunks,
                                                                              TIE_BLOCK_TARGETS)
        f.set_source_blocks(BO, BD, CM)
        table = []
        for sp in (0, 8, 16, 24, 32, 48, 64, 100, 128, 200, 256):
            f.set_config(nsplit=sp)
            ms = _time(lambda: f.evaluate(None, S[0], S[1], S[2]), reps)
            table.append(dict(nsplit=sp, ms=ms))
            log(f"  {mode} block sources nsplit={sp}: {ms:.3f} ms")
        out[blocks_key(mode, "blocksrc", ns, nblk, f.column_waves())] = pick(table, _chunks)
    return out


def apply_blocks(field, direction, ns, nblk, cache=None):
    cache = load_cache() if cache is None else cache
    mode = {v: k for k, v in MODES.items() if k != "plain_fp64"}[field.mode]
    e = cache.get(blocks_key(mode, direction, ns, nblk, field.column_waves()))
    if e:
        field.set_config(nsplit=e["nsplit"])
    return e


def apply_eval(field, ns, nt, cache=None):
    """Configure `field` from the cache for an evaluation of this size; returns the entry."""
    cache = load_cache() if cache is None else cache
    mode = {v: k for k, v in MODES.items() if k != "plain_fp64"}[field.mode]
    e = cache.get(eval_key(mode, ns, nt))
    if e:
        field.set_config(variant=e["variant"], nsplit=e["nsplit"])
    return e


def apply_batch(field, problems, cache=None):
    cache = load_cache() if cache is None else cache
    mode = {v: k for k, v in MODES.items() if k != "plain_fp64"}[field.mode]
    e = cache.get(batch_key(mode, problems, field.column_waves()))
    if e:
        field.set_config(variant=e["variant"])
        _lib.check(field.lib.tgrav_set_batch_config(field.handle, float(e["item"]),
                                                     int(e["wg_per_cu"])), "tgrav_set_batch_config")
        _lib.check(field.lib.tgrav_set_batch_scales(field.handle, float(e.get("col_scale", 1.0)),
                                                     float(e.get("gs_scale", 1.0))), "tgrav_set_batch_scales")
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--eval", nargs=2, type=int, action="append", default=[],
                    metavar=("NS", "NT"))
    ap.add_argument("--batch", action="append", default=[], choices=["c4", "c5"])
    ap.add_argument("--blocks", action="append", default=[], choices=["c2", "c5"])
    ap.add_argument("--modes", default="mixed,fp64")
    ap.add_argument("--cache", default=CACHE)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    cache = load_cache(args.cache)
    t0 = time.time()
    for mode in args.modes.split(","):
        for ns, nt in args.eval:
            best, _ = tune_eval(mode, ns, nt, reps=args.reps)
            cache[eval_key(mode, ns, nt)] = dict(variant=best["variant"], nsplit=best["nsplit"],
                                                 ms=best["ms"], k=best["k"], threads=best["threads"])
            print(f"best {mode} eval ns={ns} nt={nt}: {cache[eval_key(mode, ns, nt)]}", flush=True)
            save_cache(cache, args.cache)
        for name in args.blocks:
            res = tune_blocks(mode, name, reps=args.reps)
            cache.update(res)
            print(f"best {mode} blocks {name}: {res}", flush=True)
            save_cache(cache, args.cache)
        for name in args.batch:
            from . import workloads as W
            if name == "c4":
                probs = W.c4(1004)
                best, _ = tune_batch(mode, probs, reps=max(2, args.reps // 2))
            else:
                probs = c5_plan_problems()
                best, _ = tune_batch(mode, probs, reps=2, items=(0.0, 1e7, 2.5e7, 5e7, 1e8))
            key = batch_key(mode, probs, best["waves"])
            cache[key] = dict(variant=best["variant"], item=best["item"], wg_per_cu=best["wg_per_cu"],
                              col_scale=best["col_scale"], gs_scale=best["gs_scale"], ms=best["ms"], k=best["k"],
                              threads=best["threads"])
            print(f"best {mode} batch {name}: {cache[key]}", flush=True)
            save_cache(cache, args.cache)
    print(f"tuned in {time.time() - t0:.1f} s -> {args.cache}", flush=True)


if __name__ == "__main__":
    main()
