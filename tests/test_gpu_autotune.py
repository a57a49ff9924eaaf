"""The autotuner itself on the GPU (needs an MI355X): tune_eval at a tiny size writes an entry, and a
context created with that cache applies it -- the tuned launch equals the same shape set by hand,
bit for bit -- and the batch tuner runs on a small plan."""
import json

import numpy as np
import pytest

from torch_amd import GravityField, _lib, autotune
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_tune_eval_entry_is_applied(tmp_path):
    ns, nt = 1500, 2500
    best, table = autotune.tune_eval("fp64", ns, nt, reps=2, nsplits=(0, 1, 3), log=lambda *a: None)
    assert len(table) >= 3 and best["ms"] > 0
    cache = {autotune.eval_key("fp64", ns, nt): dict(variant=best["variant"], nsplit=best["nsplit"],
                                                    ms=best["ms"], k=best["k"], threads=best["threads"])}
    path = tmp_path / "cache.json"
    path.write_text(json.dumps(cache))
    assert _lib.load().tgrav_tune_cache_entries(str(path).encode()) == 1
    rng = W.rng_for(5)
    st = W.star_cluster(rng, ns)
    t = rng.uniform(-2, 2, (3, nt))
    res = []
    for tune, cfg in ((str(path), None), (None, (best["variant"], best["nsplit"]))):
        with GravityField("fp64", tune_cache=tune) as f:
            if cfg:
                f.set_config(variant=cfg[0], nsplit=cfg[1])
            f.set_sources(st["x"], st["y"], st["z"], st["m"], st["eps"])
            res.append(f.evaluate(None, t[0], t[1], t[2]))
    for q in range(4):
        np.testing.assert_array_equal(res[0][q], res[1][q])


def test_tune_batch_small_plan():
    rng = W.rng_for(6)
    probs = []
    for p in range(3):
        st = W.star_cluster(rng, 300 + 200 * p)
        t = rng.uniform(-2, 2, (3, 1000 + 500 * p))
        probs.append(dict(sources={k: st[k] for k in ("x", "y", "z", "m", "eps")},
                          targets=dict(x=t[0], y=t[1], z=t[2])))
    best, table = autotune.tune_batch("mixed", probs, reps=1, items=(0.0, 2e5), wg_per_cu=(0, 1),
                                      log=lambda *a: None)
    assert len(table) >= 4 and best["ms"] > 0 and best["item"] in (0.0, 2e5)


def test_tune_batch_block_lists_and_apply():
    """A plan with a block-target and a block-source problem: the batch tuner also times the block
    lists' item scales, its choice carries them, and apply_batch sets that shape on a field (the
    plan built there writes the partial bytes the tuner saw for it)."""
    from torch_amd.batch import BatchPlan, problems_from_workload
    rng = W.rng_for(7)
    nblk = 6
    o = rng.uniform(-1.5, 1.0, (nblk, 3))
    dx = np.full(nblk, 0.25 / 8)
    st = W.star_cluster(rng, 2000)
    cm = rng.lognormal(0.0, 1.0, nblk * 512) * 1e-3
    probs = [dict(sources={k: st[k] for k in ("x", "y", "z", "m", "eps")},
                  targets=dict(block_origin=o, block_dx=dx)),
             dict(sources=dict(block_origin=o, block_dx=dx, block_eps=dx, m=cm),
                  targets=dict(x=st["x"], y=st["y"], z=st["z"]))]
    best, table = autotune.tune_batch("fp64", probs, reps=1, items=(0.0, 2e5), log=lambda *a: None)
    scaled = [r for r in table if (r["col_scale"], r["gs_scale"]) != (1.0, 1.0)]
    assert len(scaled) == len(autotune.LIST_SCALES) - 1
    assert best["col_scale"] in (1.0, 2.5) and best["gs_scale"] in (1.0, 1.5)
    cache = {autotune.batch_key("fp64", probs, best["waves"]): dict(
        variant=best["variant"], item=best["item"], wg_per_cu=best["wg_per_cu"], col_scale=best["col_scale"],
        gs_scale=best["gs_scale"], ms=best["ms"], k=best["k"], threads=best["threads"])}
    with GravityField("fp64", tune_cache=None) as f:
        e = autotune.apply_batch(f, probs, cache=cache)
        assert e is not None
        plan = BatchPlan(f, problems_from_workload(probs))
        assert plan.partial_bytes == best["partial_bytes"]
        plan.close()
