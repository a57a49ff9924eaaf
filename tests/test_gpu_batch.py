"""Batched persistent evaluation and its hipGraph replay (needs an MI355X).

Each problem of a batch must agree with the oracle (FP64: 1e-12; mixed: 1e-6 median,
1e-4 max) and with a standalone evaluation of the same problem; graph replays must be
bit-identical to the eager (non-graph) run, including after unrelated eager work between
replays; the fused kick must equal v + dt*a bit for bit.
"""
import numpy as np
import pytest

import oracle
from tests import _util
from torch_amd import GravityField, TgravError, _lib
from torch_amd import workloads as W
from torch_amd.batch import BatchPlan, problems_from_workload

pytestmark = pytest.mark.gpu
G = W.G_PC_MSUN_KMS


def small_batch(seed=31, n=12):
    """Uneven problems incl. an empty-source and an empty-target one."""
    rng = W.rng_for(seed)
    probs = []
    for p in range(n):
        ns = int(10 ** rng.uniform(0, 4.3))
        nt = int(10 ** rng.uniform(0, 4.5))
        if p == 3:
            ns = 0
        if p == 5:
            nt = 0
        st = W.star_cluster(rng, max(ns, 1))
        src = dict(x=st["x"][:ns], y=st["y"][:ns], z=st["z"][:ns], m=st["m"][:ns],
                   eps=rng.uniform(0.005, 0.1, ns))
        t = rng.uniform(-2, 2, (3, nt))
        probs.append(dict(sources=src, targets=dict(x=t[0], y=t[1], z=t[2], eps=np.zeros(nt))))
    return probs


def to_np(outs):
    return [tuple(None if a is None else a.cpu().numpy() for a in o) for o in outs]


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_batch_matches_oracle_and_standalone(mode):
    import torch
    probs = small_batch()
    with GravityField(mode) as f:
        plan = BatchPlan(f, problems_from_workload(probs))
        got = to_np(plan.run(graph=False))
        torch.cuda.synchronize()
        for i, p in enumerate(probs):
            S, T = p["sources"], p["targets"]
            nt = len(T["x"])
            if nt == 0:
                continue
            ref = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], T["x"], T["y"], T["z"],
                                      T["eps"], G, mode=oracle.NEUMAIER, threads=16)
            if len(S["x"]) == 0:
                assert all(np.all(a == 0) for a in got[i])
                continue
            s = _util.summary(got[i], ref)
            if mode == "fp64":
                assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (i, s)
            else:
                assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (i, s)
            f.set_sources(S["x"], S["y"], S["z"], S["m"], S["eps"])
            one = f.evaluate(T["eps"], T["x"], T["y"], T["z"])
            s1 = _util.summary(got[i], one)
            assert s1["acc_max"] <= (1e-12 if mode == "fp64" else 1e-5), (i, s1)
        plan.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_graph_replays_are_identical(mode):
    import torch
    probs = small_batch(seed=32)
    with GravityField(mode) as f:
        f.set_config(nsplit=0)
        lib = _lib.load()
        # small work items force chunked problems (partials + reduction inside the graph)
        assert lib.tgrav_set_batch_config(f.handle, 2.0e5, 0) == 0
        plan = BatchPlan(f, problems_from_workload(probs))
        eager = to_np(plan.run(graph=False))
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device="cuda")
        for rep in range(4):
            for o in plan.outputs:
                for a in o:
                    if a is not None:
                        a.fill_(np.nan)
            out = to_np(plan.run(graph=True))
            torch.cuda.synchronize()
            for i, (a, b) in enumerate(zip(out, eager)):
                for p, q in zip(a, b):
                    if p is not None:
                        assert np.array_equal(p, q), (rep, i)
            junk = (junk * 1.0001).sum() + junk      # eager work between replays
            torch.cuda.synchronize()
        plan.close()


def test_batch_kick_in_graph():
    import torch
    probs = small_batch(seed=33, n=6)
    dev = problems_from_workload(probs)
    rng = W.rng_for(5)
    vel0 = []
    for p in dev:
        nt = p["targets"]["x"].numel()
        v = [torch.as_tensor(rng.normal(size=nt), device="cuda") for _ in range(3)]
        vel0.append([a.clone() for a in v])
        p["velocities"] = v
    dt = 6.5e-4
    with GravityField("fp64") as f:
        plan = BatchPlan(f, dev, dt=dt)
        for rep in range(3):
            acc = to_np(plan.run(graph=True))
            torch.cuda.synchronize()
        for i, p in enumerate(dev):
            if p["targets"]["x"].numel() == 0:
                continue
            v = [a.cpu().numpy() for a in vel0[i]]
            for _ in range(3):
                v = list(oracle.kick(dt, acc[i][0], acc[i][1], acc[i][2], *v))
            for q in range(3):
                assert np.array_equal(p["velocities"][q].cpu().numpy(), v[q]), (i, q)
        plan.close()


def test_batch_argument_errors():
    with GravityField("fp64") as f:
        bad = [dict(sources=dict(x=np.zeros(2), y=np.zeros(2), z=np.zeros(2), m=np.ones(2)),
                    targets=dict(x=np.zeros(3), y=np.zeros(3), z=np.zeros(3)))]
        with pytest.raises(TgravError) as e:
            BatchPlan(f, problems_from_workload(bad), dt=0.1)     # dt != 0 needs velocities
        assert e.value.code in (_lib.EINVAL,) or isinstance(e.value, TgravError)


def test_c4_batch_mixed_against_fp64():
    """The 64-problem uneven batch (c4, graph replay): mixed within the error gate of the FP64 batch
    over every target, and both modes against the oracle on 64 seeded targets of every problem
    (FP64 within 1e-12; mixed within the gate)."""
    import torch
    probs = W.c4(1004)
    dev = problems_from_workload(probs)
    res = {}
    for mode in ("fp64", "mixed"):
        with GravityField(mode) as f:
            plan = BatchPlan(f, dev)
            res[mode] = to_np(plan.run(graph=True))
            torch.cuda.synchronize()
            plan.close()
    ea = np.concatenate([_util.rel_acc_err(m, r) for m, r in zip(res["mixed"], res["fp64"])])
    print("c4 mixed vs fp64: median", np.median(ea), "max", ea.max(), flush=True)
    assert np.median(ea) <= 1e-6 and ea.max() <= 1e-4
    # both modes against the oracle on 64 targets of every problem
    em = []
    for i in range(len(probs)):
        S, T = probs[i]["sources"], probs[i]["targets"]
        idx = _util.subsample(len(T["x"]), 64, i)
        ref = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], T["x"][idx], T["y"][idx],
                                  T["z"][idx], None, G, mode=oracle.NEUMAIER, threads=16)
        s = _util.summary(tuple(a[idx] for a in res["fp64"][i]), ref)
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (i, s)
        em.append(_util.rel_acc_err(tuple(a[idx] for a in res["mixed"][i]), ref))
    em = np.concatenate(em)
    print("c4 mixed vs oracle: median", np.median(em), "max", em.max(), flush=True)
    assert np.median(em) <= 1e-6 and em.max() <= 1e-4


def block_batch(seed=41, n=9):
    """Problems whose targets are 8^3 leaf blocks (incl. an empty-source one), given both ways."""
    rng = W.rng_for(seed)
    probs = []
    for p in range(n):
        ns = 0 if p == 4 else int(10 ** rng.uniform(0, 4.0))
        nblk = int(rng.integers(1, 12))
        o = rng.uniform(-2.0, 1.5, (nblk, 3))
        s = 10.0 ** rng.uniform(np.log10(0.03125), np.log10(0.5), nblk)
        tx, ty, tz, _ = W.block_cell_centres(o, s)
        st = W.star_cluster(rng, max(ns, 1))
        src = dict(x=st["x"][:ns], y=st["y"][:ns], z=st["z"][:ns], m=st["m"][:ns],
                   eps=rng.uniform(0.005, 0.1, ns))
        probs.append(dict(sources=src, targets=dict(x=tx, y=ty, z=tz, block_origin=o, block_dx=s / 8)))
    return probs


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_block_target_problems(mode):
    """Block-target problems (column kernel inside the batch) against the same problems with
    explicit cell centres, against the oracle, over graph replays, with the fused kick.  The
    explicit centres are in the blocks' C order, so the plan finds their column structure and
    gives them the same kernel: the results are the same bits."""
    import torch
    probs = block_batch()
    explicit = [dict(sources=p["sources"], targets={k: p["targets"][k] for k in ("x", "y", "z")})
                for p in probs]
    dev_b, dev_e = problems_from_workload(probs), problems_from_workload(explicit)
    rng = W.rng_for(6)
    vel0 = []
    for p in dev_b:
        nt = p["targets"]["x"].numel()
        v = [torch.as_tensor(rng.normal(size=nt), device="cuda") for _ in range(3)]
        vel0.append([a.clone() for a in v])
        p["velocities"] = v
    dt = 3e-4
    with GravityField(mode) as f:
        pe = BatchPlan(f, dev_e)
        ref = to_np(pe.run(graph=False))
        pb = BatchPlan(f, dev_b, dt=dt)
        assert pb.n_items > 0
        eager = to_np(pb.run(graph=False))
        torch.cuda.synchronize()
        for rep in range(3):
            again = to_np(pb.run(graph=True))
            torch.cuda.synchronize()
            for i in range(len(probs)):
                for q in range(4):
                    assert np.array_equal(again[i][q], eager[i][q]), (rep, i, q)
        pe.close()
        pb.close()
    for i in range(len(probs)):
        for q in range(4):
            assert np.array_equal(ref[i][q], eager[i][q]), (i, q)
    for i, p in enumerate(probs):
        S, T = p["sources"], p["targets"]
        s = _util.summary(eager[i], ref[i]) if len(S["x"]) else None
        if len(S["x"]) == 0:
            assert all(np.all(a == 0) for a in eager[i])
            continue
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (i, s)
            o = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], T["x"], T["y"], T["z"], None,
                                    G, mode=oracle.NEUMAIER, threads=16)
            s2 = _util.summary(eager[i], o)
            assert s2["acc_max"] <= 1e-12 and s2["pot_max"] <= 1e-12, (i, s2)
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (i, s)
        # 4 runs (1 eager + 3 replays), each kicking by dt with the same acceleration
        v = [a.cpu().numpy() for a in vel0[i]]
        for _ in range(4):
            v = list(oracle.kick(dt, eager[i][0], eager[i][1], eager[i][2], *v))
        for q in range(3):
            assert np.array_equal(dev_b[i]["velocities"][q].cpu().numpy(), v[q]), (i, q)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_column_items_scaled(mode):
    """tgrav_set_batch_scales: the column list's items at 2.5x the item size give fewer work items
    and partial sums, the block and explicit forms stay bit-identical, graph replays equal the
    eager run, and the results agree with the unscaled plan (FP64 within 1e-12, mixed within the
    gate)."""
    import torch
    probs = block_batch(seed=45)
    explicit = [dict(sources=p["sources"], targets={k: p["targets"][k] for k in ("x", "y", "z")})
                for p in probs]
    dev_b, dev_e = problems_from_workload(probs), problems_from_workload(explicit)
    lib = _lib.load()
    res = {}
    with GravityField(mode) as f:
        set_item(f, 2e5)
        for scale in (1.0, 2.5):
            assert lib.tgrav_set_batch_scales(f.handle, scale, 1.0) == 0
            pb, pe = BatchPlan(f, dev_b), BatchPlan(f, dev_e)
            eager = to_np(pb.run(graph=False))
            again = to_np(pb.run(graph=True))
            expl = to_np(pe.run(graph=True))
            torch.cuda.synchronize()
            for i in range(len(probs)):
                for q in range(4):
                    assert np.array_equal(again[i][q], eager[i][q]), (scale, i, q)
                    assert np.array_equal(expl[i][q], eager[i][q]), (scale, i, q)
            res[scale] = (eager, pb.n_items, pb.partial_bytes)
            pb.close()
            pe.close()
        assert lib.tgrav_set_batch_scales(f.handle, 0.0, 0.0) == 0
        assert lib.tgrav_set_batch_scales(f.handle, -1.0, 1.0) == _lib.EINVAL
    assert res[2.5][1] < res[1.0][1] and res[2.5][2] < res[1.0][2]
    for i in range(len(probs)):
        if len(probs[i]["sources"]["x"]) == 0:
            continue
        s = _util.summary(res[2.5][0][i], res[1.0][0][i])
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (i, s)
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (i, s)


def test_column_chunks_stay_within_the_l2():
    """A block-target problem with 60,000 sources and an item size that would give one chunk is
    cut into two (a chunk's sources stay within 2 MB): two work items, FP64 within 1e-12 of the
    oracle."""
    rng = W.rng_for(46)
    o = rng.uniform(-1.0, 0.5, (1, 3))
    s = np.array([0.25])
    tx, ty, tz, _ = W.block_cell_centres(o, s)
    st = W.star_cluster(rng, 60000)
    eps = rng.uniform(0.005, 0.05, 60000)
    probs = [dict(sources=dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=eps),
                  targets=dict(x=tx, y=ty, z=tz, block_origin=o, block_dx=s / 8))]
    with GravityField("fp64") as f:
        set_item(f, 1e30)
        plan = BatchPlan(f, problems_from_workload(probs))
        assert plan.n_items == 2
        got = to_np(plan.run(graph=True))[0]
        plan.close()
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], eps, tx, ty, tz, None, G,
                              mode=oracle.NEUMAIER, threads=16)
    s2 = _util.summary(got, ref)
    assert s2["acc_max"] <= 1e-12 and s2["pot_max"] <= 1e-12, s2


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_explicit_targets_without_column_structure(mode):
    """Explicit targets whose count is a multiple of 512 but which are not in the blocks' C order
    (shuffled cell centres) stay on the generic work list: against the oracle (FP64) and the
    FP64 plan (mixed)."""
    probs = block_batch(seed=43, n=5)
    rng = W.rng_for(9)
    shuffled = []
    for p in probs:
        perm = rng.permutation(len(p["targets"]["x"]))
        shuffled.append(dict(sources=p["sources"], targets={k: p["targets"][k][perm] for k in ("x", "y", "z")}))
    res = {}
    for m in ("fp64", mode):
        with GravityField(m) as f:
            plan = BatchPlan(f, problems_from_workload(shuffled))
            res[m] = to_np(plan.run(graph=True))
            plan.close()
    for i, p in enumerate(shuffled):
        S, T = p["sources"], p["targets"]
        if len(S["x"]) == 0:
            continue
        o = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], T["x"], T["y"], T["z"], None,
                                G, mode=oracle.NEUMAIER, threads=16)
        s = _util.summary(res["fp64"][i], o)
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (i, s)
        if mode == "mixed":
            s = _util.summary(res["mixed"][i], res["fp64"][i])
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (i, s)


def cell_blocks(rng, nblk, lo=-2.0, hi=1.5):
    o = rng.uniform(lo, hi, (nblk, 3))
    s = 10.0 ** rng.uniform(np.log10(0.03125), np.log10(0.5), nblk)
    tx, ty, tz, _ = W.block_cell_centres(o, s)
    m = rng.lognormal(0.0, 1.0, nblk * 512) * np.repeat((s / 8) ** 3, 512)
    return o, s / 8, (tx, ty, tz), m


def set_item(f, item):
    _lib.check(_lib.load().tgrav_set_batch_config(f.handle, float(item), 0), "tgrav_set_batch_config")


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_block_source_problems(mode):
    """Problems whose sources are 8^3 leaf blocks (the reverse direction: cells as point masses onto
    stars), in one plan with a block-target problem (the forward direction) and a generic one.
    Unsplit, each equals the eager block kernels (k_eval_gridsrc, k_eval_grid) bit for bit; with the
    default item sizes, FP64 is within 1e-12 of the oracle (binary members included) and mixed meets
    the gate; graph replays equal the eager run of the plan."""
    import torch
    rng = W.rng_for(71)
    o, dx, (cx, cy, cz), cm = cell_blocks(rng, 14)
    ce = np.repeat(dx, 512)
    st = W.star_cluster(rng, 900)
    st["x"][:2] = [0.1, 0.1 + 5e-6]            # a 1 AU pair
    st["y"][:2] = [0.2, 0.2]
    st["z"][:2] = [-0.3, -0.3]
    seps = np.full(900, 0.0)
    t2 = rng.uniform(-2, 2, (3, 700))
    probs = [dict(sources=dict(block_origin=o, block_dx=dx, block_eps=dx, m=cm),
                  targets=dict(x=st["x"], y=st["y"], z=st["z"], eps=seps)),
             dict(sources=dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=np.full(900, 0.01)),
                  targets=dict(block_origin=o, block_dx=dx)),
             dict(sources=dict(x=st["x"][:300], y=st["y"][:300], z=st["z"][:300], m=st["m"][:300]),
                  targets=dict(x=t2[0], y=t2[1], z=t2[2]))]
    dev = problems_from_workload(probs)
    with GravityField(mode) as f:
        f.set_config(nsplit=1)
        f.set_source_blocks(o, dx, cm, eps=dx)
        e_rev = f.evaluate(seps, st["x"], st["y"], st["z"])
        f.set_sources(st["x"], st["y"], st["z"], st["m"], np.full(900, 0.01))
        e_fwd = f.evaluate_blocks(o, dx)
        set_item(f, 1e30)                       # one chunk per problem, as the eager launches
        plan = BatchPlan(f, dev)
        assert plan.n_items > 0
        got = to_np(plan.run(graph=True))
        torch.cuda.synchronize()
        for q in range(4):
            assert np.array_equal(got[0][q], e_rev[q]), q
            assert np.array_equal(got[1][q], e_fwd[q]), q
        plan.close()
        set_item(f, 0)                          # default item sizes: the sources are split
        f.set_config(nsplit=0)
        plan = BatchPlan(f, dev)
        eager = to_np(plan.run(graph=False))
        torch.cuda.synchronize()
        for rep in range(2):
            again = to_np(plan.run(graph=True))
            torch.cuda.synchronize()
            for i in range(3):
                for q in range(4):
                    assert np.array_equal(again[i][q], eager[i][q]), (rep, i, q)
        n0 = plan.partial_bytes
        plan.close()
        # the block-source list's items at 2.5x (tgrav_set_batch_scales): fewer partial sums, the
        # same field (FP64 within 1e-12, mixed within the gate)
        set_item(f, 2e5)
        plan = BatchPlan(f, dev)
        base = to_np(plan.run(graph=True))
        nb0 = plan.partial_bytes
        plan.close()
        _lib.check(_lib.load().tgrav_set_batch_scales(f.handle, 1.0, 2.5), "tgrav_set_batch_scales")
        plan = BatchPlan(f, dev)
        scaled = to_np(plan.run(graph=True))
        assert plan.partial_bytes < nb0, (plan.partial_bytes, nb0, n0)
        plan.close()
        _lib.check(_lib.load().tgrav_set_batch_scales(f.handle, 0.0, 0.0), "tgrav_set_batch_scales")
        set_item(f, 0)
    s_sc = _util.summary(scaled[0], base[0])
    if mode == "fp64":
        assert s_sc["acc_max"] <= 1e-12 and s_sc["pot_max"] <= 1e-12, s_sc
    else:
        assert s_sc["acc_median"] <= 1e-6 and s_sc["acc_max"] <= 1e-4, s_sc
    refs = [oracle.eval_plummer(cx, cy, cz, cm, ce, st["x"], st["y"], st["z"], None, G, mode=oracle.NEUMAIER,
                                threads=16),
            oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], np.full(900, 0.01), cx, cy, cz, None, G,
                                mode=oracle.NEUMAIER, threads=16),
            oracle.eval_plummer(st["x"][:300], st["y"][:300], st["z"][:300], st["m"][:300], None, t2[0], t2[1],
                                t2[2], None, G, mode=oracle.NEUMAIER, threads=16)]
    for i in range(3):
        s = _util.summary(eager[i], refs[i])
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (i, s)
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (i, s)
    s = _util.summary(tuple(a[:2] for a in eager[0]), tuple(a[:2] for a in refs[0]))   # the binary
    assert s["acc_max"] <= (1e-12 if mode == "fp64" else 1e-4), s


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_column_structure_checked_at_every_run(mode):
    """A plan created on zero-filled explicit targets (every aligned 8 then share x and y) whose
    targets are written afterwards: each run tests the column structure again, inside the graph,
    so shuffled centres take the generic kernel and C-ordered ones the column kernel."""
    import torch
    rng = W.rng_for(81)
    o, dx, (cx, cy, cz), _ = cell_blocks(rng, 6)
    st = W.star_cluster(rng, 500)
    S = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=np.full(500, 0.01))
    n = len(cx)
    tgt = {k: torch.zeros(n, dtype=torch.float64, device="cuda") for k in ("x", "y", "z")}
    src = {k: torch.as_tensor(v, device="cuda") for k, v in S.items()}
    perm = rng.permutation(n)
    with GravityField(mode) as f:
        plan = BatchPlan(f, [dict(sources=src, targets=tgt)])
        blk = BatchPlan(f, [dict(sources=src, targets=dict(block_origin=torch.as_tensor(o, device="cuda"),
                                                          block_dx=torch.as_tensor(dx, device="cuda")))])
        ref_blk = to_np(blk.run(graph=False))[0]
        for order in ("shuffled", "c_order", "shuffled"):
            p = perm if order == "shuffled" else np.arange(n)
            for k, v in zip(("x", "y", "z"), (cx, cy, cz)):
                tgt[k].copy_(torch.as_tensor(v[p], device="cuda"))
            got = to_np(plan.run(graph=True))[0]
            torch.cuda.synchronize()
            ref = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], cx[p], cy[p], cz[p], None, G,
                                      mode=oracle.NEUMAIER, threads=16)
            s = _util.summary(got, ref)
            if mode == "fp64":
                assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (order, s)
            else:
                assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (order, s)
            if order == "c_order":              # the column kernel: the block form's bits
                for q in range(4):
                    assert np.array_equal(got[q], ref_blk[q]), q
        plan.close()
        blk.close()


def test_plan_left_open_is_closed_with_its_field():
    """A plan the caller never closed is closed by its field before the field's context goes
    (the plan uses the context's device and stream), so dropping it afterwards is harmless."""
    import gc
    probs = block_batch(seed=47, n=2)
    f = GravityField("fp64")
    plan = BatchPlan(f, problems_from_workload(probs))
    plan.run(graph=False)
    f.close()
    assert plan._h is None
    plan.close()
    del plan
    gc.collect()
