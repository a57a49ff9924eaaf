"""Block kernels with several waves per workgroup (tgrav_set_column_waves).  Column kernel (block
targets): the W waves of a workgroup hold the same 8^3 leaf block, split its sources in rounds of
64 W, and add their partial sums in LDS in wave order.  Block-source kernel: the W waves hold the
same tile of targets (2 per lane mixed, 4 FP64) and split the source blocks (wave v takes every
W-th block).

For every W in (1, 2, 4): FP64 within 1e-12 of the FP64 oracle and mixed within the parity bar
(<= 1e-6 median, <= 1e-4 max relative acceleration error), with source counts below one round
(waves with nothing to do), ragged last rounds and source chunks split over workgroups.  For a
fixed W the block form, the explicit C-order form, the fused kick and the unsplit batch plan give
the same bits, and a repeated run is bit-identical."""
import numpy as np
import pytest

import oracle
from tests import _util
from tests.test_gpu_batch import cell_blocks, set_item, to_np
from torch_amd import GasGravityField, GravityField, StellarGravityField, TgravError
from torch_amd import workloads as W
from torch_amd.batch import BatchPlan, problems_from_workload

pytestmark = pytest.mark.gpu
G = W.G_PC_MSUN_KMS


@pytest.fixture(scope="module")
def c2():
    return W.c2(1002)


def check_bar(mode, s):
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
    else:
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, s


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("waves", [1, 2, 4])
def test_column_waves_match_oracle(mode, waves, c2):
    st, ce = c2["stars"], c2["cells"]
    o, dx = ce["block_origin"][:24], ce["block_size"][:24] / 8
    bx, by, bz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    te = np.repeat(dx, 512)
    with StellarGravityField(mode) as f:
        f.set_column_waves(waves)
        assert f.column_waves() == waves
        for ns in (1, 100, 257, 1000):          # below one round of 64 W sources, ragged rounds
            sl = slice(0, ns)
            S = [st[k][sl] for k in ("x", "y", "z", "m", "eps")]
            f.set_stars(*S)
            ref = oracle.eval_plummer(*S, bx, by, bz, te, G, mode=oracle.NEUMAIER, threads=16)
            for nsplit in (1, 3):
                f.set_config(nsplit=nsplit)
                got = f.evaluate_blocks(o, dx, eps=dx)
                s = _util.summary(got, ref)
                print(mode, "waves", waves, "ns", ns, "nsplit", nsplit, s, flush=True)
                check_bar(mode, s)
                again = f.evaluate_blocks(o, dx, eps=dx)
                explicit = f.evaluate(te, bx, by, bz)          # detected as columns
                for p, q, r in zip(got, again, explicit):
                    assert np.array_equal(p, q)
                    assert np.array_equal(p, r)
                v = [np.zeros(len(bx)) for _ in range(3)]
                f.evaluate_blocks(o, dx, eps=dx, dt=0.5, vel=v)
                for p, q in zip(v, got[:3]):
                    assert np.array_equal(p, 0.5 * q)
            f.set_config(nsplit=0)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_column_waves_batch_equals_eager(mode):
    """Four waves per column workgroup: block-target and explicit C-order problems in one plan,
    unsplit, equal the eager column kernels bit for bit; split (default item sizes), the plan meets
    the bar against the oracle and graph replays equal its eager run."""
    import torch
    rng = W.rng_for(73)
    o, dx, (cx, cy, cz), cm = cell_blocks(rng, 11)
    st = W.star_cluster(rng, 1500)
    seps = np.full(1500, 0.02)
    probs = [dict(sources=dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=seps),
                  targets=dict(block_origin=o, block_dx=dx)),
             dict(sources=dict(x=st["x"][:150], y=st["y"][:150], z=st["z"][:150], m=st["m"][:150]),
                  targets=dict(x=cx, y=cy, z=cz))]
    dev = problems_from_workload(probs)
    with GravityField(mode) as f:
        f.set_column_waves(4)
        f.set_config(nsplit=1)
        f.set_sources(st["x"], st["y"], st["z"], st["m"], seps)
        e0 = f.evaluate_blocks(o, dx)
        f.set_sources(st["x"][:150], st["y"][:150], st["z"][:150], st["m"][:150])
        e1 = f.evaluate(None, cx, cy, cz)
        set_item(f, 1e30)                       # one chunk per problem, as the eager launches
        plan = BatchPlan(f, dev)
        got = to_np(plan.run(graph=True))
        torch.cuda.synchronize()
        for q in range(4):
            assert np.array_equal(got[0][q], e0[q]), q
            assert np.array_equal(got[1][q], e1[q]), q
        plan.close()
        set_item(f, 0)
        f.set_config(nsplit=0)
        plan = BatchPlan(f, dev)
        eager = to_np(plan.run(graph=False))
        torch.cuda.synchronize()
        for rep in range(2):
            again = to_np(plan.run(graph=True))
            torch.cuda.synchronize()
            for i in range(2):
                for q in range(4):
                    assert np.array_equal(again[i][q], eager[i][q]), (rep, i, q)
        plan.close()
    refs = [oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], seps, cx, cy, cz, None, G,
                                mode=oracle.NEUMAIER, threads=16),
            oracle.eval_plummer(st["x"][:150], st["y"][:150], st["z"][:150], st["m"][:150], None, cx, cy, cz,
                                None, G, mode=oracle.NEUMAIER, threads=16)]
    for i in range(2):
        s = _util.summary(tuple(eager[i]), refs[i])
        print(mode, "batch waves 4 problem", i, s, flush=True)
        check_bar(mode, s)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("waves", [1, 4])
def test_shared_and_per_lane_dz_agree(mode, waves, c2):
    """Explicit column targets whose 64 columns of a group hold the same 8 z take the shared dz (each
    source's 8 dz formed once, by the lane that stages the source); a group whose columns differ in z
    forms them per lane.  Reversing the 8 z of one column per group keeps every group's bounding box
    (the mixed anchor) and sends each group down the per-lane path: every other column gives the same
    bits as before, the reversed column the same values in reverse order, eagerly and in a batch plan
    (the persistent kernel), and both layouts meet the bar against the oracle."""
    import torch
    st, ce = c2["stars"], c2["cells"]
    o, dx = ce["block_origin"][:16], ce["block_size"][:16] / 8
    bx, by, bz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    te = np.repeat(dx, 512)
    rz = bz.reshape(-1, 64, 8).copy()
    rz[:, 37, :] = rz[:, 37, ::-1]                  # column 37 of every group, z reversed
    rz = rz.ravel()
    col = np.zeros(len(bz), dtype=bool)
    col.reshape(-1, 64, 8)[:, 37, :] = True
    perm = np.arange(len(bz)).reshape(-1, 64, 8)
    perm[:, 37, :] = perm[:, 37, ::-1]               # b[t] corresponds to a[perm[t]]
    perm = perm.ravel()
    S = [st[k][:2000] for k in ("x", "y", "z", "m", "eps")]
    with StellarGravityField(mode) as f:
        f.set_column_waves(waves)
        f.set_stars(*S)
        f.set_config(nsplit=1)
        a = f.evaluate(te, bx, by, bz)
        b = f.evaluate(te, bx, by, rz)
        for p, q in zip(a, b):
            assert np.array_equal(p[~col], q[~col])
            assert np.array_equal(p[perm[col]], q[col])
        ref = oracle.eval_plummer(*S, bx, by, rz, te, G, mode=oracle.NEUMAIER, threads=16)
        s = _util.summary(b, ref)
        print(mode, "waves", waves, "per-lane dz vs oracle", s, flush=True)
        check_bar(mode, s)
        if waves == 4:
            dev = problems_from_workload([dict(sources=dict(zip(("x", "y", "z", "m", "eps"), S)),
                                               targets=dict(x=bx, y=by, z=z, eps=te)) for z in (bz, rz)])
            set_item(f, 1e30)                       # one chunk per problem, as the eager launches
            plan = BatchPlan(f, dev)
            got = to_np(plan.run(graph=True))
            torch.cuda.synchronize()
            plan.close()
            set_item(f, 0)
            for q in range(4):
                assert np.array_equal(got[0][q], a[q]), q
                assert np.array_equal(got[1][q], b[q]), q
        f.set_config(nsplit=0)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("waves", [1, 2, 4])
def test_block_source_waves_match_oracle(mode, waves, c2):
    """13 source blocks (ragged rounds of W blocks, chunks of 1, 2 and 5 blocks) onto 700 stars
    (a partial last tile of 256): the oracle's field to the bar; repeated runs and explicit
    cells in the blocks' C order (detected as blocks) give the same bits."""
    st, ce = c2["stars"], c2["cells"]
    nb = 13
    sl = slice(0, nb * 512)
    o, dxb = ce["block_origin"][:nb], ce["block_size"][:nb] / 8
    tx, ty, tz = st["x"][:700], st["y"][:700], st["z"][:700]
    ref = oracle.eval_plummer(ce["x"][sl], ce["y"][sl], ce["z"][sl], ce["m"][sl], ce["eps"][sl], tx, ty, tz,
                              None, G, mode=oracle.NEUMAIER, threads=16)
    with GasGravityField(mode) as g:
        g.set_column_waves(waves)
        for nsplit in (13, 7, 3, 1):            # chunks of 1, 2, 5 and 13 blocks
            g.set_config(nsplit=nsplit)
            g.set_blocks(o, dxb, ce["m"][sl])
            got = g.evaluate(None, tx, ty, tz)
            s = _util.summary(got, ref)
            print(mode, "block sources waves", waves, "nsplit", nsplit, s, flush=True)
            check_bar(mode, s)
            again = g.evaluate(None, tx, ty, tz)
            g.set_sources(ce["x"][sl], ce["y"][sl], ce["z"][sl], ce["m"][sl], ce["eps"][sl])
            detected = g.evaluate(None, tx, ty, tz)
            for p, q, r in zip(got, again, detected):
                assert np.array_equal(p, q)
                assert np.array_equal(p, r)


def test_column_waves_argument_errors():
    with GravityField("fp64") as f:
        for bad in (0, 3, 8, -1):
            with pytest.raises(TgravError):
                f.set_column_waves(bad)
        assert f.column_waves() == 4           # the default
