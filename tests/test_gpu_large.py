"""The largest configuration, c5 (100,000 stars <-> 4,194,304 cells in 8,192 leaf blocks), on
the GPU (needs an MI355X): FP64 within 1e-12 of the oracle on seeded subsamples of both
directions, mixed precision within the 1e-6 median / 1e-4 max gate against the oracle on seeded
subsamples (binary members included) and against FP64 over every target, the graph-replayed plan
the bench times against the oracle, and the explicit-array form (block C order, detected on the
device) bit-identical to the block form."""
import numpy as np
import pytest

import oracle
from tests import _util
from torch_amd import GasGravityField, StellarGravityField
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu
G = W.G_PC_MSUN_KMS


@pytest.fixture(scope="module")
def c5():
    return W.c5(1005)


@pytest.fixture(scope="module")
def fields(c5):
    import torch
    st, ce = c5["stars"], c5["cells"]
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    S = [d(st[k]) for k in ("x", "y", "z", "m", "eps")]
    BO, BD, CM = d(ce["block_origin"]), d(ce["block_size"] / 8), d(ce["m"])
    out = {}
    for mode in ("fp64", "mixed"):
        with StellarGravityField(mode, validate=False) as f, GasGravityField(mode, validate=False) as g:
            f.set_stars(*S)
            fwd = f.evaluate_blocks(BO, BD)
            g.set_blocks(BO, BD, CM)
            rev = g.evaluate(None, S[0], S[1], S[2])
            torch.cuda.synchronize()
            out[mode] = (tuple(a.cpu().numpy() for a in fwd), tuple(a.cpu().numpy() for a in rev))
    return out


def test_c5_fp64_matches_oracle(c5, fields):
    st, ce = c5["stars"], c5["cells"]
    fwd, rev = fields["fp64"]
    it = _util.subsample(len(ce["x"]), 64, 51)
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][it], ce["y"][it],
                              ce["z"][it], None, G, mode=oracle.NEUMAIER, threads=16)
    s = _util.summary(tuple(a[it] for a in fwd), ref)
    print("c5 fwd fp64 vs oracle", s, flush=True)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    js = _util.subsample(len(st["x"]), 32, 52)
    ref = oracle.eval_plummer(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"], st["x"][js], st["y"][js],
                              st["z"][js], None, G, mode=oracle.NEUMAIER, threads=16)
    s = _util.summary(tuple(a[js] for a in rev), ref)
    print("c5 rev fp64 vs oracle", s, flush=True)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12


def test_c5_mixed_matches_oracle(c5, fields):
    """Mixed precision against the oracle itself: 256 seeded cells (forward), 64 seeded stars and
    128 seeded binary members (reverse), within the 1e-6 median / 1e-4 max gate."""
    st, ce = c5["stars"], c5["cells"]
    fwd, rev = fields["mixed"]
    it = _util.subsample(len(ce["x"]), 256, 53)
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][it], ce["y"][it],
                              ce["z"][it], None, G, mode=oracle.NEUMAIER, threads=16)
    s = _util.summary(tuple(a[it] for a in fwd), ref)
    print("c5 fwd mixed vs oracle", s, flush=True)
    assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
    members = np.flatnonzero(st["binary"] >= 0)
    js = np.union1d(_util.subsample(len(st["x"]), 64, 54), members[_util.subsample(len(members), 128, 55)])
    ref = oracle.eval_plummer(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"], st["x"][js], st["y"][js],
                              st["z"][js], None, G, mode=oracle.NEUMAIER, threads=16)
    s = _util.summary(tuple(a[js] for a in rev), ref)
    print("c5 rev mixed vs oracle", s, flush=True)
    assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4


def test_c5_plan_matches_oracle(c5):
    """C5 as the bench runs it: one graph-replayed plan (stars -> 8^3 blocks, block sources -> stars),
    tuned shape, both modes, against the oracle on seeded subsamples of both directions."""
    import torch
    from torch_amd import GravityField, autotune
    from torch_amd.batch import BatchPlan
    from torch_amd.profile_driver import plan_problems
    st, ce = c5["stars"], c5["cells"]
    probs = plan_problems("c5")
    it = _util.subsample(len(ce["x"]), 64, 56)
    js = _util.subsample(len(st["x"]), 32, 57)
    ref_f = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][it], ce["y"][it],
                                ce["z"][it], None, G, mode=oracle.NEUMAIER, threads=16)
    ref_r = oracle.eval_plummer(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"], st["x"][js], st["y"][js],
                                st["z"][js], None, G, mode=oracle.NEUMAIER, threads=16)
    for mode in ("fp64", "mixed"):
        with GravityField(mode, validate=False) as f:
            autotune.apply_batch(f, probs)
            plan = BatchPlan(f, probs)
            out = plan.run(graph=True)
            torch.cuda.synchronize()
            fwd = tuple(a.cpu().numpy() for a in out[0])
            rev = tuple(a.cpu().numpy() for a in out[1])
            plan.close()
        for name, got, ref in (("fwd", tuple(a[it] for a in fwd), ref_f), ("rev", tuple(a[js] for a in rev), ref_r)):
            s = _util.summary(got, ref)
            print("c5 plan", mode, name, s, flush=True)
            if mode == "fp64":
                assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (name, s)
            else:
                assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (name, s)


def test_c5_mixed_against_fp64(c5, fields):
    st = c5["stars"]
    for q, name in ((0, "fwd"), (1, "rev")):
        s = _util.summary(fields["mixed"][q], fields["fp64"][q])
        print("c5", name, "mixed vs fp64", s, flush=True)
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
    b = st["binary"] >= 0
    sb = _util.summary(tuple(a[b] for a in fields["mixed"][1]), tuple(a[b] for a in fields["fp64"][1]))
    print("c5 binary members", sb, flush=True)
    assert sb["acc_median"] <= 1e-6 and sb["acc_max"] <= 1e-4


def test_c5_explicit_arrays_equal_block_form(c5, fields):
    """The cells as explicit arrays (targets of the forward, sources of the reverse) take the
    same column kernels as the block form, so the results are the same bits."""
    import torch
    st, ce = c5["stars"], c5["cells"]
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    S = [d(st[k]) for k in ("x", "y", "z", "m", "eps")]
    C = [d(ce[k]) for k in ("x", "y", "z", "m", "eps")]
    with StellarGravityField("mixed", validate=False) as f, GasGravityField("mixed", validate=False) as g:
        f.set_stars(*S)
        fwd = f.evaluate(None, C[0], C[1], C[2])
        g.set_sources(*C)
        rev = g.evaluate(None, S[0], S[1], S[2])
        torch.cuda.synchronize()
        for p, q in zip(fwd, fields["mixed"][0]):
            np.testing.assert_array_equal(p.cpu().numpy(), q)
        for p, q in zip(rev, fields["mixed"][1]):
            np.testing.assert_array_equal(p.cpu().numpy(), q)
