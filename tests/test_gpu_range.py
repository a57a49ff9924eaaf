"""FP64 parity in cgs, where r^2 leaves the FP32 range (needs an MI355X).

FLASH works in cgs (the hydro side's converter, src/torch_mainloop.py:126): positions of
1e19-1e20 cm, so r^2 reaches 1e41, beyond FLT_MAX = 3.4e38.  The FP64 kernels seed 1/sqrt with
v_rsq_f64 (built with TG_FP64_SEED32=1, in FP32 for the tiles whose bound on r^2, from the source
statistics and the tile's targets, stays below 1e36); these inputs span both sides of that bound
(+-30 pc boxes, and a compact 0.1 pc cluster),
through every FP64 kernel family: explicit points, 8^3 block targets (column kernel), block
sources, explicit points in block order, the batch plan, and FLASH's kernels isolated and
periodic.  Bar: 1e-12 of each target's sum of |terms| against the oracle in cgs.  The mixed
mode refuses such inputs with TGRAV_ELIMIT when validation is on.
"""
import numpy as np
import pytest

import oracle
from torch_amd import GasGravityField, GravityField, StellarGravityField, TgravError, _lib
from torch_amd import units as U

pytestmark = pytest.mark.gpu
G = U.G_CGS
PC = 3.0856775814913673e18
MSUN = 1.32712440018e26 / 6.6743e-8      # g, so that G M_sun is the IAU value in cgs
TOL = 1e-12


def term_scales(s, m, se, t, te=None):
    d = s[None, :, :] - t[:, None, :]
    r2 = (d ** 2).sum(-1) + se[None, :] ** 2 + (0.0 if te is None else te[:, None] ** 2)
    r = np.sqrt(np.where(r2 > 0, r2, 1.0))
    gm = G * np.abs(m)[None, :]
    return (gm * np.sqrt((d ** 2).sum(-1)) / r ** 3).sum(1), (gm / r).sum(1)


def check(got, ref, sa, sp, tol=TOL):
    got = [np.asarray(g.cpu().numpy() if hasattr(g, "cpu") else g) for g in got]
    da = np.sqrt(sum((got[c] - ref[c]) ** 2 for c in range(3)))
    dp = np.abs(got[3] - ref[3])
    assert np.all(da <= tol * sa), f"acc: worst {np.max(da / sa):.3e}"
    assert np.all(dp <= tol * sp), f"pot: worst {np.max(dp / sp):.3e}"


def stars(seed, n, half_pc):
    rng = np.random.default_rng(seed)
    s = rng.uniform(-half_pc, half_pc, size=(n, 3)) * PC
    m = rng.uniform(0.1, 50.0, n) * MSUN
    e = np.full(n, 0.01 * PC)
    return s, m, e


def blocks(seed, nblk, half_pc):
    """Leaf blocks inside [-half_pc, half_pc]^3 pc (cell widths up to half_pc / 10)."""
    rng = np.random.default_rng(seed)
    dx = rng.choice([0.025, 0.05, 0.1], nblk) * half_pc * PC
    o = rng.uniform(-half_pc, 0.2 * half_pc, size=(nblk, 3)) * PC
    return o, dx


@pytest.mark.parametrize("half_pc", [30.0, 0.05])
def test_plummer_fp64_kernels_in_cgs(half_pc):
    print(f"cgs plummer, half box {half_pc} pc", flush=True)
    s, m, e = stars(11, 700, half_pc)
    o, dx = blocks(12, 3, half_pc)
    cx, cy, cz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    t = np.stack([cx, cy, cz], 1)
    ref = oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, e, cx, cy, cz, None, G, mode=oracle.NEUMAIER)
    sa, sp = term_scales(s, m, e, t)
    with StellarGravityField("fp64", G=G) as f:
        f.set_stars(s[:, 0], s[:, 1], s[:, 2], m, e)
        check(f.evaluate(None, cx, cy, cz), ref, sa, sp)                 # explicit, block C order
        check(f.evaluate_blocks(o, dx), ref, sa, sp)                     # column kernel
        f.set_block_kernel(False)
        check(f.evaluate(None, cx, cy, cz), ref, sa, sp)                 # generic kernel
        perm = np.random.default_rng(3).permutation(len(cx))
        got = f.evaluate(None, cx[perm], cy[perm], cz[perm])
        check(got, tuple(r[perm] for r in ref), sa[perm], sp[perm])
    # reverse: the cells as sources (m = rho dV, eps = dx) onto the stars, as blocks and explicit
    rng = np.random.default_rng(13)
    cm = rng.uniform(1e-23, 1e-21, len(cx)) * np.repeat(dx, 512) ** 3
    ce = np.repeat(dx, 512)
    ref_r = oracle.eval_plummer(cx, cy, cz, cm, ce, s[:, 0], s[:, 1], s[:, 2], None, G, mode=oracle.NEUMAIER)
    sa_r, sp_r = term_scales(t, cm, ce, s)
    with GasGravityField("fp64", G=G) as g:
        g.set_blocks(o, dx, cm, eps=dx)
        check(g.evaluate(None, s[:, 0], s[:, 1], s[:, 2]), ref_r, sa_r, sp_r)
        g.set_sources(cx, cy, cz, cm, ce)                                  # detected as blocks
        check(g.evaluate(None, s[:, 0], s[:, 1], s[:, 2]), ref_r, sa_r, sp_r)


def test_batch_plan_fp64_in_cgs():
    import torch
    from torch_amd.batch import BatchPlan
    probs, refs = [], []
    for q, half in enumerate((30.0, 0.05, 10.0)):
        s, m, e = stars(20 + q, 300 + 100 * q, half)
        o, dx = blocks(30 + q, 2, half)
        cx, cy, cz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
        tgt = ({"block_origin": o, "block_dx": dx} if q != 1 else {"x": cx, "y": cy, "z": cz})
        probs.append({"sources": {"x": s[:, 0], "y": s[:, 1], "z": s[:, 2], "m": m, "eps": e}, "targets": tgt})
        ref = oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, e, cx, cy, cz, None, G, mode=oracle.NEUMAIER)
        refs.append((ref, term_scales(s, m, e, np.stack([cx, cy, cz], 1))))
    f = GravityField("fp64", G=G, validate=False)
    plan = BatchPlan(f, probs)
    for graph in (False, True, True):
        out = plan.run(graph=graph)
        torch.cuda.synchronize()
        for o_, (ref, (sa, sp)) in zip(out, refs):
            check(o_, ref, sa, sp)
    plan.close()
    f.close()


@pytest.mark.parametrize("kernel", ["spline", "linear"])
def test_flash_isolated_and_periodic_in_cgs(kernel):
    print(f"cgs flash {kernel}", flush=True)
    rng = np.random.default_rng(41)
    o, dx = blocks(42, 2, 20.0)
    cx, cy, cz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    cm = rng.uniform(1e-23, 1e-21, len(cx)) * np.repeat(dx, 512) ** 3
    st = rng.uniform(-20.0, 20.0, size=(200, 3)) * PC
    rs = 0.5 * PC
    with GasGravityField("fp64", G=G) as g:
        g.set_softening(kernel, rs, quirks=True)
        ref = oracle.eval_flash(cx, cy, cz, cm, st[:, 0], st[:, 1], st[:, 2], G, rs, kernel=kernel, quirks=True)
        sa, sp = term_scales(np.stack([cx, cy, cz], 1), cm, np.full(len(cm), rs / 4), st)
        g.set_blocks(o, dx, cm)
        check(g.evaluate(None, st[:, 0], st[:, 1], st[:, 2]), ref, sa, sp)
        g.set_sources(cx, cy, cz, cm)
        check(g.evaluate(None, st[:, 0], st[:, 1], st[:, 2]), ref, sa, sp)
        # periodic: a 40 pc box, images within 2 boxes (nrep = 2), a handful of stars
        box = (40.0 * PC,) * 3
        g.set_periodic(box, nrep=2)
        few = st[:24]
        ref_p = oracle.eval_flash(cx, cy, cz, cm, few[:, 0], few[:, 1], few[:, 2], G, rs, kernel=kernel,
                                  quirks=True, box=box, nrep=2)
        got = g.evaluate(None, few[:, 0], few[:, 1], few[:, 2])
        # term scale over all 125 images (an upper bound of the images inside the cutoff)
        r = np.arange(-2, 3)
        offs = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * box[0]
        s3 = np.stack([cx, cy, cz], 1)
        sa_p = np.zeros(len(few))
        sp_p = np.zeros(len(few))
        for off in offs:
            a_, p_ = term_scales(s3 - off, cm, np.full(len(cm), rs / 4), few)
            sa_p += a_
            sp_p += p_
        check(got, ref_p, sa_p, sp_p)


def test_mixed_refuses_inputs_beyond_its_range():
    s, m, e = stars(51, 200, 30.0)
    t = np.random.default_rng(52).uniform(-30, 30, size=(100, 3)) * PC
    with StellarGravityField("mixed", G=G) as f:
        f.set_stars(s[:, 0], s[:, 1], s[:, 2], m, e)
        with pytest.raises(TgravError) as ei:
            f.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
        assert ei.value.code == _lib.ELIMIT
        assert "FP32 range" in str(ei.value)
    # the same configuration in pc, Msun is within range and meets the mixed gate
    with StellarGravityField("mixed") as f:
        f.set_stars(s[:, 0] / PC, s[:, 1] / PC, s[:, 2] / PC, m / MSUN, e / PC)
        got = f.evaluate(None, t[:, 0] / PC, t[:, 1] / PC, t[:, 2] / PC)
    ref = oracle.eval_plummer(s[:, 0] / PC, s[:, 1] / PC, s[:, 2] / PC, m / MSUN, e / PC, t[:, 0] / PC,
                              t[:, 1] / PC, t[:, 2] / PC, None, U.G_PC_MSUN_KMS, mode=oracle.NEUMAIER)
    ga, ra = np.stack(got[:3], 1), np.stack(ref[:3], 1)
    err = np.linalg.norm(ga - ra, axis=1) / np.linalg.norm(ra, axis=1)
    assert np.median(err) <= 1e-6 and err.max() <= 1e-4
