"""Fused bidirectional pass (needs an MI355X): both fields from one pass over the pairs,
checked against the oracle's restatement of Particles_sinkAccelGasOnSinksAndSinksOnGas,
against two directed evaluations, and against Newton's third law."""
import numpy as np
import pytest

import oracle
from tests import _util
from torch_amd import GravityField, _lib
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu
G = W.G_PC_MSUN_KMS


def cells_and_particles(seed, nblk=3, npart=300):
    rng = W.rng_for(seed)
    o = rng.uniform(-1.5, 1.0, (nblk, 3))
    size = rng.choice([0.125, 0.25, 0.5], nblk)
    x, y, z, dx = W.block_cell_centres(o, size)
    rho = 50.0 * np.exp(rng.normal(size=x.size))
    C = dict(x=x, y=y, z=z, m=rho * dx ** 3, eps=dx)
    st = W.star_cluster(rng, npart, binary_frac=0.2)
    P = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=st["eps"])
    return C, P


def run(f, C, P):
    return f.bidirectional(C["x"], C["y"], C["z"], C["m"], C["eps"], P["x"], P["y"], P["z"], P["m"], P["eps"])


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("nblk,npart", [(1, 1), (1, 300), (3, 257), (5, 1000)])
def test_matches_oracle(mode, nblk, npart):
    """Explicit cells (in the blocks' C order) with the block kernels on -- recognised, the column
    kernel -- and off -- the generic fused kernel -- against the oracle."""
    C, P = cells_and_particles(nblk * 100 + npart, nblk, npart)
    (pa, pb, pc, pp), (ca, cb, cc, cp) = oracle.bidirectional(
        C["x"], C["y"], C["z"], C["m"], C["eps"], P["x"], P["y"], P["z"], P["m"], P["eps"], G)
    for block_kernel in (True, False):
        with GravityField(mode) as f:
            f.set_block_kernel(block_kernel)
            co, po = run(f, C, P)
        for name, got, ref in (("cells", co, (ca, cb, cc, cp)), ("particles", po, (pa, pb, pc, pp))):
            s = _util.summary(got, ref)
            print(mode, nblk, npart, block_kernel, name, s, flush=True)
            if mode == "fp64":
                assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (name, s)
            else:
                assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (name, s)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_every_variant_and_chunking_agree(mode):
    C, P = cells_and_particles(7, 6, 2000)
    lib = _lib.load()
    ref = None
    with GravityField(mode) as f:
        f.set_block_kernel(False)          # the generic fused kernel's variants
        for v in range(4):
            for ns in (1, 0, 5):
                f.set_bidir_config(variant=v, nsplit=ns)
                co, po = run(f, C, P)
                if ref is None:
                    ref = (co, po)
                    continue
                tol = 1e-12 if mode == "fp64" else 1e-5
                for got, r in ((co, ref[0]), (po, ref[1])):
                    s = _util.summary(got, r)
                    assert s["acc_max"] <= tol and s["pot_max"] <= tol, (v, ns, s)
        assert lib.tgrav_bidir_variant_info(f.mode, 4, None, None) == _lib.EINVAL


def blocks_and_particles(seed, nblk, npart):
    rng = W.rng_for(seed)
    o = rng.uniform(-1.5, 1.0, (nblk, 3))
    size = rng.choice([0.125, 0.25, 0.5], nblk)
    x, y, z, dx = W.block_cell_centres(o, size)
    m = 50.0 * np.exp(rng.normal(size=x.size)) * dx ** 3
    st = W.star_cluster(rng, npart, binary_frac=0.2)
    pick = rng.choice(x.size, min(20, npart), replace=False)     # some particles inside cells
    for q, c in zip(("x", "y", "z"), (x, y, z)):
        st[q][:len(pick)] = c[pick] + rng.normal(0, 0.01, len(pick))
    B = dict(o=o, dx=size / 8, m=m, eps=size / 8)
    C = dict(x=x, y=y, z=z, m=m, eps=dx)
    P = dict(x=st["x"], y=st["y"], z=st["z"], m=st["m"], eps=st["eps"])
    return B, C, P


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("nblk,npart", [(1, 1), (1, 65), (3, 300), (7, 1000)])
def test_blocks_match_explicit_and_oracle(mode, nblk, npart):
    """Cells as 8^3 blocks (column kernel): the explicit-cell fused pass and the oracle.  The
    explicit cells are in the blocks' C order, so they are recognised and take the same column
    kernel (the same bits); shuffled, they take the generic fused kernel."""
    B, C, P = blocks_and_particles(nblk * 1000 + npart, nblk, npart)
    with GravityField(mode) as f:
        co, po = f.bidirectional_blocks(B["o"], B["dx"], B["m"], B["eps"], P["x"], P["y"], P["z"], P["m"],
                                        P["eps"])
        ce, pe = run(f, C, P)
        perm = W.rng_for(nblk).permutation(len(C["x"]))
        cs, ps = run(f, {k: v[perm] for k, v in C.items()}, P)
    for q in range(4):
        np.testing.assert_array_equal(ce[q], co[q])
        np.testing.assert_array_equal(pe[q], po[q])
    inv = np.argsort(perm)
    for got, ref in ((tuple(a[inv] for a in cs), co), (ps, po)):
        s = _util.summary(got, ref)
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, s
    (pa, pb, pc, pp), (ca, cb, cc, cp) = oracle.bidirectional(
        C["x"], C["y"], C["z"], C["m"], C["eps"], P["x"], P["y"], P["z"], P["m"], P["eps"], G)
    for name, got, ref, alt in (("cells", co, (ca, cb, cc, cp), ce), ("particles", po, (pa, pb, pc, pp), pe)):
        s = _util.summary(got, ref)
        print(mode, nblk, npart, name, s, flush=True)
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (name, s)
            s2 = _util.summary(got, alt)
            assert s2["acc_max"] <= 1e-12 and s2["pot_max"] <= 1e-12, (name, s2)
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (name, s)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_shared_and_per_lane_dz_agree(mode):
    """Explicit cells in the blocks' C order, whose 64 columns per block hold the same 8 z, take the
    column kernel with each staged particle's 8 dz formed once; listing the 8 cells of one column per
    block in reverse z order keeps the column structure and every block's bounding box (the mixed
    anchor) but not the common z, so those blocks form dz per lane.  Each cell's result is the same
    bits at its new place; the particles' sums (added over a column's cells in the other order)
    agree within the bar, and both meet it against the oracle."""
    B, C, P = blocks_and_particles(4242, 4, 500)
    perm = np.arange(len(C["x"])).reshape(-1, 64, 8)
    perm[:, 21, :] = perm[:, 21, ::-1]
    perm = perm.ravel()
    R = {k: v[perm] for k, v in C.items()}
    with GravityField(mode) as f:
        ca, pa = run(f, C, P)
        cb, pb = run(f, R, P)
    for q in range(4):
        np.testing.assert_array_equal(cb[q], ca[q][perm])
    (qa, qb, qc, qp), _ = oracle.bidirectional(
        C["x"], C["y"], C["z"], C["m"], C["eps"], P["x"], P["y"], P["z"], P["m"], P["eps"], G)
    for name, got, ref in (("shared vs per-lane", pb, pa), ("per-lane vs oracle", pb, (qa, qb, qc, qp))):
        s = _util.summary(got, ref)
        print(mode, "particles", name, s, flush=True)
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (name, s)
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (name, s)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_block_order_with_varied_softening_in_a_block(mode):
    """Cells in the blocks' C order whose softening is not uniform within one 512-cell block fail the
    structure test (the column kernel reads one eps per block) and must take the generic fused
    kernel: both fields still match the oracle with the varied eps, and differ from the uniform run."""
    B, C, P = blocks_and_particles(4242, 3, 300)
    Cv = dict(C)
    Cv["eps"] = C["eps"].copy()
    Cv["eps"][512 + 37] *= 1.5
    with GravityField(mode) as f:
        cu, pu = run(f, C, P)
        cv, pv = run(f, Cv, P)
    (pa, pb, pc, pp), (ca, cb, cc, cp) = oracle.bidirectional(
        Cv["x"], Cv["y"], Cv["z"], Cv["m"], Cv["eps"], P["x"], P["y"], P["z"], P["m"], P["eps"], G)
    for name, got, ref in (("cells", cv, (ca, cb, cc, cp)), ("particles", pv, (pa, pb, pc, pp))):
        s = _util.summary(got, ref)
        print(mode, name, s, flush=True)
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, (name, s)
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4, (name, s)
    assert not np.array_equal(cv[0], cu[0]) and not np.array_equal(pv[0], pu[0])


def test_blocks_chunking_deterministic_and_errors():
    B, C, P = blocks_and_particles(9, 4, 3000)
    lib = _lib.load()
    with GravityField("fp64") as f:
        outs = []
        for ns in (1, 0, 7, 0):
            f.set_bidir_config(nsplit=ns)
            outs.append(f.bidirectional_blocks(B["o"], B["dx"], B["m"], B["eps"], P["x"], P["y"], P["z"],
                                               P["m"], P["eps"]))
        for side in range(2):
            for q in range(4):
                np.testing.assert_array_equal(outs[1][side][q], outs[3][side][q])   # same shape: bitwise
                for k in (0, 2):
                    s = _util.summary(outs[k][side], outs[1][side])
                    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
        assert lib.tgrav_bidirectional_blocks(f.handle, 1, 4, *([None] * 6), 0, *([None] * 13),
                                              f.mode, 0) == _lib.EINVAL


def test_c2_fused_equals_directed_and_third_law():
    w = W.c2(1002)
    st, ce = w["stars"], w["cells"]
    with GravityField("fp64") as f:
        co, po = f.bidirectional(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"],
                                 st["x"], st["y"], st["z"], st["m"], st["eps"])
        f.set_sources(st["x"], st["y"], st["z"], st["m"], st["eps"])
        dc = f.evaluate(ce["eps"], ce["x"], ce["y"], ce["z"])
        f.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"])
        dp = f.evaluate(st["eps"], st["x"], st["y"], st["z"])
    for name, got, ref in (("cells", co, dc), ("stars", po, dp)):
        s = _util.summary(got, ref)
        print("c2 fused vs directed", name, s, flush=True)
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    for q in range(3):
        fc, fp = np.sum(ce["m"] * co[q]), np.sum(st["m"] * po[q])
        scale = np.sum(np.abs(ce["m"] * co[q])) + np.sum(np.abs(st["m"] * po[q]))
        assert abs(fc + fp) <= 1e-12 * scale


def test_c2_fused_mixed_error():
    w = W.c2(1002)
    st, ce = w["stars"], w["cells"]
    args = (ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"], st["x"], st["y"], st["z"], st["m"], st["eps"])
    with GravityField("fp64") as f:
        ref = f.bidirectional(*args)
    with GravityField("mixed") as f:
        got = f.bidirectional(*args)
    for k, name in ((0, "cells"), (1, "stars")):
        s = _util.summary(got[k], ref[k])
        print("c2 fused mixed", name, s, flush=True)
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4


def test_empty_sides():
    C, P = cells_and_particles(9, 1, 10)
    with GravityField("fp64") as f:
        co, po = f.bidirectional(C["x"], C["y"], C["z"], C["m"], C["eps"], [], [], [], [], None)
        assert all(np.all(a == 0) for a in co) and all(a.size == 0 for a in po)
        co, po = f.bidirectional([], [], [], [], None, P["x"], P["y"], P["z"], P["m"], P["eps"])
        assert all(np.all(a == 0) for a in po)
