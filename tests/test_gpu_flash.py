"""GPU parity of the FLASH softening kernels and the periodic image sum (needs an MI355X).

Bar: FP64, each target within 1e-12 of the oracle relative to the sum of the magnitudes of
its terms, G sum_j m_j / max(r_ij, rsoft/4)^k (k = 2 acceleration, 1 potential) -- the
reference's quirk potential mixes signs, and periodic image sums cancel, so a plain
relative error would measure the cancellation, not the kernel.  The branch decisions at
r = rsoft and at the image cutoff are checked on inputs placed exactly on and one ulp
around the thresholds, where a differing decision changes a result by a whole term.
"""
import numpy as np
import pytest

import oracle
from torch_amd import GasGravityField, GravityField, TgravError, _lib
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu
G = W.G_PC_MSUN_KMS
TOL = 1e-12


def term_scales(sx, sy, sz, sm, tx, ty, tz, rsoft, box=None, nrep=6):
    """Per-target G sum m / max(r, rsoft/4)^2 and G sum m / max(r, rsoft/4) (over images)."""
    s = np.stack([sx, sy, sz], 1)
    t = np.stack([tx, ty, tz], 1)
    d0 = t[:, None, :] - s[None, :, :]
    offs = [np.zeros(3)]
    if box is not None:
        r = np.arange(-nrep, nrep + 1)
        offs = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3) * np.asarray(box)
    sa = np.zeros(len(t))
    sp = np.zeros(len(t))
    for o in offs:
        rr = np.maximum(np.sqrt(((d0 + o) ** 2).sum(-1)), rsoft / 4)
        if box is not None:
            keep = rr < nrep * min(box) * (1 + 1e-9)
            sa += (G * sm[None, :] / rr ** 2 * keep).sum(1)
            sp += (G * sm[None, :] / rr * keep).sum(1)
        else:
            sa += (G * sm[None, :] / rr ** 2).sum(1)
            sp += (G * sm[None, :] / rr).sum(1)
    return sa, sp


def check(got, ref, sa, sp, tol=TOL):
    da = np.sqrt(sum((np.asarray(got[c]) - ref[c]) ** 2 for c in range(3)))
    dp = np.abs(np.asarray(got[3]) - ref[3])
    assert np.all(da <= tol * sa), f"acc: worst {np.max(da / sa):.3e}"
    assert np.all(dp <= tol * sp), f"pot: worst {np.max(dp / sp):.3e}"


@pytest.fixture(scope="module")
def c1():
    return W.c1(1001)


@pytest.fixture()
def field():
    f = GasGravityField("fp64")
    yield f
    f.close()


@pytest.mark.parametrize("kernel", ["spline", "linear"])
@pytest.mark.parametrize("quirks", [False, True])
def test_flash_isolated_reverse_c1(c1, field, kernel, quirks):
    """4,096 gas cells -> 1,024 stars, rsoft = 0.3 pc (about 8 cells inside the radius of
    each star)."""
    ce, st = c1["cells"], c1["stars"]
    rs = 0.3
    field.set_softening(kernel, rs, quirks=quirks)
    field.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"])
    got = field.evaluate(st["eps"], st["x"], st["y"], st["z"])
    ref = oracle.eval_flash(ce["x"], ce["y"], ce["z"], ce["m"], st["x"], st["y"], st["z"], G, rs,
                            kernel=kernel, quirks=quirks, threads=16)
    sa, sp = term_scales(ce["x"], ce["y"], ce["z"], ce["m"], st["x"], st["y"], st["z"], rs)
    check(got, ref, sa, sp)


@pytest.mark.parametrize("kernel", ["spline", "linear"])
def test_flash_isolated_forward_blocks_equal_centres(c1, kernel):
    """Stars -> 8^3 leaf-block targets: block-structured targets give bit-identical results to
    their explicit centres, and match the oracle."""
    st = c1["stars"]
    rng = W.rng_for(41)
    nblk = 6
    ox, oy, oz = (rng.uniform(-1, 0.5, nblk) for _ in range(3))
    bdx = rng.uniform(0.05, 0.2, nblk)
    x, y, z = oracle.block_centres(ox, oy, oz, bdx)
    with GravityField("fp64") as f:
        f.set_softening(kernel, 0.05, quirks=True)
        f.set_sources(st["x"], st["y"], st["z"], st["m"], st["eps"])
        a = f.evaluate(None, x, y, z)
        b = f.evaluate_blocks(np.stack([ox, oy, oz], 1), bdx, nb=8)
    for c in range(4):
        np.testing.assert_array_equal(a[c], b[c])
    ref = oracle.eval_flash(st["x"], st["y"], st["z"], st["m"], x, y, z, G, 0.05, kernel=kernel,
                            quirks=True, threads=16)
    sa, sp = term_scales(st["x"], st["y"], st["z"], st["m"], x, y, z, 0.05)
    check(a, ref, sa, sp)


@pytest.mark.parametrize("kernel", ["spline", "linear"])
def test_flash_block_sources(field, kernel):
    """Cells handed over as 8^3 leaf blocks: the generic kernel on them gives the same sums as
    the same cells as points, bit for bit; the column kernel (the default for block sources)
    matches the oracle, including stars placed within the softening radius of cells; the cells
    as explicit arrays in the blocks' C order are recognised and give the column kernel's bits;
    shuffled, they take the generic kernel alone (a fused kick adds dt * a once)."""
    rng = W.rng_for(42)
    cells = W.gas_cells(rng, box=4.0, base_n=2, n_leaves=3, sigma=1.0, ngrid=16, rho_mean=10.0)
    st = W.star_cluster(rng, 300)
    pick = rng.choice(len(cells["x"]), 40, replace=False)          # 40 stars next to cells
    for q, c in zip("xyz", ("x", "y", "z")):
        st[q][:40] = cells[c][pick] + rng.normal(0, 0.02, 40)
    st["x"][40] = cells["x"][pick[0]]                              # one exactly on a cell centre
    st["y"][40], st["z"][40] = cells["y"][pick[0]], cells["z"][pick[0]]
    rs = 0.2
    field.set_softening(kernel, rs, quirks=True)
    field.set_block_kernel(False)
    field.set_sources(cells["x"], cells["y"], cells["z"], cells["m"], cells["eps"])
    a = field.evaluate(None, st["x"], st["y"], st["z"])
    field.set_source_blocks(cells["block_origin"], cells["block_size"] / 8, cells["m"])
    b = field.evaluate(None, st["x"], st["y"], st["z"])
    for c in range(4):
        np.testing.assert_array_equal(a[c], b[c])
    field.set_block_kernel(True)
    col = field.evaluate(None, st["x"], st["y"], st["z"])
    ref = oracle.eval_flash(cells["x"], cells["y"], cells["z"], cells["m"], st["x"], st["y"], st["z"], G, rs,
                            kernel=kernel, quirks=True, threads=16)
    sa, sp = term_scales(cells["x"], cells["y"], cells["z"], cells["m"], st["x"], st["y"], st["z"], rs)
    check(col, ref, sa, sp)
    field.set_sources(cells["x"], cells["y"], cells["z"], cells["m"], cells["eps"])
    ex = field.evaluate(None, st["x"], st["y"], st["z"])
    for c in range(4):
        np.testing.assert_array_equal(ex[c], col[c])
    perm = rng.permutation(len(cells["x"]))
    assert len(perm) % 512 == 0
    field.set_sources(*(cells[k][perm] for k in ("x", "y", "z", "m", "eps")))
    sh = field.evaluate(None, st["x"], st["y"], st["z"])
    check(sh, ref, sa, sp)
    v0 = [rng.normal(0, 1, len(st["x"])) for _ in range(3)]
    v = [q.copy() for q in v0]
    kicked = field.evaluate_and_kick(None, st["x"], st["y"], st["z"], 0.01, *v)
    for c in range(3):
        np.testing.assert_array_equal(kicked[c], sh[c])
    for p, q in zip(v, oracle.kick(0.01, sh[0], sh[1], sh[2], *v0)):
        np.testing.assert_array_equal(p, q)


def test_flash_decisions_at_rsoft(field):
    """Points exactly at rsoft and one ulp inside/outside, along each axis, from one cell: the
    reference's quirk potential flips sign there, so a differing branch shows at once."""
    rs = 0.25
    pts = []
    for r in (np.nextafter(rs, 0), rs, np.nextafter(rs, 1)):
        for axis in range(3):
            for sgn in (1.0, -1.0):
                p = np.zeros(3)
                p[axis] = sgn * r
                pts.append(p)
    # and r2 = dx^2 + dy^2 landing within rounding of rsoft^2
    for k in range(40):
        a = rs * np.cos(0.03 * k + 0.1)
        b = np.sqrt(rs * rs - a * a)
        pts.append(np.array([a, b, 0.0]))
        pts.append(np.array([0.0, np.nextafter(a, 0), b]))
    t = np.array(pts)
    one = np.zeros(1)
    for kernel in ("spline", "linear"):
        field.set_softening(kernel, rs, quirks=True)
        field.set_sources(one, one, one, np.ones(1), None)
        got = field.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
        ref = oracle.eval_flash(one, one, one, np.ones(1), t[:, 0], t[:, 1], t[:, 2], G, rs,
                                kernel=kernel, quirks=True)
        np.testing.assert_allclose(got[3], ref[3], rtol=1e-13)
        for c in range(3):
            np.testing.assert_allclose(got[c], ref[c], rtol=1e-13, atol=1e-13 * G / rs ** 2)


@pytest.mark.parametrize("per_target", [False, True], ids=["pairs", "per_target"])
@pytest.mark.parametrize("kernel", ["spline", "linear"])
@pytest.mark.parametrize("quirks", [False, True])
def test_flash_periodic_matches_oracle(field, kernel, quirks, per_target):
    """Periodic image sum with the reference's nrep = 6 in an anisotropic box; some points lie
    outside the box.  Both image walks: the pairs grouped by separation (few pairs here, so a
    wave's separations spread widely and most images take the reference's test) and the
    per-target walk."""
    rng = W.rng_for(43)
    box = (1.0, 1.3, 0.9)
    ns, nt = 80, 70
    s = rng.uniform(0, 1, (ns, 3)) * box
    t = rng.uniform(-0.2, 1.2, (nt, 3)) * box
    t[:5] = s[:5] + rng.normal(0, 0.03, (5, 3))          # pairs inside the softening radius
    m = rng.uniform(0.1, 3.0, ns)
    rs = 0.15
    field.set_softening(kernel, rs, quirks=quirks)
    field.set_periodic(box, nrep=6, per_target=per_target)
    field.set_sources(s[:, 0], s[:, 1], s[:, 2], m, None)
    got = field.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
    ref = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rs,
                            kernel=kernel, quirks=quirks, box=box, nrep=6, threads=16)
    sa, sp = term_scales(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], rs, box=box, nrep=6)
    check(got, ref, sa, sp)


@pytest.mark.parametrize("per_target", [False, True], ids=["pairs", "per_target"])
def test_flash_periodic_decisions_at_cutoff(field, per_target):
    """Images exactly at and one ulp around the cutoff nrep * min(L) are kept or dropped as
    the reference drops them (radius < maxradius_pbc), in both image walks."""
    box, nrep, rs = (1.0, 1.0, 1.0), 2, 0.1
    xs = [0.0, 2.0 ** -50, -(2.0 ** -50), 2.0 ** -40, 0.5, np.nextafter(0.5, 1), 1e-17]
    t = np.array([[x, 0.0, 0.0] for x in xs] + [[0.0, x, 0.0] for x in xs] +
                 [[0.3, 0.4, 0.0], [0.6, 0.8, 0.0], [0.0, 0.6, 0.8]])
    one = np.zeros(1)
    field.set_softening("spline", rs, quirks=True)
    field.set_periodic(box, nrep=nrep, per_target=per_target)
    field.set_sources(one, one, one, np.ones(1), None)
    got = field.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
    ref = oracle.eval_flash(one, one, one, np.ones(1), t[:, 0], t[:, 1], t[:, 2], G, rs, quirks=True,
                            box=box, nrep=nrep)
    scale = G * 1.0 / rs ** 2 * 100
    for c in range(3):
        np.testing.assert_allclose(got[c], ref[c], rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(got[3], ref[3], rtol=1e-12, atol=1e-12 * G / rs)


@pytest.mark.parametrize("nrep,quirks", [(7, False), (8, False), (9, True), (10, False), (11, False)])
def test_flash_periodic_nrep_beyond_the_row_table(field, nrep, quirks):
    """nrep = 7 fills the grouped walk's static row table (15 x 15 rows); nrep 8 .. 10 size the table
    in LDS at launch ((2 nrep + 1)^2 rows, up to 21 KB); nrep = 11 takes the per-target walk by
    itself: all against the oracle."""
    rng = W.rng_for(47)
    box = (1.0, 1.1, 1.2)
    s = rng.uniform(0, 1, (30, 3)) * box
    t = rng.uniform(0, 1, (40, 3)) * box
    m = rng.uniform(0.5, 2.0, 30)
    rs = 0.12
    field.set_softening("spline", rs, quirks=quirks)
    field.set_periodic(box, nrep=nrep)
    field.set_sources(s[:, 0], s[:, 1], s[:, 2], m, None)
    got = field.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
    ref = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rs,
                            kernel="spline", quirks=quirks, box=box, nrep=nrep, threads=16)
    sa, sp = term_scales(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], rs, box=box, nrep=nrep)
    check(got, ref, sa, sp)


def test_flash_periodic_c1_scale(c1):
    """c1periodic (4,096 cells -> 1,024 stars, 4 pc box, nrep = 6, spline with the reference's
    potential): 4.2 M pairs, dense enough in separation that most images of a wave's rows are
    summed as certain.  Both image walks against the oracle at 1e-12 (relative to the terms'
    magnitudes) on 48 seeded stars."""
    from torch_amd import GasGravityField
    st, ce = c1["stars"], c1["cells"]
    box, rs = (4.0, 4.0, 4.0), 0.25
    out = {}
    for per_target in (False, True):
        with GasGravityField("fp64") as f:
            f.set_softening("spline", rs, quirks=True)
            f.set_periodic(box, nrep=6, per_target=per_target)
            f.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], None)
            out[per_target] = f.evaluate(None, st["x"], st["y"], st["z"])
    idx = np.sort(W.rng_for(48).choice(len(st["x"]), 48, replace=False))
    ref = oracle.eval_flash(ce["x"], ce["y"], ce["z"], ce["m"], st["x"][idx], st["y"][idx], st["z"][idx], G, rs,
                            kernel="spline", quirks=True, box=box, nrep=6, threads=16)
    sa, sp = term_scales(ce["x"], ce["y"], ce["z"], ce["m"], st["x"][idx], st["y"][idx], st["z"][idx], rs,
                         box=box, nrep=6)
    for per_target in (False, True):
        check(tuple(a[idx] for a in out[per_target]), ref, sa, sp)


def test_flash_periodic_pairs_in_target_batches(c1):
    """9,000 stars x 4,096 cells = 36.9 M pairs, more than one batch of the grouped walk (2^25
    pairs, 8,192 stars here): stars on both sides of the batch boundary and at the ends against
    the oracle, and both image walks against each other there."""
    from torch_amd import GasGravityField
    ce = c1["cells"]
    st = W.star_cluster(W.rng_for(51), 9000, r_vir=1.0, eps=0.01)
    box, rs = (4.0, 4.0, 4.0), 0.25
    out = {}
    for per_target in (False, True):
        with GasGravityField("fp64") as f:
            f.set_softening("linear", rs, quirks=False)
            f.set_periodic(box, nrep=6, per_target=per_target)
            f.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], None)
            out[per_target] = f.evaluate(None, st["x"], st["y"], st["z"])
    idx = np.array([0, 1, 2, 8189, 8190, 8191, 8192, 8193, 8194, 8997, 8998, 8999])
    ref = oracle.eval_flash(ce["x"], ce["y"], ce["z"], ce["m"], st["x"][idx], st["y"][idx], st["z"][idx], G, rs,
                            kernel="linear", quirks=False, box=box, nrep=6, threads=16)
    sa, sp = term_scales(ce["x"], ce["y"], ce["z"], ce["m"], st["x"][idx], st["y"][idx], st["z"][idx], rs,
                         box=box, nrep=6)
    for per_target in (False, True):
        check(tuple(a[idx] for a in out[per_target]), ref, sa, sp)


def test_flash_periodic_block_targets_equal_centres():
    """Periodic sums onto 8^3 leaf-block targets (the grouped walk reads their centres from the
    blocks): bit-identical to the explicit centres, and against the oracle on a subsample."""
    rng = W.rng_for(49)
    box = (2.0, 2.0, 2.0)
    s = rng.uniform(0, 2, (60, 3))
    m = rng.uniform(0.5, 2.0, 60)
    nblk = 2
    ox, oy, oz = (rng.uniform(0.0, 1.6, nblk) for _ in range(3))
    bdx = rng.uniform(0.02, 0.05, nblk)
    x, y, z = oracle.block_centres(ox, oy, oz, bdx)
    with GravityField("fp64") as f:
        f.set_softening("spline", 0.1, quirks=True)
        f.set_periodic(box, nrep=3)
        f.set_sources(s[:, 0], s[:, 1], s[:, 2], m, None)
        a = f.evaluate(None, x, y, z)
        b = f.evaluate_blocks(np.stack([ox, oy, oz], 1), bdx, nb=8)
    for c in range(4):
        np.testing.assert_array_equal(a[c], b[c])
    idx = np.sort(W.rng_for(50).choice(len(x), 64, replace=False))
    ref = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, x[idx], y[idx], z[idx], G, 0.1, kernel="spline",
                            quirks=True, box=box, nrep=3, threads=16)
    sa, sp = term_scales(s[:, 0], s[:, 1], s[:, 2], m, x[idx], y[idx], z[idx], 0.1, box=box, nrep=3)
    check(tuple(q[idx] for q in a), ref, sa, sp)


def test_flash_periodic_chunked_deterministic(field):
    """Many source chunks (few targets): repeated runs are bit-identical and the kick is
    v + dt * a rounded as the reference rounds it."""
    rng = W.rng_for(44)
    box = (2.0, 2.0, 2.0)
    s = rng.uniform(0, 2, (700, 3))
    t = rng.uniform(0, 2, (40, 3))
    m = rng.uniform(0.5, 1.5, 700)
    field.set_softening("linear", 0.1)
    field.set_periodic(box, nrep=3)
    field.set_sources(s[:, 0], s[:, 1], s[:, 2], m, None)
    a = field.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
    b = field.evaluate(None, t[:, 0], t[:, 1], t[:, 2])
    for c in range(4):
        np.testing.assert_array_equal(a[c], b[c])
    v = [rng.normal(0, 1, 40) for _ in range(3)]
    vk = [q.copy() for q in v]
    out = field.evaluate_and_kick(None, t[:, 0], t[:, 1], t[:, 2], 0.01, *vk)
    exp = oracle.kick(0.01, out[0], out[1], out[2], *v)
    for c in range(3):
        np.testing.assert_array_equal(vk[c], exp[c])
        np.testing.assert_array_equal(out[c], a[c])


def test_flash_errors_and_plummer_restored(c1):
    st, ce = c1["stars"], c1["cells"]
    with GravityField("fp64") as f:
        f.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"])
        base = f.evaluate(st["eps"], st["x"], st["y"], st["z"])
        for bad in [(5, 0.1, 0), (1, 0.0, 0), (1, -1.0, 0), (1, float("nan"), 0), (1, 0.1, 2)]:
            assert f.lib.tgrav_set_softening(f._ctx, bad[0], bad[1], bad[2]) == _lib.EINVAL
        for bad in [(1, 0.0, 1.0, 1.0, 6), (1, 1.0, 1.0, 1.0, 0), (1, 1.0, 1.0, 1.0, 65),
                    (1, 1.0, float("inf"), 1.0, 6)]:
            assert f.lib.tgrav_set_periodic(f._ctx, *bad) == _lib.EINVAL
        f.set_periodic((4.0, 4.0, 4.0))                      # Plummer + periodic: rejected
        with pytest.raises(TgravError):
            f.evaluate(st["eps"], st["x"], st["y"], st["z"])
        f.set_periodic(None)
        f.set_softening("spline", 0.2)
        with pytest.raises(TgravError):
            f.bidirectional(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"],
                            st["x"], st["y"], st["z"], st["m"], st["eps"])
        f.set_softening("plummer")
        again = f.evaluate(st["eps"], st["x"], st["y"], st["z"])
        for c in range(4):
            np.testing.assert_array_equal(base[c], again[c])
    with GravityField("mixed") as f:
        f.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"])
        f.set_softening("spline", 0.2)
        with pytest.raises(TgravError):
            f.evaluate(st["eps"], st["x"], st["y"], st["z"])
