"""Property-based GPU tests (hypothesis, derandomized): random shapes and structures against
the oracle, for every input format of the library (needs an MI355X).

Inputs live on a lattice of spacing 1/64 pc (a few lattice units apart up to a few pc), so
  * coincident source/target pairs happen (the zero-separation rule is exercised), and every
    other pair is at least 1/64 apart, which keeps the mixed kernels' first-order error bound
    valid (tests/_util.term_scale_bounds);
  * every displacement, cell centre and anchor offset is exact in FP64 and in FP32, so two
    exact properties hold bit for bit: translating everything by a lattice vector leaves the
    outputs unchanged, and scaling every mass by 4 scales every output by 4.
Bars: FP64 within 1e-12 of the sum of |terms| per target (condition-aware: random clouds
can cancel), mixed within its analytic bound; the exact properties bit for bit.  Hypothesis
draws the sizes and the structure (spread, softening pattern, mass pattern, coincidences,
host or device arrays); numpy, seeded from a drawn integer, draws the values.  25 examples per
property by default; TGRAV_PROPERTY_EXAMPLES sets another count for a longer sweep.
"""
import os

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import oracle
from tests import _util
from torch_amd import GasGravityField, GravityField, StellarGravityField
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu
G = W.G_PC_MSUN_KMS
H = 1.0 / 64
EPS = np.array([0.0, 1.0 / 256, 1.0 / 32, 1.0 / 8])
SHIFT = np.array([1.5, -2.0, 0.75])            # a lattice vector
PROP = settings(max_examples=int(os.environ.get("TGRAV_PROPERTY_EXAMPLES", 25)), derandomize=True, deadline=None, database=None,
                suppress_health_check=list(HealthCheck))


@st.composite
def cloud(draw, nmax, nmin=0):
    return dict(n=draw(st.integers(nmin, nmax)), seed=draw(st.integers(0, 2 ** 31 - 1)),
                spread=draw(st.sampled_from([2, 8, 64])),
                eps=draw(st.sampled_from(["none", "zero", "uniform", "mixed"])),
                mass=draw(st.sampled_from(["positive", "with_zeros", "signed", "powers_of_two"])))


def points(c, rng):
    return rng.integers(-c["spread"], c["spread"] + 1, size=(c["n"], 3)) * H


def masses(c, rng):
    n = c["n"]
    if c["mass"] == "powers_of_two":
        return 2.0 ** rng.integers(-4, 5, n)
    m = rng.uniform(0.05, 5.0, n)
    if c["mass"] == "with_zeros":
        m[rng.random(n) < 0.3] = 0.0
    elif c["mass"] == "signed":
        m *= rng.choice([-1.0, 1.0], n)
    return m


def softening(c, rng):
    n = c["n"]
    if c["eps"] == "none":
        return None
    if c["eps"] == "zero":
        return np.zeros(n)
    if c["eps"] == "uniform":
        return np.full(n, EPS[rng.integers(1, 4)])
    return EPS[rng.integers(0, 4, n)]


def coincide(dst, src, rng, frac=0.25):
    """Moves a fraction of the rows of dst onto rows of src (exact coincidences)."""
    if len(dst) and len(src):
        k = rng.random(len(dst)) < frac
        dst[k] = src[rng.integers(0, len(src), int(k.sum()))]
    return dst


def sq(e, n):
    return np.zeros(n) if e is None else e * e


def diag(p):
    return float(np.linalg.norm(p.max(0) - p.min(0))) if len(p) else 0.0


def as_device(arrs, on):
    if not on:
        return arrs
    import torch
    return [None if a is None else torch.as_tensor(np.ascontiguousarray(a), device="cuda") for a in arrs]


def host(out):
    return tuple(None if o is None else (o.cpu().numpy() if hasattr(o, "cpu") else o) for o in out)


def same(a, b):
    for p, q in zip(a, b):
        assert (p is None) == (q is None)
        if p is not None:
            np.testing.assert_array_equal(p, q)


def check(mode, got, ref, s, m, se2, t, te2, anchor, fold=512, potential=True):
    ok, worst = _util.within_bounds(got, ref, _util.term_scale_bounds(s, m, se2, t, te2, G, mode, anchor, fold),
                                    potential)
    assert ok, f"{mode}: worst error / bound = {worst:.3g}"


@pytest.fixture(scope="module")
def fields():
    f = {m: GravityField(m) for m in ("fp64", "mixed")}
    yield f
    for v in f.values():
        v.close()


# ------------------------------------------------------------- explicit targets ----
@PROP
@given(mode=st.sampled_from(["fp64", "mixed"]), src=cloud(700), tgt=cloud(1500),
       hit=st.booleans(), device=st.booleans())
def test_explicit_targets(fields, mode, src, tgt, hit, device):
    rng = np.random.Generator(np.random.PCG64(src["seed"]))
    s, m, se = points(src, rng), masses(src, rng), softening(src, rng)
    trng = np.random.Generator(np.random.PCG64(tgt["seed"]))
    t, te = points(tgt, trng), softening(tgt, trng)
    if hit:
        t = coincide(t, s, trng)
    f = fields[mode]

    def run(shift=0.0, scale=1.0, **kw):
        S = as_device([s[:, 0] + shift * SHIFT[0], s[:, 1] + shift * SHIFT[1], s[:, 2] + shift * SHIFT[2],
                       m * scale, se], device)
        T = as_device([te, t[:, 0] + shift * SHIFT[0], t[:, 1] + shift * SHIFT[1], t[:, 2] + shift * SHIFT[2]],
                      device)
        f.set_sources(*S)
        return host(f.evaluate(*T, **kw))

    got = run()
    ref = oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, se, t[:, 0], t[:, 1], t[:, 2], te, G,
                              mode=oracle.NEUMAIER)
    check(mode, got, ref, s, m, sq(se, len(s)), t, sq(te, len(t)), np.full(len(t), diag(t)))
    moved = run(shift=1.0, scale=4.0)
    same(moved, tuple(4.0 * g for g in got))
    acc_only = run(potential=False)
    same(acc_only[:3], got[:3])
    assert acc_only[3] is None
    pot_only = run(acceleration=False)
    same(pot_only[3:], got[3:])


# ---------------------------------------------------------------- block targets ----
@st.composite
def blocks(draw, nmax):
    return dict(n=draw(st.integers(1, nmax)), seed=draw(st.integers(0, 2 ** 31 - 1)),
                spread=draw(st.sampled_from([2, 16, 64])),
                eps=draw(st.sampled_from(["none", "cell", "table"])))


def make_blocks(b):
    rng = np.random.Generator(np.random.PCG64(b["seed"]))
    o = rng.integers(-b["spread"], b["spread"] + 1, size=(b["n"], 3)) * H
    dx = H * 2.0 ** rng.integers(-2, 2, b["n"])
    e = {"none": None, "cell": dx.copy(), "table": EPS[rng.integers(0, 4, b["n"])]}[b["eps"]]
    return o, dx, e


def centres(o, dx):
    c = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    return np.stack(c, 1)


@PROP
@given(mode=st.sampled_from(["fp64", "mixed"]), src=cloud(600), blk=blocks(6), hit=st.booleans(),
       column=st.booleans(), device=st.booleans())
def test_block_targets(fields, mode, src, blk, hit, column, device):
    o, dx, e = make_blocks(blk)
    t = centres(o, dx)
    rng = np.random.Generator(np.random.PCG64(src["seed"]))
    s, m, se = points(src, rng), masses(src, rng), softening(src, rng)
    if hit:
        s = coincide(s, t, rng)
    f = fields[mode]
    f.set_block_kernel(column)
    try:
        def run(shift=0.0, scale=1.0):
            S = as_device([s[:, 0] + shift * SHIFT[0], s[:, 1] + shift * SHIFT[1], s[:, 2] + shift * SHIFT[2],
                           m * scale, se], device)
            B = as_device([o + shift * SHIFT, dx, e], device)
            f.set_sources(*S)
            return host(f.evaluate_blocks(B[0], B[1], eps=B[2]))

        got = run()
        # the same centres as explicit targets (block C order): the same kernel, bit for bit
        te = None if e is None else np.repeat(e, 512)
        expl = host(f.evaluate(*as_device([te, t[:, 0], t[:, 1], t[:, 2]], device)))
        same(expl, got)
        moved = run(shift=1.0, scale=4.0)
    finally:
        f.set_block_kernel(True)
    ref = oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, se, t[:, 0], t[:, 1], t[:, 2], te, G,
                              mode=oracle.NEUMAIER)
    check(mode, got, ref, s, m, sq(se, len(s)), t, sq(te, len(t)), np.full(len(t), diag(t) + 2 * dx.max()))
    same(moved, tuple(4.0 * g for g in got))


# ---------------------------------------------------------------- block sources ----
@PROP
@given(mode=st.sampled_from(["fp64", "mixed"]), blk=blocks(4), tgt=cloud(1200), mass=cloud(0),
       hit=st.booleans(), device=st.booleans())
def test_block_sources(fields, mode, blk, tgt, mass, hit, device):
    o, dx, e = make_blocks(blk)
    c = centres(o, dx)
    mrng = np.random.Generator(np.random.PCG64(mass["seed"]))
    cm = masses(dict(mass, n=len(c)), mrng)
    ce = dx if e is None else e                 # the default softening of a block is its cell width
    trng = np.random.Generator(np.random.PCG64(tgt["seed"]))
    t, te = points(tgt, trng), softening(tgt, trng)
    if hit:
        t = coincide(t, c, trng)
    f = fields[mode]

    def run(shift=0.0, scale=1.0):
        B = as_device([o + shift * SHIFT, dx, cm * scale, e], device)
        T = as_device([te, t[:, 0] + shift * SHIFT[0], t[:, 1] + shift * SHIFT[1], t[:, 2] + shift * SHIFT[2]],
                      device)
        f.set_source_blocks(B[0], B[1], B[2], eps=B[3])
        return host(f.evaluate(*T))

    got = run()
    cel = np.repeat(ce, 512)
    # the same cells as explicit sources (block C order): detected as blocks, bit for bit
    f.set_sources(*as_device([c[:, 0], c[:, 1], c[:, 2], cm, cel], device))
    T0 = as_device([te, t[:, 0], t[:, 1], t[:, 2]], device)
    same(host(f.evaluate(*T0)), got)
    moved = run(shift=1.0, scale=4.0)
    ref = oracle.eval_plummer(c[:, 0], c[:, 1], c[:, 2], cm, cel, t[:, 0], t[:, 1], t[:, 2], te, G,
                              mode=oracle.NEUMAIER)
    check(mode, got, ref, c, cm, cel * cel, t, sq(te, len(t)), np.full(len(t), diag(t)))
    same(moved, tuple(4.0 * g for g in got))


# --------------------------------------------------------- fused two-way pass ----
@PROP
@given(mode=st.sampled_from(["fp64", "mixed"]), blk=blocks(3), par=cloud(400), mass=cloud(0),
       hit=st.booleans(), device=st.booleans())
def test_bidirectional_blocks(fields, mode, blk, par, mass, hit, device):
    o, dx, e = make_blocks(blk)
    c = centres(o, dx)
    mrng = np.random.Generator(np.random.PCG64(mass["seed"]))
    cm = np.abs(masses(dict(mass, n=len(c)), mrng))
    prng = np.random.Generator(np.random.PCG64(par["seed"]))
    p, pm, pe = points(par, prng), masses(par, prng), softening(par, prng)
    if hit:
        p = coincide(p, c, prng)
    f = fields[mode]

    def run(shift=0.0, scale=1.0):
        A = as_device([o + shift * SHIFT, dx, cm * scale, e, p[:, 0] + shift * SHIFT[0],
                       p[:, 1] + shift * SHIFT[1], p[:, 2] + shift * SHIFT[2], pm * scale, pe], device)
        co, po = f.bidirectional_blocks(*A)
        return host(co), host(po)

    gc, gp = run()
    mc, mp = run(shift=1.0, scale=4.0)
    cel = np.zeros(len(c)) if e is None else np.repeat(e, 512)
    # the same cells as explicit arrays (the blocks' C order): recognised, the same bits
    A = as_device([c[:, 0], c[:, 1], c[:, 2], cm, cel, p[:, 0], p[:, 1], p[:, 2], pm, pe], device)
    xc, xp = f.bidirectional(*A)
    same(host(xc), gc)
    same(host(xp), gp)
    rp, rc = oracle.bidirectional(c[:, 0], c[:, 1], c[:, 2], cm, cel, p[:, 0], p[:, 1], p[:, 2], pm, pe, G)
    ctr = np.repeat(o + 4.0 * dx[:, None], 512, axis=0)              # block centres, per cell
    half = np.repeat(np.sqrt(3.0) * 3.5 * dx, 512)
    check(mode, gc, rc, p, pm, sq(pe, len(p)), c, cel * cel, half)
    check(mode, gp, rp, c, cm, cel * cel, p, sq(pe, len(p)),
          np.linalg.norm(p[:, None, :] - ctr[None, :, :], axis=2))
    same(mc, tuple(4.0 * g for g in gc))
    same(mp, tuple(4.0 * g for g in gp))


# ------------------------------------------------------------ FLASH kernels ----
@PROP
@given(src=cloud(400), tgt=cloud(600), blk=blocks(3), kernel=st.sampled_from(["spline", "linear"]),
       quirks=st.booleans(), rsoft=st.sampled_from([1.0 / 16, 0.25, 1.0]),
       form=st.sampled_from(["points", "blocks", "points in block order", "periodic", "periodic blocks"]),
       box=st.sampled_from([0.5, 1.0, 2.0]), nrep=st.integers(1, 2), hit=st.booleans(), device=st.booleans())
def test_flash_kernels(fields, src, tgt, blk, kernel, quirks, rsoft, form, box, nrep, hit, device):
    """FLASH's spline/linear kernels (FP64) on every source form -- explicit points, 8^3 blocks,
    explicit points in block order (detected on the device), and the periodic image sum -- within
    1e-12 of the sum of the terms' magnitudes against the oracle.  Lattice inputs put pairs exactly
    on r = rsoft and on the image cutoff, where the branch decisions must match the reference."""
    from tests.test_gpu_flash import check as flash_check
    from tests.test_gpu_flash import term_scales
    rng = np.random.Generator(np.random.PCG64(src["seed"]))
    if form in ("blocks", "points in block order", "periodic blocks"):
        o, dx, _ = make_blocks(blk)
        s = centres(o, dx)
        m = masses(dict(src, n=len(s)), rng)
    else:
        s, m = points(src, rng), masses(src, rng)
    trng = np.random.Generator(np.random.PCG64(tgt["seed"]))
    t = points(tgt, trng)
    if hit:
        t = coincide(t, s, trng)
    periodic = form.startswith("periodic")
    if form == "periodic":         # keep the image sum small: (2 nrep + 1)^3 images per pair
        s, m, t = s[:150], m[:150], t[:200]
    elif periodic:                 # block sources: every cell of one or two blocks
        keep = 512 * min(2, len(s) // 512)
        o, dx, s, m, t = o[:keep // 512], dx[:keep // 512], s[:keep], m[:keep], t[:100]
    f = fields["fp64"]
    f.set_softening(kernel, rsoft, quirks=quirks)
    if periodic:
        f.set_periodic((box, box, box), nrep=nrep)
    try:
        if form in ("blocks", "periodic blocks"):
            f.set_source_blocks(*as_device([o, dx, m], device))
        else:
            f.set_sources(*as_device([s[:, 0], s[:, 1], s[:, 2], m], device), None)
        got = host(f.evaluate(None, *as_device([t[:, 0], t[:, 1], t[:, 2]], device)))
    finally:
        f.set_softening("plummer")
        f.set_periodic(None)
    B = (box, box, box) if periodic else None
    ref = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rsoft,
                            kernel=kernel, quirks=quirks, threads=16, box=B, nrep=nrep)
    sa, sp = term_scales(s[:, 0], s[:, 1], s[:, 2], np.abs(m), t[:, 0], t[:, 1], t[:, 2], rsoft, box=B, nrep=nrep)
    flash_check(got, ref, sa + 1e-300, sp + 1e-300)


# ------------------------------------------------------------- particle mesh ----
@PROP
@given(nblk=st.integers(1, 4), nb=st.sampled_from([4, 8]), ng=st.integers(1, 4), npart=st.integers(0, 300),
       seed=st.integers(0, 2 ** 31 - 1), mass=st.sampled_from(["positive", "with_zeros", "signed", "powers_of_two"]),
       device=st.booleans())
def test_particle_mesh(nblk, nb, ng, npart, seed, mass, device):
    """The PM pieces on random guarded block layouts: particles anywhere in their block's
    interior, on a lattice that puts some exactly on cell centres and faces.  Gradient, grid
    kick and interpolation bit-identical to the oracle; deposit within 1e-13 of its sequential
    sums (relative to the cell's sum of |contributions|) and order-free."""
    import torch

    from torch_amd import LeafBlocks, ParticleMesh
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.integers(-8, 9, size=(nblk, 3)) * H
    size = nb * H * 2.0 ** rng.integers(-2, 2, nblk)
    blk = rng.integers(0, nblk, npart)
    frac = rng.integers(0, 4 * nb, size=(npart, 3)) / (4.0 * nb)   # quarter-cell lattice, [0, 1)
    pos = o[blk] + frac * size[blk, None]
    m = masses(dict(n=npart, mass=mass), rng)
    d = (lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")) if device else (lambda a: a)
    B = LeafBlocks(d(o), d(size), nb=nb, ng=ng)
    Bh = LeafBlocks(o, size, nb=nb, ng=ng)
    ob = dict(origin=o, size=size, nb=nb, ng=ng)
    h1 = lambda t: t.cpu().numpy() if hasattr(t, "cpu") else t   # noqa: E731
    with ParticleMesh() as pm:
        P = [d(pos[:, 0]), d(pos[:, 1]), d(pos[:, 2])]
        rho = h1(pm.update_grid_var(*P, d(m), d(blk), B))
        ref = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], m, blk, ob)
        scale = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], np.abs(m), blk, ob)
        assert np.all(np.abs(rho - ref) <= 1e-13 * scale), "deposit"
        perm = rng.permutation(npart)
        again = h1(pm.update_grid_var(*(d(a[perm]) for a in (pos[:, 0], pos[:, 1], pos[:, 2], m, blk)), B))
        np.testing.assert_array_equal(again, rho)
        phi = rng.normal(0, 1, Bh.ncell)
        gref = oracle.pm_gradient(phi, ob)
        v0 = [rng.normal(0, 1, Bh.ncell) for _ in range(3)]
        v = [d(a.copy()) for a in v0]
        g = pm.accel_list_of_blocks(d(phi), B, dt=0.3, vel=v)
        for q in range(3):
            np.testing.assert_array_equal(h1(g[q]), gref[q])
            np.testing.assert_array_equal(h1(v[q]), v0[q] + gref[q] * 0.3)
        a = pm.map_mesh_to_particles(*P, d(blk), B, *(d(q) for q in gref))
        aref = oracle.pm_interpolate(pos[:, 0], pos[:, 1], pos[:, 2], blk, ob, *gref)
        for q in range(3):
            np.testing.assert_array_equal(h1(a[q]), aref[q])


# -------------------------------------------------------------------- batches ----
@st.composite
def batch(draw):
    k = draw(st.integers(1, 5))
    return [dict(src=draw(cloud(400)), tgt=draw(cloud(900)), blk=draw(st.one_of(st.none(), blocks(3))))
            for _ in range(k)]


@PROP
@given(mode=st.sampled_from(["fp64", "mixed"]), probs=batch())
def test_batch_plans(fields, mode, probs):
    import torch

    from torch_amd.batch import BatchPlan
    f = fields[mode]
    P, refs = [], []
    for q in probs:
        rng = np.random.Generator(np.random.PCG64(q["src"]["seed"]))
        s, m, se = points(q["src"], rng), masses(q["src"], rng), softening(q["src"], rng)
        S = dict(x=s[:, 0], y=s[:, 1], z=s[:, 2], m=m, eps=se)
        if q["blk"] is None:
            trng = np.random.Generator(np.random.PCG64(q["tgt"]["seed"]))
            t, te = points(q["tgt"], trng), softening(q["tgt"], trng)
            T = dict(x=t[:, 0], y=t[:, 1], z=t[:, 2], eps=te)
            anchor = diag(t)
        else:
            o, dx, e = make_blocks(q["blk"])
            t = centres(o, dx)
            te = None if e is None else np.repeat(e, 512)
            T = dict(block_origin=o, block_dx=dx, block_eps=e)
            anchor = diag(t) + 2 * dx.max()
        P.append(dict(sources=S, targets=T))
        refs.append((s, m, se, t, te, anchor))
    with BatchPlan(f, P) as plan:
        eager = [host(o) for o in plan.run(graph=False)]
        torch.cuda.synchronize()
        for _ in range(2):
            replay = [host(o) for o in plan.run(graph=True)]
            torch.cuda.synchronize()
            for a, b in zip(replay, eager):
                same(a, b)
    # block problems handed over as explicit centres in the blocks' C order: recognised when the
    # plan is created, the same bits
    PX = [dict(sources=p["sources"], targets=dict(x=r[3][:, 0], y=r[3][:, 1], z=r[3][:, 2], eps=r[4]))
          if "block_origin" in p["targets"] else p for p, r in zip(P, refs)]
    with BatchPlan(f, PX) as plan:
        for a, b in zip([host(o) for o in plan.run(graph=True)], eager):
            same(a, b)
    for got, (s, m, se, t, te, anchor) in zip(eager, refs):
        ref = oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, se, t[:, 0], t[:, 1], t[:, 2], te, G,
                                  mode=oracle.NEUMAIER)
        check(mode, got, ref, s, m, sq(se, len(s)), t, sq(te, len(t)), np.full(len(t), anchor))


# --------------------------------------------------------------- the wrappers ----
@PROP
@given(src=cloud(300, nmin=1), tgt=cloud(300))
def test_reference_method_names_agree_with_evaluate(src, tgt):
    """The AMUSE-named methods are the same evaluation: get_gravity_at_point and
    get_potential_at_point (stellar side), get_accel_gas_on_particles (gas side)."""
    rng = np.random.Generator(np.random.PCG64(src["seed"]))
    s, m, se = points(src, rng), np.abs(masses(src, rng)), softening(src, rng)
    trng = np.random.Generator(np.random.PCG64(tgt["seed"]))
    t = points(tgt, trng)
    te = np.zeros(len(t))
    with StellarGravityField("fp64") as sf, GasGravityField("fp64") as gf:
        sf.set_stars(s[:, 0], s[:, 1], s[:, 2], m, se)
        gf.set_sources(s[:, 0], s[:, 1], s[:, 2], m, se)
        full = sf.evaluate(te, t[:, 0], t[:, 1], t[:, 2])
        same(sf.get_gravity_at_point(te, t[:, 0], t[:, 1], t[:, 2]), full[:3])
        same((sf.get_potential_at_point(te, t[:, 0], t[:, 1], t[:, 2]),), full[3:])
        same(gf.get_accel_gas_on_particles(te, t[:, 0], t[:, 1], t[:, 2]), full[:3])
