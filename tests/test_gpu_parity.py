"""GPU parity of libtgrav.so against the CPU oracle (needs an MI355X).

Bars: plain FP64 within 1e-12 normwise relative error of the oracle per target (or, for
the brute-force KATs with cancellation, within the condition-aware FP64 summation bound);
mixed precision <= 1e-6 median and <= 1e-4 max relative acceleration error against the
FP64 oracle.  Integer-exact properties (kick, block centres, host vs device pointers,
determinism) are checked bit for bit.
"""
import os

import numpy as np
import pytest

import oracle
from tests import _util
from tests.test_oracle import kat_error_ok, load_kats
from torch_amd import GasGravityField, GravityField, StellarGravityField, TgravError, _lib
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
G = W.G_PC_MSUN_KMS
THREADS = 16


@pytest.fixture(scope="module")
def fp64():
    f = GravityField("fp64")
    yield f
    f.close()


@pytest.fixture(scope="module")
def mixed():
    f = GravityField("mixed")
    yield f
    f.close()


@pytest.fixture(scope="module")
def c1():
    return np.load(os.path.join(HERE, "golden", "c1_oracle.npz"))


@pytest.fixture(scope="module")
def c2():
    return W.c2(1002)


def gpu_eval(f, S, T, eps_t=None):
    f.set_sources(S["x"], S["y"], S["z"], S["m"], S.get("eps"))
    return f.evaluate(eps_t, T["x"], T["y"], T["z"])


def stars(d, pre):
    return dict(x=d[pre + "x"], y=d[pre + "y"], z=d[pre + "z"], m=d[pre + "m"], eps=d[pre + "eps"])


# ------------------------------------------------------------------ KATs ----
@pytest.mark.parametrize("mode,k", [("fp64", 64.0), ("mixed", None)])
def test_brute_force_kats(mode, k):
    kats, Gk = load_kats()
    with GravityField(mode, G=Gk) as f:
        for i, c in enumerate(kats):
            s, t = c["s"], c["t"]
            f.set_sources(s[:, 0], s[:, 1], s[:, 2], c["m"], c["se"])
            got = np.stack(f.evaluate(c["te"], t[:, 0], t[:, 1], t[:, 2]), 1)
            if k is not None:
                ok, worst = kat_error_ok(got, c, Gk, k)
                assert ok, f"case {i}: {worst:.2f}x the FP64 bound"
            else:
                ok, worst = _util.within_mixed_bound(got, c["expected"], s, c["m"], c["se"], t,
                                                     c["te"], Gk)
                print(f"KAT {i} mixed: error / analytic bound = {worst:.3f}", flush=True)
                assert ok, f"case {i}: {worst:.2f}x the analytic mixed-precision bound"


def test_single_pair_sweep_fp64(fp64):
    """One source, targets over 12 decades of distance: the FP64 rsqrt refinement keeps each
    pair within a few ulp of the closed form."""
    r = np.geomspace(1e-6, 1e6, 4001)
    fp64.set_sources([0.0], [0.0], [0.0], [1.5], [0.0])
    ax, ay, az, phi = fp64.evaluate(0.0, r, np.zeros_like(r), np.zeros_like(r))
    np.testing.assert_allclose(ax, -G * 1.5 / r ** 2, rtol=8e-16 * 4, atol=0)
    np.testing.assert_allclose(phi, -G * 1.5 / r, rtol=8e-16 * 4, atol=0)
    assert np.all(ay == 0) and np.all(az == 0)


# ------------------------------------------------------------------- c1 ----
def test_c1_fp64_matches_oracle(fp64, c1):
    S, C = stars(c1, "star_"), stars(c1, "cell_")
    fwd = gpu_eval(fp64, S, C)
    s = _util.summary(fwd, tuple(c1["fwd"]))
    print("c1 fwd fp64", s, flush=True)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    rev = gpu_eval(fp64, C, S)
    s = _util.summary(rev, tuple(c1["rev"]))
    print("c1 rev fp64", s, flush=True)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12


def test_c1_mixed_error(mixed, c1):
    S, C = stars(c1, "star_"), stars(c1, "cell_")
    for name, got, ref in (("fwd", gpu_eval(mixed, S, C), c1["fwd"]),
                           ("rev", gpu_eval(mixed, C, S), c1["rev"])):
        s = _util.summary(got, tuple(ref))
        print(f"c1 {name} mixed", s, flush=True)
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
        assert s["pot_median"] <= 1e-6 and s["pot_max"] <= 1e-4


# ------------------------------------------------------------------- c2 ----
def test_c2_fp64_matches_oracle_on_subsample(fp64, c2):
    st, ce = c2["stars"], c2["cells"]
    fwd = gpu_eval(fp64, st, ce)
    idx = _util.subsample(len(ce["x"]), 2048, 21)
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][idx], ce["y"][idx],
                              ce["z"][idx], None, G, mode=oracle.NEUMAIER, threads=THREADS)
    s = _util.summary(tuple(a[idx] for a in fwd), ref)
    print("c2 fwd fp64", s, flush=True)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    rev = gpu_eval(fp64, ce, st)
    idx = _util.subsample(len(st["x"]), 384, 22)
    idx = np.union1d(idx, np.where(st["binary"] >= 0)[0][:128])
    ref = oracle.eval_plummer(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"], st["x"][idx], st["y"][idx],
                              st["z"][idx], None, G, mode=oracle.NEUMAIER, threads=THREADS)
    s = _util.summary(tuple(a[idx] for a in rev), ref)
    print("c2 rev fp64", s, flush=True)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12


def test_c2_mixed_against_fp64(fp64, mixed, c2):
    """Full-size mixed error against the (oracle-pinned) FP64 kernel; binary members apart."""
    st, ce = c2["stars"], c2["cells"]
    for name, S, T in (("fwd", st, ce), ("rev", ce, st)):
        ref = gpu_eval(fp64, S, T)
        got = gpu_eval(mixed, S, T)
        s = _util.summary(got, ref)
        print(f"c2 {name} mixed", s, flush=True)
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
        if name == "rev":
            b = st["binary"] >= 0
            sb = _util.summary(tuple(a[b] for a in got), tuple(a[b] for a in ref))
            print("c2 rev mixed binary members", sb, flush=True)
            assert sb["acc_median"] <= 1e-6 and sb["acc_max"] <= 1e-4


# ------------------------------------------------------------ edge cases ----
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("ns,nt", [(1, 1), (3, 63), (257, 65), (1000, 1), (5, 1029), (777, 2049)])
def test_ragged_sizes(mode, ns, nt):
    rng = W.rng_for(ns * 7 + nt)
    S = dict(x=rng.normal(size=ns), y=rng.normal(size=ns), z=rng.normal(size=ns),
             m=rng.uniform(0.1, 2, ns), eps=rng.uniform(0.001, 0.05, ns))
    t = rng.normal(size=(3, nt))
    te = rng.uniform(0, 0.02, nt)
    with GravityField(mode) as f:
        got = gpu_eval(f, S, dict(x=t[0], y=t[1], z=t[2]), te)
    ref = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], t[0], t[1], t[2], te, G,
                              mode=oracle.NEUMAIER)
    s = _util.summary(got, ref)
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s
    else:
        assert s["acc_max"] <= 1e-4 and s["pot_max"] <= 1e-4, s


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_self_exclusion_and_zero_softening(mode):
    """eps = 0 and coincident source/target: the pair contributes nothing."""
    x = np.array([0.0, 1.0, 0.0, 0.3])
    y = np.array([0.0, 0.0, 2.0, -0.4])
    z = np.zeros(4)
    m = np.array([1.0, 2.0, 3.0, 0.5])
    with GravityField(mode) as f:
        f.set_sources(x, y, z, m, None)
        got = f.evaluate(None, x, y, z)
    ref = oracle.eval_plummer(x, y, z, m, None, x, y, z, None, G, mode=oracle.LONGDOUBLE)
    assert all(np.all(np.isfinite(g)) for g in got)
    if mode == "fp64":
        for g, r in zip(got, ref):
            np.testing.assert_allclose(g, r, rtol=1e-12, atol=0)
    else:
        P = np.stack([x, y, z], 1)
        ok, worst = _util.within_mixed_bound(got, ref, P, m, None, P, None, G)
        assert ok, worst


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_empty_inputs(mode):
    with GravityField(mode) as f:
        f.set_sources([], [], [], [], None)
        ax, ay, az, phi = f.evaluate(0.0, np.ones(5), np.ones(5), np.ones(5))
        assert np.all(ax == 0) and np.all(phi == 0)
        f.set_sources([0.0], [0.0], [0.0], [1.0], [0.1])
        out = f.evaluate(None, np.zeros(0), np.zeros(0), np.zeros(0))
        assert all(o.shape == (0,) for o in out)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("device", [False, True])
def test_empty_inputs_every_entry_point(mode, device):
    """Zero targets, zero sources, zero particles or zero cells, host or device arrays, with
    input validation on: every entry point returns empty or zero fields (an empty launch is a
    no-op, not an invalid configuration)."""
    import torch
    rng = W.rng_for(58)
    d = (lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device="cuda")) if device else \
        (lambda a: np.asarray(a, dtype=np.float64))
    h = (lambda t: t.cpu().numpy()) if device else (lambda t: t)
    z = d(np.zeros(0))
    o, dx = d(np.zeros((1, 3))), d([0.1])
    cm, ce = d(rng.uniform(0.1, 1, 512)), d([0.01])
    p = [d(rng.uniform(-1, 1, 7)) for _ in range(3)]
    pm, pe = d(rng.uniform(0.1, 1, 7)), d(np.full(7, 0.02))
    with GravityField(mode) as f:
        co, po = f.bidirectional_blocks(o, dx, cm, ce, z, z, z, z, z)     # no particles
        assert all(np.all(h(a) == 0) for a in co) and all(h(a).shape == (0,) for a in po)
        co, po = f.bidirectional_blocks(d(np.zeros((0, 3))), d(np.zeros(0)), z, z, *p, pm, pe)   # no cells
        assert all(h(a).shape == (0,) for a in co) and all(np.all(h(a) == 0) for a in po)
        cx = [d(rng.uniform(-1, 1, 512)) for _ in range(3)]
        co, po = f.bidirectional(*cx, cm, d(np.full(512, 0.01)), z, z, z, z, z)
        assert all(np.all(h(a) == 0) for a in co) and all(h(a).shape == (0,) for a in po)
        co, po = f.bidirectional(z, z, z, z, z, *p, pm, pe)
        assert all(h(a).shape == (0,) for a in co) and all(np.all(h(a) == 0) for a in po)
        f.set_sources(z, z, z, z, None)
        out = f.evaluate(None, *p)
        assert all(np.all(h(a) == 0) for a in out)
        out = f.evaluate_blocks(o, dx)
        assert all(np.all(h(a) == 0) for a in out)
        f.set_sources(*p, pm, pe)
        out = f.evaluate(None, z, z, z)
        assert all(h(a).shape == (0,) for a in out)
        out = f.evaluate_blocks(d(np.zeros((0, 3))), d(np.zeros(0)))
        assert all(h(a).shape == (0,) for a in out)


def test_error_codes():
    f = GravityField("fp64")
    with pytest.raises(TgravError) as e:
        f.evaluate(0.0, [0.0], [0.0], [0.0])
    assert e.value.code == _lib.ENOSOURCES
    with pytest.raises(TgravError) as e:
        f.set_sources([0.0, np.nan], [0.0, 0.0], [0.0, 0.0], [1.0, 1.0])
    assert e.value.code == _lib.ENONFINITE
    f.set_sources([0.0], [0.0], [0.0], [1.0])
    with pytest.raises(TgravError) as e:
        f.evaluate(0.0, [np.inf], [0.0], [0.0])
    assert e.value.code == _lib.ENONFINITE
    with pytest.raises(TgravError) as e:
        f.evaluate(0.0, [0.0, 1.0], [0.0], [0.0])
    assert e.value.code == _lib.EINVAL
    lib = _lib.load()
    assert lib.tgrav_eval(f.handle, 1, None, None, None, None, None, None, None, None, 0, 0) == _lib.EINVAL
    assert lib.tgrav_eval(f.handle, 0, None, None, None, None, None, None, None, None, 5, 0) == _lib.EINVAL
    assert lib.tgrav_eval(f.handle, -1, None, None, None, None, None, None, None, None, 0, 0) == _lib.EINVAL
    f.close()


# --------------------------------------------------- launch-shape choices ----
def test_chunked_sources_match_single_pass(fp64, c2):
    """Splitting the source range over workgroups (chunk sums + ordered reduction) agrees
    with one pass per target to rounding."""
    ce, st = c2["cells"], c2["stars"]
    sl = slice(0, 200_000)
    S = {k: ce[k][sl] for k in ("x", "y", "z", "m", "eps")}
    T = {k: st[k][:3000] for k in ("x", "y", "z")}
    outs = []
    for nsplit in (1, 0, 7, 64):
        fp64.set_config(nsplit=nsplit)
        outs.append(gpu_eval(fp64, S, T))
    fp64.set_config(nsplit=0)
    for o in outs[1:]:
        s = _util.summary(o, outs[0])
        print("chunked vs single pass", s, flush=True)
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_every_variant_agrees(mode, c2):
    st, ce = c2["stars"], c2["cells"]
    T = {k: ce[k][:40_000] for k in ("x", "y", "z")}
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], T["x"][:1024], T["y"][:1024],
                              T["z"][:1024], None, G, mode=oracle.NEUMAIER, threads=THREADS)
    lib = _lib.load()
    with GravityField(mode) as f:
        f.set_block_kernel(False)      # the cells are in block C order: keep the generic kernels
        for v in range(lib.tgrav_num_variants(_lib.PLAIN_FP64 if mode == "fp64" else _lib.MIXED)):
            f.set_config(variant=v)
            got = gpu_eval(f, st, T)
            s = _util.summary(tuple(a[:1024] for a in got), ref)
            print(mode, "variant", v, s, flush=True)
            if mode == "fp64":
                assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
            else:
                assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_deterministic(mode, c2):
    st, ce = c2["stars"], c2["cells"]
    T = {k: ce[k][:100_000] for k in ("x", "y", "z")}
    with GravityField(mode) as f:
        a = gpu_eval(f, st, T)
        b = gpu_eval(f, st, T)
    for p, q in zip(a, b):
        assert np.array_equal(p, q)


def test_device_pointers_match_host_pointers(fp64, c2):
    import torch
    st, ce = c2["stars"], c2["cells"]
    T = {k: ce[k][:50_000] for k in ("x", "y", "z")}
    host = gpu_eval(fp64, st, T)
    d = lambda a: torch.as_tensor(a, device="cuda")  # noqa: E731
    fp64.set_sources(d(st["x"]), d(st["y"]), d(st["z"]), d(st["m"]), d(st["eps"]))
    dev = fp64.evaluate(None, d(T["x"]), d(T["y"]), d(T["z"]))
    torch.cuda.synchronize()
    for p, q in zip(host, dev):
        assert np.array_equal(p, q.cpu().numpy())


# ------------------------------------------------------------- kick/blocks ----
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("layout", ["c1 grid", "random cloud", "c2 cells shuffled"])
def test_kick_with_unblocked_sources_multiple_of_512(mode, layout, c2):
    """Explicit sources whose count is a multiple of 512 but which are not 8^3 leaf blocks in C
    order: the block-source kernel is launched as a candidate and must stand down, so the fused
    kick adds dt * a exactly once."""
    rng = W.rng_for(57)
    if layout == "c1 grid":
        w = W.c1(1001)
        ce, st = w["cells"], w["stars"]
        src = [ce[k] for k in ("x", "y", "z", "m", "eps")]
        tgt = [st[k] for k in ("x", "y", "z")]
    elif layout == "random cloud":
        src = list(rng.uniform(-1, 1, (3, 1024))) + [rng.uniform(0.1, 1, 1024), np.full(1024, 0.02)]
        tgt = list(rng.uniform(-1, 1, (3, 3000)))
    else:
        ce, st = c2["cells"], c2["stars"]
        perm = rng.permutation(len(ce["x"]))
        src = [ce[k][perm] for k in ("x", "y", "z", "m", "eps")]
        tgt = [st[k][:2000] for k in ("x", "y", "z")]
    assert len(src[0]) % 512 == 0
    dt = 0.01
    with GasGravityField(mode) as f:
        f.set_sources(*src)
        b = f.evaluate(None, *tgt)
        v0 = [rng.normal(0, 1, len(tgt[0])) for _ in range(3)]
        v = [q.copy() for q in v0]
        a = f.evaluate_and_kick(None, *tgt, dt, *v)
    for p, q in zip(a[:3], b[:3]):
        assert np.array_equal(p, q)
    for p, q in zip(v, oracle.kick(dt, b[0], b[1], b[2], *v0)):
        assert np.array_equal(p, q)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_fused_kick_equals_eval_then_kick(mode, c2):
    st, ce = c2["stars"], c2["cells"]
    dt = 0.5 * 1.3e-3
    with GasGravityField(mode) as f:
        f.set_sources(ce["x"], ce["y"], ce["z"], ce["m"], ce["eps"])
        vx, vy, vz = st["vx"].copy(), st["vy"].copy(), st["vz"].copy()
        a = f.evaluate_and_kick(None, st["x"], st["y"], st["z"], dt, vx, vy, vz)
        b = f.evaluate(None, st["x"], st["y"], st["z"])
        for p, q in zip(a[:3], b[:3]):
            assert np.array_equal(p, q)
        ex = oracle.kick(dt, b[0], b[1], b[2], st["vx"], st["vy"], st["vz"])
        for p, q in zip((vx, vy, vz), ex):
            assert np.array_equal(p, q)
        # standalone kick
        wx, wy, wz = st["vx"].copy(), st["vy"].copy(), st["vz"].copy()
        f.kick(dt, b[0], b[1], b[2], wx, wy, wz)
        for p, q in zip((wx, wy, wz), ex):
            assert np.array_equal(p, q)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_block_targets_equal_explicit_centres(mode, c2):
    """Block-structured targets: with the generic kernel the centres rebuilt on the device
    give the explicit-centre results bit for bit; the 8^3 column kernel agrees with them to
    the parity bar (its r^2 is summed as (dx^2 + dy^2 + eps^2) + dz^2).  Explicit centres in
    the blocks' C order take the column kernel too (detected on the device) and then equal
    the block form bit for bit."""
    st, ce = c2["stars"], c2["cells"]
    o, size = ce["block_origin"][:300], ce["block_size"][:300]
    dx = size / 8
    bx, by, bz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    with StellarGravityField(mode) as f:
        f.set_stars(st["x"], st["y"], st["z"], st["m"], st["eps"])
        f.set_block_kernel(False)
        exp = f.evaluate(np.repeat(dx, 512), bx, by, bz)
        blk = f.evaluate_blocks(o, dx, eps=dx)
        for p, q in zip(blk, exp):
            assert np.array_equal(p, q)
        f.set_block_kernel(True)
        col = f.evaluate_blocks(o, dx, eps=dx)
        s = _util.summary(col, exp)
        print(mode, "column kernel vs generic", s, flush=True)
        if mode == "fp64":
            assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
        else:
            assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
        cols = f.evaluate(np.repeat(dx, 512), bx, by, bz)
        for p, q in zip(cols, col):
            assert np.array_equal(p, q)
        v = [np.zeros(len(bx)) for _ in range(3)]
        f.evaluate_blocks(o, dx, eps=dx, dt=0.25, vel=v)
        for p, q in zip(v, col[:3]):
            assert np.array_equal(p, 0.25 * q)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_explicit_column_detection(mode, c2):
    """Explicit targets take the column kernel exactly when every aligned group of 8 shares x
    and y and every aligned group of 512 shares eps (checked on the device); otherwise the
    generic kernel runs, bit-identical to the generic kernel forced on.  A ragged tail (not a
    multiple of 512) and the fused kick take the same path."""
    st, ce = c2["stars"], c2["cells"]
    o, dx = ce["block_origin"][:40], ce["block_size"][:40] / 8
    bx, by, bz = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx)
    e = np.repeat(dx, 512)
    n = 39 * 512 + 203                                       # ragged: last group partial
    cases = {"structured": (e, bx, by, bz)}
    px = bx.copy()
    px[5 * 512 + 8 * 3 + 5] += 1e-9                          # one target leaves its column
    cases["perturbed_x"] = (e, px, by, bz)
    pe = e.copy()
    pe[7 * 512 + 100] *= 1.5                                 # eps varies inside a group of 512
    cases["perturbed_eps"] = (pe, bx, by, bz)
    with StellarGravityField(mode) as f:
        f.set_stars(st["x"], st["y"], st["z"], st["m"], st["eps"])
        for name, (ee, x, y, z) in cases.items():
            for m in (len(x), n):
                args = (ee[:m], x[:m], y[:m], z[:m])
                got = f.evaluate(*args)
                f.set_block_kernel(False)
                gen = f.evaluate(*args)
                f.set_block_kernel(True)
                if name != "structured":
                    for p, q in zip(got, gen):
                        assert np.array_equal(p, q), name
                    continue
                s = _util.summary(got, gen)
                print(mode, name, m, "columns vs generic", s, flush=True)
                if mode == "fp64":
                    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
                else:
                    assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
                blk = f.evaluate_blocks(o, dx, eps=dx)
                for p, q in zip(got, blk):
                    if m == len(x) or mode == "fp64":     # the mixed anchor of a partial group differs
                        assert np.array_equal(p, q[:m])
                dt = 0.37
                v = [np.zeros(m) for _ in range(3)]
                kicked = f.evaluate_and_kick(*args, dt, *v)
                for p, q, w in zip(kicked[:3], got[:3], v):
                    assert np.array_equal(p, q)
                    assert np.array_equal(w, q * dt)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_column_kernel_matches_oracle_full_c2(mode, c2):
    """All 2,000 c2 leaf blocks through the column kernel; oracle on a subsample."""
    st, ce = c2["stars"], c2["cells"]
    dxb = ce["block_size"] / 8
    with StellarGravityField(mode) as f:
        f.set_stars(st["x"], st["y"], st["z"], st["m"], st["eps"])
        got = f.evaluate_blocks(ce["block_origin"], dxb)
        for ns in (1, 3):
            f.set_config(nsplit=ns)
            again = f.evaluate_blocks(ce["block_origin"], dxb)
            s = _util.summary(again, got)
            assert s["acc_max"] <= (1e-12 if mode == "fp64" else 1e-5), s
    idx = _util.subsample(len(ce["x"]), 2048, 23)
    ref = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], ce["x"][idx], ce["y"][idx],
                              ce["z"][idx], None, G, mode=oracle.NEUMAIER, threads=THREADS)
    s = _util.summary(tuple(a[idx] for a in got), ref)
    print(mode, "column kernel c2", s, flush=True)
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    else:
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4


# ---------------------------------------------------- full-size properties ----
def test_third_law_at_c2_scale(fp64, c2):
    """With symmetric softening (stars eps_*, cells eps = dx), the stars' total force on the
    gas cancels the gas's total force on the stars (the reference's force_error check,
    interface.F90:1467-1522) -- over all 1.02e10 pairs."""
    st, ce = c2["stars"], c2["cells"]
    a_c = gpu_eval(fp64, st, ce, eps_t=ce["eps"])
    a_s = gpu_eval(fp64, ce, st, eps_t=st["eps"])
    for q in range(3):
        f_c = np.sum(ce["m"] * a_c[q])
        f_s = np.sum(st["m"] * a_s[q])
        scale = np.sum(np.abs(ce["m"] * a_c[q])) + np.sum(np.abs(st["m"] * a_s[q]))
        print("third law", q, f_c, f_s, scale, flush=True)
        assert abs(f_c + f_s) <= 1e-12 * scale


def test_superposition_at_c2_scale(fp64, c2):
    st, ce = c2["stars"], c2["cells"]
    full = gpu_eval(fp64, ce, st)
    half = len(ce["x"]) // 2
    parts = [gpu_eval(fp64, {k: ce[k][sl] for k in ("x", "y", "z", "m", "eps")}, st)
             for sl in (slice(0, half), slice(half, None))]
    summed = tuple(p + q for p, q in zip(*parts))
    s = _util.summary(summed, full)
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12, s


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_acceleration_only_path_is_bit_identical(mode, c2):
    """get_gravity_at_point asks for the acceleration alone: that path skips the potential
    accumulator and must give the same accelerations bit for bit."""
    st, ce = c2["stars"], c2["cells"]
    T = {k: ce[k][:200_000] for k in ("x", "y", "z")}
    with StellarGravityField(mode) as f:
        f.set_stars(st["x"], st["y"], st["z"], st["m"], st["eps"])
        full = f.evaluate(None, T["x"], T["y"], T["z"])
        acc = f.get_gravity_at_point(0.0, T["x"], T["y"], T["z"])
        pot = f.get_potential_at_point(0.0, T["x"], T["y"], T["z"])
    for p, q in zip(full[:3], acc):
        assert np.array_equal(p, q)
    assert np.array_equal(full[3], pot)


def test_reference_method_names_give_the_oracle_field(c1):
    """GasGravityField.get_accel_gas_on_particles / get_potential_at_point and
    StellarGravityField.get_gravity_at_point on the c1 fixture."""
    S, C = stars(c1, "star_"), stars(c1, "cell_")
    with GasGravityField("fp64") as g:
        g.set_sources(C["x"], C["y"], C["z"], C["m"], C["eps"])
        ax, ay, az = g.get_accel_gas_on_particles(np.zeros(len(S["x"])), S["x"], S["y"], S["z"])
        phi = g.get_potential_at_point(np.zeros(len(S["x"])), S["x"], S["y"], S["z"])
    s = _util.summary((ax, ay, az, phi), tuple(c1["rev"]))
    assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    with StellarGravityField("fp64") as f:
        f.set_stars(S["x"], S["y"], S["z"], S["m"], S["eps"])
        gx, gy, gz = f.get_gravity_at_point(0.0, C["x"], C["y"], C["z"])
    s = _util.summary((gx, gy, gz, c1["fwd"][3]), tuple(c1["fwd"]))
    assert s["acc_max"] <= 1e-12


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_block_sources(mode, c2):
    """Cells given as 8^3 source blocks (tgrav_set_source_blocks): with the generic kernel
    they reproduce the explicit-cell results bit for bit; the column kernel agrees to the
    parity bar and with the oracle.  Explicit cells in the blocks' C order are detected as
    blocks and take the same column kernel, bit for bit."""
    st, ce = c2["stars"], c2["cells"]
    nb = 400
    sl = slice(0, nb * 512)
    o, dxb = ce["block_origin"][:nb], ce["block_size"][:nb] / 8
    with GasGravityField(mode) as g:
        g.set_sources(ce["x"][sl], ce["y"][sl], ce["z"][sl], ce["m"][sl], ce["eps"][sl])
        detected = g.evaluate(None, st["x"], st["y"], st["z"])
        g.set_block_kernel(False)
        exp = g.evaluate(None, st["x"], st["y"], st["z"])
        g.set_blocks(o, dxb, ce["m"][sl])
        gen = g.evaluate(None, st["x"], st["y"], st["z"])
        for p, q in zip(gen, exp):
            assert np.array_equal(p, q)
        g.set_block_kernel(True)
        col = g.evaluate(None, st["x"], st["y"], st["z"])
        for p, q in zip(detected, col):
            assert np.array_equal(p, q)
        for ns in (1, 7):
            g.set_config(nsplit=ns)
            again = g.evaluate(None, st["x"], st["y"], st["z"])
            s = _util.summary(again, col)
            assert s["acc_max"] <= (1e-12 if mode == "fp64" else 1e-5), (ns, s)
        g.set_config(nsplit=0)
        few = g.evaluate(0.001, st["x"][:37], st["y"][:37], st["z"][:37])   # ragged, eps_t > 0
    s = _util.summary(col, exp)
    print(mode, "block sources: column vs explicit", s, flush=True)
    if mode == "fp64":
        assert s["acc_max"] <= 1e-12 and s["pot_max"] <= 1e-12
    else:
        assert s["acc_median"] <= 1e-6 and s["acc_max"] <= 1e-4
    idx = np.union1d(_util.subsample(len(st["x"]), 200, 24), np.arange(37))
    ref = oracle.eval_plummer(ce["x"][sl], ce["y"][sl], ce["z"][sl], ce["m"][sl], ce["eps"][sl],
                              st["x"][idx], st["y"][idx], st["z"][idx], None, G, mode=oracle.NEUMAIER,
                              threads=THREADS)
    s = _util.summary(tuple(a[idx] for a in col), ref)
    assert s["acc_max"] <= (1e-12 if mode == "fp64" else 1e-4), s
    ref37 = oracle.eval_plummer(ce["x"][sl], ce["y"][sl], ce["z"][sl], ce["m"][sl], ce["eps"][sl],
                                st["x"][:37], st["y"][:37], st["z"][:37], np.full(37, 0.001), G,
                                mode=oracle.NEUMAIER, threads=THREADS)
    s = _util.summary(few, ref37)
    assert s["acc_max"] <= (1e-12 if mode == "fp64" else 1e-4), s


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_explicit_block_source_detection(mode, c2):
    """Explicit sources are taken as 8^3 blocks exactly when every aligned group of 512 is a
    C-ordered tensor-product block (x from i only, y from j only, z from k only) with one eps;
    otherwise the generic kernel runs, bit-identical to the generic kernel forced on.  With
    explicit block-ordered targets as well (cells onto cells), the block-source kernel wins."""
    st, ce = c2["stars"], c2["cells"]
    nb = 64
    sl = slice(0, nb * 512)
    o, dxb = ce["block_origin"][:nb], ce["block_size"][:nb] / 8
    S = [ce[k][sl].copy() for k in ("x", "y", "z", "m", "eps")]
    T = (None, st["x"], st["y"], st["z"])
    with GasGravityField(mode) as g:
        g.set_blocks(o, dxb, S[3])
        blk = g.evaluate(*T)
        cells_t = (ce["eps"][:4096], ce["x"][:4096], ce["y"][:4096], ce["z"][:4096])
        blk_cells = g.evaluate(*cells_t)
        g.set_sources(*S)
        for p, q in zip(g.evaluate(*T), blk):
            assert np.array_equal(p, q)
        for p, q in zip(g.evaluate(*cells_t), blk_cells):
            assert np.array_equal(p, q)
        bad = [a.copy() for a in S]
        bad[1][17 * 512 + 3 * 64 + 2 * 8 + 5] += 1e-9        # one y leaves its row
        bad_eps = [a.copy() for a in S]
        bad_eps[4][9 * 512 + 300] *= 2.0                     # eps varies inside a block
        for srcs in (bad, bad_eps, [a[:-512 + 3] for a in S]):   # the last: not whole blocks
            g.set_sources(*srcs)
            got = g.evaluate(*T)
            g.set_block_kernel(False)
            gen = g.evaluate(*T)
            g.set_block_kernel(True)
            for p, q in zip(got, gen):
                assert np.array_equal(p, q)


def test_block_sources_errors():
    with GasGravityField("fp64") as g:
        lib = _lib.load()
        assert lib.tgrav_set_source_blocks(g.handle, 2, 4, None, None, None, None, None, None, 0) == _lib.EINVAL
        assert lib.tgrav_set_source_blocks(g.handle, -1, 8, None, None, None, None, None, None, 0) == _lib.EINVAL
        with pytest.raises(TgravError):
            g.set_blocks(np.zeros((1, 3)), np.array([0.1]), np.full(512, np.nan))


def test_library_applies_the_tune_cache(c2):
    """tgrav_create with the autotuner's cache: the block forward launch takes the cached source
    split (bit-identical to setting it explicitly); an explicit tgrav_set_config wins over the
    cache; and another split gives other bits, so the comparisons can tell.  FP64, where the
    split changes the order of the FP64 sums (the mixed kernel's FP32 tile sums add exactly in
    FP64, so its results do not depend on the split at all)."""
    import json
    from torch_amd import autotune
    cache = os.path.join(HERE, "..", "torch_amd", "tune_cache.json")
    with GravityField("fp64") as f0:          # the block kernels' entries are keyed by their waves
        suf = autotune.waves_suffix(f0.column_waves())
    fwd, src = f"|blockfwd{suf}|fp64|s13|b11", f"|blocksrc{suf}|fp64|s13|b11"
    tuned = [v["nsplit"] for k, v in json.load(open(cache)).items() if fwd in k][0]
    other = 5 if tuned != 5 else 7
    st, ce = c2["stars"], c2["cells"]
    bo, bd = ce["block_origin"], ce["block_size"] / 8
    res = []
    for tune, split in ((cache, None), (None, tuned), (None, other), (cache, other)):
        with GravityField("fp64", tune_cache=tune) as f:
            if split is not None:
                f.set_config(nsplit=split)
            f.set_sources(st["x"], st["y"], st["z"], st["m"], st["eps"])
            res.append(f.evaluate_blocks(bo, bd))
    for q in range(4):
        np.testing.assert_array_equal(res[0][q], res[1][q])          # cache == explicit tuned split
        np.testing.assert_array_equal(res[2][q], res[3][q])          # explicit split wins over the cache
    assert any(not np.array_equal(res[0][q], res[2][q]) for q in range(4))
    # block sources: the entry is keyed by the target (star) count and the block count
    tuned_s = [v["nsplit"] for k, v in json.load(open(cache)).items() if src in k][0]
    res = []
    for tune, split in ((cache, None), (None, tuned_s), (None, 64)):
        with GasGravityField("fp64", tune_cache=tune) as g:
            if split is not None:
                g.set_config(nsplit=split)
            g.set_blocks(bo, bd, ce["m"])
            res.append(g.evaluate(None, st["x"], st["y"], st["z"]))
    for q in range(4):
        np.testing.assert_array_equal(res[0][q], res[1][q])
    assert any(not np.array_equal(res[0][q], res[2][q]) for q in range(4))
    with pytest.raises(TgravError):
        GravityField("fp64", tune_cache=os.path.join(HERE, "conftest.py"))
