"""Pins the CPU oracle (oracle/tgrav_oracle.c) -- CPU only.

There are no golden vectors for this path in the reference, so the oracle is pinned by
(1) a 60-digit decimal brute force (committed KATs), (2) closed forms, (3) invariants and
(4) special cases with textbook answers.  Each check is chosen so that a plausible
mistake -- a dropped term, a wrong sign, a wrong index or a transposed operand -- fails at
least one of them; test_planted_mistakes_are_caught shows that for mistakes planted in
a copy of the oracle's output.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import brute
from torch_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
U = 2.0 ** -53


def load_kats():
    d = np.load(os.path.join(HERE, "golden", "plummer_kat.npz"))
    out = []
    for i in range(int(d["n_cases"])):
        out.append({k: d[f"c{i}_{k}"] for k in ("s", "m", "se", "t", "te", "expected")})
    return out, float(d["G"])


def abs_terms(s, m, se, t, te, G):
    """Per target: sum_j |G m_j w_ij d_ij| (3 components) and sum_j G m_j / r_ij."""
    d = s[None, :, :] - t[:, None, :]
    r2 = (d ** 2).sum(-1) + se[None, :] ** 2 + te[:, None] ** 2
    with np.errstate(divide="ignore"):
        rinv = np.where(r2 > 0, 1.0 / np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0)
    w = G * m[None, :] * rinv ** 3
    return np.abs(w[..., None] * d).sum(1), (G * m[None, :] * rinv).sum(1)


def run_oracle(c, G, mode=oracle.SEQ):
    s, t = c["s"], c["t"]
    return np.stack(oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], c["m"], c["se"],
                                        t[:, 0], t[:, 1], t[:, 2], c["te"], G, mode=mode), 1)


def kat_error_ok(got, c, G, k_abs):
    """|got - exact| <= k_abs * u * (sum of |terms|) + 4 u |exact|, per component."""
    exp = c["expected"]
    sa, sp = abs_terms(c["s"], c["m"], c["se"], c["t"], c["te"], G)
    scale = np.concatenate([sa, sp[:, None]], 1)
    bound = k_abs * U * scale + 4 * U * np.abs(exp) + 1e-300
    return np.all(np.abs(got - exp) <= bound), np.max(np.abs(got - exp) / bound)


@pytest.mark.parametrize("mode,k", [(oracle.SEQ, 64.0), (oracle.NEUMAIER, 8.0),
                                    (oracle.LONGDOUBLE, 2.0)])
def test_oracle_matches_decimal_brute_force(mode, k):
    kats, G = load_kats()
    for i, c in enumerate(kats):
        ok, worst = kat_error_ok(run_oracle(c, G, mode), c, G, k)
        assert ok, f"KAT case {i}: error {worst:.2f}x the bound (mode {mode})"


def test_brute_force_is_reproducible():
    """The committed KATs are what oracle/brute.py computes (guards the fixture itself)."""
    kats, G = load_kats()
    c = kats[1]
    s, t = c["s"], c["t"]
    res = brute.eval_plummer(s[:4, 0], s[:4, 1], s[:4, 2], c["m"][:4], c["se"][:4],
                             t[:3, 0], t[:3, 1], t[:3, 2], c["te"][:3], G)
    full = run_oracle(dict(c, s=s[:4], m=c["m"][:4], se=c["se"][:4], t=t[:3], te=c["te"][:3]),
                      G, oracle.LONGDOUBLE)
    np.testing.assert_allclose(np.array([[float(v) for v in r] for r in res]), full,
                               rtol=1e-15, atol=0)


def test_planted_mistakes_are_caught():
    """Mistakes planted in the oracle's result must fail the KAT check."""
    kats, G = load_kats()
    c = kats[1]
    good = run_oracle(c, G, oracle.NEUMAIER)
    assert kat_error_ok(good, c, G, 8.0)[0]
    # dropped term: leave out source 7
    keep = np.arange(len(c["m"])) != 7
    dropped = run_oracle(dict(c, s=c["s"][keep], m=c["m"][keep], se=c["se"][keep]), G)
    assert not kat_error_ok(dropped, c, G, 64.0)[0]
    # wrong sign of the displacement (x_t - x_s)
    assert not kat_error_ok(good * np.array([-1, -1, -1, 1]), c, G, 64.0)[0]
    # wrong sign of the potential
    assert not kat_error_ok(good * np.array([1, 1, 1, -1]), c, G, 64.0)[0]
    # transposed operand: axes swapped
    assert not kat_error_ok(good[:, [1, 0, 2, 3]], c, G, 64.0)[0]
    # eps applied per target instead of per source
    swapped = run_oracle(dict(c, se=np.resize(c["te"], len(c["m"])), te=np.resize(c["se"], len(c["te"]))), G)
    assert not kat_error_ok(swapped, c, G, 64.0)[0]


def test_single_source_closed_form(G):
    rng = W.rng_for(3)
    n = 500
    s = rng.normal(size=(n, 3))
    t = rng.normal(size=(n, 3)) * 3
    m = rng.uniform(0.1, 50, n)
    es = rng.uniform(0, 0.3, n)
    et = rng.uniform(0, 0.3, n)
    for i in range(n):
        ax, ay, az, ph = oracle.eval_plummer(s[i:i + 1, 0], s[i:i + 1, 1], s[i:i + 1, 2], m[i:i + 1],
                                             es[i:i + 1], t[i:i + 1, 0], t[i:i + 1, 1],
                                             t[i:i + 1, 2], et[i:i + 1], G)
        d = s[i] - t[i]
        r2 = d @ d + es[i] ** 2 + et[i] ** 2
        a = G * m[i] * d / r2 ** 1.5
        np.testing.assert_allclose([ax[0], ay[0], az[0]], a, rtol=4e-15, atol=0)
        np.testing.assert_allclose(ph[0], -G * m[i] / np.sqrt(r2), rtol=4e-15)


def _cluster(seed, ns=300, nt=200):
    rng = W.rng_for(seed)
    st = W.star_cluster(rng, ns, binary_frac=0.2)
    t = rng.uniform(-2, 2, (nt, 3))
    te = rng.uniform(0, 0.02, nt)
    return st, t, te, rng


def test_linearity_in_mass_is_exact(G):
    st, t, te, _ = _cluster(11)
    a1 = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], t[:, 0], t[:, 1], t[:, 2], te, G)
    a2 = oracle.eval_plummer(st["x"], st["y"], st["z"], 2 * st["m"], st["eps"], t[:, 0], t[:, 1], t[:, 2], te, G)
    for p, q in zip(a1, a2):
        assert np.array_equal(2 * p, q)      # scaling by 2 is exact in binary FP


def test_superposition(G):
    st, t, te, _ = _cluster(12)
    n = len(st["m"])
    full = np.stack(oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"],
                                        t[:, 0], t[:, 1], t[:, 2], te, G, mode=oracle.NEUMAIER))
    parts = []
    for sl in (slice(0, n // 3), slice(n // 3, n)):
        parts.append(np.stack(oracle.eval_plummer(st["x"][sl], st["y"][sl], st["z"][sl], st["m"][sl],
                                                  st["eps"][sl], t[:, 0], t[:, 1], t[:, 2], te, G,
                                                  mode=oracle.NEUMAIER)))
    np.testing.assert_allclose(parts[0] + parts[1], full, rtol=1e-13, atol=1e-13 * np.abs(full).max())


def test_translation_and_permutation_invariance(G):
    st, t, te, rng = _cluster(13)
    base = np.stack(oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"],
                                        t[:, 0], t[:, 1], t[:, 2], te, G, mode=oracle.NEUMAIER))
    off = 0.75  # shifts by a power-of-two multiple keep the differences exact here
    sh = np.stack(oracle.eval_plummer(st["x"] + off, st["y"] - off, st["z"] + 2 * off, st["m"], st["eps"],
                                      t[:, 0] + off, t[:, 1] - off, t[:, 2] + 2 * off, te, G,
                                      mode=oracle.NEUMAIER))
    scale = np.abs(base).max(axis=1, keepdims=True)
    assert np.all(np.abs(sh - base) <= 1e-12 * scale)
    p = rng.permutation(len(st["m"]))
    pr = np.stack(oracle.eval_plummer(st["x"][p], st["y"][p], st["z"][p], st["m"][p], st["eps"][p],
                                      t[:, 0], t[:, 1], t[:, 2], te, G, mode=oracle.NEUMAIER))
    assert np.all(np.abs(pr - base) <= 1e-14 * scale)


def test_newton_third_law(G):
    """sum_t m_t a_t(S) = -sum_s m_s a_s(T) with the symmetric eps2 = eps_s^2 + eps_t^2 (the
    reference's force-balance diagnostic, interface.F90:1467-1522)."""
    rng = W.rng_for(14)
    S = W.star_cluster(rng, 400)
    T = W.star_cluster(rng, 300, r_vir=0.7)
    T["eps"] = rng.uniform(0.001, 0.05, 300)
    aT = oracle.eval_plummer(S["x"], S["y"], S["z"], S["m"], S["eps"], T["x"], T["y"], T["z"], T["eps"], G,
                             mode=oracle.NEUMAIER)
    aS = oracle.eval_plummer(T["x"], T["y"], T["z"], T["m"], T["eps"], S["x"], S["y"], S["z"], S["eps"], G,
                             mode=oracle.NEUMAIER)
    for q in range(3):
        fT = np.sum(T["m"] * aT[q])
        fS = np.sum(S["m"] * aS[q])
        scale = np.sum(np.abs(T["m"] * aT[q])) + np.sum(np.abs(S["m"] * aS[q]))
        assert abs(fT + fS) <= 1e-13 * scale


def test_acceleration_is_minus_grad_potential(G):
    st, t, te, _ = _cluster(15, ns=200, nt=50)
    h = 1e-5
    a = np.stack(oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"],
                                     t[:, 0], t[:, 1], t[:, 2], te, G, mode=oracle.LONGDOUBLE))
    for q in range(3):
        tp, tm = t.copy(), t.copy()
        tp[:, q] += h
        tm[:, q] -= h
        pp = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], tp[:, 0], tp[:, 1], tp[:, 2],
                                 te, G, mode=oracle.LONGDOUBLE)[3]
        pm = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"], tm[:, 0], tm[:, 1], tm[:, 2],
                                 te, G, mode=oracle.LONGDOUBLE)[3]
        grad = (pp - pm) / (2 * h)
        np.testing.assert_allclose(a[q], -grad, rtol=1e-5, atol=1e-6 * np.abs(a[:3]).max())


def test_far_field_and_zero_softening_limits(G):
    rng = W.rng_for(16)
    s = rng.uniform(-0.01, 0.01, (50, 3))
    m = rng.uniform(1, 2, 50)
    t = np.array([[1000.0, 0.0, 0.0], [0.0, -2000.0, 0.0]])
    a = np.stack(oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, np.full(50, 0.01),
                                     t[:, 0], t[:, 1], t[:, 2], None, G))
    com = (m[:, None] * s).sum(0) / m.sum()
    for i in range(2):
        d = com - t[i]
        r = np.linalg.norm(d)
        exp = G * m.sum() * d / r ** 3
        np.testing.assert_allclose(a[:3, i], exp, rtol=0, atol=1e-6 * np.linalg.norm(exp))
        np.testing.assert_allclose(a[3, i], -G * m.sum() / r, rtol=1e-9)
    # eps -> 0: converges to the unsoftened sum
    t2 = rng.uniform(-1, 1, (20, 3))
    a0 = np.stack(oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, None, t2[:, 0], t2[:, 1], t2[:, 2], None, G))
    ae = np.stack(oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, np.full(50, 1e-9),
                                      t2[:, 0], t2[:, 1], t2[:, 2], None, G))
    np.testing.assert_allclose(ae, a0, rtol=1e-12)


def test_symmetric_pair_and_self_exclusion(G):
    a = oracle.eval_plummer(np.array([-1.0, 1.0]), np.zeros(2), np.zeros(2), np.array([3.0, 3.0]), None,
                            np.array([0.0, -1.0]), np.zeros(2), np.zeros(2), None, G)
    assert a[0][0] == 0.0 and a[1][0] == 0.0 and a[2][0] == 0.0
    # target 1 coincides with source 0 (eps = 0): only source 1 acts, at distance 2
    np.testing.assert_allclose(a[0][1], G * 3.0 / 4.0, rtol=1e-15)
    np.testing.assert_allclose(a[3][1], -G * 3.0 / 2.0, rtol=1e-15)


def test_plummer_sphere_potential_statistical(G):
    """A many-body Plummer sphere approaches Phi = -G M / sqrt(r^2 + b^2)."""
    rng = W.rng_for(17)
    n = 20000
    pos, b = W.plummer_positions(rng, n, r_vir=1.0, mfrac=1.0 - 1e-9)
    m = np.full(n, 1.0 / n)
    r = np.array([0.0, 0.3, 1.0, 3.0])
    t = np.stack([r, np.zeros(4), np.zeros(4)], 1)
    phi = oracle.eval_plummer(pos[:, 0], pos[:, 1], pos[:, 2], m, None, t[:, 0], t[:, 1], t[:, 2], None, 1.0,
                              threads=4)[3]
    np.testing.assert_allclose(phi, -1.0 / np.sqrt(r ** 2 + b ** 2), rtol=0.03)


def test_flash_kernel_limits(G):
    rs = 0.05
    s = np.zeros((1, 3))
    m = np.array([2.0])
    r = np.array([1e-9, 0.2 * rs, 0.5 * rs, 0.7 * rs, rs * (1 - 1e-12), rs, 2 * rs])
    t = np.stack([r, np.zeros_like(r), np.zeros_like(r)], 1)
    for kernel in ("spline", "linear"):
        ax, _, _, phi = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rs,
                                          kernel=kernel, quirks=False)
        # r >= rsoft: Newtonian; continuous at rsoft from inside
        np.testing.assert_allclose(ax[5:], -G * 2.0 / r[5:] ** 2, rtol=1e-14)
        np.testing.assert_allclose(phi[5:], -G * 2.0 / r[5:], rtol=1e-14)
        np.testing.assert_allclose(ax[4], ax[5], rtol=1e-9)
        np.testing.assert_allclose(phi[4], phi[5], rtol=1e-9)
        assert abs(ax[0]) < 1e-6 * abs(ax[5])            # r -> 0: force -> 0
        # inside: a = -dphi/dr (central differences)
        h = 1e-7
        tp, tm = t[1:4].copy(), t[1:4].copy()
        tp[:, 0] += h
        tm[:, 0] -= h
        pp = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, tp[:, 0], tp[:, 1], tp[:, 2], G, rs, kernel=kernel)[3]
        pm = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, tm[:, 0], tm[:, 1], tm[:, 2], G, rs, kernel=kernel)[3]
        np.testing.assert_allclose(ax[1:4], -(pp - pm) / (2 * h), rtol=1e-6)
    # quirks: the reference's spline potential flips sign inside the radius,
    # and its linear branch returns prefactor * slope
    _, _, _, pq = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rs,
                                    kernel="spline", quirks=True)
    assert np.all(pq[1:4] > 0) and np.all(pq[5:] < 0)
    _, _, _, pl = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rs,
                                    kernel="linear", quirks=True)
    np.testing.assert_allclose(pl[1:4], -G * 2.0 / rs ** 3, rtol=1e-15)


def _quirk_pin_error(eval_flash, G, rs, m=2.0):
    """Largest violation, in units of the term scale G m hinv, of the identity between the
    reference's spline potential inside rsoft (Gravity_getPotentialAtPoint.F90:202-211) and the
    clean one: quirk + clean kernel = hinv q^4 / 5 for q < 1 and 0 for 1 <= q < 2, where the
    clean potential is pinned independently (a = -grad phi, continuity with -1/r at rsoft)."""
    h = rs / 2.0
    hinv = 1.0 / h
    q = np.array([0.01, 0.1, 0.3, 0.5, 0.8, 0.99, 1.0, 1.2, 1.5, 1.9, 1.99])
    r = q * h
    t = np.stack([r, np.zeros_like(r), np.zeros_like(r)], 1)
    s = np.zeros((1, 3))
    mm = np.array([m])
    pq = eval_flash(s[:, 0], s[:, 1], s[:, 2], mm, t[:, 0], t[:, 1], t[:, 2], G, rs, kernel="spline",
                    quirks=True)[3]
    pc = eval_flash(s[:, 0], s[:, 1], s[:, 2], mm, t[:, 0], t[:, 1], t[:, 2], G, rs, kernel="spline",
                    quirks=False)[3]
    # phi = -G m * kernel for one source, so the kernels' identity reads on -phi / (G m)
    got = -(pq + pc) / (G * m)
    want = np.where(q < 1.0, 0.2 * hinv * q ** 4, 0.0)
    return float(np.max(np.abs(got - want)) / hinv)


def test_flash_quirk_potential_pinned(G, tmp_path):
    """The quirk potential inside rsoft is pinned by its sum with the clean one (not only by its
    sign); a mistyped coefficient planted in a compiled copy of the oracle must fail the pin."""
    assert _quirk_pin_error(oracle.eval_flash, G, 0.05) < 1e-14
    assert _quirk_pin_error(oracle.eval_flash, G, 3.0) < 1e-14
    src = open(os.path.join(os.path.dirname(oracle.__file__), "tgrav_oracle.c")).read()
    quirk_line = "pk = hinv * (2.0 / 3.0 * q * q - 0.1 * pow(q, 4) + 0.1 * pow(q, 5) - 7.0 / 5.0);"
    assert src.count(quirk_line) == 1
    plants = {"q4 coefficient 0.1 -> 0.3": quirk_line.replace("0.1 * pow(q, 4)", "0.3 * pow(q, 4)"),
              "q5 coefficient 0.1 -> 0.3": quirk_line.replace("0.1 * pow(q, 5)", "0.3 * pow(q, 5)"),
              "constant 7/5 -> 6/5": quirk_line.replace("7.0 / 5.0", "6.0 / 5.0")}
    import subprocess
    for i, (name, bad) in enumerate(plants.items()):
        c_path = tmp_path / f"planted{i}.c"
        so_path = tmp_path / f"libplanted{i}.so"
        c_path.write_text(src.replace(quirk_line, bad))
        subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-o", str(so_path),
                        str(c_path), "-lm"], check=True)
        planted = oracle.load(str(so_path))
        saved = oracle._LIB
        oracle._LIB = planted
        try:
            err = _quirk_pin_error(oracle.eval_flash, G, 0.05)
        finally:
            oracle._LIB = saved
        assert err > 1e-3, f"planted mistake '{name}' passed the quirk pin ({err})"


def test_flash_spline_matches_closed_form_table(G):
    """Spline force kernel at q = 0.5 and 1.5 written out by hand (Monaghan & Lattanzio)."""
    rs = 0.1
    h = rs / 2
    for q in (0.5, 1.5):
        r = q * h
        if q < 1:
            f = (4 / 3 * q - 1.2 * q ** 3 + 0.5 * q ** 4) / h ** 2
        else:
            f = (8 / 3 * q - 3 * q ** 2 + 1.2 * q ** 3 - q ** 4 / 6 - 1 / (15 * q ** 2)) / h ** 2
        ax = oracle.eval_flash(np.zeros(1), np.zeros(1), np.zeros(1), np.ones(1), np.array([r]), np.zeros(1),
                               np.zeros(1), 1.0, rs)[0]
        np.testing.assert_allclose(ax[0], -f, rtol=1e-14)


def _images_in_cutoff(d_inside, box, nrep):
    """Image offsets n*L (n in [-nrep, nrep]^3, nz fastest) of one pair whose radius is below
    nrep * min(box), and the margin of the closest one to the cutoff."""
    r = np.arange(-nrep, nrep + 1)
    n = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    d = d_inside[None, :] + n * np.asarray(box)[None, :]
    rad = np.sqrt((d ** 2).sum(1))
    rmax = nrep * min(box)
    return n[rad < rmax], np.min(np.abs(rad - rmax)) / rmax


@pytest.mark.parametrize("kernel", ["spline", "linear"])
@pytest.mark.parametrize("quirks", [False, True])
def test_flash_periodic_equals_isolated_over_replicated_images(G, kernel, quirks):
    """The periodic branch = the (closed-form-pinned) isolated kernel applied to every image
    of every cell inside the cutoff, built independently with numpy; the quirk potential
    keeps only the last such image in loop order."""
    rng = W.rng_for(31)
    box, nrep, rs = (1.0, 1.3, 0.9), 3, 0.35
    ns, nt = 5, 4
    s = rng.uniform(0, 1, (ns, 3)) * box
    t = rng.uniform(0, 1, (nt, 3)) * box
    t[0] = s[0] + np.array([0.05, -0.02, 0.03])          # one pair inside the softening radius
    m = rng.uniform(0.5, 2.0, ns)
    got = oracle.eval_flash(s[:, 0], s[:, 1], s[:, 2], m, t[:, 0], t[:, 1], t[:, 2], G, rs,
                            kernel=kernel, quirks=quirks, box=box, nrep=nrep)
    for i in range(nt):
        img_x, img_m, last_x, last_m = [], [], [], []
        for j in range(ns):
            n, margin = _images_in_cutoff(t[i] - s[j], box, nrep)
            assert margin > 1e-9
            # an image at offset n*L is the cell moved to x_s - n*L
            pos = s[j][None, :] - n * np.asarray(box)[None, :]
            img_x.append(pos)
            img_m.append(np.full(len(n), m[j]))
            last_x.append(pos[-1:])
            last_m.append(m[j:j + 1])
        X, M = np.concatenate(img_x), np.concatenate(img_m)
        ref = oracle.eval_flash(X[:, 0], X[:, 1], X[:, 2], M, t[i:i + 1, 0], t[i:i + 1, 1], t[i:i + 1, 2],
                                G, rs, kernel=kernel, quirks=quirks)
        for c in range(3):
            np.testing.assert_allclose(got[c][i], ref[c][0], rtol=1e-11, atol=1e-14 * abs(ref[c][0]) + 1e-300)
        if quirks:
            LX, LM = np.concatenate(last_x), np.concatenate(last_m)
            ref_p = oracle.eval_flash(LX[:, 0], LX[:, 1], LX[:, 2], LM, t[i:i + 1, 0], t[i:i + 1, 1],
                                      t[i:i + 1, 2], G, rs, kernel=kernel, quirks=True)[3]
            np.testing.assert_allclose(got[3][i], ref_p[0], rtol=1e-12)
        else:
            np.testing.assert_allclose(got[3][i], ref[3][0], rtol=1e-11)


def test_flash_periodic_symmetry_points(G):
    """Half a box length from a lone cell the image lattice (cut by a sphere about the target)
    is mirror-symmetric: the field vanishes; a huge box reduces to the isolated field."""
    ax, ay, az, _ = oracle.eval_flash(np.zeros(1), np.zeros(1), np.zeros(1), np.ones(1),
                                      np.array([0.5, 0.5]), np.array([0.0, 0.5]), np.array([0.0, 0.5]),
                                      G, 0.1, box=(1.0, 1.0, 1.0), nrep=4)
    assert np.max(np.abs([ax, ay, az])) < 1e-13 * G
    s = np.array([0.2]), np.array([0.1]), np.array([-0.3])
    t = np.array([0.25, 1.0]), np.array([0.1, -0.4]), np.array([-0.3, 0.7])
    iso = oracle.eval_flash(*s, np.array([3.0]), *t, G, 0.1)
    per = oracle.eval_flash(*s, np.array([3.0]), *t, G, 0.1, box=(1e7, 1e7, 1e7), nrep=2)
    scale = max(np.max(np.abs(iso[c])) for c in range(3))
    for c in range(3):
        np.testing.assert_allclose(per[c], iso[c], rtol=1e-9, atol=1e-12 * scale)


def test_bidirectional_equals_two_directed_sums_and_third_law(G):
    rng = W.rng_for(18)
    cells = W.gas_cells(rng, box=4.0, base_n=2, n_leaves=2, sigma=1.0, ngrid=16, rho_mean=10.0)
    st = W.star_cluster(rng, 150)
    (pa, pb, pc, pp), (ca, cb, cc, cp) = oracle.bidirectional(
        cells["x"], cells["y"], cells["z"], cells["m"], cells["eps"],
        st["x"], st["y"], st["z"], st["m"], st["eps"], G)
    on_p = oracle.eval_plummer(cells["x"], cells["y"], cells["z"], cells["m"], cells["eps"],
                               st["x"], st["y"], st["z"], st["eps"], G, mode=oracle.NEUMAIER)
    on_c = oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"],
                               cells["x"], cells["y"], cells["z"], cells["eps"], G, mode=oracle.NEUMAIER)
    for got, exp in zip((pa, pb, pc, pp, ca, cb, cc, cp), on_p + on_c):
        np.testing.assert_allclose(got, exp, rtol=1e-11, atol=1e-12 * np.abs(exp).max())
    for gp, gc in ((pa, ca), (pb, cb), (pc, cc)):
        scale = np.abs(st["m"] * gp).sum() + np.abs(cells["m"] * gc).sum()
        assert abs(np.sum(st["m"] * gp) + np.sum(cells["m"] * gc)) <= 1e-13 * scale


def test_kick_is_exact():
    a = np.array([1.0, -2.0, 0.5])
    v = oracle.kick(0.25, a, 2 * a, -a, np.ones(3), np.zeros(3), np.full(3, 3.0))
    np.testing.assert_array_equal(v[0], 1.0 + 0.25 * a)
    np.testing.assert_array_equal(v[1], 0.5 * a)
    np.testing.assert_array_equal(v[2], 3.0 - 0.25 * a)


def test_block_centres_match_meshgrid_c_order():
    rng = W.rng_for(19)
    o = rng.uniform(-1, 1, (3, 3))
    dx = np.array([0.1, 0.05, 0.2])
    x, y, z = oracle.block_centres(o[:, 0], o[:, 1], o[:, 2], dx, nb=8)
    for b in range(3):
        c = [o[b, q] + (np.arange(8) + 0.5) * dx[b] for q in range(3)]
        X, Y, Z = np.meshgrid(*c, indexing="ij")
        sl = slice(b * 512, (b + 1) * 512)
        np.testing.assert_array_equal(x[sl], np.ravel(X, "C"))
        np.testing.assert_array_equal(y[sl], np.ravel(Y, "C"))
        np.testing.assert_array_equal(z[sl], np.ravel(Z, "C"))


def test_c1_golden_is_reproduced(G):
    """The committed c1 fixture: inputs regenerate bit for bit from seed 1001 and the plain
    sequential oracle agrees with the compensated one to <= 1e-12."""
    d = np.load(os.path.join(HERE, "golden", "c1_oracle.npz"))
    w = W.c1(int(d["seed"]))
    st, ce = w["stars"], w["cells"]
    for k, v in (("star_x", st["x"]), ("star_m", st["m"]), ("cell_z", ce["z"]), ("cell_m", ce["m"])):
        assert np.array_equal(d[k], v), k
    fwd = np.stack(oracle.eval_plummer(st["x"], st["y"], st["z"], st["m"], st["eps"],
                                       ce["x"], ce["y"], ce["z"], None, G, threads=4))
    err = np.abs(fwd - d["fwd"]).max(axis=0) / np.abs(d["fwd"]).max(axis=0)
    assert err.max() <= 1e-12
