"""Property-based tests of the oracle (hypothesis, derandomized; CPU only).

The same lattice inputs as the GPU property tests (spacing 1/64 pc): displacements are exact,
so translation by a lattice vector and scaling the masses by 4 are exact properties of the
sequential FP64 sum; superposition, the summation modes and Newton's third law hold to the
rounding of the sums, measured against the sum of |terms|.
"""
import os

import numpy as np
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

import oracle
from tests import _util
from tests.test_gpu_properties import SHIFT, cloud, masses, points, softening, sq

PROP = settings(max_examples=int(os.environ.get("TGRAV_PROPERTY_EXAMPLES", 40)), derandomize=True, deadline=None, database=None,
                suppress_health_check=list(HealthCheck))


def draw_problem(src, tgt):
    rng = np.random.Generator(np.random.PCG64(src["seed"]))
    s, m, se = points(src, rng), masses(src, rng), softening(src, rng)
    trng = np.random.Generator(np.random.PCG64(tgt["seed"]))
    t, te = points(tgt, trng), softening(tgt, trng)
    return s, m, se, t, te


def ev(s, m, se, t, te, G, mode=oracle.SEQ):
    return oracle.eval_plummer(s[:, 0], s[:, 1], s[:, 2], m, se, t[:, 0], t[:, 1], t[:, 2], te, G, mode=mode)


@PROP
@given(src=cloud(200), tgt=cloud(200))
def test_translation_and_mass_scaling_are_exact(G, src, tgt):
    s, m, se, t, te = draw_problem(src, tgt)
    base = ev(s, m, se, t, te, G)
    moved = ev(s + SHIFT, 4.0 * m, se, t + SHIFT, te, G)
    for p, q in zip(moved, base):
        np.testing.assert_array_equal(p, 4.0 * q)


@PROP
@given(src=cloud(200, nmin=1), tgt=cloud(200), cut=st.floats(0.0, 1.0))
def test_superposition_and_summation_modes(G, src, tgt, cut):
    s, m, se, t, te = draw_problem(src, tgt)
    k = int(cut * len(s))
    e = np.zeros(len(s)) if se is None else se
    whole = ev(s, m, e, t, te, G, oracle.LONGDOUBLE)
    parts = [ev(s[a:b], m[a:b], e[a:b], t, te, G, oracle.NEUMAIER) for a, b in ((0, k), (k, len(s)))]
    summed = tuple(p + q for p, q in zip(*parts))
    bounds = _util.term_scale_bounds(s, m, sq(se, len(s)), t, sq(te, len(t)), G, "fp64", None)
    for mode in (oracle.SEQ, oracle.NEUMAIER):
        ok, worst = _util.within_bounds(ev(s, m, e, t, te, G, mode), whole, bounds)
        assert ok, (mode, worst)
    ok, worst = _util.within_bounds(summed, whole, bounds)
    assert ok, worst


@PROP
@given(cel=cloud(150, nmin=1), par=cloud(150, nmin=1))
def test_fused_pass_obeys_the_third_law(G, cel, par):
    """Symmetric softening eps2 = ce^2 + pe^2: sum m_p a_p + sum m_c a_c = 0 to rounding, and
    each side equals the directed sum."""
    rng = np.random.Generator(np.random.PCG64(cel["seed"]))
    c, cm, ce = points(cel, rng), np.abs(masses(cel, rng)), softening(cel, rng)
    prng = np.random.Generator(np.random.PCG64(par["seed"]))
    p, pm, pe = points(par, prng), np.abs(masses(par, prng)), softening(par, prng)
    ce = np.zeros(len(c)) if ce is None else ce
    pe = np.zeros(len(p)) if pe is None else pe
    rp, rc = oracle.bidirectional(c[:, 0], c[:, 1], c[:, 2], cm, ce, p[:, 0], p[:, 1], p[:, 2], pm, pe, G)
    bp = _util.term_scale_bounds(c, cm, ce * ce, p, pe * pe, G, "fp64", None)
    bc = _util.term_scale_bounds(p, pm, pe * pe, c, ce * ce, G, "fp64", None)
    assert _util.within_bounds(rp, ev(c, cm, ce, p, pe, G, oracle.LONGDOUBLE), bp)[0]
    assert _util.within_bounds(rc, ev(p, pm, pe, c, ce, G, oracle.LONGDOUBLE), bc)[0]
    fp = np.stack(rp[:3], 1) * pm[:, None]
    fc = np.stack(rc[:3], 1) * cm[:, None]
    scale = (bp[0] * pm).sum() + (bc[0] * cm).sum()
    assert np.all(np.abs(fp.sum(0) + fc.sum(0)) <= 1e3 * scale + 1e-300)
