"""Shared helpers of the parity tests: error metrics against the oracle."""
import numpy as np


def rel_acc_err(got, ref):
    """Normwise relative acceleration error per target: |a - a_ref| / |a_ref|."""
    ga = np.stack(got[:3], 1)
    ra = np.stack(ref[:3], 1)
    num = np.linalg.norm(ga - ra, axis=1)
    den = np.linalg.norm(ra, axis=1)
    return num / np.where(den > 0, den, 1.0)


def rel_pot_err(got, ref):
    den = np.abs(ref[3])
    return np.abs(got[3] - ref[3]) / np.where(den > 0, den, 1.0)


def summary(got, ref):
    ea, ep = rel_acc_err(got, ref), rel_pot_err(got, ref)
    return dict(acc_median=float(np.median(ea)), acc_max=float(ea.max()),
                pot_median=float(np.median(ep)), pot_max=float(ep.max()))


def subsample(n, k, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.sort(rng.choice(n, size=min(k, n), replace=False))


def mixed_error_bound(s, m, se, t, te, G, fold=64):
    """Analytic per-target bound of the mixed-precision kernel's error when all targets share
    one tile (anchor c = centre of their bounding box):

    * coordinates are stored as fp32(x - c), each off by <= 2^-24 |x - c|, so a displacement
      is off by delta_ij <= 2^-24 (|x_s - c| + |x_t - c|) (the FP32 difference itself is exact
      for close pairs, Sterbenz, and within 2^-24 |d| otherwise);
    * the Plummer acceleration kernel m d / r^3 has Jacobian norm <= 2 m / r^3 and the potential
      gradient m |d| / r^3 <= m / r^2;
    * FP32 rsqrt/products/fold sums add <= (8 + fold) 2^-24 of the sum of |terms|.
    Returns (acc_bound[nt], pot_bound[nt])."""
    u = 2.0 ** -24
    c = 0.5 * (t.min(0) + t.max(0))
    d = s[None, :, :] - t[:, None, :]
    dd = (d ** 2).sum(-1)
    r2 = dd + (se[None, :] ** 2 if se is not None else 0) + (te[:, None] ** 2 if te is not None else 0)
    ok = r2 > 0
    r = np.sqrt(np.where(ok, r2, 1.0))
    delta = u * (np.linalg.norm(s - c, axis=1)[None, :] + np.linalg.norm(t - c, axis=1)[:, None]) + \
        u * np.sqrt(dd)
    gm = G * m[None, :]
    acc = np.where(ok, gm * (2.0 / r ** 3 * delta + (8 + fold) * u * np.sqrt(dd) / r ** 3), 0).sum(1)
    pot = np.where(ok, gm * (delta / r ** 2 + (8 + fold) * u / r), 0).sum(1)
    return acc, pot


def within_mixed_bound(got, exp, s, m, se, t, te, G):
    ab, pb = mixed_error_bound(s, m, se, t, te, G)
    ga = np.stack(got[:3], 1) if isinstance(got, tuple) else got[:, :3]
    ea = np.stack(exp[:3], 1) if isinstance(exp, tuple) else exp[:, :3]
    gp = got[3] if isinstance(got, tuple) else got[:, 3]
    ep = exp[3] if isinstance(exp, tuple) else exp[:, 3]
    err_a = np.linalg.norm(ga - ea, axis=1)
    err_p = np.abs(gp - ep)
    ratio = max(float(np.max(err_a / (ab + 1e-300))), float(np.max(err_p / (pb + 1e-300))))
    return ratio <= 1.0, ratio


def term_scale_bounds(s, m, se2, t, te2, G, mode, anchor, fold=512):
    """Per-target error bounds, condition-aware (relative to the sum of |terms|, so cancelling
    fields are compared fairly), for property tests on small inputs.

    s[ns, 3], m[ns], se2[ns] (source eps^2), t[nt, 3], te2[nt]; anchor: per-target (nt) or
    per-pair (nt, ns) bound on the distance from the target to the FP64 anchor its mixed
    kernel subtracts (any point of the target tile's bounding box, or of its leaf block).
    fp64: 1e-12 of sum_j G|m_j| |d_ij| / r_ij^3 (acceleration) and of sum_j G|m_j| / r_ij
    (potential).  mixed: the first-order bound of _util.mixed_error_bound with the anchor
    distance given, the FP32 partial sums over `fold` terms, and 1% slack for second order
    (inputs keep |d| >= 1/64, so the displacement error is below 3e-5 |d|).
    Returns (acc_bound[nt], pot_bound[nt])."""
    d = s[None, :, :] - t[:, None, :]
    dd = (d ** 2).sum(-1)
    r2 = dd + se2[None, :] + te2[:, None]
    ok = r2 > 0
    r = np.sqrt(np.where(ok, r2, 1.0))
    dn = np.sqrt(dd)
    gm = G * np.abs(m)[None, :]
    if mode == "fp64":
        acc = np.where(ok, gm * dn / r ** 3, 0).sum(1)
        pot = np.where(ok, gm / r, 0).sum(1)
        return 1e-12 * acc + 1e-300, 1e-12 * pot + 1e-300
    u = 2.0 ** -24
    A = anchor if np.ndim(anchor) == 2 else np.asarray(anchor)[:, None]
    delta = u * (2.0 * dn + 2.0 * A)
    acc = np.where(ok, gm * (2.0 * delta / r ** 3 + (8 + fold) * u * dn / r ** 3), 0).sum(1)
    pot = np.where(ok, gm * (delta / r ** 2 + (8 + fold) * u / r), 0).sum(1)
    return 1.01 * acc + 1e-300, 1.01 * pot + 1e-300


def within_bounds(got, ref, bounds, potential=True):
    """True if |a - a_ref| <= acc bound and |phi - phi_ref| <= pot bound for every target;
    also returns the worst ratio."""
    ab, pb = bounds
    ea = np.linalg.norm(np.stack(got[:3], 1) - np.stack(ref[:3], 1), axis=1)
    worst = float(np.max(ea / ab, initial=0.0))
    if potential:
        worst = max(worst, float(np.max(np.abs(got[3] - ref[3]) / pb, initial=0.0)))
    return worst <= 1.0, worst
