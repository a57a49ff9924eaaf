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
