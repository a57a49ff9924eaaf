"""Pins of the oracle's particle-mesh pieces (CPU): closed forms of cloud-in-cell weights and
central differences, conservation, exactness on linear / quadratic fields, and the adjoint
identity between deposit and interpolation."""
import numpy as np
import pytest

import oracle
from torch_amd import workloads as W


def blocks_of(nblk=3, nb=8, ng=2, seed=5):
    rng = W.rng_for(seed)
    return dict(origin=rng.uniform(-1, 1, (nblk, 3)), size=rng.uniform(0.5, 1.5, nblk), nb=nb, ng=ng)


def centre(B, b, i, j, k):
    """Centre of guarded cell (i, j, k) of block b."""
    d = B["size"][b] / B["nb"]
    return B["origin"][b] + (np.array([i, j, k]) - B["ng"] + 0.5) * d


def index(B, b, i, j, k):
    n = B["nb"] + 2 * B["ng"]
    return b * n ** 3 + (k * n + j) * n + i


def test_deposit_on_cell_centre_and_split():
    B = blocks_of()
    d = B["size"][1] / 8
    c = centre(B, 1, 3, 5, 7)
    rho = oracle.pm_deposit([c[0]], [c[1]], [c[2]], [2.0], [1], B)
    assert rho[index(B, 1, 3, 5, 7)] == pytest.approx(2.0 / d ** 3, rel=1e-14)
    assert np.count_nonzero(rho) == 1
    # a quarter cell along +x: 3/4 stays, 1/4 goes to i+1 (x is the fastest index)
    rho = oracle.pm_deposit([c[0] + 0.25 * d], [c[1]], [c[2]], [2.0], [1], B)
    np.testing.assert_allclose(rho[index(B, 1, 3, 5, 7)], 0.75 * 2.0 / d ** 3, rtol=1e-13)
    np.testing.assert_allclose(rho[index(B, 1, 4, 5, 7)], 0.25 * 2.0 / d ** 3, rtol=1e-13)
    # a quarter cell along -z: 1/4 to k-1 (stride nbg^2)
    rho = oracle.pm_deposit([c[0]], [c[1]], [c[2] - 0.25 * d], [2.0], [1], B)
    np.testing.assert_allclose(rho[index(B, 1, 3, 5, 6)], 0.25 * 2.0 / d ** 3, rtol=1e-13)


def test_deposit_conserves_mass_and_is_linear():
    B = blocks_of()
    rng = W.rng_for(6)
    n = 500
    blk = rng.integers(0, 3, n)
    pos = B["origin"][blk] + rng.uniform(0, 1, (n, 3)) * B["size"][blk][:, None]
    m = rng.uniform(0.1, 2.0, n)
    rho = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], m, blk, B)
    nbg = 8 + 2 * B["ng"]
    dv = np.repeat((B["size"] / 8) ** 3, nbg ** 3)
    for b in range(3):
        sl = slice(b * nbg ** 3, (b + 1) * nbg ** 3)
        np.testing.assert_allclose((rho[sl] * dv[sl]).sum(), m[blk == b].sum(), rtol=1e-13)
    rho2 = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], 3.0 * m, blk, B)
    np.testing.assert_allclose(rho2, 3.0 * rho, rtol=1e-14, atol=0)
    with pytest.raises(RuntimeError):       # cloud outside its guarded block
        oracle.pm_deposit([B["origin"][0, 0] + 5.0], [B["origin"][0, 1]], [B["origin"][0, 2]], [1.0], [0], B)


def test_gradient_exact_on_quadratics_and_end_rule():
    B = blocks_of(ng=3)
    n = 8 + 2 * 3
    nblk = 3
    phi = np.zeros(nblk * n ** 3)
    exp = [np.zeros_like(phi) for _ in range(3)]
    a, b2, c2 = 0.7, -1.3, 0.4
    for b in range(nblk):
        d = B["size"][b] / 8
        idx = np.arange(n) - 3 + 0.5
        X = B["origin"][b][0] + idx * d
        Y = B["origin"][b][1] + idx * d
        Z = B["origin"][b][2] + idx * d
        zz, yy, xx = np.meshgrid(Z, Y, X, indexing="ij")          # k slowest, i fastest
        sl = slice(b * n ** 3, (b + 1) * n ** 3)
        phi[sl] = (a * xx ** 2 + b2 * yy ** 2 + c2 * zz).ravel()
        exp[0][sl] = (-2 * a * xx).ravel()
        exp[1][sl] = (-2 * b2 * yy).ravel()
        exp[2][sl] = np.full(n ** 3, -c2)
    g = oracle.pm_gradient(phi, B)
    for ax in range(3):
        G = g[ax].reshape(nblk, n, n, n)
        E = exp[ax].reshape(nblk, n, n, n)
        inner = [slice(None)] * 4
        inner[3 - ax] = slice(1, n - 1)                             # axis 0 (x) is the last index
        np.testing.assert_allclose(G[tuple(inner)], E[tuple(inner)], rtol=1e-12, atol=1e-12)
        lo = [slice(None)] * 4
        lo[3 - ax] = 0
        lo1 = list(lo)
        lo1[3 - ax] = 1
        np.testing.assert_array_equal(G[tuple(lo)], G[tuple(lo1)])   # ends copy their neighbour


def test_interpolation_exact_on_linear_fields_and_adjoint():
    B = blocks_of(ng=2)
    n = 12
    rng = W.rng_for(7)
    nblk = 3
    fields = []
    coef = rng.normal(0, 1, (3, 4))
    for q in range(3):
        f = np.zeros(nblk * n ** 3)
        for b in range(nblk):
            d = B["size"][b] / 8
            idx = np.arange(n) - 2 + 0.5
            zz, yy, xx = np.meshgrid(B["origin"][b][2] + idx * d, B["origin"][b][1] + idx * d,
                                     B["origin"][b][0] + idx * d, indexing="ij")
            f[b * n ** 3:(b + 1) * n ** 3] = (coef[q, 0] + coef[q, 1] * xx + coef[q, 2] * yy +
                                              coef[q, 3] * zz).ravel()
        fields.append(f)
    np_ = 300
    blk = rng.integers(0, nblk, np_)
    pos = B["origin"][blk] + rng.uniform(0, 1, (np_, 3)) * B["size"][blk][:, None]
    a = oracle.pm_interpolate(pos[:, 0], pos[:, 1], pos[:, 2], blk, B, *fields)
    for q in range(3):
        exp = coef[q, 0] + pos @ coef[q, 1:]
        np.testing.assert_allclose(a[q], exp, rtol=1e-11, atol=1e-12)
    # at a cell centre: that cell's value
    c = centre(B, 2, 4, 6, 3)
    v = oracle.pm_interpolate([c[0]], [c[1]], [c[2]], [2], B, *fields)
    assert v[0][0] == pytest.approx(fields[0][index(B, 2, 4, 6, 3)], rel=1e-13)
    # adjoint: sum_cells rho g dV = sum_p m_p a_p(g) for any cell field g
    m = rng.uniform(0.5, 1.5, np_)
    rho = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], m, blk, B)
    g = rng.normal(0, 1, nblk * n ** 3)
    dv = np.repeat((B["size"] / 8) ** 3, n ** 3)
    lhs = (rho * g * dv).sum()
    rhs = (m * oracle.pm_interpolate(pos[:, 0], pos[:, 1], pos[:, 2], blk, B, g, g, g)[0]).sum()
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
