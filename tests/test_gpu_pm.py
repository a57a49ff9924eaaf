"""GPU parity of the particle-mesh pieces (needs an MI355X) against the oracle.

Bars: the finite-difference acceleration, the grid kick and the CIC interpolation use the
oracle's operations in the oracle's order and must match it bit for bit; the CIC deposit
sums in exact fixed point (order-free), so it matches the oracle's sequential FP64 sums to
1e-13 per cell and is bit-identical under any permutation of the particles.
"""
import numpy as np
import pytest

import oracle
from torch_amd import LeafBlocks, ParticleMesh, TgravError
from torch_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2pm():
    """c2's 2,000 leaf blocks with 4 guard layers and the stars inside the 4 pc box."""
    w = W.c2(1002)
    st, ce = w["stars"], w["cells"]
    blk = W.containing_block(st["x"], st["y"], st["z"], ce["block_origin"], ce["block_size"])
    keep = blk >= 0
    B = LeafBlocks(ce["block_origin"], ce["block_size"], nb=8, ng=4)
    P = dict(x=st["x"][keep], y=st["y"][keep], z=st["z"][keep], m=st["m"][keep], blk=blk[keep])
    return B, P


def oblocks(B):
    return dict(origin=B.origin, size=B.size, nb=B.nb, ng=B.ng)


def plummer_phi(B, b=0.6, M=500.0):
    x, y, z = B.centres()
    return -M / np.sqrt(x * x + y * y + z * z + b * b)


@pytest.fixture()
def pm():
    f = ParticleMesh()
    yield f
    f.close()


def test_deposit_matches_oracle_and_is_order_free(pm, c2pm):
    B, P = c2pm
    got = pm.update_grid_var(P["x"], P["y"], P["z"], P["m"], P["blk"], B)
    ref = oracle.pm_deposit(P["x"], P["y"], P["z"], P["m"], P["blk"], oblocks(B))
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-300)
    dv = np.repeat((B.size / B.nb) ** 3, B.nbg ** 3)
    np.testing.assert_allclose((got * dv).sum(), P["m"].sum(), rtol=1e-13)
    perm = W.rng_for(3).permutation(len(P["x"]))
    again = pm.update_grid_var(P["x"][perm], P["y"][perm], P["z"][perm], P["m"][perm], P["blk"][perm], B)
    np.testing.assert_array_equal(again, got)
    base = np.full(B.ncell, 0.25)
    acc = pm.update_grid_var(P["x"], P["y"], P["z"], P["m"], P["blk"], B, rho=base.copy(), accumulate=True)
    np.testing.assert_array_equal(acc, 0.25 + got)


def test_deposit_signed_and_tiny_masses(pm):
    """Mixed-sign masses spanning 30 decades still sum to the oracle's value per cell."""
    rng = W.rng_for(4)
    B = LeafBlocks(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), np.array([1.0, 1.0]), nb=8, ng=1)
    n = 4000
    blk = rng.integers(0, 2, n)
    pos = B.origin[blk] + rng.uniform(0, 1, (n, 3))
    m = rng.normal(0, 1, n) * 10.0 ** rng.uniform(-15, 15, n)
    got = pm.update_grid_var(pos[:, 0], pos[:, 1], pos[:, 2], m, blk, B)
    ref = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], m, blk, oblocks(B))
    # the sequential FP64 sum itself carries ~n eps of the largest |term|
    scale = oracle.pm_deposit(pos[:, 0], pos[:, 1], pos[:, 2], np.abs(m), blk, oblocks(B))
    assert np.all(np.abs(got - ref) <= 1e-13 * scale + 1e-300)


def test_deposit_errors(pm, c2pm):
    B, P = c2pm
    bad = P["blk"].copy()
    bad[7] = (bad[7] + 1) % B.nblk if B.nblk > 1 else 5       # another block: cloud outside it
    with pytest.raises(TgravError):
        pm.update_grid_var(P["x"], P["y"], P["z"], P["m"], bad, B)
    bad[7] = B.nblk + 3
    with pytest.raises(TgravError):
        pm.update_grid_var(P["x"], P["y"], P["z"], P["m"], bad, B)
    assert pm.lib.tgrav_pm_deposit(pm.handle, 1, None, None, None, None, None, 1, 8, 0, None, None, None,
                                   None, None, 0, 0) == -1


def test_gradient_and_kick_bit_exact(pm, c2pm):
    B, _ = c2pm
    phi = plummer_phi(B)
    g = pm.accel_list_of_blocks(phi, B)
    ref = oracle.pm_gradient(phi, oblocks(B))
    for q in range(3):
        np.testing.assert_array_equal(g[q], ref[q])
    rng = W.rng_for(8)
    v = [rng.normal(0, 1, B.ncell) for _ in range(3)]
    v0 = [a.copy() for a in v]
    dt = 0.37
    g2 = pm.accel_list_of_blocks(phi, B, dt=dt, vel=v)
    for q in range(3):
        np.testing.assert_array_equal(g2[q], ref[q])
        np.testing.assert_array_equal(v[q], v0[q] + ref[q] * dt)
    # kick only
    v1 = [a.copy() for a in v0]
    assert pm.accel_list_of_blocks(phi, B, dt=dt, vel=v1, store=False) is None
    for q in range(3):
        np.testing.assert_array_equal(v1[q], v0[q] + ref[q] * dt)


def test_interpolation_and_long_range_force_bit_exact(pm, c2pm):
    B, P = c2pm
    phi = plummer_phi(B)
    g = oracle.pm_gradient(phi, oblocks(B))
    got = pm.map_mesh_to_particles(P["x"], P["y"], P["z"], P["blk"], B, *g)
    ref = oracle.pm_interpolate(P["x"], P["y"], P["z"], P["blk"], oblocks(B), *g)
    for q in range(3):
        np.testing.assert_array_equal(got[q], ref[q])
    lrf = pm.long_range_force(P["x"], P["y"], P["z"], P["blk"], B, phi)
    for q in range(3):
        np.testing.assert_array_equal(lrf[q], ref[q])
    # physical sanity: the FD + CIC field of the Plummer potential points inwards
    r = np.stack([P["x"], P["y"], P["z"]], 1)
    a = np.stack(lrf, 1)
    assert np.mean(np.sum(a * r, 1) < 0) > 0.99


def test_device_tensors_match_host(pm, c2pm):
    import torch
    B, P = c2pm
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    BD = LeafBlocks(d(B.origin), d(B.size), nb=B.nb, ng=B.ng)
    rho_h = pm.update_grid_var(P["x"], P["y"], P["z"], P["m"], P["blk"], B)
    rho_d = pm.update_grid_var(d(P["x"]), d(P["y"]), d(P["z"]), d(P["m"]), d(P["blk"]), BD)
    phi = plummer_phi(B)
    g_h = pm.accel_list_of_blocks(phi, B)
    g_d = pm.accel_list_of_blocks(d(phi), BD)
    a_h = pm.map_mesh_to_particles(P["x"], P["y"], P["z"], P["blk"], B, *g_h)
    a_d = pm.map_mesh_to_particles(d(P["x"]), d(P["y"]), d(P["z"]), d(P["blk"]), BD, *g_d)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(rho_d.cpu().numpy(), rho_h)
    for q in range(3):
        np.testing.assert_array_equal(g_d[q].cpu().numpy(), g_h[q])
        np.testing.assert_array_equal(a_d[q].cpu().numpy(), a_h[q])
