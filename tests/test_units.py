"""AMUSE quantities at the Python boundary (CPU; AMUSE is absent, tests/amuse_stub.py stands in).

The reference's legacy functions are declared (length, length, length, length) -> (acc, acc,
acc) and -> potential (src/amuse/torch_amuse_flash/interface.py:1436-1446, 1616-1620); the
bridge calls them with `0.0 | units.m` and `array | units.cm` and reads the results with
`.value_in(units.cm * units.s**-2)` (.deprecated/bridge_multiples.py:566-578, 605).  These tests
drive the Python field through that call shape with a stand-in library object that evaluates
with the oracle, so the unit handling and the pointer marshalling are checked without a GPU.
"""
import ctypes

import numpy as np
import pytest

import oracle
from tests.amuse_stub import Quantity, units
from torch_amd import GasGravityField, StellarGravityField, TgravError
from torch_amd import units as U

PC_CM = 3.0856775814913673e18


def _arr(p, n):
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), (n,)) if p else None


class OracleLib:
    """Stands in for libtgrav.so on the CPU: the calls the field makes, evaluated by the oracle."""

    def __init__(self, G):
        self.G = G
        self.src = None
        self.calls = []

    def tgrav_use_own_stream(self, ctx):
        return 0

    def tgrav_set_sources(self, ctx, n, x, y, z, m, e, kind):
        self.calls.append("set_sources")
        self.src = [np.array(_arr(p, n)) if p else None for p in (x, y, z, m, e)]
        return 0

    def tgrav_eval(self, ctx, n, e, x, y, z, ax, ay, az, phi, mode, kind):
        self.calls.append("eval")
        s = self.src
        res = oracle.eval_plummer(s[0], s[1], s[2], s[3], s[4], _arr(x, n), _arr(y, n), _arr(z, n),
                                  _arr(e, n), self.G, mode=oracle.NEUMAIER)
        for p, r in zip((ax, ay, az, phi), res):
            if p:
                _arr(p, n)[:] = r
        return 0


def make_field(cls, units_system=None, G=None):
    f = cls.__new__(cls)
    f._ctx = None
    f.mode = 0
    f.units = units_system
    f.G = G if G is not None else (units_system.G if units_system else U.G_PC_MSUN_KMS)
    f.lib = OracleLib(f.G)
    return f


def c2_like(n_s=64, n_t=40, seed=5):
    rng = np.random.default_rng(seed)
    s = rng.normal(size=(n_s, 3)) * 0.7
    m = rng.uniform(0.1, 20.0, n_s)
    eps = np.full(n_s, 0.01)
    t = rng.uniform(-2.0, 2.0, size=(n_t, 3))
    return s, m, eps, t


def test_stellar_side_call_shape_in_cgs():
    """get_gravity_at_point(eps | m, xx | cm, ...) as bridge_multiples calls it, read back with
    .value_in(cm s^-2): equals the oracle evaluated in cgs numbers, and the same physics through
    the default pc / Msun / km/s system."""
    s, m, eps, t = c2_like()
    cgs = U.cgs(units)
    f = make_field(StellarGravityField, cgs)
    f.set_stars(s[:, 0] * PC_CM | units.cm, s[:, 1] * PC_CM | units.cm, s[:, 2] * PC_CM | units.cm,
                m | units.MSun, eps * PC_CM | units.cm)
    xx, yy, zz = (t[:, q] * PC_CM | units.cm for q in range(3))
    gx, gy, gz = f.get_gravity_at_point(0.0 | units.m, xx, yy, zz)
    assert all(isinstance(g, Quantity) for g in (gx, gy, gz))
    a_cgs = [g.value_in(units.cm * units.s ** -2) for g in (gx, gy, gz)]
    msun_g = units.MSun.scale / units.g.scale
    ref = oracle.eval_plummer(s[:, 0] * PC_CM, s[:, 1] * PC_CM, s[:, 2] * PC_CM, m * msun_g, eps * PC_CM,
                              *(t[:, q] * PC_CM for q in range(3)), None, cgs.G, mode=oracle.NEUMAIER)
    for q in range(3):
        np.testing.assert_allclose(a_cgs[q], ref[q], rtol=1e-12, atol=0)
    # potential: (length^4) -> potential, cm^2 s^-2
    phi = f.get_potential_at_point(0.0 | units.m, xx, yy, zz)
    np.testing.assert_allclose(phi.value_in(units.cm ** 2 * units.s ** -2), ref[3], rtol=1e-12)

    # the same stars in the default unit system (pc, Msun, km/s): same accelerations in cm s^-2, to
    # the 10 digits the library's G = 4.300917270e-3 pc (km/s)^2 / Msun carries
    g = make_field(StellarGravityField, U.pc_msun_kms(units))
    g.set_stars(s[:, 0] | units.parsec, s[:, 1] | units.parsec, s[:, 2] | units.parsec, m | units.MSun,
                eps | units.parsec)
    hx, hy, hz = g.get_gravity_at_point(0.0 | units.m, xx, yy, zz)
    for q, h in enumerate((hx, hy, hz)):
        np.testing.assert_allclose(h.value_in(units.cm * units.s ** -2), ref[q], rtol=5e-10)
    assert g.lib.src[0].max() < 10.0                          # the library saw pc numbers


def test_gas_side_call_shape():
    """hydro.get_accel_gas_on_particles(0.0 | units.m, loc[:,0], loc[:,1], loc[:,2])
    (.deprecated/bridge_multiples.py:605) with cells from densities in g cm^-3."""
    rng = np.random.default_rng(9)
    cells = rng.uniform(-1, 1, size=(50, 3))
    dx = np.full(50, 0.125)
    rho = rng.uniform(1e-22, 1e-20, 50)                       # g / cm^3
    stars = rng.normal(size=(12, 3))
    cgs = U.cgs(units)
    gas = make_field(GasGravityField, cgs)
    gas.set_cells(*(cells[:, q] * PC_CM | units.cm for q in range(3)), rho | units.g / units.cm ** 3,
                  dx * PC_CM | units.cm)
    loc = [stars[:, q] * PC_CM | units.cm for q in range(3)]
    acc = gas.get_accel_gas_on_particles(0.0 | units.m, *loc)
    assert len(acc) == 3 and all(isinstance(a, Quantity) for a in acc)
    m_g = rho * (dx * PC_CM) ** 3
    ref = oracle.eval_plummer(*(cells[:, q] * PC_CM for q in range(3)), m_g, dx * PC_CM,
                              *(stars[:, q] * PC_CM for q in range(3)), None, cgs.G, mode=oracle.NEUMAIER)
    for q in range(3):
        np.testing.assert_allclose(acc[q].value_in(units.cm * units.s ** -2), ref[q], rtol=1e-12)
    # units are checked: an acceleration cannot be read as a length
    with pytest.raises(ValueError):
        acc[0].value_in(units.cm)


def test_plain_arrays_stay_plain_and_custom_G_needs_units():
    s, m, eps, t = c2_like(8, 5)
    f = make_field(StellarGravityField)
    f.set_stars(s[:, 0], s[:, 1], s[:, 2], m, eps)
    out = f.get_gravity_at_point(0.0, t[:, 0], t[:, 1], t[:, 2])
    assert all(isinstance(a, np.ndarray) for a in out)
    g = make_field(StellarGravityField, None, G=6.6743e-8)
    with pytest.raises(TgravError):
        g.set_stars(s[:, 0] | units.cm, s[:, 1] | units.cm, s[:, 2] | units.cm, m | units.g)


def test_unit_system_derived_units():
    us = U.pc_msun_kms(units)
    assert us.acceleration.dims == (1, 0, -2) and us.potential.dims == (2, 0, -2)
    assert us.time.dims == (0, 0, 1) and us.density.dims == (-3, 1, 0)
    # pc / (km/s) = 0.9778 Myr
    assert abs(us.time.scale / (3.15576e13) - 0.97779) < 1e-4
    # G M_sun / pc in (km/s)^2 equals the library's constant
    gm = 1.32712440018e20 / 3.0856775814913673e16 / 1e6
    assert abs(gm / us.G - 1.0) < 1e-9
