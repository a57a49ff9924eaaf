"""AMUSE quantities at the Python boundary.

The reference's legacy functions take and return quantities, not bare numbers:
``get_accel_gas_on_particles`` is declared ``(length, length, length, length) -> (acc, acc,
acc, ERROR_CODE)`` and ``get_potential_at_point`` ``(length, length, length, length) ->
(potential, ERROR_CODE)`` (src/amuse/torch_amuse_flash/interface.py:1436-1446, 1616-1620).
The bridge hands them ``0.0 | units.m`` and ``array | units.cm`` and reads the results with
``.value_in(units.cm * units.s**-2)`` (.deprecated/bridge_multiples.py:566-578, 605).

A :class:`UnitSystem` names the units the library's plain FP64 arrays are in and the matching
G.  Any argument with a ``value_in`` method is read in those units; when a call received a
quantity, its results come back as quantities (``array | unit``, the idiom the reference's
own code uses).  AMUSE itself is not a dependency: the unit objects are whatever the caller's
quantities understand, by default those of ``amuse.units.units`` imported on first use.
"""

G_PC_MSUN_KMS = 4.300917270e-3     # G M_sun / (pc (km/s)^2): IAU G M_sun = 1.32712440018e20 m^3 s^-2
G_CGS = 6.6743e-8                  # cm^3 g^-1 s^-2 (CODATA 2018), in place of FLASH's `Newton`


def is_quantity(a):
    return hasattr(a, "value_in")


class UnitSystem:
    """Base units of the library's arrays: length, mass and speed (time = length / speed), and G
    in those units.  Derived: acceleration = speed^2 / length (the unit of G M / r^2), potential
    = speed^2, density = mass / length^3."""

    def __init__(self, length, mass, speed, G, name=""):
        self.length, self.mass, self.speed, self.G, self.name = length, mass, speed, float(G), name
        self.time = length / speed
        self.acceleration = speed ** 2 / length
        self.potential = speed ** 2
        self.density = mass / length ** 3

    def unit(self, kind):
        return getattr(self, kind)

    def value(self, a, kind):
        """The number(s) of `a` in this system's `kind` unit; plain numbers pass through."""
        if a is None or not is_quantity(a):
            return a
        return a.value_in(self.unit(kind))

    def quantity(self, a, kind):
        return None if a is None else a | self.unit(kind)

    def __repr__(self):
        return f"UnitSystem({self.name or 'custom'}, G={self.G!r})"


def _amuse_units():
    try:
        from amuse.units import units
    except ImportError as e:      # quantities without AMUSE: the caller must name the units
        raise ImportError("quantities need a unit system: pass units=UnitSystem(...) built from the "
                          "unit objects the quantities use (amuse.units is not importable)") from e
    return units


def pc_msun_kms(units=None):
    """pc, M_sun, km/s (time unit pc / (km/s) = 0.9778 Myr): the library's default numbers."""
    u = _amuse_units() if units is None else units
    return UnitSystem(u.parsec, u.MSun, u.kms, G_PC_MSUN_KMS, "pc, MSun, km/s")


def cgs(units=None, G=G_CGS):
    """cm, g, cm/s: FLASH's units (the hydro side's converter, src/torch_mainloop.py:126)."""
    u = _amuse_units() if units is None else units
    return UnitSystem(u.cm, u.g, u.cm / u.s, G, "cgs")
