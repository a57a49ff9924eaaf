"""A minimal stand-in for AMUSE quantities (AMUSE is not installed): units with an SI scale
and dimension exponents, `value | unit` building a quantity, `quantity.value_in(unit)`
reading it -- the surface the reference's bridge uses (.deprecated/bridge_multiples.py:566-605).
"""
import numpy as np

_DIMS = ("L", "M", "T")


class Unit:
    __array_ufunc__ = None          # `ndarray | unit` defers to Unit.__ror__, as with AMUSE units

    def __init__(self, scale, dims, name=""):
        self.scale, self.dims, self.name = float(scale), tuple(dims), name

    def __mul__(self, o):
        return Unit(self.scale * o.scale, [a + b for a, b in zip(self.dims, o.dims)])

    def __truediv__(self, o):
        return Unit(self.scale / o.scale, [a - b for a, b in zip(self.dims, o.dims)])

    def __pow__(self, p):
        return Unit(self.scale ** p, [a * p for a in self.dims])

    def __ror__(self, value):
        return Quantity(value, self)

    def __repr__(self):
        return self.name or f"Unit({self.scale:g}, {self.dims})"


class Quantity:
    def __init__(self, value, unit):
        self.number = np.asarray(value, dtype=np.float64) if np.ndim(value) else float(value)
        self.unit = unit

    def value_in(self, unit):
        if unit.dims != self.unit.dims:
            raise ValueError(f"cannot express {self.unit.dims} in {unit.dims}")
        return self.number * (self.unit.scale / unit.scale)

    def __getitem__(self, i):
        return Quantity(self.number[i], self.unit)

    def __len__(self):
        return len(self.number)


class units:
    m = Unit(1.0, (1, 0, 0), "m")
    cm = Unit(1e-2, (1, 0, 0), "cm")
    km = Unit(1e3, (1, 0, 0), "km")
    parsec = Unit(3.0856775814913673e16, (1, 0, 0), "parsec")
    kg = Unit(1.0, (0, 1, 0), "kg")
    g = Unit(1e-3, (0, 1, 0), "g")
    # M_sun such that G M_sun is the IAU 1.32712440018e20 m^3 s^-2 with G = 6.6743e-11
    MSun = Unit(1.32712440018e20 / 6.6743e-11, (0, 1, 0), "MSun")
    s = Unit(1.0, (0, 0, 1), "s")
    kms = Unit(1e3, (1, 0, -1), "kms")
