"""Component and direction tags and physical constants.

Mirrors the public names of the reference's ``gmes.constant`` module
(src/constant.i:48-218, src/constant.hh:11-47, src/constant.cc:8-13).
Components are used as classes (``Ex``, not ``Ex()``), exactly as in the
reference; ``index`` is this engine's component number on the C-ABI.
"""

from math import atan, sqrt

pi = 4 * atan(1)
h = 6.6260695729e-34
hbar = h / (2 * pi)
c0 = 299792458.0
mu0 = 4 * pi * 1e-7
eps0 = 1 / (c0 * c0 * mu0)
Z0 = sqrt(mu0 / eps0)

YOTTA = 1e24
ZETTA = 1e21
EXA = 1e18
PETA = 1e15
TERA = 1e12
GIGA = 1e9
MEGA = 1e6
KILO = 1e3
MILLI = 1e-3
MICRO = 1e-6
NANO = 1e-9
PICO = 1e-12
FEMTO = 1e-15
ATTO = 1e-18
ZEPTO = 1e-21
YOCTO = 1e-24

inf = float("inf")


class Component(object):
    tag = 0
    index = -1


class Electric(Component):
    tag = 1


class Ex(Electric):
    tag = 3
    index = 0

    @classmethod
    def str(cls):
        return "ex"


class Ey(Electric):
    tag = 4
    index = 1

    @classmethod
    def str(cls):
        return "ey"


class Ez(Electric):
    tag = 5
    index = 2

    @classmethod
    def str(cls):
        return "ez"


class Magnetic(Component):
    tag = 2


class Hx(Magnetic):
    tag = 6
    index = 3

    @classmethod
    def str(cls):
        return "hx"


class Hy(Magnetic):
    tag = 7
    index = 4

    @classmethod
    def str(cls):
        return "hy"


class Hz(Magnetic):
    tag = 8
    index = 5

    @classmethod
    def str(cls):
        return "hz"


class ElectricCurrent(Component):
    tag = 9


class Jx(ElectricCurrent):
    tag = 10

    @classmethod
    def str(cls):
        return "jx"


class Jy(ElectricCurrent):
    tag = 11

    @classmethod
    def str(cls):
        return "jy"


class Jz(ElectricCurrent):
    tag = 12

    @classmethod
    def str(cls):
        return "jz"


class MagneticCurrent(Component):
    tag = 13


class Mx(MagneticCurrent):
    tag = 14

    @classmethod
    def str(cls):
        return "mx"


class My(MagneticCurrent):
    tag = 15

    @classmethod
    def str(cls):
        return "my"


class Mz(MagneticCurrent):
    tag = 16

    @classmethod
    def str(cls):
        return "mz"


class Directional(object):
    tag = 17


class X(Directional):
    tag = 18

    @classmethod
    def str(cls):
        return "x"


class Y(Directional):
    tag = 19

    @classmethod
    def str(cls):
        return "y"


class Z(Directional):
    tag = 20

    @classmethod
    def str(cls):
        return "z"


class PlusX(X):
    tag = 21
    vector = (1.0, 0.0, 0.0)

    @classmethod
    def str(cls):
        return "+x"


class MinusX(X):
    tag = 22
    vector = (-1.0, 0.0, 0.0)

    @classmethod
    def str(cls):
        return "-x"


class PlusY(Y):
    tag = 23
    vector = (0.0, 1.0, 0.0)

    @classmethod
    def str(cls):
        return "+y"


class MinusY(Y):
    tag = 24
    vector = (0.0, -1.0, 0.0)

    @classmethod
    def str(cls):
        return "-y"


class PlusZ(Z):
    tag = 25
    vector = (0.0, 0.0, 1.0)

    @classmethod
    def str(cls):
        return "+z"


class MinusZ(Z):
    tag = 26
    vector = (0.0, 0.0, -1.0)

    @classmethod
    def str(cls):
        return "-z"


COMPONENTS = (Ex, Ey, Ez, Hx, Hy, Hz)
E_COMPONENTS = (Ex, Ey, Ez)
H_COMPONENTS = (Hx, Hy, Hz)
