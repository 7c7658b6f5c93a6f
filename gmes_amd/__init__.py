"""gmes_amd -- MI355X-native FDTD engine with the GMES Python API.

Public names mirror the reference's ``gmes/__init__.py:50-144``; a
script written for the reference runs after ``import gmes_amd as gmes``
(or ``from gmes_amd import *``).
"""

from . import constant, fdtd, geometry, material, source
from .constant import (ATTO, FEMTO, GIGA, KILO, MEGA, MICRO, MILLI, NANO, PETA, PICO, TERA, Z0, c0,
                       eps0, inf, mu0, pi)
from .constant import (Ex, Ey, Ez, Hx, Hy, Hz, Jx, Jy, Jz, MinusX, MinusY, MinusZ, Mx, My, Mz,
                       PlusX, PlusY, PlusZ, X, Y, Z)
from .fdtd import (FDTD, TEMxFDTD, TEMyFDTD, TEMzFDTD, TExFDTD, TEyFDTD, TEzFDTD, TimeStep,
                   TMxFDTD, TMyFDTD, TMzFDTD)
from .geometry import (Block, Cartesian, Cone, Cylinder, DefaultMedium, Ellipsoid, Shell, Sphere)
from .material import (Const, Cpml, CriticalPoint, DcpAde, DcpPlrc, DcpRc, Drude, DrudePole, Dm2,
                       Dielectric, Dummy, Lorentz, LorentzPole, Upml)
from .source import (Bandpass, Continuous, DifferentiatedGaussian, GaussianBeam, PointSource,
                     TotalFieldScatteredField)

__version__ = "0.1.0"
