"""Synthetic benchmark workloads of the engine.

* ``air2d``: the plain FP64 method -- 2-D TMz in air, 10 x 10 at
  resolution 20 (200 x 200 cells), Dielectric default medium, one-unit
  CPML shell, a continuous Ez point source at frequency 0.8 at the
  origin, run to t = 10 (examples/air2d.py:35-45 of the reference).
* ``spheres3d``: 3-D full-vector box of n^3 cells (n = 256 for the
  headline run) with 40 seed-placed dielectric spheres (radius 4-20
  cells, eps uniform in [1.5, 12]) outside a 12-cell CPML on all faces, a
  Gaussian-pulse (Bandpass) Ex point source at the centre and Courant
  ratio 0.5; 2000 steps.  ``scale`` shrinks every length for parity runs.
"""

import numpy as np

from . import fdtd as _fdtd
from . import geometry as _geo
from . import material as _mat
from . import source as _src
from .constant import Ex, Ez


def air2d(precision="fp64", _engine=None, verbose=False):
    space = _geo.Cartesian(size=(10, 10, 0), resolution=20)
    geom = [_geo.DefaultMedium(material=_mat.Dielectric()),
            _geo.Shell(material=_mat.Cpml(), thickness=1)]
    src = [_src.PointSource(src_time=_src.Continuous(freq=0.8), center=(0, 0, 0), component=Ez)]
    sim = _fdtd.TMzFDTD(space, geom, src, verbose=verbose, precision=precision, _engine=_engine)
    sim.init()
    return sim


def spheres3d_geometry(n=256, nsph=40, pml=12, seed=2):
    """Geometry list of the spheres box (draw order = list order)."""
    rng = np.random.default_rng(seed)
    half = n / 2.0
    rmin, rmax = 4.0 * n / 256.0, 20.0 * n / 256.0
    geom = [_geo.DefaultMedium(material=_mat.Dielectric(1))]
    for _ in range(nsph):
        r = rng.uniform(rmin, rmax)
        lim = half - pml - r
        center = rng.uniform(-lim, lim, 3)
        eps = rng.uniform(1.5, 12.0)
        geom.append(_geo.Sphere(_mat.Dielectric(eps), center=tuple(center), radius=r))
    geom.append(_geo.Shell(_mat.Cpml(), thickness=pml))
    return geom


def spheres3d(n=256, precision="fp32", nsph=40, pml=None, seed=2, _engine=None, verbose=False):
    pml = 12 * n / 256.0 if pml is None else pml
    space = _geo.Cartesian(size=(n, n, n), resolution=1)
    geom = spheres3d_geometry(n, nsph, pml, seed)
    src = [_src.PointSource(_src.Bandpass(freq=0.05, fwidth=0.025), center=(0, 0, 0), component=Ex)]
    sim = _fdtd.FDTD(space, geom, src, courant_ratio=0.5, verbose=verbose, precision=precision,
                     _engine=_engine)
    sim.init()
    return sim
