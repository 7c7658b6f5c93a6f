"""Small simulation cases shared by the parity tests.

Each builder returns an initialized simulation driven by the given C-ABI
library (the HIP engine or one of the CPU oracles) at the given precision.
Sizes are chosen so the CPU oracles finish in a second or two.
"""

import numpy as np

import gmes_amd as g
from gmes_amd import constant as const


def rel_l2(a, b):
    """Relative L2 error of field dict a against reference field dict b."""
    num = 0.0
    den = 0.0
    for c in const.COMPONENTS:
        x = np.asarray(a[c], np.complex128)
        y = np.asarray(b[c], np.complex128)
        num += float(np.sum(np.abs(x - y) ** 2))
        den += float(np.sum(np.abs(y) ** 2))
    return np.sqrt(num) / np.sqrt(den) if den > 0 else np.sqrt(num)


def fields(sim):
    return {c: np.array(sim.field[c]) for c in const.COMPONENTS}


def kat_tmz(lib, precision="fp64"):
    """tests/test_fdtd.py:55-79 of the reference."""
    space = g.Cartesian(size=(2, 2, 0), resolution=5)
    geometry = [g.DefaultMedium(material=g.Dielectric()), g.Shell(material=g.Cpml())]
    sources = [g.PointSource(src_time=g.Continuous(freq=0.8, width=0.5), center=(0, 0, 0),
                             component=g.Ez)]
    sim = g.TMzFDTD(space, geometry, sources, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def air2d(lib, precision="fp64"):
    """examples/air2d.py:35-45 of the reference (the plain-method FP64 workload)."""
    space = g.Cartesian(size=(10, 10, 0), resolution=20)
    geometry = [g.DefaultMedium(material=g.Dielectric()), g.Shell(material=g.Cpml(), thickness=1)]
    sources = [g.PointSource(src_time=g.Continuous(freq=0.8), center=(0, 0, 0), component=g.Ez)]
    sim = g.TMzFDTD(space, geometry, sources, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def spheres3d(lib, precision="fp64", n=24, nsph=6, pml=4, seed=2, courant=0.5, bound=None):
    """A reduced 3-D dielectric-spheres box with a CPML (or ``bound``) on
    every face and a Gaussian-pulse Ex point source at the centre."""
    rng = np.random.default_rng(seed)
    half = n / 2
    space = g.Cartesian(size=(n, n, n), resolution=1)
    geometry = [g.DefaultMedium(material=g.Dielectric(1))]
    for _ in range(nsph):
        r = rng.uniform(2, max(2.5, n / 8))
        lim = half - pml - r
        c = rng.uniform(-lim, lim, 3)
        geometry.append(g.Sphere(g.Dielectric(rng.uniform(1.5, 12)), center=tuple(c), radius=r))
    geometry.append(g.Shell((bound or g.Cpml)(), thickness=pml))
    src = [g.PointSource(g.Bandpass(freq=0.05, fwidth=0.025), center=(0, 0, 0), component=g.Ex)]
    sim = g.FDTD(space, geometry, src, courant_ratio=courant, verbose=False, precision=precision,
                 _engine=lib)
    sim.init()
    return sim


def mode_case(cls, lib, precision="fp64", pml=g.Cpml):
    """A 2-D (or 1-D) mode class with a dielectric block, a PML shell and a
    point source on one of its E components."""
    if cls in (g.TEMxFDTD, g.TEMyFDTD, g.TEMzFDTD):
        axis = {g.TEMxFDTD: 0, g.TEMyFDTD: 1, g.TEMzFDTD: 2}[cls]
        size = [0, 0, 0]
        size[axis] = 8
        comp = {g.TEMxFDTD: g.Ey, g.TEMyFDTD: g.Ez, g.TEMzFDTD: g.Ex}[cls]
    else:
        axis = {g.TExFDTD: 0, g.TEyFDTD: 1, g.TEzFDTD: 2, g.TMxFDTD: 0, g.TMyFDTD: 1,
                g.TMzFDTD: 2}[cls]
        size = [6, 6, 6]
        size[axis] = 0
        comp = {g.TExFDTD: g.Ey, g.TEyFDTD: g.Ez, g.TEzFDTD: g.Ex, g.TMxFDTD: g.Ex,
                g.TMyFDTD: g.Ey, g.TMzFDTD: g.Ez}[cls]
    space = g.Cartesian(size=tuple(size), resolution=4)
    center = [0.5, 0.5, 0.5]
    center[axis] = 0.0 if size[axis] == 0 else 0.5
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Block(g.Dielectric(4.0), center=tuple(center), size=(1.5, 1.5, 1.5)),
                g.Shell(pml(), thickness=1)]
    src = [g.PointSource(g.Continuous(freq=0.5), center=(0, 0, 0), component=comp)]
    sim = cls(space, geometry, src, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def dispersive2d(material, lib, precision="fp64", mode=None, bloch=None):
    """A 2-D box with one dispersive block, a CPML and a CW point source."""
    mode = mode or g.TMzFDTD
    comp = {g.TMzFDTD: g.Ez, g.TEzFDTD: g.Ex}[mode]
    space = g.Cartesian(size=(6, 6, 0), resolution=6)
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Block(material, center=(0.8, 0.3, 0), size=(2, 1.5, 1)),
                g.Shell(g.Cpml(), thickness=1)]
    src = [g.PointSource(g.Continuous(freq=0.6), center=(-0.5, 0, 0), component=comp)]
    sim = mode(space, geometry, src, verbose=False, precision=precision, bloch=bloch, _engine=lib)
    sim.init()
    return sim


def drude():
    return g.Drude(eps_inf=1.5, dps=(g.DrudePole(omega=3.0, gamma=0.2),
                                     g.DrudePole(omega=1.0, gamma=0.05)))


def lorentz():
    return g.Lorentz(eps_inf=2.0, lps=(g.LorentzPole(amp=1.5, omega=0.8, gamma=0.05),
                                       g.LorentzPole(amp=0.5, omega=1.6, gamma=0.1)))


def dcp_ade():
    return g.DcpAde(eps_inf=1.2, sigma=0.05, dps=(g.DrudePole(omega=2.0, gamma=0.1),),
                    cps=(g.CriticalPoint(amp=0.8, phi=-0.4, omega=1.2, gamma=0.2),
                         g.CriticalPoint(amp=0.3, phi=0.7, omega=2.2, gamma=0.3)))


def dcp_plrc():
    return g.DcpPlrc(eps_inf=1.2, sigma=0.05, dps=(g.DrudePole(omega=2.0, gamma=0.1),),
                     cps=(g.CriticalPoint(amp=0.8, phi=-0.4, omega=1.2, gamma=0.2),))


def dcp_rc():
    return g.DcpRc(eps_inf=1.2, sigma=0.05, dps=(g.DrudePole(omega=2.0, gamma=0.1),),
                   cps=(g.CriticalPoint(amp=0.8, phi=-0.4, omega=1.2, gamma=0.2),))


def const_case(lib, precision="fp64"):
    space = g.Cartesian(size=(4, 4, 4), resolution=3)
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Block(g.Const(0.25), center=(0.5, 0, 0), size=(1, 1, 1)),
                g.Shell(g.Upml(), thickness=1)]
    src = [g.PointSource(g.Continuous(freq=0.4), center=(-0.5, 0, 0), component=g.Ey)]
    sim = g.FDTD(space, geometry, src, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def dm2_1d(lib, precision="fp64", rtol=1e-10):
    """A 1-D (TEMz) two-level-medium slab driven by a pulse."""
    space = g.Cartesian(size=(0, 0, 8), resolution=10)
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Block(g.Dm2(omega=(2.0, 2.2), n_atom=(0.3, 0.2), rho30=-1, gamma=0.5, t1=20,
                              t2=10, rtol=rtol), center=(0, 0, 1), size=(1, 1, 2)),
                g.Shell(g.Cpml(), thickness=1)]
    src = [g.PointSource(g.DifferentiatedGaussian(tw=0.5, t0=2.0), center=(0, 0, -1.5),
                         component=g.Ex)]
    sim = g.TEMzFDTD(space, geometry, src, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def currents(lib, precision="fp64"):
    """J and M current sources plus two hard sources sharing a cell."""
    space = g.Cartesian(size=(4, 4, 4), resolution=3)
    geometry = [g.DefaultMedium(material=g.Dielectric(2.0, 1.5)), g.Shell(g.Cpml(), thickness=1)]
    src = [g.PointSource(g.Continuous(freq=0.5), center=(0.2, 0, 0), component=g.Jz, amp=2.0),
           g.PointSource(g.Bandpass(freq=0.5, fwidth=0.2), center=(-0.3, 0.3, 0), component=g.My),
           g.PointSource(g.Continuous(freq=0.3), center=(0, -0.4, 0.3), component=g.Ex),
           g.PointSource(g.Continuous(freq=0.7), center=(0, -0.4, 0.3), component=g.Ex, amp=0.5)]
    sim = g.FDTD(space, geometry, src, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def bloch2d(lib, precision="fp64"):
    """Complex fields with a Bloch wave vector on x and y."""
    space = g.Cartesian(size=(4, 3, 0), resolution=5)
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Cylinder(g.Dielectric(6.0), center=(0.3, 0, 0), radius=0.6, height=10)]
    src = [g.PointSource(g.Bandpass(freq=0.4, fwidth=0.2), center=(-0.7, 0.2, 0), component=g.Ez)]
    sim = g.TMzFDTD(space, geometry, src, verbose=False, bloch=(0.3, -0.5, 0), precision=precision,
                    _engine=lib)
    sim.init()
    return sim


def bloch3d(lib, precision="fp64"):
    space = g.Cartesian(size=(4, 4, 4), resolution=2)
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Sphere(g.Dielectric(3.0), center=(0.4, 0, 0), radius=0.8),
                g.Shell(g.Cpml(), thickness=1, plus_x=False, minus_x=False, plus_y=False,
                        minus_y=False)]
    src = [g.PointSource(g.Bandpass(freq=0.4, fwidth=0.2), center=(-0.4, 0.2, 0), component=g.Ey)]
    sim = g.FDTD(space, geometry, src, verbose=False, bloch=(0.4, -0.2, 0), precision=precision,
                 _engine=lib)
    sim.init()
    return sim


def tfsf2d(lib, precision="fp64", scatterer=True, bloch=None, res=20):
    """examples/tfsf_with_scatterer.py of the reference (examples/tfsf.py
    without the cylinder): a CW plane wave along (1,-1,0) through a 3x3
    TFSF box in a 5x5 TMz cell."""
    space = g.Cartesian(size=(5, 5, 0), resolution=res)
    geometry = [g.DefaultMedium(g.Dielectric())]
    if scatterer:
        geometry.append(g.Cylinder(g.Dielectric(3), center=(0, 0, 0), radius=1, axis=(0, 0, 1)))
    geometry.append(g.Shell(material=g.Cpml(), thickness=0.5))
    sources = [g.TotalFieldScatteredField(src_time=g.Continuous(freq=0.8), center=(0, 0, 0),
                                          size=(3, 3, 1), direction=(1, -1, 0),
                                          polarization=(0, 0, 1))]
    sim = g.TMzFDTD(space, geometry, sources, bloch=bloch, verbose=False, precision=precision,
                    _engine=lib)
    sim.init()
    return sim


def tfsf3d(lib, precision="fp64", bloch=None, scatterer=True, direction=(1.0, 0.6, 0.35),
           polarization=None, center=(0.125, 0, -0.125), size=(1.4375, 1.1875, 1.4375),
           src_time=None):
    """3-D TFSF box.  By default the wave vector and polarization are
    oblique, so every one of the 24 faces carries a correction and edge
    cells merge two."""
    space = g.Cartesian(size=(4, 4, 4), resolution=8)
    geometry = [g.DefaultMedium(g.Dielectric())]
    if scatterer:
        geometry.append(g.Sphere(g.Dielectric(4), center=(0.1, -0.05, 0.05), radius=0.4))
    geometry.append(g.Shell(material=g.Cpml(), thickness=0.5))
    k = np.array(direction, np.double)
    e = np.cross(k, (0.2, -1.0, 0.4)) if polarization is None else polarization
    st = g.Bandpass(freq=0.8, fwidth=1.5, s=3) if src_time is None else src_time
    sources = [g.TotalFieldScatteredField(src_time=st, center=tuple(center), size=tuple(size),
                                          direction=tuple(k), polarization=tuple(e))]
    sim = g.FDTD(space, geometry, sources, bloch=bloch, verbose=False, precision=precision,
                 _engine=lib)
    sim.init()
    return sim


def gaussian_beam2d(lib, precision="fp64"):
    """A Gaussian beam entering a TMz cell through the -x face of its box."""
    space = g.Cartesian(size=(4, 4, 0), resolution=10)
    geometry = [g.DefaultMedium(g.Dielectric()), g.Shell(material=g.Cpml(), thickness=0.5)]
    sources = [g.GaussianBeam(src_time=g.Continuous(freq=0.8), directivity=g.PlusX,
                              center=(-0.5, 0, 0), size=(0, 2.5, 0), direction=(1, 0, 0),
                              polarization=(0, 0, 1), waist=0.6)]
    sim = g.TMzFDTD(space, geometry, sources, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


def many_records3d(lib, precision="fp64", nobj=320, n=40, seed=7):
    """A 3-D box with more distinct material records than a byte can index
    (each small block has its own eps), a CPML and a pulsed Ex source."""
    rng = np.random.default_rng(seed)
    space = g.Cartesian(size=(n, n, n), resolution=1)
    geometry = [g.DefaultMedium(material=g.Dielectric(1))]
    for i in range(nobj):
        c = rng.uniform(-n / 2 + 6, n / 2 - 6, 3)
        geometry.append(g.Block(g.Dielectric(1.5 + 0.01 * i), center=tuple(c), size=(2, 2, 2)))
    geometry.append(g.Shell(g.Cpml(), thickness=4))
    src = [g.PointSource(g.Bandpass(freq=0.1, fwidth=0.05), center=(0.5, 0, 0), component=g.Ex)]
    sim = g.FDTD(space, geometry, src, courant_ratio=0.5, verbose=False, precision=precision,
                 _engine=lib)
    sim.init()
    return sim
