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
* ``batch_instances``: 64 heterogeneous instances advanced together -- 32
  2-D TMz of side log-uniform in 128-1024 cells and 32 3-D of side
  log-uniform in 64-192 cells, at resolution 16; background Dielectric
  eps in [1, 4] plus 4-12 random blocks/cylinders (2-D) or blocks/spheres
  (3-D) of Dielectric, Drude or Lorentz material; a UPML or CPML boundary
  of 8-16 cells; a Bandpass point source of frequency in [0.3, 1.2] with
  fwidth f/4; Courant ratio 0.5; 1000 steps in FP32.
"""

import numpy as np

from . import fdtd as _fdtd
from . import geometry as _geo
from . import material as _mat
from . import source as _src
from .constant import Ex, Ez
from . import pygeom as _pg


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


def spheres3d(n=256, precision="fp32", nsph=40, pml=None, seed=2, _engine=None, verbose=False, amp=1):
    pml = 12 * n / 256.0 if pml is None else pml
    space = _geo.Cartesian(size=(n, n, n), resolution=1)
    geom = spheres3d_geometry(n, nsph, pml, seed)
    src = [_src.PointSource(_src.Bandpass(freq=0.05, fwidth=0.025), center=(0, 0, 0), component=Ex, amp=amp)]
    sim = _fdtd.FDTD(space, geom, src, courant_ratio=0.5, verbose=verbose, precision=precision,
                     _engine=_engine)
    sim.init()
    return sim


def _random_material(rng):
    kind = rng.integers(3)
    if kind == 0:
        return _mat.Dielectric(float(rng.uniform(1.5, 12.0)))
    if kind == 1:
        return _mat.Drude(eps_inf=float(rng.uniform(1.0, 4.0)),
                          dps=(_mat.DrudePole(omega=float(rng.uniform(1.0, 10.0)),
                                              gamma=float(rng.uniform(0.01, 0.1))),))
    poles = tuple(_mat.LorentzPole(amp=float(rng.uniform(0.5, 3.0)), omega=float(rng.uniform(0.2, 1.0)),
                                   gamma=float(rng.uniform(0.01, 0.1)))
                  for _ in range(int(rng.integers(1, 4))))
    return _mat.Lorentz(eps_inf=float(rng.uniform(1.0, 4.0)), lps=poles)


def batch_instance(rng, three_d, side, precision="fp32", _engine=None, res=16, init=True):
    """One instance of the batched workload: a box of ``side`` cells per
    axis (2-D TMz or 3-D) at resolution ``res``.  init=False returns it
    unallocated (the random draws are the same)."""
    L = side / float(res)
    pml = int(rng.integers(8, 17)) / float(res)
    inner = L / 2.0 - pml
    size = (L, L, L) if three_d else (L, L, 0)
    space = _geo.Cartesian(size=size, resolution=res)
    geom = [_geo.DefaultMedium(material=_mat.Dielectric(float(rng.uniform(1.0, 4.0))))]
    for _ in range(int(rng.integers(4, 13))):
        ext = rng.uniform(0.1, 0.3) * 2 * inner
        c = rng.uniform(-inner + ext / 2, inner - ext / 2, 3)
        if not three_d:
            c[2] = 0.0
        mat = _random_material(rng)
        if rng.integers(2) == 0:
            sz = rng.uniform(0.5, 1.0, 3) * ext
            geom.append(_pg.Block(mat, center=tuple(c), size=tuple(sz) if three_d else (sz[0], sz[1], 1)))
        elif three_d:
            geom.append(_pg.Sphere(mat, center=tuple(c), radius=ext / 2))
        else:
            geom.append(_pg.Cylinder(mat, center=tuple(c), radius=ext / 2, axis=(0, 0, 1)))
    bound = _mat.Upml() if rng.integers(2) == 0 else _mat.Cpml()
    geom.append(_geo.Shell(material=bound, thickness=pml))
    f = float(rng.uniform(0.3, 1.2))
    pos = tuple(rng.uniform(-inner / 2, inner / 2, 3) * (1, 1, 1 if three_d else 0))
    src = [_src.PointSource(_src.Bandpass(freq=f, fwidth=f / 4), center=pos, component=Ez)]
    cls = _fdtd.FDTD if three_d else _fdtd.TMzFDTD
    sim = cls(space, geom, src, courant_ratio=0.5, verbose=False, precision=precision, _engine=_engine)
    if init:
        sim.init()
    return sim


def batch_sizes(n2d=32, n3d=32, seed=4, lo2=128, hi2=1024, lo3=64, hi3=192):
    """(three_d, side) of every instance, in a fixed seeded order."""
    rng = np.random.default_rng(seed)
    out = [(False, int(round(np.exp(rng.uniform(np.log(lo2), np.log(hi2)))))) for _ in range(n2d)]
    out += [(True, int(round(np.exp(rng.uniform(np.log(lo3), np.log(hi3)))))) for _ in range(n3d)]
    return out


def batch_instances(n2d=32, n3d=32, seed=4, precision="fp32", _engine=None, only=None, **sizes):
    """The instances of the batched workload (seeded; list order = draw order).
    only: indices of the instances to return (the others are drawn and
    dropped, so each instance is the same whatever is selected)."""
    rng = np.random.default_rng(seed + 1000)
    out = []
    for i, (t, s) in enumerate(batch_sizes(n2d, n3d, seed, **sizes)):
        keep = only is None or i in only
        sim = batch_instance(rng, t, s, precision=precision, _engine=_engine, init=keep)
        if keep:
            out.append(sim)
    return out


def z_cpml_shell(lx, ly, lz, d):
    """A CPML of thickness d on the two z faces of an lx x ly x lz cell
    that is periodic along x and y.

    The reference grades a PML material along every axis of its shell,
    whichever faces the shell has (src/pygeom.pyx:1061-1079 hands the
    shell's half sizes to Pml.init, src/material.pyx:441-470), so a z-only
    boundary shell also carries x and y grading in the corners of its z
    slabs next to the periodic seams.  Those unterminated patches support
    a growing surface mode on the slab interfaces (the fields of a point
    source in such a cell grow about 1e5-fold per 300 steps in FP64).  This
    shell is wider than the cell along x and y by more than 2 d, so that
    grading lies outside the grid and only the z grading acts, within the
    reference's own API."""
    return _geo.Shell(material=_mat.Cpml(), size=(lx + 2.5 * d, ly + 2.5 * d, lz), thickness=d,
                      plus_x=False, minus_x=False, plus_y=False, minus_y=False)


def _tables_of(sim_args):
    """Material tables of a slab, for a worker process (host only)."""
    side, radius, res, precision = sim_args
    sim = pc_slab(side, radius, precision=precision, res=res, _engine=False)
    return sim.host_tables()


def pc_slab(side, radius, steps=None, precision="fp32", res=16, nfreq=50, decimation=4,
            _engine=None, tiles=None):
    """One served solve: a photonic-crystal slab (eps 12, 0.5 a thick) with a
    triangular lattice of air holes of the given radius (units of a), in a
    cube of ``side`` cells at ``res`` cells per a; CPML (1 a) on z, Bloch
    periodic x/y at k = 0 (complex fields); a TFSF plane wave along -z
    (E along x) whose box is wider than the cell, so only its z faces act;
    a running-DFT flux plane below the slab at ``nfreq`` frequencies in
    [0.2, 0.6] c/a.  The lattice is not commensurate with the square cell,
    so the holes are cut at the periodic seam.  Returns (sim, flux)."""
    L = side / float(res)
    dz = 1.0 / res
    space = _geo.Cartesian(size=(L, L, L), resolution=res)
    geom = [_geo.DefaultMedium(material=_mat.Dielectric()),
            _pg.Block(_mat.Dielectric(12.0), center=(0, 0, 0), size=(2 * L, 2 * L, 0.5))]
    a2 = np.array((0.5, np.sqrt(3) / 2))
    n = int(np.ceil(L)) + 2
    for u in range(-n, n + 1):
        for v in range(-n, n + 1):
            x, y = u * 1.0 + v * a2[0], v * a2[1]
            if abs(x) <= L / 2 + radius and abs(y) <= L / 2 + radius:
                geom.append(_pg.Cylinder(_mat.Dielectric(1.0), center=(x, y, 0), radius=radius, height=0.5,
                                         axis=(0, 0, 1)))
    geom.append(z_cpml_shell(L, L, L, 1.0))
    # TFSF z faces between the CPML layers: low edge on a node, high edge
    # half way between nodes (the rounding rule of the face index ranges)
    zlo = -L / 2 + 1.25 + 0.0 * dz
    zhi = L / 2 - 1.25 + 0.5 * dz
    src = [_src.TotalFieldScatteredField(_src.Bandpass(freq=0.4, fwidth=0.4, s=5), center=(0, 0, (zlo + zhi) / 2),
                                         size=(2 * L, 2 * L, zhi - zlo), direction=(0, 0, -1),
                                         polarization=(1, 0, 0))]
    if _engine is False:
        # host tables only: no device, no source set-up
        return _fdtd.FDTD(space, geom, [], courant_ratio=0.5, bloch=(0, 0, 0), verbose=False,
                          precision=precision)
    sim = _fdtd.FDTD(space, geom, src, courant_ratio=0.5, bloch=(0, 0, 0), verbose=False,
                     precision=precision, _engine=_engine)
    sim.init(tiles=tiles)
    flux = sim.add_dft_flux(np.linspace(0.2, 0.6, nfreq), z=-0.5 * (0.25 + (L / 2 - 1.25)),
                            decimation=decimation)
    return sim, flux


def serve_requests(n=48, rate=20.0, seed=5, lo=64, hi=256):
    """(arrival time, side, radius, steps) of the served stream: Poisson
    arrivals at ``rate`` per second, side log-uniform in [lo, hi] cells,
    hole radius uniform in [0.2, 0.4] a, steps uniform in 1500..4000."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        side = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        out.append((float(t[i]), side, float(rng.uniform(0.2, 0.4)), int(rng.integers(1500, 4001))))
    return out


def pc_slab_footprint(side, nfreq=50, complex_fp32=True):
    """Device bytes of a served slab solve, estimated before it is built:
    six fields, six id arrays, full-grid CPML state of the six components,
    and the four DFT boxes of the flux plane (sums and compensations)."""
    es = 8 if complex_fp32 else 16
    cells = float(side) ** 3
    fields = cells * 6 * es
    ids = cells * 6 * 2
    cpml = cells * 6 * 2 * es
    dft = nfreq * 6 * float(side) ** 2 * 32  # compensated sums: two (re, im) pairs per cell
    return int(1.05 * (fields + ids + cpml + dft)) + (64 << 20)


def serve_request_list(n=48, rate=20.0, seed=5, nfreq=50, decimation=4, precision="fp32",
                       _engine=None, lo=64, hi=256, steps=None, host_pool=None):
    """The served stream as gmes_amd.serve.Request objects; each result is
    the flux spectrum of its slab.  host_pool: an executor (processes) that
    builds the material tables, so that set-up runs in parallel."""
    from .serve import Request

    def builder(side, radius):
        def build():
            tiles = None
            if host_pool is not None:
                tiles = host_pool.submit(_tables_of, (side, radius, 16, precision)).result()
            sim, flux = pc_slab(side, radius, precision=precision, nfreq=nfreq, decimation=decimation,
                                _engine=_engine, tiles=tiles)
            return sim, flux.flux
        return build

    out = []
    for i, (t, side, radius, nsteps) in enumerate(serve_requests(n, rate, seed, lo, hi)):
        n = nsteps if steps is None else steps
        out.append(Request(i, builder(side, radius), n, arrival=t, footprint=pc_slab_footprint(side, nfreq),
                           work=float(side) ** 3 * n))
        out[-1].side = side
    return out


def dispersive3d(scale=1.0, seed=3, precision="fp32", _engine=None, courant=0.3, src_time=None, amp=1):
    """3-D dispersive Bloch cell: 384 x 384 x 192 cells (times ``scale``) at
    dx = 1; a periodic 8 x 8 array of Drude blocks (24 cells, pitch 48;
    wp in [1, 10], gamma in [0.01, 0.1]), 16 seed-placed Lorentz spheres
    (radius 4-16 cells, 1-3 poles), a 16-cell two-critical-point DCP-ADE
    slab, a 10-cell CPML on z, Bloch-periodic complex fields on x/y at a
    random k, and a Bandpass plane wave along -z (E along x) through a
    TFSF box that spans x/y (only its z faces act).  Courant 0.3: the
    reference's Drude recursion grows without bound at dx = 1 for
    wp * dt above about 2 (wp = 10 at Courant 0.5), and 0.3 keeps the
    whole wp range stable without changing the method."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = int(round(384 * scale)), int(round(384 * scale)), int(round(192 * scale))
    space = _geo.Cartesian(size=(nx, ny, nz), resolution=1)
    kx, ky = rng.uniform(-np.pi / 384, np.pi / 384, 2)
    geom = [_geo.DefaultMedium(material=_mat.Dielectric())]
    pitch, blk = 48 * scale, 24 * scale
    nb = max(1, int(round(8 * scale * 384 / 384)))
    for u in range(nb):
        for v in range(nb):
            c = (-nx / 2 + (u + 0.5) * pitch, -ny / 2 + (v + 0.5) * pitch, 0.0)
            mat = _mat.Drude(eps_inf=1.0, dps=(_mat.DrudePole(omega=float(rng.uniform(1, 10)),
                                                              gamma=float(rng.uniform(0.01, 0.1))),))
            geom.append(_pg.Block(mat, center=c, size=(blk, blk, blk)))
    zin = nz / 2 - 10 * max(scale, 0.5) - 4
    for _ in range(16):
        r = float(rng.uniform(4, 16) * scale)
        c = (float(rng.uniform(-nx / 2, nx / 2)), float(rng.uniform(-ny / 2, ny / 2)),
             float(rng.uniform(-zin + r, zin - r)))
        poles = tuple(_mat.LorentzPole(amp=float(rng.uniform(0.5, 3.0)), omega=float(rng.uniform(0.2, 1.0)),
                                       gamma=float(rng.uniform(0.01, 0.1)))
                      for _ in range(int(rng.integers(1, 4))))
        geom.append(_pg.Sphere(_mat.Lorentz(eps_inf=float(rng.uniform(1.0, 3.0)), lps=poles), center=c,
                               radius=r))
    dcp = _mat.DcpAde(eps_inf=1.5, sigma=0.01, dps=(_mat.DrudePole(omega=1.2, gamma=0.05),),
                      cps=(_mat.CriticalPoint(amp=0.8, phi=-0.4, omega=0.6, gamma=0.1),
                           _mat.CriticalPoint(amp=0.4, phi=0.7, omega=0.9, gamma=0.2)))
    geom.append(_pg.Block(dcp, center=(0, 0, -zin + 8 * scale + 4), size=(2 * nx, 2 * ny, 16 * scale)))
    geom.append(z_cpml_shell(nx, ny, nz, max(1.0, round(10 * scale))))
    ztop = zin + 2
    st = _src.Bandpass(freq=0.05, fwidth=0.025) if src_time is None else src_time
    src = [_src.TotalFieldScatteredField(st, center=(0, 0, 0.25),
                                         size=(2 * nx, 2 * ny, 2 * ztop - 0.5), direction=(0, 0, -1),
                                         polarization=(1, 0, 0), amp=amp)]
    sim = _fdtd.FDTD(space, geom, src, courant_ratio=courant, bloch=(kx, ky, 0.0), verbose=False,
                     precision=precision, _engine=_engine)
    sim.init()
    return sim
