"""Periodic and Bloch boundaries pinned by what the mathematics fixes.

The ghost exchange (gmes/fdtd.py:811-1097) copies the first transverse
plane of an E component onto its last plane, and the last plane of an H
component onto its first, times exp(i k_a (x_dest - x_src)) for complex
fields (:822-836).  The reference's own test holds only a deepcopy
self-check (tests/test_fdtd.py:81-121), so these checks use two
properties of the continuous and the discrete problem alike:

* Bloch: a cell of length L at wave vector k is the first half of a 2L
  supercell at the same k whose second half holds the same objects and a
  source multiplied by exp(i k L).  A wrong sign or length in the phase
  breaks it.
* Periodic, real fields: in a medium periodic along x with no absorbing
  layer on x, moving the source by m cells moves every field word by m
  cells, bit for bit (every cell runs the same arithmetic on the same
  values).
"""

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd import constant as const

from cases import fields


def _libs():
    return [pytest.param("oracle", id="oracle"),
            pytest.param("engine", id="gpu", marks=pytest.mark.gpu)]


@pytest.fixture
def lib(request, oracle_lib):
    if request.param == "oracle":
        return oracle_lib
    return request.getfixturevalue("engine_lib")


def _bloch_tmz(lib, copies, k, steps):
    """TMz, x period L = 2 (resolution 10), y period 1.5, CPML none, a
    dielectric cylinder per period and one Bandpass Ez source per period;
    the source of period p carries phase k_x L p."""
    L = 2.0
    nx = L * copies
    space = g.Cartesian(size=(nx, 1.5, 0), resolution=10)
    geometry = [g.DefaultMedium(g.Dielectric())]
    sources = []
    for p in range(copies):
        x0 = -nx / 2 + L * p + L / 2  # centre of period p
        geometry.append(g.Cylinder(g.Dielectric(5.0), center=(x0 + 0.3, 0.1, 0), radius=0.45,
                                   height=10))
        sources.append(g.PointSource(g.Bandpass(freq=0.6, fwidth=0.3, phase=k[0] * L * p),
                                     center=(x0 - 0.5, 0.2, 0), component=g.Ez))
    sim = g.TMzFDTD(space, geometry, sources, bloch=k, verbose=False, _engine=lib)
    sim.init()
    sim._advance(steps)
    out = fields(sim)
    sim.close()
    return out


@pytest.mark.parametrize("lib", _libs(), indirect=True)
@pytest.mark.parametrize("k", [(0.7, -0.4, 0), (-1.9, 0.25, 0)])
def test_bloch_cell_equals_half_of_supercell(lib, k):
    steps = 150
    one = _bloch_tmz(lib, 1, k, steps)
    two = _bloch_tmz(lib, 2, k, steps)
    nx = one[g.Ez].shape[0] - 1   # Ez: nx + 1 planes, the last a ghost
    assert two[g.Ez].shape[0] == 2 * nx + 1
    # the supercell's first planes along x are the cell's arrays, ghost
    # planes included: the cell's ghost holds the Bloch continuation, which
    # the supercell computes directly
    num = den = 0.0
    for c in (g.Ez, g.Hx, g.Hy):
        a = one[c]
        b = two[c][:a.shape[0]]
        num += float(np.sum(np.abs(a - b) ** 2))
        den += float(np.sum(np.abs(a) ** 2))
    assert den > 0
    assert (num / den) ** 0.5 < 1e-11, (num / den) ** 0.5
    # the supercell's second period is the first times exp(i k L)
    e = two[g.Ez]
    np.testing.assert_allclose(e[nx:2 * nx], e[:nx] * np.exp(1j * k[0] * 2.0),
                               rtol=0, atol=1e-11 * np.abs(e).max())


def _periodic_real(lib, mode, shift, steps):
    """Real fields, periodic on every axis (no absorbing layer), a medium
    periodic along x with period 1 (resolution 8), and a hard source moved
    by ``shift`` cells along x."""
    if mode is g.TMzFDTD:
        space, comp, center = g.Cartesian(size=(4, 3, 0), resolution=8), g.Ez, (-1.0, 0.25, 0)
    else:
        space, comp, center = g.Cartesian(size=(4, 2, 2), resolution=4), g.Ey, (-1.0, 0.25, 0.25)
    geometry = [g.DefaultMedium(g.Dielectric())]
    for p in range(4):
        geometry.append(g.Block(g.Dielectric(3.0), center=(-1.5 + p + 0.125, 0, 0),
                                size=(0.375, 10, 10)))
    dx = space.dr[0]
    src = [g.PointSource(g.Continuous(freq=0.5), center=(center[0] + shift * dx,) + center[1:],
                         component=comp)]
    sim = mode(space, geometry, src, verbose=False, _engine=lib)
    sim.init()
    sim._advance(steps)
    out = fields(sim)
    sim.close()
    return out


@pytest.mark.parametrize("lib", _libs(), indirect=True)
@pytest.mark.parametrize("mode", [g.TMzFDTD, g.FDTD], ids=["tmz", "3d"])
def test_periodic_shift_of_the_source_shifts_every_field(lib, mode):
    steps, shift = 60, 8        # one material period (8 cells at resolution 8)
    if mode is g.FDTD:
        shift = 4               # resolution 4
    a = _periodic_real(lib, mode, 0, steps)
    b = _periodic_real(lib, mode, shift, steps)
    moved = 0
    for c in const.COMPONENTS:
        if a[c].size == 1:
            continue
        if c in (g.Ex, g.Hx):        # x is their own axis: no x ghost plane
            inner_a, inner_b = a[c], b[c]
        elif c in const.E_COMPONENTS:
            inner_a, inner_b = a[c][:-1], b[c][:-1]   # E: the last x plane is a ghost
        else:
            inner_a, inner_b = a[c][1:], b[c][1:]     # H: the first x plane is a ghost
        np.testing.assert_array_equal(inner_b, np.roll(inner_a, shift, axis=0), err_msg=c.__name__)
        moved += int(np.count_nonzero(inner_a))
    assert moved > 0
