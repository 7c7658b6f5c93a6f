"""The reference's own known answers for the host code that feeds the
engine and the oracle alike (oscillators, coefficient setup, geometry and
the voxeliser), restated here and run through this build's code paths.

GPU-vs-oracle parity cannot see a mistake in code both of them consume;
these checks pin that code against the reference's tests
(tests/test_source_time.py, test_pw_upml.py, test_pw_dcp.py,
test_pw_source.py, test_pw_cpml.py, test_pw_dm2.py, test_pygeom.py,
test_geometry.py), against brute force, and against the closed forms of
the physics (the discrete Drude and Lorentz permittivities).
"""

from math import pi, sin

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd import _abi
from gmes_amd import constant as const
from gmes_amd.geometry import Cartesian
from gmes_amd.pygeom import GeomBoxTree
from gmes_amd.tiles import owner_map


# ---------------------------------------------------------------- oscillators
# tests/test_source_time.py:54-98

def test_continuous_default_width():
    assert g.Continuous(freq=2).width == 2.5


def test_continuous_window_and_complex_phase():
    source = g.Continuous(freq=0.5, phase=pi / 2, width=1)
    source.init(cmplx=True)
    assert source.oscillator(-1) == 0
    assert source.oscillator(0) == 0j
    assert abs(source.oscillator(1).real - 0.0) < 1e-12
    assert abs(source.oscillator(1).imag + 1.0) < 1e-12


@pytest.mark.parametrize("end,width,time,expected", [
    (4, 1, 0.5, 0.5), (4, 1, 2, 1.0), (4, 1, 3.5, 0.5), (2, 1, 1, 1.0),
    (1.5, 1, 0.75, sin(0.375 * pi) ** 4), (1, 2, 0.5, sin(0.125 * pi) ** 4)])
def test_continuous_overlapping_ramps(end, width, time, expected):
    source = g.Continuous(freq=1, start=0, end=end, width=width)
    source.init(cmplx=True)
    assert abs(abs(source.oscillator(time)) - expected) < 1e-7
    assert source.oscillator(end) == 0j
    assert source.oscillator(end + 1e-12) == 0


def test_bandpass_is_zero_outside_cutoff():
    source = g.Bandpass(freq=1, fwidth=0.5)
    source.init(cmplx=False)
    assert source.oscillator(source.peak_time + source.cutoff + 1) == 0
    # the Gaussian envelope at the peak: |1 / (-2 pi i f)| (gmes/source.py:135-178)
    source.init(cmplx=True)
    assert abs(abs(source.oscillator(source.peak_time)) - 1 / (2 * pi)) < 1e-15


def test_differentiated_gaussian_is_antisymmetric():
    source = g.DifferentiatedGaussian(tw=2, t0=5)
    source.init(cmplx=False)
    assert abs(source.oscillator(4) + source.oscillator(6)) < 1e-7
    assert source.oscillator(5) == 0.0


# ------------------------------------------------------------- CPML grading
# tests/test_pw_cpml.py:60-71: the grading stays finite just outside the
# layer's outer boundary, where rounding can put a coordinate

def test_cpml_grading_finite_at_rounded_outer_boundaries():
    space = Cartesian((0, 0, 0))
    space.dt = 1
    cpml = g.Cpml()
    cpml.init(space, ((0, 0, 0), (1, 1, 1), 0.5))
    eps = np.finfo(float).eps
    for x in (-1 - eps, 1 + eps):
        vals = (cpml.sigma(x, 0), cpml.kappa(x, 0), cpml.a(x, 0), cpml.b(x, 0), cpml.c(x, 0))
        assert np.isfinite(vals).all()


# -------------------------------------------------------- geometry (grid)
# tests/test_geometry.py:9-81 (the MPI partition test is out of scope: one
# process drives one GPU)

def test_serial_grid_sizes_are_integral():
    space = Cartesian(size=(4, 6, 0), resolution=(2, 3, 4))
    for v in (space.whole_field_size, space.general_field_size, space.my_field_size):
        np.testing.assert_array_equal(v, (8, 18, 1))
        assert v.dtype.kind == "i"


def test_component_coordinate_round_trips():
    space = Cartesian(size=(4, 6, 2), resolution=2)
    for comp in ("ex", "ey", "ez", "hx", "hy", "hz"):
        point = getattr(space, comp + "_index_to_space")(2, 3, 1)
        assert getattr(space, "space_to_" + comp + "_index")(*point) == (2, 3, 1)


def test_component_indices_floor_negative_nearest_grid_values():
    space = Cartesian(size=(2, 2, 2), resolution=2)
    assert space.space_to_ex_index(-1.3, 0, 0)[0] == -1
    for comp in ("ex", "ey", "ez", "hx", "hy", "hz"):
        first = np.array(getattr(space, comp + "_index_to_space")(0, 0, 0))
        below = first - 0.6 * space.dr
        assert getattr(space, "space_to_" + comp + "_index")(*below) == (-1, -1, -1)


@pytest.mark.parametrize("comp", ["ex", "ey", "ez", "hx", "hy", "hz"])
def test_component_indices_handle_both_nearest_grid_boundaries(comp):
    space = Cartesian(size=(4, 4, 4), resolution=2)
    base_index = np.array((2, 2, 2))
    base = np.array(getattr(space, comp + "_index_to_space")(*base_index))
    to_index = getattr(space, "space_to_" + comp + "_index")
    for axis in range(3):
        for offset, delta in ((-0.500001, -1), (-0.5, 0), (-0.499999, 0), (0.499999, 0),
                              (0.5, 1), (0.500001, 1)):
            point = base.copy()
            point[axis] += offset * space.dr[axis]
            expected = base_index.copy()
            expected[axis] += delta
            assert to_index(*point) == tuple(expected)


# -------------------------------------------------------- geometry (pygeom)
# tests/test_pygeom.py:19-141

@pytest.mark.parametrize("shape_class,radius,radius2", [(g.Cylinder, 1.2, 1.2), (g.Cone, 1.2, 0.4)])
@pytest.mark.parametrize("axis", [(1, 1, 0), (1, -2, 3), (-3, 1, 2)])
def test_rotated_end_caps_define_cartesian_bounds(shape_class, radius, radius2, axis):
    center = np.array((0.25, -0.5, 0.75))
    if shape_class is g.Cylinder:
        shape = shape_class(g.Dielectric(), center=center, axis=axis, radius=radius, height=1.5)
    else:
        shape = shape_class(g.Dielectric(), center=center, axis=axis, radius=radius,
                            radius2=radius2, height=1.5)
    box = shape.geom_box()
    radial = np.sqrt(1 - shape.axis * shape.axis)
    low_cap = center - 0.75 * shape.axis
    high_cap = center + 0.75 * shape.axis
    np.testing.assert_allclose(box.low, np.minimum(low_cap - radius * radial,
                                                   high_cap - radius2 * radial))
    np.testing.assert_allclose(box.high, np.maximum(low_cap + radius * radial,
                                                    high_cap + radius2 * radial))
    perpendicular = np.cross(shape.axis, (0, 0, 1))
    if np.linalg.norm(perpendicular) < 1e-12:
        perpendicular = np.cross(shape.axis, (0, 1, 0))
    perpendicular /= np.linalg.norm(perpendicular)
    point = center - 0.7 * shape.axis + 0.9 * radius * perpendicular
    assert shape.in_object(tuple(point))
    assert box.in_box(tuple(point))


def test_tree_keeps_rotated_cylinder_interior_point():
    default = g.DefaultMedium(g.Dielectric(1))
    cylinder = g.Cylinder(g.Dielectric(4), axis=(1, 1, 0), radius=1, height=0.1)
    default.init(None)
    cylinder.init(None)
    tree = GeomBoxTree((default, cylinder))
    shape, _ = tree.object_of_point((0.9 / np.sqrt(2), -0.9 / np.sqrt(2), 0))
    assert shape is cylinder


@pytest.mark.parametrize("center", [(0, 0, 0), (5, 0, 0)])
def test_shell_bounds_and_tree_follow_center(center):
    space = Cartesian(size=(12, 4, 4), resolution=10)
    default = g.DefaultMedium(g.Dielectric(1))
    shell = g.Shell(g.Dielectric(4), center=center, size=(2, 2, 2), thickness=0.2)
    default.init(space)
    shell.init(space)
    tree = GeomBoxTree((default, shell))
    point = (center[0] + 0.9, center[1], center[2])
    np.testing.assert_allclose(shell.box.low, np.array(center) - 1)
    np.testing.assert_allclose(shell.box.high, np.array(center) + 1)
    assert shell.in_object(point) and shell.box.in_box(point)
    assert tree.object_of_point(point)[0] is shell


def test_block_and_ellipsoid_use_inverse_basis_coordinates():
    block = g.Block(g.Dielectric(), center=(0.25, -0.5, 0.75), e1=(1, 0, 0), e2=(1, 1, 0),
                    e3=(0, 0, 1), size=(2, 2, 2))
    assert block.in_object(tuple(block.center + 0.9 * block.e1 + 0.9 * block.e2))
    assert not block.in_object(tuple(block.center + 1.1 * block.e1))
    ell = g.Ellipsoid(g.Dielectric(), center=(-0.25, 0.5, -0.75), e1=(1, 0, 0), e2=(1, 1, 0),
                      e3=(0, 0, 1), size=(2, 2, 2))
    assert ell.in_object(tuple(ell.center + 0.6 * ell.e1 + 0.6 * ell.e2))
    assert not ell.in_object(tuple(ell.center + 0.8 * ell.e1 + 0.8 * ell.e2))


# ------------------------------------------------------------ voxeliser
# tiles.owner_map decides, per staggered point of every component, the last
# object of the geometry list containing the point (else object 0) and,
# for a compound (PML) owner, the last earlier object containing it
# (src/pygeom.pyx:143-156,425-441).  Checked against that rule evaluated
# point by point with each primitive's in_object.

def _brute_owner(space, comp, shape, geom_list):
    owner = np.zeros(shape, np.int32)
    under = np.zeros(shape, np.int32)
    for idx in np.ndindex(*shape):
        p = tuple(float(v) for v in space.index_to_space(comp, *idx))
        o = len(geom_list) - 1
        while o > 0 and not geom_list[o].in_object(p):
            o -= 1
        owner[idx] = o
        if isinstance(geom_list[o].material, (g.Cpml, g.Upml)):
            u = o - 1
            while u > 0 and not geom_list[u].in_object(p):
                u -= 1
            under[idx] = u
    return owner, under


def _geometries():
    skew = dict(e1=(1, 0.2, 0), e2=(0.3, 1, 0.1), e3=(0, -0.2, 1))
    # 3-D: radii and centres on the 0.25 grid, so many points lie exactly
    # on a surface, where the inclusive tests decide
    yield "3d_primitives", g.FDTD, Cartesian(size=(6, 6, 6), resolution=4), [
        g.DefaultMedium(g.Dielectric()),
        g.Sphere(g.Dielectric(2), center=(0.5, 0, 0.25), radius=1.25),
        g.Cylinder(g.Dielectric(3), center=(-1, 0.5, 0), axis=(1, 1, 0), radius=0.75, height=2),
        g.Cone(g.Dielectric(4), center=(1, -1, -0.5), axis=(1, -2, 3), radius=1.2, radius2=0.4,
               height=1.5),
        g.Block(g.Dielectric(5), center=(0.25, 0.75, -1), size=(1.5, 1, 1), **skew),
        g.Ellipsoid(g.Dielectric(6), center=(-0.75, -1, 1), size=(2, 1, 1.5), **skew),
        g.Shell(g.Upml(), center=(0.25, 0, 0), size=(4, 4.5, 4), thickness=0.75),
        g.Shell(g.Cpml(), thickness=0.5)]
    yield "tmz_cylinder", g.TMzFDTD, Cartesian(size=(5, 5, 0), resolution=10), [
        g.DefaultMedium(g.Dielectric()),
        g.Cylinder(g.Dielectric(3), center=(0, 0, 0), radius=1, axis=(0, 0, 1)),
        g.Shell(material=g.Cpml(), thickness=0.5)]
    rng = np.random.default_rng(2)
    spheres = [g.DefaultMedium(g.Dielectric(1))]
    for _ in range(12):
        r = float(rng.integers(2, 7))
        c = tuple(float(v) for v in rng.integers(-10 + r, 10 - r, 3))
        spheres.append(g.Sphere(g.Dielectric(float(rng.uniform(1.5, 12))), center=c, radius=r))
    spheres.append(g.Shell(g.Cpml(), thickness=4))
    yield "spheres_cpml", g.FDTD, Cartesian(size=(24, 24, 24), resolution=1), spheres


@pytest.mark.parametrize("case", list(_geometries()), ids=lambda c: c[0])
def test_voxeliser_matches_per_point_rule(case, oracle_lib):
    name, mode, space, geom = case
    sim = mode(space, geom, [], verbose=False, _engine=oracle_lib)
    comps = tuple(sim.e_field_compnt) + tuple(sim.h_field_compnt)
    points = 0
    for comp in comps:
        shape = space.storage_shape(comp, comps)
        owner, under, _ = owner_map(space, comp, shape, sim.geom_list)
        bo, bu = _brute_owner(space, comp, shape, sim.geom_list)
        np.testing.assert_array_equal(owner, bo, err_msg="%s %s owner" % (name, comp.__name__))
        compound = np.array([isinstance(o.material, (g.Cpml, g.Upml)) for o in sim.geom_list])
        m = compound[owner]
        np.testing.assert_array_equal(under[m], bu[m], err_msg="%s %s under" % (name, comp.__name__))
        points += owner.size
        owned = np.bincount(owner.ravel(), minlength=len(sim.geom_list))
        assert (owned[1:] > 0).all(), owned  # every object owns points of this component
    assert points > 10000


# ------------------------------------------------------------- discrete ADE
# The Drude and Lorentz updates (src/pw_drude.hh:90-101, src/pw_lorentz.hh:
# 96-107) with z = exp(i w dt):
#   q (z - a1 - a0/z) = a2 E          (polarisation recursion)
#   E z = c0 C z^(1/2) + c1 (z - 1) q + c2 E   (field update)
# so the discrete permittivity seen by the curl, C = i w eps_d E in the
# limit, is eps_d = (dt / c0) [(z - c2) - c1 (z - 1) q/E] / (z - 1) * ... ;
# it must approach eps_inf + chi(w) + sigma/(i w) at second order in dt:
#   Drude   chi = wp^2 / (-w^2 + i gamma w)
#   Lorentz chi = amp w0^2 / (w0^2 - w^2 + i gamma w)

def _eps_discrete(mat, dt, w):
    space = Cartesian((0, 0, 0))
    space.dt = dt
    mat.init(space)
    z = np.exp(1j * w * dt)
    chi = sum(a[2] / (z - a[1] - a[0] / z) for a in mat.a)
    c0, c1, c2 = mat.c
    # E (z - c2) - c1 (z - 1) q = c0 C z^(1/2); the continuum C = i w D
    # corresponds to (z - 1)/dt z^(-1/2) D
    return ((z - c2) - c1 * (z - 1) * chi) * dt / (c0 * (z - 1))


@pytest.mark.parametrize("kind", ["drude", "lorentz"])
def test_discrete_dispersion_converges_to_the_model(kind):
    if kind == "drude":
        poles = (g.DrudePole(omega=2.0, gamma=0.3), g.DrudePole(omega=0.7, gamma=0.05))
        mat = lambda: g.Drude(eps_inf=1.7, sigma=0.2, dps=poles)
        chi = lambda w: sum(p.omega ** 2 / (-w ** 2 + 1j * p.gamma * w) for p in poles)
    else:
        poles = (g.LorentzPole(amp=1.5, omega=0.8, gamma=0.05), g.LorentzPole(amp=0.5, omega=1.6, gamma=0.2))
        mat = lambda: g.Lorentz(eps_inf=2.2, sigma=0.1, lps=poles)
        chi = lambda w: sum(p.amp * p.omega ** 2 / (p.omega ** 2 - w ** 2 + 1j * p.gamma * w) for p in poles)
    for w in (0.3, 0.9, 1.7):
        exact = mat().eps_inf + chi(w) + mat().sigma / (1j * w)
        errs = [abs(_eps_discrete(mat(), dt, w) - exact) / abs(exact) for dt in (0.02, 0.01, 0.005)]
        assert errs[0] < 0.05
        # second order: halving dt quarters the error
        assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5, errs


# DCP-ADE (src/pw_dcp.hh:105-167,200-225), same z = exp(i w dt):
#   q (z - a1 - a0/z) = a2 (z + 2 + 1/z) E
#   p (z - b1 - b0/z) = (b2/z + b3 + b4 z) E
#   E (z - c3 - c2/z) - c1 [sum_D ((1 - a1) - a0/z) q + sum_C ((1 - b1) - b0/z) p] = c0 C z^(1/2)
# and the continuum model, for the exp(+i w t) convention used here:
#   eps_inf + sigma/(i w) + sum_D wD^2 / (-w^2 + i gamma w)
#     + sum_C 2 A W [(W cos phi - G sin phi) - i w sin phi] / (W^2 + G^2 - w^2 + 2 i G w)
# (the critical-point term is A W [e^(-i phi)/(W - w + i G) + e^(i phi)/(W + w - i G)]).

def _eps_discrete_dcp(mat, dt, w):
    space = Cartesian((0, 0, 0))
    space.dt = dt
    mat.init(space)
    z = np.exp(1j * w * dt)
    c0, c1, c2, c3 = mat.c
    s = 0
    for a0, a1, a2 in mat.a:
        s += ((1 - a1) - a0 / z) * a2 * (z + 2 + 1 / z) / (z - a1 - a0 / z)
    for b0, b1, b2, b3, b4 in mat.b:
        s += ((1 - b1) - b0 / z) * (b2 / z + b3 + b4 * z) / (z - b1 - b0 / z)
    return ((z - c3 - c2 / z) - c1 * s) * dt / (c0 * (z - 1))


def test_dcp_ade_dispersion_converges_to_the_model():
    dps = (g.DrudePole(omega=2.0, gamma=0.3),)
    cps = (g.CriticalPoint(amp=1.2, phi=-0.6, omega=1.3, gamma=0.4),
           g.CriticalPoint(amp=0.7, phi=0.9, omega=0.8, gamma=0.15))
    mat = lambda: g.DcpAde(eps_inf=2.1, sigma=0.15, dps=dps, cps=cps)

    def model(w):
        e = 2.1 + 0.15 / (1j * w)
        e += sum(p.omega ** 2 / (-w ** 2 + 1j * p.gamma * w) for p in dps)
        e += sum(2 * c.amp * c.omega * ((c.omega * np.cos(c.phi) - c.gamma * np.sin(c.phi)) - 1j * w * np.sin(c.phi))
                 / (c.omega ** 2 + c.gamma ** 2 - w ** 2 + 2j * c.gamma * w) for c in cps)
        return e
    for w in (0.3, 0.9, 1.7):
        exact = model(w)
        errs = [abs(_eps_discrete_dcp(mat(), dt, w) - exact) / abs(exact) for dt in (0.02, 0.01, 0.005)]
        assert errs[0] < 0.01
        assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5, errs


# ------------------------------------------------- simulation-level KATs
# run on the oracle here and on the GPU engine (gpu marker)

def _libs():
    return [pytest.param("oracle", id="oracle"),
            pytest.param("engine", id="gpu", marks=pytest.mark.gpu)]


@pytest.fixture
def lib(request, oracle_lib):
    if request.param == "oracle":
        return oracle_lib
    return request.getfixturevalue("engine_lib")


class _Constant(g.Continuous):
    """oscillator(t) = value (the reference tests' SimpleNamespace source)."""

    def __init__(self, value):
        g.Continuous.__init__(self, freq=1.0)
        self.value = value

    def oscillator(self, time):
        return self.value


@pytest.mark.parametrize("lib", _libs(), indirect=True)
@pytest.mark.parametrize("comp,field", [(g.Jx, "ex"), (g.Jy, "ey"), (g.Jz, "ez"), (g.Mx, "hx"),
                                        (g.My, "hy"), (g.Mz, "hz")])
def test_current_source_applies_material_scaled_update(lib, comp, field):
    """tests/test_pw_source.py:92-118: F -= dt s / eps (or mu) with s = 2,
    dt = 0.5, eps = mu = 4: -0.25."""
    space = g.Cartesian(size=(2, 2, 2), resolution=2)
    sim = g.FDTD(space, [g.DefaultMedium(g.Dielectric(4, 4))],
                 [g.PointSource(_Constant(2.0), (0, 0, 0), comp)], dt=0.5, verbose=False, _engine=lib)
    sim.init()
    sim.step()
    idx = getattr(space, "space_to_" + field + "_index")(0, 0, 0)
    assert getattr(sim, field)[idx] == -0.25
    sim.close()


@pytest.mark.parametrize("lib", _libs(), indirect=True)
@pytest.mark.parametrize("comp,field,center,boundary,dense_high", [
    (g.Jx, "ex", (-0.01, 0, 0), -0.1, False), (g.Jx, "ex", (0.01, 0, 0), 0.1, True),
    (g.Mx, "hx", (-0.01, 0, 0), -0.005, True), (g.Mx, "hx", (0.01, 0, 0), 0.005, False)])
def test_current_source_uses_material_at_snapped_point(lib, comp, field, center, boundary,
                                                       dense_high):
    """tests/test_pw_source.py:120-144: the eps (mu) of a J (M) source is the
    material at the snapped Yee point, not at the requested point."""
    space = g.Cartesian(size=(2, 2, 2), resolution=2)
    lo, hi = (boundary, 5.0) if dense_high else (-5.0, boundary)
    dense = g.Block(g.Dielectric(4, 4), center=(0.5 * (lo + hi), 0, 0), size=(hi - lo, 10, 10))
    sim = g.FDTD(space, [g.DefaultMedium(g.Dielectric(1, 1)), dense],
                 [g.PointSource(_Constant(2.0), center, comp)], dt=0.5, verbose=False, _engine=lib)
    sim.init()
    request_material, _ = sim.geom_tree.material_of_point(center)
    assert request_material.eps_inf == 1
    sim.step()
    idx = getattr(space, "space_to_" + field + "_index")(*center)
    assert getattr(sim, field)[idx] == -0.25
    sim.close()


@pytest.mark.parametrize("lib", _libs(), indirect=True)
@pytest.mark.parametrize("kind", ["plrc", "rc"])
def test_plrc_and_rc_conductive_decay(lib, kind):
    """tests/test_pw_dcp.py:81-109: with no poles, E decays by
    (eps - s)/(eps + s), s = dt sigma / 2, per step; c has the closed form.
    A uniform E in a periodic box keeps the curl zero."""
    eps_inf, sigma, dt = 2.0, 3.0, 0.1
    cond = 0.5 * dt * sigma
    decay = (eps_inf - cond) / (eps_inf + cond)
    cls = g.DcpPlrc if kind == "plrc" else g.DcpRc
    mat = cls(eps_inf=eps_inf, sigma=sigma)
    space = g.Cartesian(size=(3, 3, 3), resolution=1)
    sim = g.FDTD(space, [g.DefaultMedium(mat)], [], dt=dt, verbose=False, _engine=lib)
    sim.init()
    np.testing.assert_allclose(mat.c, (dt / (eps_inf + cond), decay, 1 / (eps_inf + cond)),
                               rtol=1e-15)
    for f in (sim.ex, sim.ey, sim.ez):
        f[...] = 1.0
    for step in range(1, 5):
        sim.step()
        for f in (sim.ex, sim.ey, sim.ez):
            np.testing.assert_allclose(f, decay ** step, rtol=1e-14)
        for f in (sim.hx, sim.hy, sim.hz):
            assert not f.any()
    sim.close()


def _cell_sim(lib, target, rec, profile_row, curl_comp, curl_idx, dtype=_abi.F64):
    """A 2x2x2-cell sim driven through the C-ABI: every cell Dummy except
    (1, 1, 1) of ``target``, which gets record ``rec``; PML profile rows of
    ``target`` all equal ``profile_row``; the curl input is 1 at
    ``curl_idx`` of ``curl_comp`` (the Dummy cells keep it constant)."""
    desc = _abi.GridDesc()
    desc.nx = desc.ny = desc.nz = 2
    desc.dtype = dtype
    desc.active = 0x3F
    desc.dx = desc.dy = desc.dz = 1.0
    desc.dt = 1.0
    for i in range(18):
        desc.phase[2 * i] = 1.0
    h = lib.sim_create(desc)
    recs = np.zeros(2, _abi.MAT_REC_DTYPE)
    recs[0] = (_abi.CLS_DUMMY, 0, 0, 0, 0, 1.0, (0.0, 0.0))
    recs[1] = rec
    lib.set_materials(h, recs, np.zeros(1))
    shapes = {}
    for c in range(6):
        shapes[c] = tuple(lib_shape(lib, h, c))
        ids = np.zeros(shapes[c], np.uint16)
        if c == target:
            ids[1, 1, 1] = 1
        lib.set_material_ids(h, c, ids)
    if profile_row is not None:
        for a in range(3):
            lib.set_profile(h, 0, target, a, np.tile(np.asarray(profile_row, np.float64),
                                                     (shapes[target][a], 1)))
    lib.commit(h)
    f = np.zeros(shapes[curl_comp], _abi.NUMPY_DTYPE[dtype])
    f[curl_idx] = 1
    lib.write_field(h, curl_comp, f)
    return h, shapes


def lib_shape(lib, h, c):
    import ctypes
    out = (ctypes.c_int64 * 3)()
    lib.call("gmes_sim_shape", h, c, out)
    return [int(v) for v in out]


@pytest.mark.parametrize("lib", _libs(), indirect=True)
@pytest.mark.parametrize("target,curl_comp,curl_idx,sign", [
    (0, 5, (2, 2, 1), 1), (1, 3, (1, 2, 2), 1), (2, 4, (2, 1, 2), 1),
    (3, 2, (1, 1, 0), -1), (4, 0, (0, 1, 1), -1), (5, 1, (1, 0, 1), -1)])
def test_upml_auxiliary_state_persists(lib, target, curl_comp, curl_idx, sign):
    """tests/test_pw_upml.py:52-88: c1 = 0.5, c2 = 1, c3 = 0, c4 = 1,
    c5 = 1, c6 = 0 and a unit curl give F = +-(1, 1.5, 1.75) over three
    updates (c1, c2 from the first curl axis, c3, c4 from the second, c5, c6
    from the component's own axis)."""
    rec = (_abi.CLS_UPML, 0, 0, 0, 0, 1.0, (0.0, 0.0))
    h, shapes = _cell_sim(lib, target, rec, (0.5, 1, 0, 1, 1, 0), curl_comp, curl_idx)
    values = []
    out = np.zeros(shapes[target], np.float64)
    for _ in range(3):
        lib.step(h, 1)
        lib.read_field(h, target, out)
        values.append(out[1, 1, 1])
    lib.sim_destroy(h)
    np.testing.assert_allclose(values, sign * np.array((1, 1.5, 1.75)), rtol=1e-15)


@pytest.mark.parametrize("lib", _libs(), indirect=True)
def test_dm2_corrector_starts_from_zero_reference(lib):
    """tests/test_pw_dm2.py:103-112: from zero fields and state, a unit D2
    curl gives a finite, nonzero Ex (the relative error of the first
    corrector pass has a zero reference)."""
    space = g.Cartesian(size=(2, 2, 2), resolution=1)
    sim = g.FDTD(space, [g.DefaultMedium(g.Dm2())], [], dt=0.1, verbose=False, _engine=lib)
    sim.init()
    sim.hy[2, 1, 2] = 1
    sim.step()
    sim.sync()
    assert np.isfinite(sim.ex).all()
    assert sim.ex[1, 1, 1] != 0
    sim.close()


# ------------------------------------------------ PML grading of a z-only shell
# The reference hands a shell's half sizes to its PML material
# (src/pygeom.pyx:1061-1079, src/material.pyx:441-470), which grades every
# axis whose half size is at least the thickness, whatever faces the shell
# has.  Reproduced for parity; the periodic workloads use a shell wider than
# the cell along x and y (workloads.z_cpml_shell) so that only z is graded.

def test_z_only_shell_grades_x_near_the_seams_as_the_reference_does():
    from gmes_amd import workloads
    space = Cartesian(size=(3, 3, 3), resolution=16)
    space.dt = 0.01
    for shell, graded in ((g.Shell(g.Cpml(), thickness=1.0, plus_x=False, minus_x=False,
                                   plus_y=False, minus_y=False), True),
                          (workloads.z_cpml_shell(3, 3, 3, 1.0), False)):
        shell.init(space)
        m = shell.material
        edge = 1.5 - 0.5 / 16              # outermost interior Ex coordinate along x
        assert (m.sigma(edge, 0) > 0) == graded
        assert m.sigma(0.0, 0) == 0 and m.sigma(1.45, 2) > 0
        assert [tuple(b.low[2:]) for b in shell.box_list] == [(0.5,), (-1.5,)]
