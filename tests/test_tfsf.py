"""TFSF plane-wave and Gaussian-beam sources on the CPU oracles.

What pins this path (the reference's Python TFSF pipeline cannot be run
here): the reference's own compiled updaters agree bitwise with the C
oracle on the same source tables; a plane wave in vacuum does not leak
into the scattered-field region (which only holds when the auxiliary grid's
dispersion, the face index ranges, the sampling offsets and the signs are
all right); the incident wave has unit amplitude and free-space impedance;
the reference's recorded run of examples/tfsf_with_scatterer.py
(examples/VERIFICATION.md:15) is reproduced to the extent explained in
test_verification_peaks; and the reference's unit tests of the source
(tests/test_source_time.py:100-135) are restated on our op tables.
"""

import math

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd import constant as const
from cases import fields, gaussian_beam2d, rel_l2, tfsf2d, tfsf3d


def _sf_tf_max(sim, comp, margin):
    """max |F| strictly inside the TFSF box (less margin) and in the band
    between the box (plus margin) and the CPML."""
    src = sim.src_list[0]
    a = np.abs(np.asarray(sim.field[comp]))
    idx = np.indices(a.shape).reshape(3, -1).T
    pts = np.array([sim.space.index_to_space(comp, *i) for i in idx])
    rel = np.abs(pts - src.center)
    active = [ax for ax in range(3) if sim.space.half_size[ax] > sim.space.dr[ax]]
    inside = np.all(rel[:, active] < src.half_size[active] - margin, axis=1)
    lim = sim.space.half_size - 0.5 - margin  # CPML thickness 0.5
    outside = (np.any(rel[:, active] > src.half_size[active] + margin, axis=1)
               & np.all(np.abs(pts[:, active]) < lim[active], axis=1))
    flat = a.reshape(-1)
    return flat[inside].max(), flat[outside].max()


@pytest.mark.parametrize("builder,kw,steps", [
    (tfsf2d, dict(res=10), 150),
    (tfsf2d, dict(res=10, bloch=(0, 0, 0)), 100),
    (tfsf3d, {}, 60),
    (tfsf3d, dict(bloch=(0.3, 0, 0.1)), 40),
    (gaussian_beam2d, {}, 150),
], ids=["2d", "2d_complex", "3d", "3d_bloch", "gaussian_beam"])
def test_oracle_matches_reference_updaters(oracle_lib, ref_lib, builder, kw, steps):
    a = builder(oracle_lib, **kw)
    b = builder(ref_lib, **kw)
    a._advance(steps)
    b._advance(steps)
    fa, fb = fields(a), fields(b)
    assert max(np.abs(v).max() for v in fa.values()) > 0.1
    for c in fa:
        np.testing.assert_array_equal(fa[c], fb[c])


def test_vacuum_leakage_2d_and_incident_wave(oracle_lib):
    """Plane wave along (1,-1,0) with Ez polarization in vacuum: the
    scattered-field region stays empty and the total-field region carries
    a unit wave with |Hx| = |Hy| = 1/sqrt(2)."""
    sim = tfsf2d(oracle_lib, scatterer=False, res=10)
    sim.step_until_t(30)
    tf, sf = _sf_tf_max(sim, g.Ez, 0.15)
    assert abs(tf - 1) < 2e-3
    assert sf < 1e-8 * tf
    for c in (g.Hx, g.Hy):
        tf, sf = _sf_tf_max(sim, c, 0.15)
        assert abs(tf - 1 / math.sqrt(2)) < 2e-3
        assert sf < 1e-8 * tf


@pytest.mark.parametrize("direction", [(1, 0, 0), (0, 1, 0)], ids=["kx", "ky"])
def test_vacuum_leakage_3d(oracle_lib, direction):
    """3-D box, CW wave along an axis, E along z: Ez, Ex/Ey, Hx and Hy
    faces over full 3-D ranges.  The box's low edges sit on grid nodes and
    its high edges half way between nodes (8 cells per unit): the face
    ranges come from rounding exact indices with floor(x + 1/2), which for
    the staggered components only brackets the same cells as the
    integer-indexed ones for such edges (gmes/source.py:705-1406)."""
    sim = tfsf3d(oracle_lib, scatterer=False, direction=direction, polarization=(0, 0, 1),
                 center=(0.03125,) * 3, size=(1.5625,) * 3, src_time=g.Continuous(freq=0.8))
    sim._advance(120)
    for c in const.COMPONENTS:
        tf, sf = _sf_tf_max(sim, c, 0.2)
        assert sf < 1e-9, c
        if c in (g.Ez, g.Hx if direction[1] else g.Hy):
            assert abs(tf - 1) < 0.05


def test_oblique_3d_runs_all_faces(oracle_lib):
    """An oblique wave corrects all 24 faces; the incident pulse fills the
    total-field region with a finite field of the expected size."""
    sim = tfsf3d(oracle_lib, scatterer=False)
    ops = sum(v.idx_size() for d in sim.pw_source.values() for v in d.values())
    assert ops > 1000
    sim._advance(60)
    for c in const.COMPONENTS:
        tf, _ = _sf_tf_max(sim, c, 0.2)
        assert 0.01 < tf < 1, c


def test_verification_peaks(oracle_lib):
    """examples/VERIFICATION.md:15 records peak |Ez|, |Hx|, |Hy| of 2.0727,
    1.9535 and 1.9615 for examples/tfsf_with_scatterer.py at t = 200
    (5714 steps).  Twelve Ez points of that grid lie exactly on the
    cylinder's surface; whether each counts as inside rests on the last bit
    of sqrt(x**2 + y**2) <= 1 and moves these peaks by several percent
    (2.07-2.13 for Ez).  With the reference's own test in IEEE double on
    this machine the run gives 2.07298, 1.98895, 1.99336: Ez agrees to
    1.5e-4 and Hx, Hy to 2 percent."""
    sim = tfsf2d(oracle_lib)
    sim.step_until_t(200)
    assert sim.time_step.n == 5714.0
    pk = {c: float(np.abs(np.asarray(sim.field[c])).max()) for c in (g.Ez, g.Hx, g.Hy)}
    assert abs(pk[g.Ez] - 2.0727) / 2.0727 < 2e-4
    assert abs(pk[g.Hx] - 1.9535) / 1.9535 < 0.025
    assert abs(pk[g.Hy] - 1.9615) / 1.9615 < 0.025


def test_dispersion_matched_aux_grid(oracle_lib):
    """The auxiliary grid's cell size solves the 1-D relation for the
    wavenumber that the 3-D relation gives along k (Taflove & Hagness eqs.
    5.65, 5.67; gmes/source.py:506-574)."""
    sim = tfsf3d(oracle_lib, scatterer=False)
    src = sim.src_list[0]
    aux = src.aux_fdtd
    dt = sim.time_step.dt
    omega = 2 * math.pi * src.src_time.freq
    ds = np.array(sim.space.dr)
    delta = 1 / aux.space.res[2]
    # the 3-D numerical wavenumber along k
    lhs = (math.sin(0.5 * dt * omega) / dt) ** 2
    zeta = None
    lo, hi = 0.0, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        f = lambda z: lhs - np.sum((np.sin(0.5 * z * ds * omega * src.k) / ds) ** 2)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    zeta = 0.5 * (lo + hi)
    k_num = zeta * omega
    # the 1-D grid propagates the same wavenumber
    assert abs(math.sin(0.5 * omega * dt) / dt - math.sin(0.5 * k_num * delta) / delta) < 1e-9
    assert aux.space.whole_field_size[2] > 100


def test_face_ops_and_edge_merge(oracle_lib):
    """Sampling indices are integers inside the auxiliary grid and E cells
    on an edge of the box carry one correction per face
    (tests/test_source_time.py:100-135)."""
    sim = tfsf3d(oracle_lib, scatterer=False)
    src = sim.src_list[0]
    aux_shape = {c: src.aux_fdtd._shape[c] for c in (g.Ex, g.Hy)}
    for comp in const.COMPONENTS:
        ops = src._pw_ops(comp, sim._shape[comp], sim.space, sim.geom_tree)
        assert ops
        per_cell = {}
        for idx, typ, kind, p in ops:
            assert typ == "tfsf"
            for key in ("i0", "i1"):
                assert isinstance(p[key], int)
            n = int(np.prod(aux_shape[g.Hy if comp in const.E_COMPONENTS else g.Ex]))
            assert 0 <= p["i0"] < n and p["i1"] == p["i0"] + 1
            assert 0 <= p["r1"] < 1 and p["r0"] + p["r1"] == 1
            per_cell.setdefault(idx, []).append(p["face"])
        faces = {f for fs in per_cell.values() for f in fs}
        assert len(faces) == 4
        # E cells on a box edge are total-field nodes next to two faces;
        # H cells are corrected across one face only
        merged = any(len(fs) == 2 for fs in per_cell.values())
        assert merged == (comp in const.E_COMPONENTS)
        assert all(len(set(fs)) == len(fs) for fs in per_cell.values())


def test_invalid_arguments():
    """gmes/source.py:391-404"""
    with pytest.raises(ValueError):
        g.TotalFieldScatteredField(g.Continuous(1), (np.inf, 0, 0), (1, 1, 1), (1, 0, 0), (0, 0, 1))
    with pytest.raises(ValueError):
        g.TotalFieldScatteredField(g.Continuous(1), (0, 0, 0), (1, np.nan, 1), (1, 0, 0), (0, 0, 1))
    with pytest.raises(TypeError):
        g.TotalFieldScatteredField(1.0, (0, 0, 0), (1, 1, 1), (1, 0, 0), (0, 0, 1))
    with pytest.raises(TypeError):
        g.GaussianBeam(g.Continuous(1), 1, (0, 0, 0), (1, 1, 1), (1, 0, 0), (0, 0, 1))


def test_gaussian_beam_envelope(oracle_lib):
    """The beam's auxiliary grid is pre-run to a steady state; its samples
    are then scaled by sin^2(pi t / 2 w) with t counted from the main grid's
    first step (gmes/source.py:1583-1603,1734-1769): the first E half step
    injects nothing, and the profile across the face is Gaussian."""
    sim = gaussian_beam2d(oracle_lib)
    src = sim.src_list[0]
    assert src.aux_fdtd.time_step.n > 0
    sim._advance(1)
    assert np.abs(np.asarray(sim.field[g.Ez])).max() == 0
    assert np.abs(np.asarray(sim.field[g.Hy])).max() > 0
    sim.step_until_t(12)
    ez = np.abs(np.asarray(sim.field[g.Ez]))[:, :, 0]
    i = int(round((0.6 + 2.0) / 0.1))
    col = ez[i, :]
    j0 = int(np.argmax(col))
    assert abs(j0 - 20) <= 1
