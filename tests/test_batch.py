"""Batched execution: many heterogeneous simulations advanced together by
one launch per kernel kind (gmes_amd.batch.Batch over gmes_batch_*).  Each
member must compute exactly what it computes alone, in FP64 and FP32, real
and complex, with and without the captured graph, and after the tile
autotuner has changed its launch shape."""

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd.batch import Batch
from cases import (bloch2d, bloch3d, const_case, dispersive2d, drude, fields, kat_tmz, lorentz, tfsf3d,
                   mode_case, rel_l2, spheres3d, tfsf2d)


def real_members(lib, precision):
    return [
        kat_tmz(lib, precision=precision),
        mode_case(g.TEzFDTD, lib, precision=precision),
        mode_case(g.TEMzFDTD, lib, precision=precision, pml=g.Upml),
        spheres3d(lib, precision=precision, n=24),
        spheres3d(lib, precision=precision, n=40, nsph=4, pml=5, seed=7),
        dispersive2d(drude(), lib, precision=precision),
        dispersive2d(lorentz(), lib, precision=precision, mode=g.TEzFDTD),
        const_case(lib, precision=precision),
        tfsf2d(lib, precision=precision, res=10),
        tfsf2d(lib, precision=precision, res=10, scatterer=False),  # a second auxiliary grid
    ]


def complex_members(lib, precision):
    return [
        bloch2d(lib, precision=precision),
        bloch3d(lib, precision=precision),
        tfsf2d(lib, precision=precision, res=10, bloch=(0, 0, 0)),
        tfsf3d(lib, precision=precision, bloch=(0.3, 0, 0.1)),
        dispersive2d(drude(), lib, precision=precision, bloch=(0.2, 0.1, 0)),
    ]


def run_alone(members, chunks):
    for n in chunks:
        for m in members:
            m._advance(n)
    return [fields(m) for m in members]


def run_batched(members, chunks, **kw):
    b = Batch(members, **kw)
    for n in chunks:
        b.advance(n)
    b.sync()
    out = [fields(m) for m in members]
    b.close()
    return out, b.tiles


CHUNKS = (7, 13, 1)


def test_oracle_batch_matches_single_steps(oracle_lib):
    alone = run_alone(real_members(oracle_lib, "fp64"), CHUNKS)
    batched, _ = run_batched(real_members(oracle_lib, "fp64"), CHUNKS)
    for a, b in zip(alone, batched):
        for c in a:
            np.testing.assert_array_equal(a[c], b[c])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
def test_batch_members_match_alone(engine_lib, precision, graph):
    alone = run_alone(real_members(engine_lib, precision), CHUNKS)
    batched, tiles = run_batched(real_members(engine_lib, precision), CHUNKS, graph=graph, autotune=True)
    assert all(t is not None for t in tiles)
    for i, (a, b) in enumerate(zip(alone, batched)):
        for c in a:
            np.testing.assert_array_equal(a[c], b[c], err_msg="member %d %s" % (i, c.__name__))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_complex_batch_members_match_alone(engine_lib, precision):
    alone = run_alone(complex_members(engine_lib, precision), CHUNKS)
    batched, _ = run_batched(complex_members(engine_lib, precision), CHUNKS)
    for i, (a, b) in enumerate(zip(alone, batched)):
        for c in a:
            np.testing.assert_array_equal(a[c], b[c], err_msg="member %d %s" % (i, c.__name__))


def tmz_many(lib, precision):
    """TMz with 140 cylinders of distinct permittivity (141+ records)."""
    rng = np.random.default_rng(12)
    space = g.Cartesian(size=(12, 12, 0), resolution=6)
    geometry = [g.DefaultMedium(material=g.Dielectric())]
    for _ in range(140):
        c = rng.uniform(-4.5, 4.5, 2)
        geometry.append(g.Cylinder(g.Dielectric(float(rng.uniform(1.5, 9))), center=(c[0], c[1], 0),
                                   radius=0.3, axis=(0, 0, 1)))
    geometry.append(g.Shell(g.Cpml(), thickness=1))
    src = [g.PointSource(g.Bandpass(freq=0.6, fwidth=0.2), center=(0.1, -0.2, 0), component=g.Ez)]
    sim = g.TMzFDTD(space, geometry, src, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_with_large_record_table(engine_lib, precision):
    """A 3-D and a TMz member with more material records than the LDS
    staging holds (140 spheres / cylinders of distinct permittivity) make
    the batch read the records through the generic path; results still
    equal the runs alone."""
    def members():
        return [spheres3d(engine_lib, precision=precision, n=32, nsph=140, seed=11),
                spheres3d(engine_lib, precision=precision, n=24),
                tmz_many(engine_lib, precision), kat_tmz(engine_lib, precision=precision)]
    m = members()
    assert len(m[0]._tiles.records) > 128 and len(m[2]._tiles.records) > 128
    alone = run_alone(members(), CHUNKS)
    batched, _ = run_batched(members(), CHUNKS)
    for i, (a, b) in enumerate(zip(alone, batched)):
        for c in a:
            np.testing.assert_array_equal(a[c], b[c], err_msg="member %d %s" % (i, c.__name__))


@pytest.mark.gpu
def test_batch_matches_oracle(engine_lib, oracle_lib):
    batched, _ = run_batched(real_members(engine_lib, "fp64"), CHUNKS)
    ref = run_alone(real_members(oracle_lib, "fp64"), CHUNKS)
    for a, b in zip(batched, ref):
        assert rel_l2(a, b) <= 1e-12


@pytest.mark.gpu
def test_autotune_keeps_results_and_sim_state(engine_lib):
    """Tuning runs trial launches on the live sim and restores its state:
    a tuned run matches an untuned one bit for bit."""
    a = spheres3d(engine_lib, precision="fp32", n=40, nsph=4, pml=5, seed=7)
    b = spheres3d(engine_lib, precision="fp32", n=40, nsph=4, pml=5, seed=7)
    a._advance(9)
    b._advance(9)
    bx, by = engine_lib.autotune(b._sim)
    assert bx * by == 256 and bx >= 32
    a._advance(11)
    b._advance(11)
    fa, fb = fields(a), fields(b)
    for c in fa:
        np.testing.assert_array_equal(fa[c], fb[c])


@pytest.mark.gpu
def test_member_and_aux_lifetimes_are_enforced(engine_lib):
    """A batch member cannot be destroyed while its batch exists, and a sim
    whose auxiliary TFSF grid was destroyed refuses to step (its source
    operations read that grid)."""
    from cases import kat_tmz, tfsf2d
    a, b = kat_tmz(engine_lib), kat_tmz(engine_lib)
    bt = engine_lib.batch_create([a._sim, b._sim])
    with pytest.raises(ValueError):
        engine_lib.sim_destroy(a._sim)
    engine_lib.batch_destroy(bt)
    a._advance(3)
    a.close()
    b.close()
    sim = tfsf2d(engine_lib, res=10)
    sim._advance(2)
    sim._aux_fdtds[0].close()
    with pytest.raises(ValueError):
        sim._advance(1)
    sim.close()
