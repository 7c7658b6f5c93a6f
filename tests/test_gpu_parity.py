"""Parity of the HIP engine (through the C-ABI) with the CPU oracle.

FP64 runs must match the oracle to a relative L2 of 1e-12 over all six
fields (the kernels keep the reference's operation order, so the match is
normally exact).  FP32 runs are checked against the engine's own FP64 run
of the same case: relative L2 <= 1e-5 unless a looser bound is stated
with the case.
"""

import numpy as np
import pytest

import gmes_amd as g
from cases import (air2d, bloch2d, bloch3d, const_case, currents, dcp_ade, dcp_plrc, dcp_rc,
                   dispersive2d, dm2_1d, drude, fields, gaussian_beam2d, kat_tmz, lorentz,
                   many_records3d, mode_case, rel_l2, spheres3d, tfsf2d, tfsf3d)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def run(builder, lib, steps, **kw):
    sim = builder(lib, **kw)
    for _ in range(steps):
        sim.step()
    out = fields(sim)
    sim.close()
    return out


def test_kat_tmz_regression(engine_lib):
    """The reference's TMz known answer (tests/test_fdtd.py:55-79)."""
    sim = kat_tmz(engine_lib)
    for _ in range(5):
        sim.step()
    assert sim.time_step.n == 5.0
    assert abs(sim.time_step.t - 0.7000357133746822) < 1e-15
    assert sim.ez.shape == (11, 11, 1)
    assert sim.ez[5, 5, 0] == -0.9996801161298625
    assert abs(np.sum(np.abs(sim.ez) ** 2) - 1.5780099423691636) < 1e-14


def test_kat_tmz_matches_oracle_bitwise(engine_lib, oracle_lib):
    a = run(kat_tmz, engine_lib, 5)
    b = run(kat_tmz, oracle_lib, 5)
    for c in a:
        np.testing.assert_array_equal(a[c], b[c])


def test_air2d_full_run(engine_lib, oracle_lib):
    """Plain method, FP64: examples/air2d.py to t = 10 (286 steps)."""
    sim = air2d(engine_lib)
    sim.step_until_t(10)
    assert sim.time_step.n == 286.0
    gpu = fields(sim)
    # examples/VERIFICATION.md:7 records the peak |Ez| of this run
    assert abs(np.max(np.abs(gpu[g.Ez])) - 0.999382775451675) < 1e-12
    ref = air2d(oracle_lib)
    ref.step_until_t(10)
    assert rel_l2(gpu, fields(ref)) <= FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spheres3d(engine_lib, oracle_lib, precision):
    steps = 60
    gpu = run(spheres3d, engine_lib, steps, precision=precision)
    if precision == "fp64":
        assert rel_l2(gpu, run(spheres3d, oracle_lib, steps)) <= FP64_TOL
    else:
        assert rel_l2(gpu, run(spheres3d, engine_lib, steps)) <= 1e-5


MODES = [g.TExFDTD, g.TEyFDTD, g.TEzFDTD, g.TMxFDTD, g.TMyFDTD, g.TMzFDTD, g.TEMxFDTD,
         g.TEMyFDTD, g.TEMzFDTD]


@pytest.mark.parametrize("mode", MODES, ids=lambda m: m.__name__)
@pytest.mark.parametrize("pml", [g.Cpml, g.Upml], ids=["cpml", "upml"])
def test_modes(engine_lib, oracle_lib, mode, pml):
    gpu = run(lambda lib: mode_case(mode, lib, pml=pml), engine_lib, 40)
    ref = run(lambda lib: mode_case(mode, lib, pml=pml), oracle_lib, 40)
    assert rel_l2(gpu, ref) <= FP64_TOL


@pytest.mark.parametrize("mat", [drude, lorentz, dcp_ade, dcp_plrc, dcp_rc],
                         ids=["drude", "lorentz", "dcp_ade", "dcp_plrc", "dcp_rc"])
@pytest.mark.parametrize("mode", [g.TMzFDTD, g.TEzFDTD], ids=["tmz", "tez"])
def test_dispersive(engine_lib, oracle_lib, mat, mode):
    steps = 50
    gpu = run(lambda lib, **kw: dispersive2d(mat(), lib, mode=mode, **kw), engine_lib, steps)
    ref = run(lambda lib, **kw: dispersive2d(mat(), lib, mode=mode, **kw), oracle_lib, steps)
    assert rel_l2(gpu, ref) <= FP64_TOL
    g32 = run(lambda lib, **kw: dispersive2d(mat(), lib, mode=mode, **kw), engine_lib, steps,
              precision="fp32")
    assert rel_l2(g32, gpu) <= 1e-5


@pytest.mark.parametrize("mat", [drude, dcp_plrc], ids=["drude", "dcp_plrc"])
def test_dispersive_bloch(engine_lib, oracle_lib, mat):
    gpu = run(lambda lib: dispersive2d(mat(), lib, bloch=(0.2, 0.1, 0)), engine_lib, 40)
    ref = run(lambda lib: dispersive2d(mat(), lib, bloch=(0.2, 0.1, 0)), oracle_lib, 40)
    assert rel_l2(gpu, ref) <= FP64_TOL


def test_const_and_upml3d(engine_lib, oracle_lib):
    gpu = run(const_case, engine_lib, 30)
    ref = run(const_case, oracle_lib, 30)
    assert rel_l2(gpu, ref) <= FP64_TOL


@pytest.mark.parametrize("builder,kw", [(spheres3d, dict(bound=g.Upml)), (const_case, {})],
                         ids=["spheres3d_upml", "const_upml"])
def test_upml3d_fp32(engine_lib, oracle_lib, builder, kw):
    """3-D UPML boxes: FP64 equals the oracle; FP32, whose 3-D kernels load
    the UPML state word with the cell's other operands, stays within 1e-5
    of the FP64 run."""
    steps = 40
    gpu = run(builder, engine_lib, steps, **kw)
    assert rel_l2(gpu, run(builder, oracle_lib, steps, **kw)) <= FP64_TOL
    assert rel_l2(run(builder, engine_lib, steps, precision="fp32", **kw), gpu) <= 1e-5


def test_dm2(engine_lib, oracle_lib):
    gpu = run(dm2_1d, engine_lib, 80)
    ref = run(dm2_1d, oracle_lib, 80)
    err = rel_l2(gpu, ref)
    print("DM2 GPU vs oracle, 80 steps: rel L2 %.3g" % err)
    assert err <= FP64_TOL


def test_dm2_failure_is_reported(engine_lib):
    """Dm2(rtol=-1) never converges (tests/test_pw_dm2.py:95-101)."""
    sim = dm2_1d(engine_lib, rtol=-1)
    with pytest.raises(RuntimeError, match="failed to converge"):
        for _ in range(40):
            sim.step()
        sim.sync()


def test_currents_and_shared_cell(engine_lib, oracle_lib):
    gpu = run(currents, engine_lib, 30)
    ref = run(currents, oracle_lib, 30)
    assert rel_l2(gpu, ref) <= FP64_TOL


@pytest.mark.parametrize("builder", [bloch2d, bloch3d], ids=["2d", "3d"])
def test_bloch(engine_lib, oracle_lib, builder):
    gpu = run(builder, engine_lib, 40)
    ref = run(builder, oracle_lib, 40)
    assert rel_l2(gpu, ref) <= FP64_TOL
    g32 = run(builder, engine_lib, 40, precision="fp32")
    assert rel_l2(g32, gpu) <= 1e-5


def test_graph_replay_matches_direct(engine_lib):
    a = spheres3d(engine_lib, n=16, nsph=3, pml=3)
    b = spheres3d(engine_lib, n=16, nsph=3, pml=3)
    engine_lib.set_graph(b._sim, 1)
    for k in range(3):
        a._advance(7)
        b._advance(7)
        fa, fb = fields(a), fields(b)
        for c in fa:
            np.testing.assert_array_equal(fa[c], fb[c])


def test_step_batches_match_single_steps(engine_lib):
    a = kat_tmz(engine_lib)
    b = kat_tmz(engine_lib)
    for _ in range(9):
        a.step()
    b._advance(9)
    for c in g.constant.COMPONENTS:
        np.testing.assert_array_equal(a.field[c], b.field[c])


def test_host_write_before_step(engine_lib, oracle_lib):
    """sim.ez[idx] = 1 before stepping reaches the device
    (tests/test_fdtd.py:28-33 of the reference)."""
    out = []
    for lib in (engine_lib, oracle_lib):
        sim = g.TMzFDTD(g.Cartesian(size=(2, 2, 0), resolution=3),
                        [g.DefaultMedium(material=g.Dielectric())], verbose=False, _engine=lib)
        sim.init()
        idx = sim.space.space_to_ez_index(0, 0, 0)
        sim.ez[idx] = 1
        sim.step_while_zero(g.Ez, (0, 0, 0))
        assert sim.time_step.n == 1
        for _ in range(3):
            sim.step()
        out.append(fields(sim))
    assert rel_l2(out[0], out[1]) <= FP64_TOL


TFSF_CASES = [
    (tfsf2d, dict(res=10), 150),
    (tfsf2d, dict(res=10, bloch=(0, 0, 0)), 100),
    (tfsf3d, {}, 60),
    (tfsf3d, dict(bloch=(0.3, 0, 0.1)), 40),
    (gaussian_beam2d, {}, 150),
]
TFSF_IDS = ["2d", "2d_complex", "3d_oblique", "3d_bloch", "gaussian_beam"]


@pytest.mark.parametrize("builder,kw,steps", TFSF_CASES, ids=TFSF_IDS)
def test_tfsf(engine_lib, oracle_lib, builder, kw, steps):
    """TFSF faces and the auxiliary 1-D grid advanced on the device."""
    gpu = run(builder, engine_lib, steps, **kw)
    assert max(np.abs(v).max() for v in gpu.values()) > 0.1
    assert rel_l2(gpu, run(builder, oracle_lib, steps, **kw)) <= FP64_TOL


@pytest.mark.parametrize("builder,kw,steps", [TFSF_CASES[0], TFSF_CASES[2]], ids=["2d", "3d"])
def test_tfsf_fp32(engine_lib, builder, kw, steps):
    gpu = run(builder, engine_lib, steps, precision="fp32", **kw)
    assert rel_l2(gpu, run(builder, engine_lib, steps, **kw)) <= 1e-5


def test_tfsf_scatterer_full_run(engine_lib, oracle_lib):
    """examples/tfsf_with_scatterer.py to t = 200 (5714 steps) in FP64."""
    sim = tfsf2d(engine_lib)
    sim.step_until_t(200)
    assert sim.time_step.n == 5714.0
    ref = tfsf2d(oracle_lib)
    ref.step_until_t(200)
    gpu = fields(sim)
    assert rel_l2(gpu, fields(ref)) <= FP64_TOL
    assert abs(np.abs(gpu[g.Ez]).max() - 2.0727) / 2.0727 < 2e-4


def _c3(lib, precision="fp64"):
    from gmes_amd import workloads
    return workloads.dispersive3d(scale=0.125, precision=precision, _engine=lib,
                                  src_time=g.Bandpass(freq=0.05, fwidth=0.1, s=3))


def test_dispersive_bloch_cell_c128(engine_lib, oracle_lib):
    """The dispersive Bloch cell (Drude blocks, Lorentz spheres, DCP slab,
    CPML on z, TFSF) at 1/8 scale, complex128, against the oracle."""
    a, b = _c3(engine_lib), _c3(oracle_lib)
    a._advance(80)
    b._advance(80)
    fa, fb = fields(a), fields(b)
    assert max(np.abs(v).max() for v in fb.values()) > 1e-3
    assert rel_l2(fa, fb) <= FP64_TOL


def test_dispersive_bloch_cell_c64(engine_lib):
    a, b = _c3(engine_lib, "fp32"), _c3(engine_lib)
    a._advance(80)
    b._advance(80)
    assert rel_l2(fields(a), fields(b)) <= 1e-5


@pytest.mark.parametrize("builder,kw", [(bloch3d, {}), (spheres3d, dict(n=16, nsph=3, pml=3)),
                                        (tfsf3d, dict(bloch=(0.3, 0, 0.1)))],
                         ids=["bloch3d", "spheres3d", "tfsf3d_bloch"])
def test_fused_e_ghosts_match_separate_copy(engine_lib, builder, kw):
    """E ghost planes written by the 3-D E kernel give the same fields, bit
    for bit, as the separate ghost copy after the E sources."""
    a, b = builder(engine_lib, **kw), builder(engine_lib, **kw)
    assert engine_lib.set_ghost_fusion(a._sim, 1)
    assert not engine_lib.set_ghost_fusion(b._sim, 0)
    for _ in range(3):
        a._advance(11)
        b._advance(11)
        fa, fb = fields(a), fields(b)
        for c in fa:
            np.testing.assert_array_equal(fa[c], fb[c])


def tez_hz_source(lib, precision="fp64", bloch=None):
    """TEz with an Hz point source: no E source operation, so the E ghost
    planes (including the z plane of Ex and Ey) can be fused."""
    space = g.Cartesian(size=(3, 4, 0), resolution=8)
    geometry = [g.DefaultMedium(material=g.Dielectric()),
                g.Block(g.Dielectric(3.0), center=(0.4, 0.2, 0), size=(0.8, 1.0, 1)),
                g.Shell(g.Cpml(), thickness=0.5)]
    src = [g.PointSource(g.Bandpass(freq=0.8, fwidth=0.2), center=(-0.3, 0.1, 0), component=g.Hz)]
    sim = g.TEzFDTD(space, geometry, src, verbose=False, precision=precision, bloch=bloch, _engine=lib)
    sim.init()
    return sim


@pytest.mark.parametrize("builder,kw", [(kat_tmz, {}), (bloch2d, {}), (bloch2d, dict(precision="fp32")),
                                        (tfsf2d, dict(res=10, bloch=(0, 0, 0))),
                                        (lambda lib, **kw: dispersive2d(drude(), lib, **kw), {}),
                                        (tez_hz_source, {}), (tez_hz_source, dict(bloch=(0.2, -0.4, 0)))],
                         ids=["kat_tmz", "bloch2d", "bloch2d_c64", "tfsf2d_bloch", "drude_tmz", "tez",
                              "tez_bloch"])
def test_fused_2d_ghosts_match_separate_copy(engine_lib, builder, kw):
    """2-D modes with a collapsed z axis: the E ghost planes written by the
    2-D E kernel, and in TMz the z ghost plane of Hx and Hy (from k = 1,
    Hy's corner line with both phases), give the same fields, bit for bit,
    as the separate ghost copies."""
    a, b = builder(engine_lib, **kw), builder(engine_lib, **kw)
    assert engine_lib.set_ghost_fusion(a._sim, 1)
    assert not engine_lib.set_ghost_fusion(b._sim, 0)
    for _ in range(3):
        a._advance(11)
        b._advance(11)
        fa, fb = fields(a), fields(b)
        assert max(np.abs(v).max() for v in fa.values()) > 0
        for c in fa:
            np.testing.assert_array_equal(fa[c], fb[c])


def test_ghost_fusion_off_for_dispersive_source_planes(engine_lib):
    """The dispersive Bloch cell's DCP slab covers the x and y source planes:
    its E ghosts keep the separate copy."""
    assert not engine_lib.set_ghost_fusion(_c3(engine_lib)._sim, 1)


@pytest.mark.parametrize("builder,kw", [(spheres3d, dict(n=16, nsph=3, pml=4)), (bloch3d, {})],
                         ids=["spheres3d_cpml_all_faces", "bloch3d_cpml_z"])
def test_cpml_psi_elision_is_bitwise(engine_lib, oracle_lib, builder, kw):
    """FP64 / complex128 CPML runs, where the kernels skip the psi words
    that stay zero (outside each axis's layers), equal the oracle bit for
    bit, corners and edges of the layers included."""
    a, b = builder(engine_lib, **kw), builder(oracle_lib, **kw)
    a._advance(40)
    b._advance(40)
    fa, fb = fields(a), fields(b)
    assert max(np.abs(v).max() for v in fb.values()) > 0
    for c in fa:
        np.testing.assert_array_equal(fa[c], fb[c])


def test_fp32_3d_with_more_than_256_records(engine_lib, oracle_lib):
    """Real-FP32 3-D sims index their records with bytes up to 256 records
    and with 16-bit ids beyond; this case has more than 256 (and more than
    the LDS staging holds).  FP64 equals the oracle; FP32 tracks FP64."""
    steps = 60
    sim = many_records3d(engine_lib, precision="fp32")
    assert len(sim._tiles.records) > 256
    sim.close()
    gpu = run(many_records3d, engine_lib, steps)
    assert rel_l2(gpu, run(many_records3d, oracle_lib, steps)) <= FP64_TOL
    err = rel_l2(run(many_records3d, engine_lib, steps, precision="fp32"), gpu)
    print("many records, FP32 vs FP64: rel L2 %.3g" % err)
    assert err <= 1e-5


@pytest.mark.parametrize("bound", [g.Cpml, g.Upml], ids=["cpml", "upml"])
def test_fp32_tile_shapes_give_the_same_bits(engine_lib, bound):
    """Real FP32, 3-D: the default 128 x 2 tile stages the records in LDS
    and, with CPML (or UPML) on every face, loads the CPML (or UPML)
    profile rows with the other operands; a 16 x 1 tile is too small to
    stage the 9 records (27 16-byte chunks) and reads them, and the rows,
    after each cell's record.  Same arithmetic per cell: the fields must
    agree bit for bit."""
    out = []
    for tile in (None, (16, 1)):
        sim = spheres3d(engine_lib, precision="fp32", n=40, bound=bound)
        if tile:
            engine_lib.set_tile(sim._sim, *tile)
        sim._advance(80)
        out.append(fields(sim))
        sim.close()
    for c in out[0]:
        np.testing.assert_array_equal(out[0][c], out[1][c], err_msg=c.__name__)
    assert np.abs(out[0][g.Ex]).max() > 0


@pytest.mark.parametrize("precision,bloch", [("fp32", None), ("fp64", None), ("fp32", (0.1, 0, 0))],
                         ids=["fp32", "fp64", "c64"])
@pytest.mark.parametrize("nz", [256, 40], ids=["padded_rows", "unpadded_rows"])
def test_field_write_read_round_trip(engine_lib, oracle_lib, precision, bloch, nz):
    """Rows of 256 cells or more keep the padded device layout, whose field
    reads and writes repack rows on the device; shorter 3-D rows are stored
    unpadded and copied as one block.  Random words written to every
    component come back unchanged, and steps taken after the write start
    from the written words (in FP64 the fields then equal the oracle's,
    given the same words)."""
    from gmes_amd import constant as const

    def build(lib):
        space = g.Cartesian(size=(5, 3, nz), resolution=1)
        sim = g.FDTD(space, [g.DefaultMedium(material=g.Dielectric(2.0))], [], verbose=False,
                     precision=precision, bloch=bloch, _engine=lib)
        sim.init()
        return sim

    def write_words(sim, lib):
        rng = np.random.default_rng(nz)
        words = {}
        for c in const.COMPONENTS:
            a = sim.field[c]
            w = rng.standard_normal(a.shape)
            if np.iscomplexobj(a):
                w = w + 1j * rng.standard_normal(a.shape)
            words[c] = np.ascontiguousarray(w.astype(a.dtype))
            lib.write_field(sim._sim, c.index, words[c])
        return words

    sim = build(engine_lib)
    words = write_words(sim, engine_lib)
    for c in const.COMPONENTS:
        out = np.empty_like(words[c])
        engine_lib.read_field(sim._sim, c.index, out)
        np.testing.assert_array_equal(out, words[c], err_msg=c.__name__)
    if precision == "fp64" and bloch is None:
        ref = build(oracle_lib)
        write_words(ref, oracle_lib)
        for s, lib in ((sim, engine_lib), (ref, oracle_lib)):
            lib.step(s._sim, 3)
            lib.sync(s._sim)
        for c in const.COMPONENTS:
            a, b = np.empty_like(words[c]), np.empty_like(words[c])
            engine_lib.read_field(sim._sim, c.index, a)
            oracle_lib.read_field(ref._sim, c.index, b)
            np.testing.assert_array_equal(a, b, err_msg=c.__name__)
            assert np.abs(a - words[c]).max() > 0
        ref.close()
    sim.close()
