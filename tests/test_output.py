"""Probe files, field dumps and point-source recordings: the reference's
file names and text formats (gmes/file_io.py:9-48, gmes/fdtd.py:607-671,
1303-1424, gmes/pw_source.py:63-67), on the CPU oracle."""

import numpy as np
import pytest

import gmes_amd as g


def _tmz(lib, filename=None):
    space = g.Cartesian(size=(2, 2, 0), resolution=8)
    geometry = [g.DefaultMedium(material=g.Dielectric()), g.Shell(g.Cpml(), thickness=0.5)]
    src = [g.PointSource(g.Bandpass(freq=1.0, fwidth=0.5), center=(0, 0, 0), component=g.Ez,
                         filename=filename)]
    sim = g.TMzFDTD(space, geometry, src, verbose=False, _engine=lib)
    sim.init()
    return sim


def test_probe_files_hold_every_half_step(oracle_lib, tmp_path):
    """set_probe writes '# location=', '# dt=' and one 'n value' line per
    half step: E components at n = k + 1/2, H components at n = k + 1."""
    sim = _tmz(oracle_lib)
    p = (0.125, -0.25, 0)
    prefix = str(tmp_path / "probe")
    sim.set_probe(p, prefix)
    want = {c: [] for c in (g.Ez, g.Hx, g.Hy)}
    idx = {c: sim.space.space_to_index(c, *p) for c in want}
    for k in range(12):
        sim.step()
        for c in want:
            n = k + (0.5 if c is g.Ez else 1.0)
            want[c].append(str(n) + " " + str(np.asarray(sim.field[c])[idx[c]]))
    for r in sim.e_recorder + sim.h_recorder:
        r.close()
    for c, suffix in ((g.Ez, "_ez.dat"), (g.Hx, "_hx.dat"), (g.Hy, "_hy.dat")):
        lines = open(prefix + suffix).read().splitlines()
        loc = sim.space.index_to_space(c, *idx[c])
        assert lines[0] == "# location=" + str(loc)
        assert lines[1] == "# dt=" + str(sim.time_step.dt)
        assert lines[2:] == want[c]
    assert any(float(v.split()[1]) != 0 for v in want[g.Ez])


def test_write_field_file_name_and_box(oracle_lib, tmp_path, monkeypatch):
    """write_field saves the component inside [low, high] as
    '<comp.str()>_t=<t>_(0,0,0).npy' ('ez' for Ez, src/constant.i:73-74),
    with optional prefix and postfix."""
    monkeypatch.chdir(tmp_path)
    sim = _tmz(oracle_lib)
    for _ in range(7):
        sim.step()
    sim.write_field(g.Ez, prefix="a_", postfix="_b")
    name = "a_ez_t=%f_(0,0,0)_b.npy" % sim.time_step.t
    data = np.load(tmp_path / name)
    ez = np.asarray(sim.field[g.Ez])
    # the whole non-ghost extent: every plane but the last along x and y
    np.testing.assert_array_equal(data, ez[:-1, :-1, :])
    sim.write_field(g.Ez, low=(-0.5, -0.5, 0), high=(0.5, 0.5, 0))
    box = np.load(tmp_path / ("ez_t=%f_(0,0,0).npy" % sim.time_step.t))
    lo = sim.space.space_to_index(g.Ez, -0.5, -0.5, 0)
    hi = sim.space.space_to_index(g.Ez, 0.5, 0.5, 0)
    np.testing.assert_array_equal(box, ez[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1])


def test_point_source_recording(oracle_lib, tmp_path):
    """A point source with a filename records 't<TAB>value' once per step
    at t = dt*(k + 1/2), the time its E component is driven."""
    rec = tmp_path / "src.dat"
    sim = _tmz(oracle_lib, filename=str(rec))
    steps = 9
    for _ in range(steps):
        sim.step()
    sim.src_list[0]._file.flush()
    lines = open(rec).read().splitlines()
    dt = sim.time_step.dt
    st = g.Bandpass(freq=1.0, fwidth=0.5)
    st.init(False)
    want = ["%f\t%f" % (dt * (k + 0.5), st.oscillator(dt * (k + 0.5))) for k in range(steps)]
    assert lines == want


@pytest.mark.gpu
def test_probe_files_gpu_match_oracle(engine_lib, oracle_lib, tmp_path):
    """Probe samples recorded on the device (a ring buffer filled inside
    each launched step and drained after the steps) write the same files,
    character for character, as the oracle's."""
    out = []
    for name, lib in (("gpu", engine_lib), ("cpu", oracle_lib)):
        sim = _tmz(lib)
        prefix = str(tmp_path / name)
        sim.set_probe((0.125, -0.25, 0), prefix)
        sim._advance(7)
        sim.step()
        sim._advance(30)
        for r in sim.e_recorder + sim.h_recorder:
            r.close()
        out.append([open(prefix + s).read() for s in ("_ez.dat", "_hx.dat", "_hy.dat")])
    assert out[0] == out[1]
    assert len(out[0][0].splitlines()) == 2 + 38


def _material_update_sizes(simulation):
    """benchmarks/field_updates.py:83-96 of the reference, restated."""
    return sorted(({"cells": u.idx_size(), "component": c.__name__, "material": type(u).__name__}
                   for c, updaters in simulation.pw_material.items() for u in updaters.values()),
                  key=lambda item: (item["component"], item["material"]))


def test_pw_material_views_carry_reference_type_names(oracle_lib):
    """tests/test_benchmarks.py:39-56 of the reference reads updater type
    names such as DielectricEzReal and CpmlEzReal; each material class names
    its own updaters for every component (src/material.pyx), ghost planes
    are Dummy, and the counts cover every storage element."""
    space = g.Cartesian(size=(4, 4, 0), resolution=5)
    geometry = [g.DefaultMedium(g.Dielectric()),
                g.Block(g.Drude(dps=(g.DrudePole(omega=1.0, gamma=0.1),)), center=(0.5, 0, 0),
                        size=(1, 1, 1)),
                g.Block(g.DcpRc(dps=(g.DrudePole(omega=1.0, gamma=0.1),),
                                cps=(g.CriticalPoint(amp=0.5, phi=0.1, omega=1.0, gamma=0.2),)),
                        center=(-0.8, 0, 0), size=(0.6, 0.6, 1)),
                g.Shell(g.Cpml(), thickness=0.6)]
    sim = g.TMzFDTD(space, geometry, [], verbose=False, _engine=oracle_lib)
    sim.init()
    items = _material_update_sizes(sim)
    names = {i["material"] for i in items}
    for want in ("DielectricEzReal", "CpmlEzReal", "DrudeEzReal", "DrudeHxReal", "DcpPlrcEzReal",
                 "DummyHxReal", "CpmlHyReal"):
        assert want in names, want
    for comp in (g.Ez, g.Hx, g.Hy):
        total = sum(i["cells"] for i in items if i["component"] == comp.__name__)
        assert total == sim.field[comp].size
    view = next(u for u in sim.pw_material[g.Ez].values() if type(u).__name__ == "DrudeEzReal")
    assert view.name() == "DrudeElectric"
    sim.close()
