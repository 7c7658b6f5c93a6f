"""Snapshot / restore and deepcopy.

The reference's floor is ``FDTD.__deepcopy__`` (gmes/fdtd.py:327-351),
tested by tests/test_fdtd.py:81-121: a copy carries the six fields and
the time step, and stepping the copy and the original gives the same
fields.  ``snapshot``/``restore`` go further, for pre-emption: the step
clock, CPML and dispersive state, running-DFT accumulators and the
auxiliary TFSF grid all come back, so a run restored to step k and run to
k + m equals, bit for bit, the first run to k + m.
"""

import ctypes
import os
from copy import deepcopy

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd import _abi
from gmes_amd import constant as const

from cases import fields, rel_l2


def tfsf_drude(lib, precision="fp64"):
    """TMz: a CW TFSF plane wave on a Drude cylinder inside a CPML shell,
    with a running DFT of Ez over the TFSF box and an Ez probe."""
    space = g.Cartesian(size=(4, 4, 0), resolution=10)
    geometry = [g.DefaultMedium(g.Dielectric()),
                g.Cylinder(g.Drude(eps_inf=1.5, dps=(g.DrudePole(omega=3.0, gamma=0.2),)),
                           center=(0.2, -0.1, 0), radius=0.5, height=10),
                g.Shell(material=g.Cpml(), thickness=0.5)]
    sources = [g.TotalFieldScatteredField(src_time=g.Continuous(freq=0.8), center=(0, 0, 0),
                                          size=(2.2, 2.2, 1), direction=(1, -0.5, 0),
                                          polarization=(0, 0, 1))]
    sim = g.TMzFDTD(space, geometry, sources, verbose=False, precision=precision, _engine=lib)
    sim.init()
    from gmes_amd.monitor import DftBox
    dft = DftBox(sim, const.Ez, (10, 10, 0), (30, 30, 1), freqs=(0.5, 0.8, 1.1))
    return sim, dft


def _aux_fields(sim):
    return [fields(aux) for aux in sim._aux_fdtds]


def _round_trip(lib, k=7, m=9):
    sim, dft = tfsf_drude(lib)
    sim.set_probe((0.3, 0.2, 0), prefix="snap")
    sim._advance(k)
    snap = sim.snapshot()
    sim._advance(m)
    first = (fields(sim), _aux_fields(sim), dft.values(), sim.time_step.n)
    sim.restore(snap)
    assert sim.time_step.n == k
    assert lib.get_step(sim._sim) == k
    sim._advance(m)
    second = (fields(sim), _aux_fields(sim), dft.values(), sim.time_step.n)
    sim.e_recorder[0].close()
    probe = open("snap_ez.dat").read().splitlines()
    sim.close()
    return first, second, probe


def _check_round_trip(first, second, probe, k, m):
    (f1, a1, d1, n1), (f2, a2, d2, n2) = first, second
    assert n1 == n2 == k + m
    for c in const.COMPONENTS:
        assert np.array_equal(f1[c], f2[c]), c
    assert len(a1) == len(a2) == 1
    for c in const.COMPONENTS:
        assert np.array_equal(a1[0][c], a2[0][c]), c
    assert np.array_equal(d1, d2)
    assert np.abs(d1).max() > 0
    # header, k + m samples, then m samples again after the restore
    body = [line for line in probe if not line.startswith("#")]
    assert len(body) == k + 2 * m
    assert body[k:k + m] == body[k + m:]


def test_snapshot_round_trip_oracle(oracle_lib, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    first, second, probe = _round_trip(oracle_lib)
    _check_round_trip(first, second, probe, 7, 9)


@pytest.mark.gpu
def test_snapshot_round_trip_gpu(engine_lib, oracle_lib, tmp_path, monkeypatch):
    """TFSF + CPML + Drude + DFT + probe, restored mid-run, bit for bit; and
    the restored GPU run equals the oracle's uninterrupted run."""
    monkeypatch.chdir(tmp_path)
    first, second, probe = _round_trip(engine_lib)
    _check_round_trip(first, second, probe, 7, 9)
    ref, dref = tfsf_drude(oracle_lib)
    ref._advance(16)
    assert rel_l2(second[0], fields(ref)) <= 1e-12
    d = dref.values()
    assert np.linalg.norm(second[2] - d) <= 1e-12 * np.linalg.norm(d)
    ref.close()


def _deepcopy_case(lib, bloch):
    """tests/test_fdtd.py:81-121 of the reference."""
    space = g.Cartesian(size=(2, 2, 0), resolution=3)
    geometry = [g.DefaultMedium(material=g.Dielectric())]
    sources = [g.PointSource(src_time=g.Continuous(freq=0.8, width=0.5), center=(0, 0, 0),
                             component=g.Ez)]
    simulation = g.TMzFDTD(space, geometry, sources, bloch=bloch, verbose=False, _engine=lib)
    simulation.init()
    simulation.step()
    copied = deepcopy(simulation)
    assert copied.space is not simulation.space
    assert copied.geom_list[0] is not simulation.geom_list[0]
    assert copied.src_list[0] is not simulation.src_list[0]
    assert copied.time_step.n == simulation.time_step.n
    assert copied.cmplx == simulation.cmplx
    if bloch is not None:
        np.testing.assert_allclose(copied.bloch, simulation.bloch)
    for comp in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert getattr(copied, comp) is not getattr(simulation, comp)
        np.testing.assert_allclose(getattr(copied, comp), getattr(simulation, comp))
    simulation.step()
    assert simulation.time_step.n == 2
    assert copied.time_step.n == 1
    copied.step()
    for comp in ("ex", "ey", "ez", "hx", "hy", "hz"):
        # the source cell is hard-set at the same time in both: equal bits
        assert np.array_equal(getattr(copied, comp), getattr(simulation, comp)), comp
    assert np.abs(simulation.ez).max() > 0
    simulation.close()
    copied.close()


@pytest.mark.parametrize("bloch", [None, (0.1, 0.2, 0)])
def test_deepcopy_oracle(oracle_lib, bloch):
    _deepcopy_case(oracle_lib, bloch)


@pytest.mark.gpu
@pytest.mark.parametrize("bloch", [None, (0.1, 0.2, 0)])
def test_deepcopy_gpu(engine_lib, bloch):
    _deepcopy_case(engine_lib, bloch)


def test_restore_rejects_foreign_blobs(oracle_lib):
    sim, _ = tfsf_drude(oracle_lib)
    blob = sim._lib.snapshot(sim._sim)
    with pytest.raises(ValueError):
        oracle_lib.restore(sim._sim, blob[:-1])
    bad = blob.copy()
    bad[0] ^= 1
    with pytest.raises(ValueError):
        oracle_lib.restore(sim._sim, bad)
    sim.close()


def test_engine_rejects_grids_beyond_32_bit_offsets():
    """Kernels index a component with 32-bit offsets: gmes_sim_create
    refuses a grid whose component reaches 2^31 device elements, before it
    touches any device (so this runs without a GPU)."""
    if not os.path.exists(_abi.ENGINE_PATH):
        pytest.skip("engine not built (run __graft_entry__.build())")
    lib = _abi.Library(_abi.ENGINE_PATH)
    for n, active in (((1290, 1290, 1290), 0x3F), ((2, 2, 1 << 29), 0x3F),
                      ((46341, 46341, 1), 0x1C), ((2901, 2901, 255), 0x3F)):  # the last: unpadded rows
        desc = _abi.GridDesc()
        desc.nx, desc.ny, desc.nz = n
        desc.dtype = _abi.F32
        desc.active = active
        desc.dx = desc.dy = desc.dz = 1.0
        desc.dt = 0.5
        h = ctypes.c_void_p()
        rc = lib.dll.gmes_sim_create(ctypes.byref(desc), ctypes.byref(h))
        assert rc == _abi.EINVAL
        assert b"2^31" in lib.dll.gmes_last_error()
