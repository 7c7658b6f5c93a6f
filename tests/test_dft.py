"""Running-DFT monitors (no reference counterpart; pinned by closed forms).

* A hard source fixes its cell's value to the oscillator at every half
  step, so the DFT of that cell is a finite sum known in closed form: this
  pins the accumulation, the box indexing and the time alignment of E
  (t = (k + 1/2) dt) and H (t = (k + 1) dt).
* Flux of a plane-wave pulse that has fully passed: equal on planes inside
  the total-field region of a lossless periodic cell, zero in the
  scattered-field region, smaller below an absorbing block."""

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd.monitor import DftBox


def hard_source_case(lib, comp, mode, precision="fp64"):
    space = g.Cartesian(size=(3, 3, 0), resolution=8)
    geometry = [g.DefaultMedium(g.Dielectric()), g.Shell(g.Cpml(), thickness=0.75)]
    st = g.Continuous(freq=0.7, width=1.5)
    src = [g.PointSource(st, center=(0.1, -0.2, 0), component=comp)]
    sim = mode(space, geometry, src, verbose=False, precision=precision, _engine=lib)
    sim.init()
    return sim, st


def closed_form(st, freqs, steps, dt, half):
    t = (np.arange(steps) + half) * dt
    osc = np.array([st.oscillator(x) for x in t])
    return np.array([np.sum(osc * np.exp(-2j * np.pi * f * t)) * dt for f in freqs])


@pytest.mark.parametrize("comp,mode,half", [(g.Ez, g.TMzFDTD, 0.5), (g.Hz, g.TEzFDTD, 1.0)],
                         ids=["E", "H"])
@pytest.mark.parametrize("decim", [1, 3])
def test_hard_source_cell_closed_form(oracle_lib, comp, mode, half, decim):
    sim, st = hard_source_case(oracle_lib, comp, mode)
    idx = sim.space.space_to_index(comp, 0.1, -0.2, 0)
    freqs = (0.3, 0.7, 1.1)
    box = DftBox(sim, comp, idx, tuple(i + 1 for i in idx), freqs, decimation=decim)
    steps = 60
    sim._advance(25)
    sim._advance(steps - 25)
    got = box.values()[:, 0, 0, 0]
    t = (np.arange(steps) + half) * sim.time_step.dt
    keep = np.arange(steps) % decim == 0
    osc = np.array([st.oscillator(x) for x in t])
    want = np.array([np.sum((osc * np.exp(-2j * np.pi * f * t))[keep]) * sim.time_step.dt * decim
                     for f in freqs])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
    if decim == 1:
        assert np.allclose(got, closed_form(st, freqs, steps, sim.time_step.dt, half), rtol=1e-12)


def plane_wave_cell(lib, precision="fp64", loss=False):
    """A periodic x-y cell (Bloch k = 0) with CPML on z and a plane-wave
    pulse along -z (E along x) through a TFSF box wider than the cell, so
    only its z faces act."""
    space = g.Cartesian(size=(0.5, 0.5, 6), resolution=10)
    geometry = [g.DefaultMedium(g.Dielectric())]
    if loss:
        geometry.append(g.Block(g.Drude(eps_inf=1.0, dps=(g.DrudePole(omega=2.0, gamma=1.0),)),
                                center=(0, 0, -0.4), size=(1, 1, 0.8)))
    geometry.append(g.Shell(g.Cpml(), thickness=1, plus_x=False, minus_x=False, plus_y=False,
                            minus_y=False))
    src = [g.TotalFieldScatteredField(g.Bandpass(freq=0.5, fwidth=0.4, s=5), center=(0, 0, 0.05),
                                      size=(2, 2, 3.15), direction=(0, 0, -1),
                                      polarization=(1, 0, 0))]
    sim = g.FDTD(space, geometry, src, bloch=(0, 0, 0), verbose=False, precision=precision,
                 _engine=lib)
    sim.init()
    return sim


def flux_run(lib, precision="fp64", loss=False, steps=900):
    sim = plane_wave_cell(lib, precision, loss)
    freqs = (0.4, 0.5, 0.6)
    planes = {z: sim.add_dft_flux(freqs, z) for z in (1.05, 0.45, -0.85, -2.05)}
    sim._advance(steps)
    return {z: p.flux() for z, p in planes.items()}


def test_plane_wave_flux_is_conserved(oracle_lib):
    f = flux_run(oracle_lib)
    inside = np.array([f[1.05], f[0.45], f[-0.85]])
    assert np.all(inside < 0)                                  # power flows towards -z
    assert np.max(np.abs(inside - inside[0])) < 2e-3 * np.max(np.abs(inside))
    assert np.max(np.abs(f[-2.05])) < 1e-6 * np.max(np.abs(inside))  # scattered field


def test_lossy_block_absorbs(oracle_lib):
    f = flux_run(oracle_lib, loss=True)
    assert np.all(np.abs(f[-0.85]) < 0.95 * np.abs(f[0.45]))   # below the absorbing block


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_flux_on_gpu_matches_oracle(engine_lib, oracle_lib, precision):
    a = flux_run(engine_lib, precision, steps=300)
    b = flux_run(oracle_lib, steps=300)
    tol = 1e-11 if precision == "fp64" else 1e-4
    scale = np.max(np.abs(b[1.05]))
    for z in b:
        assert np.max(np.abs(a[z] - b[z])) <= tol * scale, z


@pytest.mark.gpu
@pytest.mark.parametrize("comp,mode,half", [(g.Ez, g.TMzFDTD, 0.5), (g.Hz, g.TEzFDTD, 1.0)],
                         ids=["E", "H"])
def test_hard_source_cell_closed_form_gpu(engine_lib, comp, mode, half):
    sim, st = hard_source_case(engine_lib, comp, mode)
    idx = sim.space.space_to_index(comp, 0.1, -0.2, 0)
    freqs = (0.3, 0.7, 1.1)
    box = DftBox(sim, comp, idx, tuple(i + 1 for i in idx), freqs)
    sim._advance(60)
    got = box.values()[:, 0, 0, 0]
    assert np.allclose(got, closed_form(st, freqs, 60, sim.time_step.dt, half), rtol=1e-12)


def _compensated_case(lib, steps):
    """A hard-sourced Ez cell over many steps.  Compensated (Kahan)
    summation keeps each accumulator within 2 eps sum|x_i| of the exact
    sum of the FP64 terms w * F (math.fsum), whatever the number of steps,
    and within 2 ulp where the terms do not cancel (the real part at the
    source frequency);
    an uncompensated running sum drifts there by thousands of ulp over
    20000 steps."""
    import math
    from gmes_amd.monitor import dft_weights
    sim, st = hard_source_case(lib, g.Ez, g.TMzFDTD)
    idx = sim.space.space_to_index(g.Ez, 0.1, -0.2, 0)
    freqs = (0.3, 0.7, 1.1)
    box = DftBox(sim, g.Ez, idx, tuple(i + 1 for i in idx), freqs)
    sim._advance(steps)
    got = box.values()[:, 0, 0, 0]
    dt = sim.time_step.dt
    w = dft_weights(freqs, 0, steps, dt)
    v = np.array([st.oscillator((k + 0.5) * dt) for k in range(steps)])
    sim.close()
    eps = np.finfo(float).eps
    for fi, f in enumerate(freqs):
        for part, terms in ((got[fi].real, [float(a) * float(b) - float(c) * 0.0
                                            for a, b, c in zip(w[:, fi, 0, 0], v, w[:, fi, 0, 1])]),
                            (got[fi].imag, [float(a) * 0.0 + float(c) * float(b)
                                            for a, b, c in zip(w[:, fi, 0, 0], v, w[:, fi, 0, 1])])):
            exact = math.fsum(terms)
            mag = math.fsum(abs(x) for x in terms)
            err = abs(part - exact)
            assert err <= 2 * eps * mag, (f, err)
            if abs(exact) >= 0.5 * mag:  # little cancellation (the source frequency)
                assert err <= 2 * np.spacing(abs(exact)), (f, err)
    return got


def test_dft_accumulation_is_compensated_oracle(oracle_lib):
    _compensated_case(oracle_lib, 20000)


@pytest.mark.gpu
def test_dft_accumulation_is_compensated_gpu(engine_lib, oracle_lib):
    got = _compensated_case(engine_lib, 20000)
    ref = _compensated_case(oracle_lib, 20000)
    np.testing.assert_array_equal(got, ref)  # same terms, same order, same compensation
