"""The serving loop (gmes_amd.serve): continuous batching of solves that
arrive over time, admitted by estimated device footprint.  A result must
be the same whether its solve ran alone or in the mix (bitwise here: the
per-simulation arithmetic does not depend on batching)."""

import numpy as np
import pytest

import gmes_amd as g
from gmes_amd import workloads
from gmes_amd.serve import Request, Server, latency_stats


def small_request(lib, rid, freq, steps, arrival, n=24, precision="fp64"):
    def build():
        space = g.Cartesian(size=(n / 8.0, n / 8.0, 0), resolution=8)
        geometry = [g.DefaultMedium(g.Dielectric(1.5)), g.Shell(g.Cpml(), thickness=0.5)]
        src = [g.PointSource(g.Bandpass(freq=freq, fwidth=freq / 4), center=(0.1, 0, 0), component=g.Ez)]
        sim = g.TMzFDTD(space, geometry, src, verbose=False, precision=precision, _engine=lib)
        sim.init()
        return sim, lambda: np.array(sim.field[g.Ez])
    return Request(rid, build, steps, arrival=arrival, footprint=(n * n * 64))


def requests(lib, precision="fp64"):
    specs = [(0.5, 37, 0.0), (0.7, 55, 0.0), (0.3, 18, 0.01), (0.9, 70, 0.02), (0.6, 12, 0.03)]
    return [small_request(lib, i, f, s, t, precision=precision) for i, (f, s, t) in enumerate(specs)]


def alone(lib, precision="fp64"):
    out = []
    for r in requests(lib, precision):
        sim, fn = r.build()
        sim._advance(r.steps)
        out.append(fn())
        sim.close()
    return out


def test_served_results_equal_alone_oracle(oracle_lib):
    srv = Server(hbm_budget=10 ** 9, chunk=16, workers=2)
    done = srv.run(requests(oracle_lib))
    assert sorted(r.rid for r in done) == list(range(5))
    assert all(r.error is None and r.latency >= 0 for r in done)
    ref = alone(oracle_lib)
    for r in done:
        np.testing.assert_array_equal(r.result, ref[r.rid])
    assert srv.peak_active >= 2  # requests overlapped


def test_admission_respects_the_footprint_budget(oracle_lib):
    reqs = requests(oracle_lib)
    budget = 2 * reqs[0].footprint + 1   # at most two at a time
    srv = Server(hbm_budget=budget, chunk=16, workers=2)
    done = srv.run(reqs)
    assert srv.peak_reserved <= budget and srv.peak_active <= 2
    ref = alone(oracle_lib)
    for r in done:
        np.testing.assert_array_equal(r.result, ref[r.rid])


def test_oversized_request_is_rejected(oracle_lib):
    reqs = requests(oracle_lib)[:2]
    reqs[1].footprint = 10 ** 12
    done = Server(hbm_budget=10 ** 9, chunk=16).run(reqs)
    bad = [r for r in done if r.rid == 1][0]
    assert bad.error and bad.result is None
    assert latency_stats(done)["p50"] is not None


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_served_results_equal_alone_gpu(engine_lib, precision):
    srv = Server(hbm_budget=10 ** 10, chunk=16, workers=2)
    done = srv.run(requests(engine_lib, precision))
    ref = alone(engine_lib, precision)
    assert all(r.error is None for r in done)
    for r in done:
        np.testing.assert_array_equal(r.result, ref[r.rid])


@pytest.mark.gpu
def test_served_slabs_equal_alone_gpu(engine_lib):
    """Photonic-crystal slab solves (TFSF, Bloch, flux monitor) in a mix."""
    reqs = workloads.serve_request_list(n=4, rate=200.0, nfreq=8, lo=64, hi=72, steps=120,
                                        _engine=engine_lib)
    done = Server(hbm_budget=10 ** 11, chunk=50, workers=2).run(reqs)
    assert all(r.error is None for r in done)
    again = workloads.serve_request_list(n=4, rate=200.0, nfreq=8, lo=64, hi=72, steps=120,
                                         _engine=engine_lib)
    for r in again:
        sim, fn = r.build()
        sim._advance(r.steps)
        want = fn()
        sim.close()
        got = [d for d in done if d.rid == r.rid][0].result
        np.testing.assert_array_equal(got, want)
        assert np.all(np.isfinite(got)) and np.any(got != 0)


def test_tables_built_in_another_process_are_identical(oracle_lib):
    """The served slab's material tables built by the host-only builder
    (as a worker process does) give the same run as tables built in place."""
    tiles = workloads._tables_of((64, 0.3, 16, "fp64"))
    a, fa = workloads.pc_slab(64, 0.3, precision="fp64", nfreq=4, _engine=oracle_lib, tiles=tiles)
    b, fb = workloads.pc_slab(64, 0.3, precision="fp64", nfreq=4, _engine=oracle_lib)
    a._advance(12)
    b._advance(12)
    for c in g.constant.COMPONENTS:
        np.testing.assert_array_equal(np.asarray(a.field[c]), np.asarray(b.field[c]))
    np.testing.assert_array_equal(fa.flux(), fb.flux())


def test_shortest_policy_admits_least_work_first(oracle_lib):
    """With one slot and every request waiting at the start, the "shortest"
    policy serves them in order of estimated work and the "arrival" policy
    in arrival order; results do not depend on the order."""
    ref = alone(oracle_lib)
    order = {}
    for policy in ("arrival", "shortest"):
        reqs = requests(oracle_lib)
        for r in reqs:
            r.arrival = 0.0
        done = Server(hbm_budget=10 ** 9, max_active=1, chunk=16, workers=1, policy=policy).run(reqs)
        order[policy] = [r.rid for r in sorted(done, key=lambda r: r.t_done)]
        for r in done:
            np.testing.assert_array_equal(r.result, ref[r.rid])
    steps = {r.rid: r.steps for r in requests(oracle_lib)}
    assert order["arrival"] == list(range(5))
    assert order["shortest"] == sorted(range(5), key=lambda i: steps[i])
    with pytest.raises(ValueError):
        Server(hbm_budget=1, policy="fifo")


def dm2_request(lib, rid, rtol, steps, arrival):
    """A 1-D two-level-medium solve; rtol < 0 can never be met, so its DM2
    corrector reports non-convergence (tests/test_pw_dm2.py:95-101 of the
    reference)."""
    from cases import dm2_1d

    def build():
        sim = dm2_1d(lib, rtol=rtol)
        return sim, lambda: np.array(sim.field[g.Ex])
    return Request(rid, build, steps, arrival=arrival, footprint=10 ** 5)


def _failure_mix(lib):
    reqs = requests(lib)[:3]
    reqs.append(dm2_request(lib, 3, -1.0, 40, 0.0))
    reqs.append(dm2_request(lib, 4, 1e-10, 30, 0.005))
    return reqs


def _check_failure_isolated(lib):
    srv = Server(hbm_budget=10 ** 10, chunk=16, workers=2)
    done = srv.run(_failure_mix(lib))
    assert sorted(r.rid for r in done) == list(range(5))
    by = {r.rid: r for r in done}
    assert by[3].error and "failed to converge" in by[3].error and by[3].result is None
    ref = alone(lib)
    for rid in range(3):
        assert by[rid].error is None
        np.testing.assert_array_equal(by[rid].result, ref[rid])
    sim, fn = dm2_request(lib, 4, 1e-10, 30, 0.0).build()
    sim._advance(30)
    assert by[4].error is None
    np.testing.assert_array_equal(by[4].result, fn())
    sim.close()
    stats = latency_stats(done)
    assert stats["p50"] is not None


def test_failing_request_is_isolated_oracle(oracle_lib):
    """One request of a batch fails on the device: it alone is retired with
    its error, the others complete with their alone results."""
    _check_failure_isolated(oracle_lib)


@pytest.mark.gpu
def test_failing_request_is_isolated_gpu(engine_lib):
    _check_failure_isolated(engine_lib)
