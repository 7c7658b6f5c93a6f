"""The batched (64 instances) and served (photonic-crystal slab) workloads
at their stated configurations.

* Batched: every member of the 64-instance workload, advanced in one batch,
  equals the same instance advanced alone, bit for bit, in FP32; in FP64
  the batch's smallest members equal the oracle (relative L2 <= 1e-12).
* Served: a slab solve served through the request loop (in complex FP64,
  next to another request) equals the oracle's run of the same solve, its
  fields to 1e-12 and its flux spectrum to 1e-12; the complex-FP32 solve
  at its full step count tracks complex FP64 within 1e-5.
"""

import numpy as np
import pytest

from gmes_amd import constant as const
from gmes_amd import workloads
from gmes_amd.batch import Batch
from gmes_amd.serve import Request, Server

from cases import fields, rel_l2

SMALLEST = (3, 23, 51, 45)   # two smallest 2-D (151^2, 169^2), two smallest 3-D (67^3, 77^3)


def test_instance_selection_draws_the_same_instances(oracle_lib):
    """Selecting instances replays the same random draws."""
    a = workloads.batch_instances(precision="fp64", _engine=oracle_lib, only={3})
    b = workloads.batch_instances(precision="fp64", _engine=oracle_lib, only={3, 51})
    assert len(a) == 1 and len(b) == 2
    for c in a[0]._active:
        np.testing.assert_array_equal(a[0]._tiles.ids[c], b[0]._tiles.ids[c])
    assert a[0]._tiles.records == b[0]._tiles.records
    for s in a + b:
        s.close()


@pytest.mark.gpu
def test_batch64_members_equal_alone_fp32(engine_lib):
    steps = 30
    sims = workloads.batch_instances(precision="fp32")
    assert len(sims) == 64
    b = Batch(sims)
    b.advance(steps)
    b.sync()
    b.close()
    batched = [fields(s) for s in sims]
    for s in sims:
        s.close()
    del sims
    alone = workloads.batch_instances(precision="fp32")
    for i, s in enumerate(alone):
        s._advance(steps)
        f = fields(s)
        s.close()
        for c in const.COMPONENTS:
            assert np.array_equal(f[c], batched[i][c]), (i, c.__name__)
        assert any(np.abs(f[c]).max() > 0 for c in const.COMPONENTS), i


@pytest.mark.gpu
def test_batch64_fp64_smallest_members_equal_oracle(engine_lib, oracle_lib):
    steps = 30
    sims = workloads.batch_instances(precision="fp64")
    b = Batch(sims)
    b.advance(steps)
    b.sync()
    b.close()
    got = {i: fields(sims[i]) for i in SMALLEST}
    for s in sims:
        s.close()
    del sims
    refs = workloads.batch_instances(precision="fp64", _engine=oracle_lib, only=set(SMALLEST))
    order = sorted(SMALLEST)
    for i, ref in zip(order, refs):
        ref._advance(steps)
        err = rel_l2(got[i], fields(ref))
        ref.close()
        assert err <= 1e-12, (i, err)


def _slab_request(lib, rid, side, radius, steps, arrival, precision, store):
    def build():
        sim, flux = workloads.pc_slab(side, radius, precision=precision, _engine=lib)

        def result():
            store[rid] = fields(sim)
            return flux.flux()
        return sim, result
    return Request(rid, build, steps, arrival=arrival, footprint=workloads.pc_slab_footprint(side))


@pytest.mark.gpu
def test_served_slab_equals_oracle(engine_lib, oracle_lib):
    steps = 100
    store = {}
    reqs = [_slab_request(engine_lib, 0, 64, 0.3, steps, 0.0, "fp64", store),
            _slab_request(engine_lib, 1, 72, 0.25, 140, 0.0, "fp64", store)]
    done = Server(hbm_budget=8 << 30, chunk=40, workers=2).run(reqs)
    got = [r for r in done if r.rid == 0][0]
    assert got.error is None
    sim, flux = workloads.pc_slab(64, 0.3, precision="fp64", _engine=oracle_lib)
    sim._advance(steps)
    ref_fields, ref_flux = fields(sim), flux.flux()
    sim.close()
    assert rel_l2(store[0], ref_fields) <= 1e-12
    assert np.linalg.norm(got.result - ref_flux) <= 1e-12 * np.linalg.norm(ref_flux)
    assert np.linalg.norm(ref_flux) > 0


@pytest.mark.gpu
def test_served_slab_c64_tracks_c128_full_solve(engine_lib):
    """The 64^3 slab over a served solve's shortest step count (1500):
    flux spectrum and fields in complex FP32 within 1e-5 of complex FP64."""
    steps = 1500
    out = {}
    for p in ("fp32", "fp64"):
        sim, flux = workloads.pc_slab(64, 0.3, precision=p)
        sim._advance(steps)
        out[p] = (fields(sim), flux.flux())
        sim.close()
    err_f = rel_l2(out["fp32"][0], out["fp64"][0])
    err_s = np.linalg.norm(out["fp32"][1] - out["fp64"][1]) / np.linalg.norm(out["fp64"][1])
    print("64^3 slab, %d steps: rel L2 fields %.3g, flux spectrum %.3g" % (steps, err_f, err_s))
    assert err_s <= 1e-5
    assert err_f <= 1e-5
