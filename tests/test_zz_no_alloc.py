"""gmes_sim_field_sum_sq_noalloc against gmes_sim_field_sum_sq: the same
kernels with the partial sums in the library's static device array instead
of a buffer the library allocates; the results must be the same bits, on
two calls in a row (the second with a smaller box), and equal the host sum
of the downloaded field to FP64 rounding."""

import ctypes

import numpy as np
import pytest

from gmes_amd import _abi
from gmes_amd import constant as const

from cases import bloch3d, spheres3d

pytestmark = pytest.mark.gpu


def _host_sum_sq(a, lo, hi):
    box = a[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    return float(np.sum(np.abs(box.astype(np.complex128)) ** 2))


@pytest.mark.parametrize("builder,precision", [(spheres3d, "fp32"), (spheres3d, "fp64"),
                                               (bloch3d, "fp32"), (bloch3d, "fp64")],
                         ids=["f32", "f64", "c64", "c128"])
def test_noalloc_equals_allocating_entry_point(engine_lib, builder, precision):
    sim = builder(engine_lib, precision=precision)
    sim._advance(40)
    lib, h = engine_lib, sim._sim
    for comp in (const.Ex, const.Hz):
        shape = sim.field[comp].shape
        host = np.array(sim.field[comp])
        boxes = [((0, 0, 0), shape), ((1, 2, 0), (shape[0] - 1, shape[1] - 3, max(1, shape[2] // 2)))]
        for lo, hi in boxes:
            a = lib.field_sum_sq(h, comp.index, lo, hi, allocate=True)
            b = lib.field_sum_sq(h, comp.index, lo, hi, allocate=False)
            assert a == b, (comp.__name__, lo, hi, a, b)
            ref = _host_sum_sq(host, lo, hi)
            assert ref > 0
            assert abs(a - ref) <= 1e-12 * ref, (a, ref)
    sim.close()


def test_noalloc_refuses_boxes_beyond_its_limit(engine_lib):
    """A 512^3 grid's Ex holds 512 x 513 x 513 > 2^27 cells: the
    no-allocation entry point refuses it; the allocating one sums it."""
    desc = _abi.GridDesc()
    desc.nx = desc.ny = desc.nz = 512
    desc.dtype = _abi.F32
    desc.active = 0x3F
    desc.dx = desc.dy = desc.dz = 1.0
    desc.dt = 0.5
    for i in range(18):
        desc.phase[2 * i] = 1.0
    h = engine_lib.sim_create(desc)
    try:
        full = (512, 513, 513)
        with pytest.raises(ValueError, match="no-allocation limit"):
            engine_lib.field_sum_sq(h, 0, (0, 0, 0), full, allocate=False)
        assert engine_lib.field_sum_sq(h, 0, (0, 0, 0), full, allocate=True) == 0.0
        # just inside the limit: 2^27 cells
        assert engine_lib.field_sum_sq(h, 0, (0, 0, 0), (512, 512, 512), allocate=False) == 0.0
    finally:
        engine_lib.sim_destroy(h)
