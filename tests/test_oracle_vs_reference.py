"""The C oracle against the reference's own updaters (oracle/_ref).

Both libraries share the orchestration core; the per-cell arithmetic of
the oracle (oracle/oracle_update.c) is checked here against the
reference's PwMaterial<T>::update_all compiled from /root/reference/src,
for every material class, component, mode and field type.  Equality is
bitwise.
"""

import numpy as np
import pytest

import gmes_amd as g
from cases import (air2d, bloch2d, bloch3d, const_case, currents, dcp_ade, dcp_plrc, dcp_rc,
                   dispersive2d, dm2_1d, drude, fields, kat_tmz, lorentz, mode_case, spheres3d)


def both(builder, oracle_lib, ref_lib, steps):
    out = []
    for lib in (oracle_lib, ref_lib):
        sim = builder(lib)
        for _ in range(steps):
            sim.step()
        out.append(fields(sim))
    return out


def assert_same(a, b):
    for c in a:
        np.testing.assert_array_equal(a[c], b[c])


CASES = {
    "kat_tmz": (kat_tmz, 5),
    "spheres3d": (lambda lib: spheres3d(lib, n=16, nsph=4, pml=3), 25),
    "const_upml": (const_case, 15),
    "currents": (currents, 15),
    "bloch2d": (bloch2d, 20),
    "bloch3d": (bloch3d, 15),
    "dm2": (dm2_1d, 60),
}
for _name, _mat in (("drude", drude), ("lorentz", lorentz), ("dcp_ade", dcp_ade),
                    ("dcp_plrc", dcp_plrc), ("dcp_rc", dcp_rc)):
    CASES[_name + "_tmz"] = ((lambda m: lambda lib: dispersive2d(m(), lib))(_mat), 30)
    CASES[_name + "_tez"] = ((lambda m: lambda lib: dispersive2d(m(), lib, mode=g.TEzFDTD))(_mat), 30)
    CASES[_name + "_bloch"] = ((lambda m: lambda lib: dispersive2d(m(), lib, bloch=(0.2, 0.1, 0)))(_mat), 20)
for _mode in (g.TExFDTD, g.TEyFDTD, g.TEzFDTD, g.TMxFDTD, g.TMyFDTD, g.TMzFDTD, g.TEMxFDTD,
              g.TEMyFDTD, g.TEMzFDTD):
    for _pml in (g.Cpml, g.Upml):
        CASES["%s_%s" % (_mode.__name__, _pml.__name__)] = (
            (lambda m, p: lambda lib: mode_case(m, lib, pml=p))(_mode, _pml), 25)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_equals_reference(name, oracle_lib, ref_lib):
    builder, steps = CASES[name]
    a, b = both(builder, oracle_lib, ref_lib, steps)
    assert_same(a, b)
    assert any(np.any(v != 0) for v in a.values())


def test_air2d_oracle_equals_reference(oracle_lib, ref_lib):
    out = []
    for lib in (oracle_lib, ref_lib):
        sim = air2d(lib)
        sim.step_until_t(10)
        out.append(fields(sim))
    assert_same(out[0], out[1])
    # peak |Ez| of this run, examples/VERIFICATION.md:7
    assert abs(np.max(np.abs(out[0][g.Ez])) - 0.999382775451675) < 1e-12
