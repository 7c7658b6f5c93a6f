"""Writes the golden fixtures in this directory.

Each fixture is the six fields of a small case after a fixed number of
steps, computed by the reference harness: the reference's own updater
templates (src/pw_*.hh under /root/reference), compiled by oracle/Makefile
into oracle/_ref/libgmes_ref.so and driven in the reference's step order.
The case inputs (grid, materials, sources) come from tests/cases.py.
Run from the repository root, with the harness built:

    make -C oracle ref && python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import cases  # noqa: E402
import gmes_amd as g  # noqa: E402
from gmes_amd._abi import Library  # noqa: E402

# name -> (builder, kwargs, steps)
CASES = {
    "kat_tmz": (cases.kat_tmz, {}, 5),
    "tez_cpml": (cases.mode_case, dict(cls=g.TEzFDTD), 25),
    "tmx_upml": (cases.mode_case, dict(cls=g.TMxFDTD, pml=g.Upml), 25),
    "temz_cpml": (cases.mode_case, dict(cls=g.TEMzFDTD), 40),
    "drude_tmz": (cases.dispersive2d, dict(material=cases.drude()), 30),
    "lorentz_tez": (cases.dispersive2d, dict(material=cases.lorentz(), mode=g.TEzFDTD), 30),
    "dcp_ade_tmz": (cases.dispersive2d, dict(material=cases.dcp_ade()), 30),
    "dcp_plrc_tmz": (cases.dispersive2d, dict(material=cases.dcp_plrc()), 30),
    "dcp_rc_tmz": (cases.dispersive2d, dict(material=cases.dcp_rc()), 30),
    "dm2_1d": (cases.dm2_1d, {}, 40),
    "const_upml3d": (cases.const_case, {}, 12),
    "bloch2d": (cases.bloch2d, {}, 25),
    "spheres3d": (cases.spheres3d, dict(n=16, nsph=3, pml=3), 20),
    "tfsf2d": (cases.tfsf2d, dict(res=10), 60),
}


def fixture_path(name):
    return os.path.join(HERE, name + ".npz")


def run(lib, name, precision="fp64"):
    """The case's six fields after its steps, on the given C-ABI library."""
    builder, kw, steps = CASES[name]
    kw = dict(kw, precision=precision)
    if builder is cases.mode_case:
        sim = builder(kw.pop("cls"), lib, **kw)
    elif builder is cases.dispersive2d:
        sim = builder(kw.pop("material"), lib, **kw)
    else:
        sim = builder(lib, **kw)
    sim._advance(steps)
    out = {c.__name__: np.array(sim.field[c]) for c in g.constant.COMPONENTS}
    sim.close()
    return out


def main():
    lib = Library(os.path.join(ROOT, "oracle", "_ref", "libgmes_ref.so"))
    for name in CASES:
        f = run(lib, name)
        np.savez_compressed(fixture_path(name), **f)
        print(name, os.path.getsize(fixture_path(name)), "bytes")


if __name__ == "__main__":
    main()
