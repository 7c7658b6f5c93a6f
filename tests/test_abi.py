"""The C-ABI boundary, checked without a GPU: every function that
include/gmes_engine.h declares is exported by the HIP engine library and by
both CPU oracles, the ctypes table in gmes_amd/_abi.py binds exactly that
set, and the struct layouts agree.  No compute call is made on the engine
library here."""

import ctypes
import os
import re

import numpy as np
import pytest

from gmes_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gmes_engine.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gmes_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_set():
    names = declared()
    assert len(names) >= 30
    assert names == sorted(_abi.EXPORTED)


def _exports(path):
    lib = ctypes.CDLL(path)
    return [n for n in declared() if not hasattr(lib, n)]


def test_engine_library_exports_every_symbol():
    if not os.path.exists(_abi.ENGINE_PATH):
        pytest.skip("engine not built (run __graft_entry__.build())")
    assert _exports(_abi.ENGINE_PATH) == []


def test_oracles_export_every_symbol(oracle_lib):
    from conftest import ORACLE_LIB, REF_LIB
    assert _exports(ORACLE_LIB) == []
    if os.path.exists(REF_LIB):
        assert _exports(REF_LIB) == []


LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "gmes_engine.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("gmes_grid_desc %zu\n", sizeof(gmes_grid_desc));
  printf("gmes_mat_rec %zu\n", sizeof(gmes_mat_rec));
  printf("gmes_src_op %zu\n", sizeof(gmes_src_op));
  F(gmes_grid_desc, dx) F(gmes_grid_desc, phase)
  F(gmes_mat_rec, coef_off) F(gmes_mat_rec, eps_mu) F(gmes_mat_rec, value)
  F(gmes_src_op, cell) F(gmes_src_op, wave) F(gmes_src_op, i0) F(gmes_src_op, k) F(gmes_src_op, r1)
  return 0;
}
"""


def test_struct_layouts_match_the_header(tmp_path):
    """The header's structs, compiled by the C compiler, have the sizes and
    field offsets that the ctypes structures and numpy dtypes assume."""
    import subprocess
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)],
                   check=True)
    got = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
               if line)
    got = {k: int(v) for k, v in got.items()}
    assert got["gmes_grid_desc"] == ctypes.sizeof(_abi.GridDesc)
    assert got["gmes_mat_rec"] == ctypes.sizeof(_abi.MatRec) == _abi.MAT_REC_DTYPE.itemsize
    assert got["gmes_src_op"] == ctypes.sizeof(_abi.SrcOp) == _abi.SRC_OP_DTYPE.itemsize
    for name, ct, dt in (("gmes_grid_desc", _abi.GridDesc, None), ("gmes_mat_rec", _abi.MatRec,
                         _abi.MAT_REC_DTYPE), ("gmes_src_op", _abi.SrcOp, _abi.SRC_OP_DTYPE)):
        for key, off in got.items():
            if key.startswith(name + "."):
                field = key.split(".", 1)[1]
                assert getattr(ct, field).offset == off, key
                if dt is not None:
                    assert dt.fields[field][1] == off, key


def test_abi_version(oracle_lib):
    assert oracle_lib.dll.gmes_abi_version() == _abi.ABI_VERSION


def test_error_codes_map_to_reference_exceptions(oracle_lib):
    """src/pw_material.i:23-33: bad arguments -> ValueError, out of bounds ->
    IndexError."""
    from cases import kat_tmz
    sim = kat_tmz(oracle_lib)
    with pytest.raises(ValueError):
        oracle_lib.step(sim._sim, -1)
    with pytest.raises(ValueError):
        oracle_lib.write_field(sim._sim, 2, np.zeros(3, np.float64))
    ops = np.zeros(1, _abi.SRC_OP_DTYPE)
    ops["comp"] = 2
    ops["cell"] = 10 ** 9
    with pytest.raises(IndexError):
        oracle_lib.set_sources(sim._sim, ops, 0)


def test_field_sum_sq_reproduces_the_tmz_regression_value(oracle_lib):
    """gmes_sim_field_sum_sq over the whole Ez array is the reference's
    regression value sum(|ez|^2) = 1.5780099423691636 after 5 steps
    (tests/test_fdtd.py:55-79), and a sub-box equals numpy's sum."""
    from cases import kat_tmz
    sim = kat_tmz(oracle_lib)
    for _ in range(5):
        sim.step()
    ez = np.array(sim.ez)
    full = oracle_lib.field_sum_sq(sim._sim, 2, (0, 0, 0), ez.shape)
    assert abs(full - 1.5780099423691636) < 1e-14
    assert oracle_lib.field_sum_sq(sim._sim, 2, (0, 0, 0), ez.shape, allocate=False) == full
    part = oracle_lib.field_sum_sq(sim._sim, 2, (2, 3, 0), (9, 7, 1))
    assert abs(part - float(np.sum(ez[2:9, 3:7] ** 2))) <= 1e-15 * part
    with pytest.raises(IndexError):
        oracle_lib.field_sum_sq(sim._sim, 2, (0, 0, 0), (12, 11, 1))
    sim.close()
