"""ctypes view of include/gmes_engine.h.

``Library`` wraps any shared object that exports the gmes_* C-ABI and
turns error codes into the exceptions the reference raises for the same
conditions (src/pw_material.i:23-33).  ``engine()`` loads the HIP engine
built in-tree; there is no other implementation in the product, so a
missing or unloadable engine is an error.
"""

import ctypes as C
import os
import threading

import numpy as np

ABI_VERSION = 1

OK, EINVAL, ERANGE, ETYPE, ERUNTIME, ENOMEM, EDEVICE = range(7)
F32, F64, C64, C128 = range(4)

(CLS_DUMMY, CLS_CONST, CLS_DIELECTRIC, CLS_CPML, CLS_UPML, CLS_DRUDE, CLS_DCP_ADE,
 CLS_DCP_PLRC, CLS_DM2) = range(9)
NCLS = 9

SRC_SET, SRC_SUB, SRC_AUX_ADD, SRC_AUX_SUB = range(4)

NUMPY_DTYPE = {F32: np.float32, F64: np.float64, C64: np.complex64, C128: np.complex128}

_ERRORS = {EINVAL: ValueError, ERANGE: IndexError, ETYPE: TypeError, ERUNTIME: RuntimeError,
           ENOMEM: MemoryError, EDEVICE: RuntimeError}


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dtype", C.c_int32),
                ("active", C.c_uint32), ("device", C.c_int32), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("phase", C.c_double * 36)]


class MatRec(C.Structure):
    _fields_ = [("cls", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32), ("prof", C.c_int32),
                ("coef_off", C.c_int64), ("eps_mu", C.c_double), ("value", C.c_double * 2)]


class SrcOp(C.Structure):
    _fields_ = [("comp", C.c_int32), ("kind", C.c_int32), ("cell", C.c_int64),
                ("wave", C.c_int32), ("aux", C.c_int32), ("aux_comp", C.c_int32),
                ("pad_", C.c_int32), ("i0", C.c_int64), ("i1", C.c_int64), ("k", C.c_double),
                ("r0", C.c_double), ("r1", C.c_double)]


MAT_REC_DTYPE = np.dtype([("cls", np.int32), ("n1", np.int32), ("n2", np.int32),
                          ("prof", np.int32), ("coef_off", np.int64), ("eps_mu", np.float64),
                          ("value", np.float64, (2,))])
SRC_OP_DTYPE = np.dtype([("comp", np.int32), ("kind", np.int32), ("cell", np.int64),
                         ("wave", np.int32), ("aux", np.int32), ("aux_comp", np.int32),
                         ("pad_", np.int32), ("i0", np.int64), ("i1", np.int64),
                         ("k", np.float64), ("r0", np.float64), ("r1", np.float64)])
assert MAT_REC_DTYPE.itemsize == C.sizeof(MatRec)
assert SRC_OP_DTYPE.itemsize == C.sizeof(SrcOp)

_P = C.c_void_p
_SIGNATURES = {
    "gmes_last_error": (C.c_char_p, []),
    "gmes_abi_version": (C.c_int, []),
    "gmes_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "gmes_sim_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(_P)]),
    "gmes_sim_destroy": (C.c_int, [_P]),
    "gmes_sim_shape": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64)]),
    "gmes_sim_footprint": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gmes_sim_set_materials": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int64]),
    "gmes_sim_set_material_ids": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "gmes_sim_set_profile": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, C.c_int32,
                                       C.c_int32]),
    "gmes_sim_commit": (C.c_int, [_P]),
    "gmes_sim_set_sources": (C.c_int, [_P, _P, C.c_int64, C.c_int32]),
    "gmes_sim_set_waveforms": (C.c_int, [_P, C.c_int64, C.c_int64, _P]),
    "gmes_sim_attach_aux": (C.c_int, [_P, _P, C.POINTER(C.c_int32)]),
    "gmes_sim_add_probe": (C.c_int, [_P, C.c_int32, C.c_int64, C.POINTER(C.c_int32)]),
    "gmes_sim_read_probe": (C.c_int, [_P, C.c_int32, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "gmes_sim_add_dft": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "gmes_sim_set_dft_weights": (C.c_int, [_P, C.c_int64, C.c_int64, _P]),
    "gmes_sim_read_dft": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "gmes_sim_write_field": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "gmes_sim_read_field": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "gmes_sim_state_bytes": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gmes_sim_field_sum_sq": (C.c_int, [_P, C.c_int32, _P, _P, _P]),
    "gmes_sim_field_sum_sq_noalloc": (C.c_int, [_P, C.c_int32, _P, _P, _P]),
    "gmes_sim_snapshot": (C.c_int, [_P, _P, C.c_int64]),
    "gmes_sim_restore": (C.c_int, [_P, _P, C.c_int64]),
    "gmes_sim_step": (C.c_int, [_P, C.c_int64]),
    "gmes_sim_sync": (C.c_int, [_P]),
    "gmes_sim_get_step": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gmes_sim_set_graph": (C.c_int, [_P, C.c_int32]),
    "gmes_sim_set_tile": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "gmes_sim_autotune": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gmes_sim_set_ghost_fusion": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "gmes_sim_profile": (C.c_int, [_P, C.c_int32]),
    "gmes_sim_read_profile": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "gmes_batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(_P)]),
    "gmes_batch_destroy": (C.c_int, [_P]),
    "gmes_batch_step": (C.c_int, [_P, C.c_int64]),
    "gmes_batch_sync": (C.c_int, [_P]),
    "gmes_batch_set_graph": (C.c_int, [_P, C.c_int32]),
    "gmes_batch_profile": (C.c_int, [_P, C.c_int32]),
    "gmes_batch_read_profile": (C.c_int, [_P, _P, _P]),
}

EXPORTED = tuple(_SIGNATURES)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Library(object):
    """A loaded implementation of the gmes_* C-ABI."""

    def __init__(self, path):
        self.path = os.path.abspath(path)
        self.dll = C.CDLL(self.path)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(self.dll, name)
            fn.restype = res
            fn.argtypes = args
        if self.dll.gmes_abi_version() != ABI_VERSION:
            raise RuntimeError("%s: ABI version mismatch" % self.path)

    def check(self, rc):
        if rc != OK:
            msg = self.dll.gmes_last_error().decode(errors="replace")
            raise _ERRORS.get(rc, RuntimeError)(msg)

    def call(self, name, *args):
        self.check(getattr(self.dll, name)(*args))

    # -- sim --------------------------------------------------------------

    def sim_create(self, desc):
        h = C.c_void_p()
        self.call("gmes_sim_create", C.byref(desc), C.byref(h))
        return h

    def sim_destroy(self, h):
        self.call("gmes_sim_destroy", h)

    def set_materials(self, h, recs, coef):
        recs = np.ascontiguousarray(recs, MAT_REC_DTYPE)
        coef = np.ascontiguousarray(coef, np.float64)
        self.call("gmes_sim_set_materials", h, _ptr(recs), len(recs), _ptr(coef), len(coef))

    def set_material_ids(self, h, comp, ids):
        ids = np.ascontiguousarray(ids, np.uint16)
        self.call("gmes_sim_set_material_ids", h, comp, _ptr(ids), ids.size)

    def set_profile(self, h, prof, comp, axis, vals):
        vals = np.ascontiguousarray(vals, np.float64)
        self.call("gmes_sim_set_profile", h, prof, comp, axis, _ptr(vals), vals.shape[0],
                  vals.shape[1])

    def commit(self, h):
        self.call("gmes_sim_commit", h)

    def set_sources(self, h, ops, nwaves):
        ops = np.ascontiguousarray(ops, SRC_OP_DTYPE)
        self.call("gmes_sim_set_sources", h, _ptr(ops), len(ops), nwaves)

    def set_waveforms(self, h, step0, values):
        """values: complex array [nwaves][nsteps]."""
        v = np.ascontiguousarray(values, np.complex128)
        self.call("gmes_sim_set_waveforms", h, step0, v.shape[1] if v.ndim == 2 else 0, _ptr(v))

    def attach_aux(self, h, aux):
        slot = C.c_int32()
        self.call("gmes_sim_attach_aux", h, aux, C.byref(slot))
        return slot.value

    def add_probe(self, h, comp, cell):
        pid = C.c_int32()
        self.call("gmes_sim_add_probe", h, comp, cell, C.byref(pid))
        return pid.value

    def read_probe(self, h, pid, max_n=1 << 20):
        out = np.empty(max_n, np.complex128)
        n = C.c_int64()
        self.call("gmes_sim_read_probe", h, pid, _ptr(out), max_n, C.byref(n))
        return out[:n.value].copy()

    def add_dft(self, h, comp, lo, hi, nfreq, decim=1):
        lo_a = (C.c_int64 * 3)(*[int(v) for v in lo])
        hi_a = (C.c_int64 * 3)(*[int(v) for v in hi])
        i = C.c_int32()
        self.call("gmes_sim_add_dft", h, comp, lo_a, hi_a, nfreq, decim, C.byref(i))
        return i.value

    def set_dft_weights(self, h, step0, w):
        """w: float64 array [nsteps][nfreq][2][2]."""
        w = np.ascontiguousarray(w, np.float64)
        self.call("gmes_sim_set_dft_weights", h, step0, w.shape[0], _ptr(w))

    def read_dft(self, h, did, nfreq, ncells):
        out = np.empty((nfreq, ncells, 2), np.float64)
        self.call("gmes_sim_read_dft", h, did, _ptr(out), out.size)
        return out[..., 0] + 1j * out[..., 1]

    def write_field(self, h, comp, arr):
        arr = np.ascontiguousarray(arr)
        self.call("gmes_sim_write_field", h, comp, _ptr(arr), arr.nbytes)

    def read_field(self, h, comp, out):
        assert out.flags.c_contiguous
        self.call("gmes_sim_read_field", h, comp, _ptr(out), out.nbytes)
        return out

    def state_bytes(self, h):
        n = C.c_int64()
        self.call("gmes_sim_state_bytes", h, C.byref(n))
        return n.value

    def field_sum_sq(self, h, comp, lo, hi, allocate=True):
        """sum |F|^2 over the box [lo, hi) of component comp (FP64);
        allocate=False uses the no-allocation entry point."""
        lo = np.ascontiguousarray(lo, np.int64)
        hi = np.ascontiguousarray(hi, np.int64)
        out = np.zeros(1, np.float64)
        name = "gmes_sim_field_sum_sq" if allocate else "gmes_sim_field_sum_sq_noalloc"
        self.call(name, h, comp, _ptr(lo), _ptr(hi), _ptr(out))
        return float(out[0])

    def snapshot(self, h):
        buf = np.empty(self.state_bytes(h), np.uint8)
        self.call("gmes_sim_snapshot", h, _ptr(buf), buf.nbytes)
        return buf

    def restore(self, h, buf):
        buf = np.ascontiguousarray(buf, np.uint8)
        self.call("gmes_sim_restore", h, _ptr(buf), buf.nbytes)

    def step(self, h, n):
        self.call("gmes_sim_step", h, n)

    def sync(self, h):
        self.call("gmes_sim_sync", h)

    def get_step(self, h):
        n = C.c_int64()
        self.call("gmes_sim_get_step", h, C.byref(n))
        return n.value

    def footprint(self, h):
        n = C.c_int64()
        self.call("gmes_sim_footprint", h, C.byref(n))
        return n.value

    def set_graph(self, h, enable):
        self.call("gmes_sim_set_graph", h, int(enable))

    def set_tile(self, h, bx, by):
        self.call("gmes_sim_set_tile", h, bx, by)

    def autotune(self, h):
        bx, by = C.c_int32(), C.c_int32()
        self.call("gmes_sim_autotune", h, C.byref(bx), C.byref(by))
        return bx.value, by.value

    def set_ghost_fusion(self, h, enable):
        """Allow (1) or forbid (0) the fused E ghost planes; returns
        whether the fusion is in effect."""
        fused = C.c_int32()
        self.call("gmes_sim_set_ghost_fusion", h, int(enable), C.byref(fused))
        return bool(fused.value)

    def profile(self, h, enable):
        self.call("gmes_sim_profile", h, int(enable))

    def read_profile(self, h):
        ms = (C.c_double * 6)()
        n = (C.c_int64 * 6)()
        self.call("gmes_sim_read_profile", h, ms, n)
        return list(ms), list(n)

    def device_count(self):
        n = C.c_int32()
        self.call("gmes_device_count", C.byref(n))
        return n.value

    # -- batch ------------------------------------------------------------

    def batch_create(self, handles):
        arr = (C.c_void_p * len(handles))(*[h.value for h in handles])
        b = C.c_void_p()
        self.call("gmes_batch_create", arr, len(handles), C.byref(b))
        return b

    def batch_destroy(self, b):
        self.call("gmes_batch_destroy", b)

    def batch_step(self, b, n):
        self.call("gmes_batch_step", b, n)

    def batch_sync(self, b):
        self.call("gmes_batch_sync", b)

    def batch_profile(self, b, enable):
        self.call("gmes_batch_profile", b, 1 if enable else 0)

    def batch_read_profile(self, b):
        ms = np.zeros(6, np.float64)
        n = np.zeros(6, np.int64)
        self.call("gmes_batch_read_profile", b, _ptr(ms), _ptr(n))
        return ms, n

    def batch_set_graph(self, b, enable):
        self.call("gmes_batch_set_graph", b, int(enable))


_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
ENGINE_PATH = os.path.join(_LIB_DIR, "libgmes_engine.so")
_engine = None
_engine_lock = threading.Lock()


def engine():
    """The HIP engine (gmes_amd/lib/libgmes_engine.so).  Raises when it was
    not built or cannot see a GPU: the engine has no CPU path."""
    global _engine
    with _engine_lock:
        if _engine is None:
            if not os.path.exists(ENGINE_PATH):
                raise RuntimeError("HIP engine not built: %s is missing "
                                   "(run __graft_entry__.build())" % ENGINE_PATH)
            lib = Library(ENGINE_PATH)
            if lib.device_count() < 1:
                raise RuntimeError("HIP engine loaded but no GPU is visible")
            _engine = lib
        return _engine
