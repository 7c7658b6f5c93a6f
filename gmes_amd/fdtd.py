"""FDTD simulation classes: the drop-in driver over the HIP engine.

Counterpart of the reference's ``gmes.fdtd`` (gmes/fdtd.py).  The Python
side keeps the reference's API and time bookkeeping (``TimeStep``, dt and
Courant limit, Bloch handling, mode classes); the whole leapfrog step --
ghost copies, material updates, sources, TFSF corrections, the auxiliary
1-D grid and probes -- runs on the GPU behind the gmes_* C-ABI.

Field arrays (``ex`` ... ``hz``, ``field``) are host mirrors with the
reference's shapes.  Reading one downloads the device copy if the device
is newer; any array handed out is uploaded again before the next step, so
``sim.ez[idx] = 1`` before stepping works as in the reference.
"""

import sys
from cmath import exp as cexp
from copy import deepcopy
from datetime import datetime, timedelta
from math import sqrt

import numpy as np
from numpy import array, inf

from . import _abi
from . import constant as const
from .constant import Electric, Ex, Ey, Ez, Hx, Hy, Hz, Magnetic, X, Y, Z
from .file_io import ProbeFile
from .geometry import DefaultMedium, GeomBoxTree, in_range
from .pygeom import GeomBox
from .tiles import build_tiles

_PRECISIONS = {"fp64": (_abi.F64, _abi.C128), "fp32": (_abi.F32, _abi.C64)}


class TimeStep(object):
    """Current half step n and time t = n dt (gmes/fdtd.py:24-50)."""

    def __init__(self, dt, n=0.0, t=0.0):
        self.n = float(n)
        self.t = float(t)
        self.dt = float(dt)

    def half_step_up(self):
        self.n += 0.5
        self.t = self.n * self.dt


class _FieldDict(dict):
    """``FDTD.field``: component -> host array, synced on access."""

    def __init__(self, owner):
        dict.__init__(self)
        self._owner = owner

    def __getitem__(self, comp):
        return self._owner._host(comp)

    def values(self):
        return [self[c] for c in self.keys()]

    def items(self):
        return [(c, self[c]) for c in self.keys()]


class PwMaterialView(object):
    """Compatibility view of one reference updater list: idx_size(), name().
    ``view_type`` gives each list the type name of the reference's updater
    (``DielectricEzReal``, ``CpmlHxCmplx``, ...), which scripts read with
    ``type(updater).__name__`` (benchmarks/field_updates.py:83-96)."""

    def __init__(self, name, count):
        self._name = name
        self._count = count

    def name(self):
        return self._name

    def idx_size(self):
        return self._count


_VIEW_TYPES = {}


def view_type(name):
    """The PwMaterialView subclass called ``name`` (one per name)."""
    if name not in _VIEW_TYPES:
        _VIEW_TYPES[name] = type(name, (PwMaterialView,), {"__module__": __name__})
    return _VIEW_TYPES[name]



class FDTD(object):
    """Three-dimensional FDTD (gmes/fdtd.py:53-1467).

    Extra keyword arguments of this engine: ``precision`` ("fp64" -- the
    reference's float64/complex128, default -- or "fp32" for
    float32/complex64 fields and state) and ``device`` (HIP ordinal).
    """

    def __init__(self, space=None, geom_list=None, src_list=None, courant_ratio=0.99, dt=None,
                 bloch=None, verbose=True, precision="fp64", device=0, _engine=None):
        self._init_field_compnt()
        self.e_recorder = []
        self.h_recorder = []
        self.verbose = bool(verbose)
        self.space = space
        self._fig_id = int(self.space.my_id)
        self.dx, self.dy, self.dz = self.space.dr
        if precision not in _PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self.device = int(device)
        self._engine_override = _engine

        default_medium = next((i for i in geom_list if isinstance(i, DefaultMedium)))
        eps_inf = default_medium.material.eps_inf
        mu_inf = default_medium.material.mu_inf
        dt_limit = self._dt_limit(space, eps_inf, mu_inf)
        if dt is None:
            self.courant_ratio = float(courant_ratio)
            time_step_size = self.courant_ratio * dt_limit
        else:
            time_step_size = float(dt)
            self.courant_ratio = time_step_size / dt_limit
        self.space.dt = time_step_size
        self.time_step = TimeStep(time_step_size)
        if self.verbose:
            self.space.display_info()
            print("dt:", time_step_size)
            print("courant ratio:", self.courant_ratio)

        self.geom_list = geom_list
        for go in self.geom_list:
            go.init(self.space)
        self.geom_tree = GeomBoxTree(self.geom_list)

        if bloch is None:
            self.cmplx = False
            self.bloch = None
        else:
            self.cmplx = True
            self.bloch = np.array(bloch, np.double)
            if self.verbose:
                ds = np.array((self.dx, self.dy, self.dz))
                c = 1 / sqrt(eps_inf * mu_inf)
                S = c * time_step_size / ds
                numeric_bloch = 2 * sqrt(eps_inf) / ds * np.arcsin(np.sin(self.bloch * S * ds / 2) / S)
                print("Bloch wave vector is", self.bloch)
                print("numerical Bloch wave vector is", numeric_bloch)

        self.src_list = [] if src_list is None else src_list
        for so in self.src_list:
            if hasattr(so, "aux_fdtd"):
                # the auxiliary grid runs on the same engine, device and precision
                so._engine = _engine
                so._precision = precision
                so._device = self.device
            so.init(self.geom_tree, self.space, self.cmplx)
        self.pw_material = {}
        self.pw_source = {}
        self._sim = None
        self._lib = None
        self._dft_monitors = []
        self._dft_freqs = None
        self._wave_win = self._dft_win = (0, 0)  # steps covered by the uploaded windows

    # ---- setup ----------------------------------------------------------

    def _init_field_compnt(self):
        self.e_field_compnt = (Ex, Ey, Ez)
        self.h_field_compnt = (Hx, Hy, Hz)

    def _dt_limit(self, space, eps_inf, mu_inf):
        """Courant bound over the axes both curls use (gmes/fdtd.py:294-321)."""
        ds = {Ex: (1, 2), Ey: (2, 0), Ez: (0, 1), Hx: (1, 2), Hy: (2, 0), Hz: (0, 1)}
        e_ds = set()
        for i in self.e_field_compnt:
            e_ds.update(ds[i])
        h_ds = set()
        for i in self.h_field_compnt:
            h_ds.update(ds[i])
        non_inf = e_ds.intersection(h_ds)
        dr = array(space.dr, np.double)
        for i in range(3):
            if i not in non_inf:
                dr[i] = inf
        c = 1 / sqrt(eps_inf * mu_inf)
        return 1 / c / sqrt(sum(dr ** -2))

    @property
    def dtype_code(self):
        return _PRECISIONS[self.precision][1 if self.cmplx else 0]

    def _phase(self, comp, axis):
        """Ghost-copy factor exp(i k (x_dest - x_src)) (gmes/fdtd.py:822-836)."""
        if not self.cmplx:
            return 1.0
        shape = self._shape[comp]
        last = [0, 0, 0]
        last[axis] = shape[axis] - 1
        first = (0, 0, 0)
        if comp in const.E_COMPONENTS:
            dest = self.space.index_to_space(comp, *last)[axis]
            src = self.space.index_to_space(comp, *first)[axis]
        else:
            dest = self.space.index_to_space(comp, *first)[axis]
            src = self.space.index_to_space(comp, *last)[axis]
        return cexp(1j * self.bloch[axis] * (dest - src))

    def host_tables(self):
        """The material tables of this simulation, built on the host (the
        costly part of init(); may run in another process)."""
        return build_tiles(self.space, self.geom_list, self.e_field_compnt, self.h_field_compnt,
                           self.cmplx)

    def init(self, tiles=None):
        """Allocate fields, map materials and sources, build the device sim
        (gmes/fdtd.py:227-278).  ``tiles``: the result of host_tables(),
        when it was built elsewhere."""
        st = datetime.now()
        lib = self._engine_override if self._engine_override is not None else _abi.engine()
        self._lib = lib
        comps = tuple(self.e_field_compnt) + tuple(self.h_field_compnt)
        self._active = comps
        self._shape = {c: self.space.storage_shape(c, comps) for c in const.COMPONENTS}
        np_dtype = _abi.NUMPY_DTYPE[self.dtype_code]
        self._arrays = {c: np.zeros(self._shape[c], np_dtype) for c in const.COMPONENTS}
        self._dev_newer = {c: False for c in const.COMPONENTS}
        self._host_dirty = {c: False for c in const.COMPONENTS}
        self.field = _FieldDict(self)
        for c in const.COMPONENTS:
            dict.__setitem__(self.field, c, None)

        desc = _abi.GridDesc()
        n = [int(v) for v in self.space.my_field_size]
        desc.nx, desc.ny, desc.nz = n
        desc.dtype = self.dtype_code
        desc.active = sum(1 << c.index for c in comps)
        desc.device = self.device
        desc.dx, desc.dy, desc.dz = (float(v) for v in self.space.dr)
        desc.dt = float(self.time_step.dt)
        for c in const.COMPONENTS:
            for a in range(3):
                ph = complex(self._phase(c, a)) if c in comps else 1.0
                desc.phase[(c.index * 3 + a) * 2] = ph.real
                desc.phase[(c.index * 3 + a) * 2 + 1] = ph.imag
        if self._sim is not None:
            lib.sim_destroy(self._sim)
        self._sim = lib.sim_create(desc)
        self._dft_monitors = []
        self._dft_freqs = None
        self._wave_win = self._dft_win = (0, 0)  # steps covered by the uploaded windows

        if self.verbose:
            print("Mapping the piecewise material.")
        if tiles is None:
            tiles = self.host_tables()
        self._tiles = tiles
        lib.set_materials(self._sim, tiles.record_array(), tiles.coef_array())
        for c in comps:
            lib.set_material_ids(self._sim, c.index, tiles.ids[c])
        for (p, ci, a), vals in tiles.profiles.items():
            lib.set_profile(self._sim, p, ci, a, vals)
        for c in comps:
            views = {}
            for prefix, v in tiles.type_counts[c].items():
                klass = view_type(prefix + c.__name__ + ("Cmplx" if self.cmplx else "Real"))
                views[klass] = klass(prefix + ("Electric" if c in const.E_COMPONENTS else "Magnetic"), v)
            self.pw_material[c] = views
        lib.commit(self._sim)

        self._aux_fdtds = []
        for so in self.src_list:
            aux = getattr(so, "aux_fdtd", None)
            if aux is not None:
                if aux._sim is None:
                    aux.init()
                slot = lib.attach_aux(self._sim, aux._sim)
                self._aux_fdtds.append(aux)
                so._aux_slot = slot
        self.init_source()
        et = datetime.now()
        if self.verbose:
            print("Elapsed time:", (et - st))

    def init_source(self):
        """Map point sources and TFSF faces to engine source operations, in
        the reference's order (gmes/fdtd.py:505-605, gmes/pw_source.py:18-46):
        per component, per source type in order of first appearance, per
        cell in order of first appearance, and per cell its ops in order
        (the faces of a TFSF edge cell in face order)."""
        self._waves = []
        self._wave_win = (0, 0)
        env_rows = {}
        ops = []

        def wave_index(row):
            if id(row) not in env_rows:
                env_rows[id(row)] = len(self._waves)
                self._waves.append(row)
            return env_rows[id(row)]

        for comp in tuple(self.e_field_compnt) + tuple(self.h_field_compnt):
            shape = self._shape[comp]
            per_type = {}
            for so in self.src_list:
                block_getter = getattr(so, "_pw_block", None)
                if block_getter is not None:
                    blk = block_getter(comp, shape, self.space, self.geom_tree, self._tiles)
                    if blk is not None:
                        per_type.setdefault("tfsf", []).append(blk)
                    continue
                getter = getattr(so, "_pw_ops", None)
                if getter is None:
                    continue
                for idx, typ, kind, payload in getter(comp, shape, self.space, self.geom_tree):
                    per_type.setdefault(typ, {})[idx] = [(kind, payload)]
            views = {}
            for typ, table in per_type.items():
                if typ == "tfsf":
                    arr, ncells = self._tfsf_ops(comp, shape, table, wave_index)
                    ops.append(arr)
                    views[typ] = PwMaterialView(typ, ncells)
                    continue
                for idx, entries in table.items():
                    cell = (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]
                    for kind, payload in entries:
                        op = np.zeros(1, _abi.SRC_OP_DTYPE)
                        op["comp"] = comp.index
                        op["kind"] = kind
                        op["cell"] = cell
                        op["wave"] = len(self._waves)
                        self._waves.append(payload)
                        ops.append(op)
                views[typ] = PwMaterialView(typ, len(table))
            self.pw_source[comp] = views
        arr = np.concatenate(ops) if ops else np.zeros(0, _abi.SRC_OP_DTYPE)
        self._lib.set_sources(self._sim, arr, len(self._waves))

    @staticmethod
    def _tfsf_ops(comp, shape, blocks, wave_index):
        """Op table of the TFSF blocks of one component.  A cell on two
        faces of one source keeps both ops, in face order, under its first
        appearance; two sources may not share a cell (gmes/pw_source.py:
        28-35,134-147: merging different auxiliary grids is an error)."""
        cells, seen = [], None
        for blk in blocks:
            idx = blk["idx"]
            flat = (idx[:, 0] * shape[1] + idx[:, 1]) * shape[2] + idx[:, 2]
            u = np.unique(flat)
            if seen is not None and np.intersect1d(seen, u).size:
                raise ValueError("incompatible transparent source parameters")
            seen = u if seen is None else np.union1d(seen, u)
            cells.append(flat)
        out, ncells = [], 0
        for blk, flat in zip(blocks, cells):
            _, first, inverse = np.unique(flat, return_index=True, return_inverse=True)
            order = np.lexsort((np.arange(len(flat)), first[inverse]))
            op = np.zeros(len(flat), _abi.SRC_OP_DTYPE)
            op["comp"] = comp.index
            op["kind"] = blk["kind"][order]
            op["cell"] = flat[order]
            env = blk["envelope"]
            op["wave"] = -1 if env is None else wave_index(env)
            op["aux"] = blk["aux"]
            op["aux_comp"] = blk["aux_comp"]
            for key in ("i0", "i1", "k", "r0", "r1"):
                op[key] = blk[key][order]
            out.append(op)
            ncells += len(first)
        return np.concatenate(out), ncells

    # ---- host mirrors ---------------------------------------------------

    def _host(self, comp):
        if self._sim is None:
            raise AttributeError("fields are allocated by init()")
        if self._dev_newer[comp]:
            self._lib.read_field(self._sim, comp.index, self._arrays[comp])
            self._dev_newer[comp] = False
        self._host_dirty[comp] = True
        return self._arrays[comp]

    def _push_dirty(self):
        for c in const.COMPONENTS:
            if self._host_dirty[c]:
                self._lib.write_field(self._sim, c.index, self._arrays[c])
                self._host_dirty[c] = False

    ex = property(lambda self: self._host(Ex))
    ey = property(lambda self: self._host(Ey))
    ez = property(lambda self: self._host(Ez))
    hx = property(lambda self: self._host(Hx))
    hy = property(lambda self: self._host(Hy))
    hz = property(lambda self: self._host(Hz))

    # ---- stepping -------------------------------------------------------

    # steps of waveform samples and DFT weights uploaded at a time: a loop
    # of step() calls then enqueues its steps without a host round trip
    # (the upload waits for the device) except once per window
    WINDOW = 256

    def _prepare(self, nsteps):
        """Upload dirty mirrors, and waveforms and DFT weights unless the
        uploaded window already covers the next nsteps (also for the
        auxiliary grids)."""
        self._push_dirty()
        step0 = self._lib.get_step(self._sim)
        end = step0 + nsteps
        if self._waves and not (self._wave_win[0] <= step0 and end <= self._wave_win[1]):
            exact = any(w.exact for w in self._waves)
            n = nsteps if exact else max(nsteps, self.WINDOW)
            vals = np.stack([w.sample(step0, n, self.time_step.dt) for w in self._waves])
            self._lib.set_waveforms(self._sim, step0, vals)
            self._wave_win = (step0, step0 + n)
        if self._dft_monitors and not (self._dft_win[0] <= step0 and end <= self._dft_win[1]):
            from .monitor import dft_weights
            n = max(nsteps, self.WINDOW)
            self._lib.set_dft_weights(self._sim, step0,
                                      dft_weights(self._dft_freqs, step0, n, self.time_step.dt))
            self._dft_win = (step0, step0 + n)
        for aux in self._aux_fdtds:
            if aux._sim is not None:  # a closed grid: the engine refuses the step
                aux._prepare(nsteps)

    def _advance(self, nsteps):
        """Run nsteps whole steps on the device and update the bookkeeping."""
        nsteps = int(nsteps)
        if nsteps < 0:
            raise ValueError("cannot advance by a negative number of steps (%d)" % nsteps)
        if nsteps == 0:
            return
        if self._sim is None:
            raise RuntimeError("init() must be called before stepping")
        self._prepare(nsteps)
        self._lib.step(self._sim, nsteps)
        self._after_steps(nsteps)

    def _after_steps(self, nsteps):
        for _ in range(nsteps):
            self.time_step.half_step_up()
            self.time_step.half_step_up()
        for c in const.COMPONENTS:
            self._dev_newer[c] = True
        for aux in self._aux_fdtds:
            if aux._sim is not None:
                aux._after_steps(nsteps)
        if self.e_recorder or self.h_recorder:
            # drain the device ring buffer: E samples are at n = k + 1/2,
            # H samples at n = k + 1 after step k
            self._lib.sync(self._sim)
            n_last = self.time_step.n
            for probe in self.e_recorder + self.h_recorder:
                vals = self._lib.read_probe(self._sim, probe.pid)
                half = 0.5 if probe in self.e_recorder else 0.0
                steps = [n_last - len(vals) + s + 1 - half for s in range(len(vals))]
                probe.append(steps, vals if self.cmplx else vals.real)

    def add_dft_flux(self, freqs, z, decimation=1):
        """A running-DFT flux monitor on the plane z (whole x-y extent) at
        the given frequencies; see gmes_amd.monitor.DftFlux."""
        from .monitor import DftFlux
        return DftFlux(self, freqs, z, decimation)

    def sync(self):
        """Wait for the device and raise a deferred device error."""
        self._lib.sync(self._sim)

    def step(self):
        """One leapfrog step (gmes/fdtd.py:1099-1122)."""
        self._advance(1)

    def _steps_until_t(self, t):
        n, count = self.time_step.n, 0
        tt = self.time_step.t
        while tt < t:
            n += 0.5
            n += 0.5
            tt = n * self.time_step.dt
            count += 1
        return count

    def _run_printing(self, nsteps, modulus):
        """Advance nsteps, printing n and t whenever n % modulus == 0."""
        if modulus == inf or not np.isfinite(modulus):
            self._advance(nsteps)
            return
        done = 0
        while done < nsteps:
            n = self.time_step.n
            chunk = 0
            while done + chunk < nsteps:
                chunk += 1
                if (n + chunk) % modulus == 0:
                    break
            self._advance(chunk)
            done += chunk
            if self.time_step.n % modulus == 0:
                print("n:", self.time_step.n, "t:", self.time_step.t)

    def step_until_t(self, t=0, modulus=inf):
        """Step while time < t (gmes/fdtd.py:1171-1215)."""
        st = datetime.now()
        sn = self.time_step.n
        total = self._steps_until_t(t)
        if total > 0:
            self._run_printing(1, modulus)
            et = datetime.now()
            num_of_steps = (t - self.time_step.t) / self.time_step.dt
            print("Estimated time of completion:",
                  timedelta(seconds=num_of_steps * (et - st).total_seconds()))
            self._run_printing(total - 1, modulus)
        self.sync()
        et = datetime.now()
        print("Elapsed time:", (et - st), end=" ")
        print("(%d timesteps)" % (self.time_step.n - sn))

    def step_until_n(self, n=0, modulus=inf):
        """Step while the step count < n (gmes/fdtd.py:1124-1169)."""
        st = datetime.now()
        total = 0
        nn = self.time_step.n
        while nn < n:
            nn += 1.0
            total += 1
        if total > 0:
            self._run_printing(1, modulus)
            et = datetime.now()
            print("Estimated time of completion:",
                  timedelta(seconds=(n - 1) * (et - st).total_seconds()))
            self._run_printing(total - 1, modulus)
        self.sync()
        print("Elapsed time:", (datetime.now() - st))

    def step_while_zero(self, component, point, modulus=inf):
        """Step while the field at point is zero (gmes/fdtd.py:1124-1169)."""
        idx = self.space.space_to_index(component, *point)
        if not in_range(idx, self._shape[component], component):
            raise ValueError("observation point is outside the simulation domain")
        flag = True
        while flag:
            self._advance(1)
            if self.time_step.n % modulus == 0:
                print("n:", self.time_step.n, "t:", self.time_step.t)
            if self._host(component)[idx] != 0:
                flag = False

    # ---- probes and output ----------------------------------------------

    def set_probe(self, p, prefix=None):
        """Record the field at point p every half step (gmes/fdtd.py:607-671)."""
        postfix = {Ex: "_ex.dat", Ey: "_ey.dat", Ez: "_ez.dat", Hx: "_hx.dat", Hy: "_hy.dat",
                   Hz: "_hz.dat"}
        for comp in tuple(self.e_field_compnt) + tuple(self.h_field_compnt):
            idx = self.space.space_to_index(comp, *p)
            if not in_range(idx, self._shape[comp], comp):
                continue
            filename = (prefix + postfix[comp]) if prefix else postfix[comp]
            loc = self.space.index_to_space(comp, *idx)
            recorder = ProbeFile(filename, loc, self.time_step.dt, self._arrays[comp].dtype)
            shape = self._shape[comp]
            cell = (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]
            recorder.pid = self._lib.add_probe(self._sim, comp.index, cell)
            if comp in const.E_COMPONENTS:
                self.e_recorder.append(recorder)
            else:
                self.h_recorder.append(recorder)

    def write_field(self, comp, low=(-inf, -inf, -inf), high=(inf, inf, inf), prefix=None,
                    postfix=None):
        """Dump a box of a component as .npy (gmes/fdtd.py:1303-1424)."""
        low = array(low, np.double)
        high = array(high, np.double)
        for i, diff in enumerate(high - low):
            if 0 <= diff < self.space.dr[i]:
                high[i] += self.space.dr[i] / 2
                low[i] -= self.space.dr[i] / 2
            elif diff < 0:
                print("ERROR")
        dump_volume = GeomBox(low, high)
        shape = self._shape[comp]
        if issubclass(comp, Electric):
            low_bndry_idx = (0, 0, 0)
            hi = [s - 2 for s in shape]
            hi[comp.index % 3] = shape[comp.index % 3] - 1
            high_bndry_idx = tuple(hi)
        elif issubclass(comp, Magnetic):
            high_bndry_idx = [i - 1 for i in shape]
            lo = [1, 1, 1]
            lo[comp.index % 3] = 0
            low_bndry_idx = tuple(lo)
        else:
            raise ValueError("component should be of class constant.Component.")
        low_bndry_pnt = self.space.index_to_space(comp, *low_bndry_idx)
        high_bndry_pnt = self.space.index_to_space(comp, *high_bndry_idx)
        field_volume = GeomBox(low_bndry_pnt, high_bndry_pnt)
        if field_volume.overlap(dump_volume):
            dump_volume.intersection(field_volume)
        else:
            return None
        low_idx = self.space.space_to_index(comp, *(dump_volume.low))
        high_idx = [i + 1 for i in self.space.space_to_index(comp, *(dump_volume.high))]
        topo = self.space.cart_comm.Get_topo()[2]
        name = "%s_t=%f_(%d,%d,%d)" % (comp.str(), self.time_step.t, topo[0], topo[1], topo[2])
        if prefix is not None:
            name = prefix + name
        if postfix is not None:
            name = name + postfix
        data = self._host(comp)
        np.save(name, data[low_idx[0]:high_idx[0], low_idx[1]:high_idx[1], low_idx[2]:high_idx[2]])

    def write_field_all(self, low=(-inf, -inf, -inf), high=(inf, inf, inf), prefix=None,
                        postfix=None):
        for comp in tuple(self.e_field_compnt) + tuple(self.h_field_compnt):
            self.write_field(comp, low, high, prefix, postfix)

    def snapshot_field(self, comp, axis, cut):
        """PNG snapshot of a plane (gmes/fdtd.py:1434-1467); needs matplotlib."""
        idx = {X: 0, Y: 1, Z: 2}[axis]
        pnt = [0, 0, 0]
        pnt[idx] = cut
        cut_idx = self.space.space_to_index(comp, *pnt)[idx]
        data = np.take(self._host(comp), cut_idx, axis=idx)
        from .file_io import save_png
        save_png(data, "t=" + str(self.time_step.t), comp.str())

    def show_field(self, *args, **kwargs):
        raise NotImplementedError("live plotting is not part of this engine")

    show_field_line = show_field
    show_permittivity = show_field

    # ---- state ----------------------------------------------------------

    def snapshot(self):
        """Fields, time and all auxiliary state, for pre-emption and
        restart: the engine's blob holds the step clock, CPML/UPML and
        dispersive state, running-DFT accumulators and the auxiliary TFSF
        grids (their fields and state).  ``restore`` puts all of it back."""
        self._push_dirty()
        for aux in self._aux_fdtds:
            aux._push_dirty()
        self._lib.sync(self._sim)
        return dict(fields={c: self._host(c).copy() for c in const.COMPONENTS},
                    state=self._lib.snapshot(self._sim), n=self.time_step.n,
                    aux_n=[aux.time_step.n for aux in self._aux_fdtds])

    def restore(self, snap):
        """Return to a snapshot taken by ``snapshot`` of this simulation."""
        if len(snap["aux_n"]) != len(self._aux_fdtds):
            raise ValueError("snapshot of a different simulation")
        self._lib.restore(self._sim, snap["state"])
        for c in const.COMPONENTS:
            self._arrays[c][...] = snap["fields"][c]
            self._host_dirty[c] = True
            self._dev_newer[c] = False
        self.time_step.n = snap["n"]
        self.time_step.t = self.time_step.n * self.time_step.dt
        for aux, n in zip(self._aux_fdtds, snap["aux_n"]):
            aux.time_step.n = n
            aux.time_step.t = n * aux.time_step.dt
            for c in const.COMPONENTS:
                aux._host_dirty[c] = False
                aux._dev_newer[c] = True

    def __deepcopy__(self, memo=None):
        """Rebuild and copy the six fields and the time; auxiliary state is
        not copied, as in the reference (gmes/fdtd.py:327-351)."""
        if memo is None:
            memo = {}
        newcopy = self.__class__(
            space=deepcopy(self.space, memo), geom_list=deepcopy(self.geom_list, memo),
            src_list=deepcopy(self.src_list, memo), courant_ratio=self.courant_ratio,
            dt=self.time_step.dt, bloch=deepcopy(self.bloch, memo), verbose=self.verbose,
            precision=self.precision, device=self.device, _engine=self._engine_override)
        memo[id(self)] = newcopy
        newcopy.init()
        for comp in ("ex", "ey", "ez", "hx", "hy", "hz"):
            np.copyto(getattr(newcopy, comp), getattr(self, comp))
        newcopy.time_step = deepcopy(self.time_step, memo)
        # the copy's sources must inject at the copied time: move its engine
        # clock (zero state, just initialised) to the same step
        newcopy._restart_clock(int(self._lib.get_step(self._sim)))
        return newcopy

    def _restart_clock(self, step):
        """Set the engine's step clock of a freshly initialised simulation
        (its auxiliary state is still zero) to ``step``."""
        self._push_dirty()
        blob = self._lib.snapshot(self._sim)
        blob[8:16].view(np.int64)[0] = step
        self._lib.restore(self._sim, blob)

    def close(self):
        if self._sim is not None and self._lib is not None:
            self._lib.sim_destroy(self._sim)
            self._sim = None
        for aux in getattr(self, "_aux_fdtds", ()):
            aux.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TExFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ey, Ez)
        self.h_field_compnt = (Hx,)


class TEyFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ex, Ez)
        self.h_field_compnt = (Hy,)


class TEzFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ex, Ey)
        self.h_field_compnt = (Hz,)


class TMxFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ex,)
        self.h_field_compnt = (Hy, Hz)


class TMyFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ey,)
        self.h_field_compnt = (Hz, Hx)


class TMzFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ez,)
        self.h_field_compnt = (Hx, Hy)


class TEMxFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ey,)
        self.h_field_compnt = (Hz,)


class TEMyFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ez,)
        self.h_field_compnt = (Hx,)


class TEMzFDTD(FDTD):
    def _init_field_compnt(self):
        self.e_field_compnt = (Ex,)
        self.h_field_compnt = (Hy,)
