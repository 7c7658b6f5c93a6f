"""Source time functions and sources.

Drop-in counterpart of the reference's ``gmes.source`` / ``gmes.pw_source``
(gmes/source.py:87-1769, gmes/pw_source.py).  Oscillators are evaluated
on the host in FP64 with the reference's expressions (they are cheap and
this keeps every injected value bit-identical); the engine receives, per
source operation, a waveform row sampled at the half step the operation
runs at (n = k + 1/2 for E, n = k + 1 for H).
"""

from cmath import exp as cexp
from copy import deepcopy
from math import exp, floor, pi, sin, sqrt

import numpy as np
from numpy import inf

from . import constant as const
from ._abi import SRC_AUX_ADD, SRC_AUX_SUB, SRC_SET, SRC_SUB
from .geometry import in_range


class SrcTime(object):
    """Time-dependent part of a source (gmes/source.py:44-55)."""

    def init(self, cmplx):
        raise NotImplementedError

    def oscillator(self, time):
        raise NotImplementedError

    def display_info(self, indent=0):
        raise NotImplementedError


class Src(object):
    """Space-dependent part of a source (gmes/source.py:58-84)."""

    def display_info(self, indent=0):
        raise NotImplementedError

    def init(self, geom_tree, space, cmplx):
        raise NotImplementedError

    def step(self):
        raise NotImplementedError


class Continuous(SrcTime):
    """CW source with sin^2 turn-on/turn-off (gmes/source.py:87-132)."""

    def __init__(self, freq, phase=0, start=0, end=inf, width=None):
        self.freq = float(freq)
        self.phase = float(phase)
        self.start = float(start)
        self.end = float(end)
        self.width = 5 / self.freq if width is None else float(width)
        self.cmplx = False

    def init(self, cmplx):
        self.cmplx = cmplx

    def display_info(self, indent=0):
        print(" " * indent, "continuous source", "frequency:", self.freq)

    def oscillator(self, time):
        ts = time - self.start
        te = self.end - time
        if ts < 0 or te < 0:
            return 0
        rise = sin(0.5 * pi * ts / self.width) ** 2 if ts < self.width else 1
        fall = sin(0.5 * pi * te / self.width) ** 2 if te < self.width else 1
        env = rise * fall
        osc = env * cexp(2j * pi * self.freq * time + 1j * self.phase)
        return osc if self.cmplx else osc.real


class Bandpass(SrcTime):
    """Gaussian-envelope pulse (gmes/source.py:135-178)."""

    def __init__(self, freq, fwidth, s=10, phase=0):
        self.freq = float(freq)
        self.phase = float(phase)
        self.fwidth = float(fwidth)
        self.width = 1 / self.fwidth
        self.peak_time = self.width * s
        self.cutoff = 2 * self.width * s
        while exp(-0.5 * (self.cutoff / self.width) ** 2) == 0:
            self.cutoff *= 0.9
        self.cmplx = False

    def init(self, cmplx):
        self.cmplx = cmplx

    def display_info(self, indent=0):
        print(" " * indent, "bandpass source", "frequency:", self.freq)

    def oscillator(self, time):
        tt = time - self.peak_time
        if abs(tt) > self.cutoff:
            return 0
        cfactor = 1.0 / (-2j * pi * self.freq)
        osc = (cfactor * exp(-0.5 * (tt / self.width) ** 2)
               * cexp(2j * pi * self.freq * time + 1j * self.phase))
        return osc if self.cmplx else osc.real


class DifferentiatedGaussian(SrcTime):
    """-2((t-t0)/tw) exp(-((t-t0)/tw)^2) (gmes/source.py:181-213)."""

    def __init__(self, tw, t0):
        self.tw = float(tw)
        self.t0 = float(t0)
        self.cmplx = False

    def init(self, cmplx):
        self.cmplx = bool(cmplx)

    def oscillator(self, time):
        exponent = -(((time - self.t0) / self.tw) ** 2)
        osc = -2 * (time - self.t0) / self.tw * cexp(exponent)
        return osc if self.cmplx else osc.real

    def display_info(self, indent=0):
        print(" " * indent, "differentiated Gaussian", "tw:", self.tw, "t0:", self.t0)


def _record_source_value(output, time, value):
    """gmes/pw_source.py:63-67"""
    if np.iscomplexobj(value):
        output.write("%f\t%f\t%f\n" % (time, value.real, value.imag))
    else:
        output.write("%f\t%f\n" % (time, value))


class WaveRow(object):
    """Host generator of one waveform row: value(n) for the half step n.
    exact: the row has side effects per evaluation (a source that records
    its values), so it is sampled for the steps about to run only, never
    ahead."""

    def __init__(self, fn, electric, exact=False):
        self.fn = fn
        self.electric = electric
        self.exact = exact

    def sample(self, step0, nsteps, dt):
        out = np.empty(nsteps, np.complex128)
        half = 0.5 if self.electric else 1.0
        for s in range(nsteps):
            n = float(step0 + s) + half
            out[s] = self.fn(dt, n)
        return out


class PointSource(Src):
    """Hard (E/H) or current (J/M) point source (gmes/source.py:216-367,
    gmes/pw_source.py:49-131)."""

    _TARGET = {const.Ex: const.Ex, const.Jx: const.Ex, const.Ey: const.Ey, const.Jy: const.Ey,
               const.Ez: const.Ez, const.Jz: const.Ez, const.Hx: const.Hx, const.Mx: const.Hx,
               const.Hy: const.Hy, const.My: const.Hy, const.Hz: const.Hz, const.Mz: const.Hz}

    def __init__(self, src_time, center, component, amp=1, filename=None):
        self.center = np.array(center, np.double)
        self.comp = component
        self.src_time = src_time
        self.amp = float(amp)
        self.filename = str(filename) if filename else None
        self._file = None

    def init(self, geom_tree, space, cmplx):
        self.geom_tree = geom_tree
        self.src_time.init(cmplx)

    def step(self):
        pass

    def display_info(self, indent=0):
        print(" " * indent, "point source", "center:", self.center, "component:", self.comp.str())
        self.src_time.display_info(4)

    def _pw_ops(self, comp, field_shape, space, geom_tree):
        """Ops of this source on ``comp``: [(idx, kind, WaveRow, extra)]."""
        if self._TARGET.get(self.comp) is not comp:
            return []
        idx = space.space_to_index(comp, *self.center)
        if not in_range(idx, field_shape, comp):
            return []
        coords = space.index_to_space(comp, *idx)
        mat_obj, _ = geom_tree.material_of_point(coords)
        eps_inf, mu_inf = float(mat_obj.eps_inf), float(mat_obj.mu_inf)
        amp, st = self.amp, self.src_time
        if self.filename and self._file is None:
            self._file = open(self.filename, "w")
        rec = self._file
        electric = comp in const.E_COMPONENTS
        hard = issubclass(self.comp, (const.Electric, const.Magnetic))
        material = eps_inf if electric else mu_inf

        def value(dt, n):
            src_t = amp * st.oscillator(dt * n)
            if rec is not None:
                _record_source_value(rec, dt * n, src_t)
            if hard:
                return src_t
            return dt * src_t / material

        kind = SRC_SET if hard else SRC_SUB
        return [(tuple(int(v) for v in idx), "point", kind, WaveRow(value, electric, exact=rec is not None))]


# ---------------------------------------------------------------------------
# Total-field / scattered-field plane-wave source
#
# Face table (gmes/source.py:705-1406): for every updated component, the
# faces in the reference's order, each with
#   face         : the Directional of the face,
#   axis, plus   : the face normal and whether it is the + side,
#   cos_dir, cos : which unit vector is dotted with h_direction (E faces) or
#                  e_direction (H faces) to give the cosine,
#   idx_comp     : component whose space_to_<c>_index gives the index range,
#   lo_off, hi_off: offsets added to that range's low and high ends,
#   samp_comp, samp_off: the component/offset whose staggered position is
#                  projected on the beam axis to sample the incident field,
#   sign         : +1 for F += ..., -1 for F -= ... (gmes/pw_source.py:197-486).
_X, _Y, _Z = (1, 0, 0), (0, 1, 0), (0, 0, 1)
_C = const
_TFSF_FACES = {
    _C.Ex: [(_C.MinusY, 1, False, _Z, _C.Ex, (0, 0, 0), (0, 0, 0), _C.Hz, (1, 0, 0), -1),
            (_C.PlusY, 1, True, _Z, _C.Ex, (0, 0, 0), (0, 0, 0), _C.Hz, (1, 1, 0), +1),
            (_C.MinusZ, 2, False, _Y, _C.Ex, (0, 0, 0), (0, 0, 0), _C.Hy, (1, 0, 0), +1),
            (_C.PlusZ, 2, True, _Y, _C.Ex, (0, 0, 0), (0, 0, 0), _C.Hy, (1, 0, 1), -1)],
    _C.Ey: [(_C.MinusZ, 2, False, _X, _C.Ey, (0, 0, 0), (0, 0, 0), _C.Hx, (0, 1, 0), -1),
            (_C.PlusZ, 2, True, _X, _C.Ey, (0, 0, 0), (0, 0, 0), _C.Hx, (0, 1, 1), +1),
            (_C.MinusX, 0, False, _Z, _C.Ey, (0, 0, 0), (0, 0, 0), _C.Hz, (0, 1, 0), +1),
            (_C.PlusX, 0, True, _Z, _C.Ey, (0, 0, 0), (0, 0, 0), _C.Hz, (1, 1, 0), -1)],
    _C.Ez: [(_C.MinusX, 0, False, _Y, _C.Ez, (0, 0, 0), (0, 0, 0), _C.Hy, (0, 0, 1), -1),
            (_C.PlusX, 0, True, _Y, _C.Ez, (0, 0, 0), (0, 0, 0), _C.Hy, (1, 0, 1), +1),
            (_C.MinusY, 1, False, _X, _C.Ez, (0, 0, 0), (0, 0, 0), _C.Hx, (0, 0, 1), +1),
            (_C.PlusY, 1, True, _X, _C.Ez, (0, 0, 0), (0, 0, 0), _C.Hx, (0, 1, 1), -1)],
    _C.Hx: [(_C.MinusY, 1, False, _Z, _C.Ez, (0, 0, 1), (0, 0, 1), _C.Ez, (0, 0, -1), +1),
            (_C.PlusY, 1, True, _Z, _C.Ez, (0, 1, 1), (0, 1, 1), _C.Ez, (0, -1, -1), -1),
            (_C.MinusZ, 2, False, _Y, _C.Ey, (0, 1, 0), (0, 1, 0), _C.Ey, (0, -1, -1), -1),
            (_C.PlusZ, 2, True, _Y, _C.Ey, (0, 1, 1), (0, 1, 1), _C.Ey, (0, -1, 0), +1)],
    _C.Hy: [(_C.MinusZ, 2, False, _X, _C.Ex, (1, 0, 0), (1, 0, 0), _C.Ex, (-1, 0, 0), +1),
            (_C.PlusZ, 2, True, _X, _C.Ex, (1, 0, 1), (1, 0, 1), _C.Ex, (-1, 0, -1), -1),
            (_C.MinusX, 0, False, _Z, _C.Ez, (0, 0, 1), (0, 0, 1), _C.Ez, (0, 0, -1), -1),
            (_C.PlusX, 0, True, _Z, _C.Ez, (1, 0, 1), (1, 0, 1), _C.Ez, (-1, 0, -1), +1)],
    _C.Hz: [(_C.MinusX, 0, False, _Y, _C.Ey, (0, 1, 0), (0, 1, 0), _C.Ey, (0, -1, 0), +1),
            (_C.PlusX, 0, True, _Y, _C.Ey, (1, 1, 0), (1, 1, 0), _C.Ey, (-1, -1, 0), -1),
            (_C.MinusY, 1, False, _X, _C.Ex, (1, 0, 0), (1, 0, 0), _C.Ex, (0, -1, 0), -1),
            (_C.PlusY, 1, True, _X, _C.Ex, (1, 1, 0), (1, 1, 0), _C.Ex, (-1, -1, 0), +1)],
}


class TotalFieldScatteredField(Src):
    """Plane wave launched through a total-field/scattered-field box
    (gmes/source.py:370-1522).  The incident field comes from a 1-D
    auxiliary TEMz grid with matched numerical dispersion
    (gmes/source.py:549-630) that runs on the device between the E and H
    half steps of the main grid."""

    def __init__(self, src_time, center, size, direction, polarization, amp=1):
        if not isinstance(src_time, SrcTime):
            raise TypeError("src_time must be an instance of SrcTime.")
        self.src_time = src_time
        self.k = np.array(direction, np.double) / np.linalg.norm(direction)
        self.center = np.array(center, np.double)
        if not np.isfinite(self.center).all():
            raise ValueError("center must contain only finite values")
        self.size = np.array(size, np.double)
        if not np.isfinite(self.size).all():
            raise ValueError("size must contain only finite values")
        self.half_size = 0.5 * self.size
        self.e_direction = np.array(polarization, np.double) / np.linalg.norm(polarization)
        self.h_direction = np.cross(self.k, self.e_direction)
        self.amp = float(amp)
        self.on_axis_k = self._axis_in_k()
        self.aux_fdtd = None
        self._engine = None
        self._precision = "fp64"
        self._device = 0

    def init(self, geom_tree, space, cmplx):
        self.geom_tree = geom_tree
        self.src_time.init(cmplx)
        self.aux_fdtd = self._get_aux_fdtd(space, geom_tree, cmplx)
        self.aux_fdtd.init()

    def step(self):
        pass  # the engine advances the auxiliary grid with the main one

    def display_info(self, indent=0):
        print(" " * indent, "plane-wave source:")
        print(" " * indent, "propagation direction:", self.k)
        print(" " * indent, "center:", self.center)
        print(" " * indent, "source plane size:", self.size)
        print(" " * indent, "polarization direction:", self.e_direction)
        print(" " * indent, "amplitude:", self.amp)
        self.src_time.display_info(4)

    def mode_function(self, x, y, z):
        return 1.0

    def _dist_from_center(self, point):
        return np.linalg.norm(self.center - point)

    def _metric_from_center_along_beam_axis(self, point):
        return np.dot(self.k, point - self.center)

    def _dist_from_beam_axis(self, x, y, z):
        return np.linalg.norm(np.cross(self.k, (x, y, z) - self.center))

    def _axis_in_k(self):
        dot_with_axis = {}
        for d in (const.PlusX, const.PlusY, const.PlusZ, const.MinusX, const.MinusY, const.MinusZ):
            dot_with_axis[np.dot(d.vector, self.k)] = d
        return dot_with_axis[max(dot_with_axis)]

    def _3d_dispersion_relation(self, zeta, v, omega, ds, dt, k):
        """Taflove & Hagness eq. 5.65 (gmes/source.py:506-526)."""
        lhs = (sin(0.5 * dt * omega) / v / dt) ** 2
        rhs = sum((np.sin(0.5 * zeta * np.array(ds) * k) / ds) ** 2)
        return lhs - rhs

    def _1d_dispersion_relation(self, ds, zeta, v, omega, dt, k):
        """Taflove & Hagness eq. 5.67 (gmes/source.py:528-547)."""
        lhs = sin(0.5 * omega * dt) / v / dt
        if ds == 0:
            rhs = 0.5 * k * zeta
        else:
            rhs = sin(0.5 * k * zeta * ds) / ds
        return lhs - rhs

    def _get_aux_fdtd(self, space, geom_tree, cmplx):
        """The 1-D incident grid (gmes/source.py:549-630): its cell size
        matches the main grid's numerical dispersion along k; a 50-cell
        CPML and a hard Ex source sit beyond the farthest face."""
        from scipy.optimize import bisect
        from .fdtd import TEMzFDTD
        from .geometry import Cartesian
        from .material import Cpml
        from .pygeom import DefaultMedium, Shell
        default_medium = geom_tree.object_of_point((inf, inf, inf))[0]
        eps_inf = default_medium.material.eps_inf
        mu_inf = default_medium.material.mu_inf
        v = 1 / sqrt(eps_inf * mu_inf)
        ds = tuple(space.dr)
        dt = space.dt
        omega = 2 * pi * self.src_time.freq
        wave_vector = omega * self.k / v
        zeta = bisect(self._3d_dispersion_relation, 0, 2, (v, omega, ds, dt, wave_vector))
        wave_number = omega / v
        delta_1d = bisect(self._1d_dispersion_relation, 0, 2 * max(ds),
                          (zeta, v, omega, dt, wave_number))
        pml_thickness = 50 * delta_1d
        vertices = []
        for x in (0.5 * self.size[0], -0.5 * self.size[0]):
            for y in (0.5 * self.size[1], -0.5 * self.size[1]):
                for z in (0.5 * self.size[2], -0.5 * self.size[2]):
                    vertices.append(self.center + (x, y, z))
        dist = list(map(abs, list(map(self._metric_from_center_along_beam_axis, vertices))))
        max_dist = max(dist)
        longitudinal_size = 2 * (max_dist + pml_thickness + 2 * delta_1d)
        aux_size = (0, 0, longitudinal_size)
        mat_objs = self.geom_tree.material_of_point((inf, inf, inf))[0]
        aux_space = Cartesian(size=aux_size, resolution=1 / delta_1d, parallel=False)
        aux_geom_list = (DefaultMedium(material=mat_objs),
                         Shell(material=Cpml(), thickness=pml_thickness))
        src_pnt = (0, 0, -max_dist - delta_1d)
        aux_src_list = (PointSource(src_time=deepcopy(self.src_time), component=const.Ex,
                                    center=src_pnt),)
        return TEMzFDTD(aux_space, list(aux_geom_list), list(aux_src_list), dt=space.dt,
                        bloch=(0, 0, 0) if cmplx else None, verbose=False,
                        precision=self._precision, device=self._device, _engine=self._engine)

    def _faces(self, comp):
        """[(face entry, cosine)] of ``comp`` in the reference's order; faces
        whose cosine is zero are left out (gmes/source.py:705-1406)."""
        electric = comp in const.E_COMPONENTS
        direction = self.h_direction if electric else self.e_direction
        out = []
        for entry in _TFSF_FACES.get(comp, ()):
            cosine = np.dot(direction, entry[3])
            if cosine != 0:
                out.append((entry, cosine))
        return out

    def _envelope_row(self, electric):
        return None

    def _mode_values(self, pts):
        """mode_function at each row of pts (1 for a plane wave)."""
        return np.ones(len(pts))

    def _beam_metric(self, pts):
        """dot(k, p - center) for each row of pts, with the reference's
        np.dot per point (gmes/source.py:449-459); for an axis-aligned k
        the product is exact and is taken vectorised."""
        nz = np.flatnonzero(self.k)
        if len(nz) == 1:
            a = int(nz[0])
            return (pts[:, a] - self.center[a]) * self.k[a]
        return np.array([np.dot(self.k, p - self.center) for p in pts])

    def _face_blocks(self, comp, field_shape, space, geom_tree, tiles=None):
        """TFSF face corrections on ``comp`` (gmes/source.py:632-1406,
        gmes/pw_source.py:197-486), one block of arrays per face in the
        reference's face order: idx (M,3), kind, i0/i1 (flat aux indices),
        r0, r1, k.  eps/mu come from the material tiles when given (the
        record of a non-ghost cell holds the underlying medium's value,
        as the reference's material_of_point/underneath rule gives)."""
        faces = self._faces(comp)
        if not faces:
            return []
        electric = comp in const.E_COMPONENTS
        aux = self.aux_fdtd
        aux_comp = const.Hy if electric else const.Ex
        aux_shape = aux._shape[aux_comp]
        d = space.dr
        own = comp.index % 3
        lo_ok = np.array([0 if (a == own or electric) else 1 for a in range(3)])
        hi_ok = np.array([field_shape[a] if (a == own or not electric) else field_shape[a] - 1
                          for a in range(3)])
        off_c = np.array(_geometry_stagger(comp))
        blocks = []
        for (face, axis, plus, _, idx_comp, lo_off, hi_off, samp_comp, samp_off, sign), cosine \
                in faces:
            if not 2 * space.half_size[axis] > space.dr[axis]:
                continue
            flip = np.ones(3)
            flip[axis] = -1
            if plus:
                low = self.center - self.half_size * flip
                high = self.center + self.half_size
            else:
                low = self.center - self.half_size
                high = self.center + self.half_size * flip
            low_idx = np.array(space.space_to_index(idx_comp, *low)) + lo_off
            high_idx = np.array([x + 1 for x in space.space_to_index(idx_comp, *high)]) + hi_off
            # the part of the face range that is in range (non-ghost storage)
            a0 = np.maximum(low_idx, lo_ok)
            a1 = np.minimum(high_idx, hi_ok)
            if np.any(a1 <= a0):
                continue
            grids = np.meshgrid(*[np.arange(a0[a], a1[a]) for a in range(3)], indexing="ij")
            idx = np.stack([g.reshape(-1) for g in grids], axis=1).astype(np.intp)
            pnt = _index_to_space_rows(space, off_c, idx)
            if tiles is not None:
                rec = tiles.record_array()
                eps_mu = rec["eps_mu"][tiles.ids[comp][idx[:, 0], idx[:, 1], idx[:, 2]]]
            else:
                eps_mu = np.empty(len(idx))
                for n, p_ in enumerate(pnt):
                    mat_obj, underneath = geom_tree.material_of_point(tuple(p_))
                    src = mat_obj if underneath is None else underneath
                    eps_mu[n] = float(src.eps_inf if electric else src.mu_inf)
            amp = cosine * self.amp * self._mode_values(pnt)
            samp_pt = _index_to_space_rows(space, np.array(_geometry_stagger(samp_comp)),
                                           idx + np.array(samp_off))
            z = self._beam_metric(samp_pt)
            g = (z + aux.space.half_size[2]) / aux.space.dr[2]
            so = _geometry_stagger(aux_comp)[2]
            samp = g - so if so != 0 else g
            fl = np.floor(samp)
            i0z = fl.astype(np.intp)
            r1 = samp - fl
            r0 = 1 - r1
            col = (0, 0) if aux_comp in const.E_COMPONENTS else _collapsed_index(aux_comp)[:2]
            if i0z.min() < 0 or i0z.max() + 1 >= aux_shape[2]:
                raise ValueError("TFSF sample point outside the auxiliary grid")
            base = (col[0] * aux_shape[1] + col[1]) * aux_shape[2]
            blocks.append(dict(face=face, idx=idx, kind=SRC_AUX_ADD if sign > 0 else SRC_AUX_SUB,
                               i0=base + i0z, i1=base + i0z + 1, r0=r0, r1=r1,
                               k=dt_over(space.dt, eps_mu, d[axis]) * amp))
        return blocks

    def _pw_block(self, comp, field_shape, space, geom_tree, tiles=None):
        """All face ops of ``comp`` as arrays, in face order (cells on two
        faces appear twice, once per face)."""
        blocks = self._face_blocks(comp, field_shape, space, geom_tree, tiles)
        if not blocks:
            return None
        cat = lambda key: np.concatenate([b[key] for b in blocks])
        return dict(idx=cat("idx"), kind=np.concatenate([np.full(len(b["idx"]), b["kind"], np.int32)
                                                         for b in blocks]),
                    i0=cat("i0"), i1=cat("i1"), r0=cat("r0"), r1=cat("r1"), k=cat("k"),
                    aux=self._aux_slot, aux_comp=(const.Hy if comp in const.E_COMPONENTS else const.Ex).index,
                    aux_owner=self.aux_fdtd, envelope=self._envelope_row(comp in const.E_COMPONENTS))

    def _pw_ops(self, comp, field_shape, space, geom_tree, tiles=None):
        """The face ops of ``comp`` one by one: [(idx, "tfsf", kind, payload)]."""
        out = []
        env = self._envelope_row(comp in const.E_COMPONENTS)
        aux_comp = const.Hy if comp in const.E_COMPONENTS else const.Ex
        for b in self._face_blocks(comp, field_shape, space, geom_tree, tiles):
            for n in range(len(b["idx"])):
                payload = dict(aux=self._aux_slot, aux_comp=aux_comp.index, i0=int(b["i0"][n]),
                               i1=int(b["i1"][n]), k=float(b["k"][n]), r0=float(b["r0"][n]),
                               r1=float(b["r1"][n]), face=b["face"], aux_owner=self.aux_fdtd,
                               envelope=env)
                out.append((tuple(int(v) for v in b["idx"][n]), "tfsf", b["kind"], payload))
        return out


def _geometry_stagger(comp):
    from .geometry import STAGGER
    return STAGGER[comp]


def _collapsed_index(comp):
    from .geometry import COLLAPSED_INDEX
    return COLLAPSED_INDEX[comp]


def _index_to_space_rows(space, off, idx):
    """index_to_space for each row of idx, with the same expression per
    axis: i*dr - half_size, or (i + off)*dr - half_size."""
    out = np.empty(idx.shape, np.float64)
    gi = idx + space.general_field_size * space.my_cart_idx
    for a in range(3):
        if off[a] == 0:
            out[:, a] = gi[:, a] * space.dr[a] - space.half_size[a]
        else:
            out[:, a] = (gi[:, a] + off[a]) * space.dr[a] - space.half_size[a]
    return out


def dt_over(dt, eps_mu, d):
    """dt / (eps_inf * d): the face-correction factor of
    gmes/pw_source.py:212-486."""
    return dt / (eps_mu * d)


class GaussianBeam(TotalFieldScatteredField):
    """Transparent Gaussian beam launched through one face of the box
    (gmes/source.py:1525-1769).  The auxiliary grid is first run to a
    steady state; afterwards the incident samples are scaled by a sin^2
    turn-on envelope of the source's width, evaluated per half step."""

    def __init__(self, src_time, directivity, center, size, direction, polarization,
                 waist=inf, amp=1):
        TotalFieldScatteredField.__init__(self, src_time, center, size, direction,
                                          polarization, amp)
        if isinstance(directivity, type) and issubclass(directivity, const.Directional):
            self.directivity = directivity
        else:
            raise TypeError("directivity must be a Directional type.")
        self.waist = float(waist)

    def init(self, geom_tree, space, cmplx):
        """gmes/source.py:1583-1603"""
        self.geom_tree = geom_tree
        self.src_time.init(cmplx)
        aux_fdtd = self._get_aux_fdtd(space, geom_tree, cmplx)
        aux_fdtd.init()
        raising = aux_fdtd.src_list[0].src_time.width
        dist = 2 * aux_fdtd.space.half_size[2]
        from .pygeom import DefaultMedium
        default_medium = next(i for i in aux_fdtd.geom_list if isinstance(i, DefaultMedium))
        v_p = 1 / sqrt(default_medium.material.eps_inf * default_medium.material.mu_inf)
        passby = raising + dist / v_p
        aux_fdtd.step_until_t(2 * passby)
        self.aux_fdtd = aux_fdtd
        self._env_width = raising

    def display_info(self, indent=0):
        print(" " * indent, "Gaussian beam source:")
        print(" " * indent, end=" ")
        print("propagation direction:", self.k, end=" ")
        print("center:", self.center)
        print("source plane size:", self.size)
        print("polarization direction:", self.e_direction)
        print("beam waist:", self.waist)
        print("maximum amp.:", self.amp)
        self.src_time.display_info(indent + 4)

    def mode_function(self, x, y, z):
        r = self._dist_from_beam_axis(x, y, z)
        return exp(-((r / self.waist) ** 2))

    def _mode_values(self, pts):
        return np.array([self.mode_function(*p) for p in pts])

    def _faces(self, comp):
        """The one face of ``comp`` whose normal is the directivity axis and
        which lies on the side the beam enters from; zero cosines are kept
        (gmes/source.py:1613-1732)."""
        electric = comp in const.E_COMPONENTS
        direction = self.h_direction if electric else self.e_direction
        axis = int(np.argmax(np.abs(self.directivity.vector)))
        plus_face = self.directivity.vector[axis] < 0
        for entry in _TFSF_FACES.get(comp, ()):
            if entry[1] == axis and entry[2] == plus_face:
                return [(entry, np.dot(direction, entry[3]))]
        return []

    def _envelope_row(self, electric):
        """env(t) with t = (number of auxiliary steps taken since the
        pre-run) * dt: k*dt at the E half of main step k, (k+1)*dt at its H
        half (gmes/source.py:1734-1769)."""
        width = self._env_width

        def value(dt, n):
            m = n - 0.5 if electric else n
            t = int(m) * dt
            if t < width:
                return sin(0.5 * pi * t / width) ** 2
            return 1

        return WaveRow(value, electric)
