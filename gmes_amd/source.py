"""Source time functions and sources.

Drop-in counterpart of the reference's ``gmes.source`` / ``gmes.pw_source``
(gmes/source.py:87-1769, gmes/pw_source.py).  Oscillators are evaluated
on the host in FP64 with the reference's expressions (they are cheap and
this keeps every injected value bit-identical); the engine receives, per
source operation, a waveform row sampled at the half step the operation
runs at (n = k + 1/2 for E, n = k + 1 for H).
"""

from cmath import exp as cexp
from math import exp, pi, sin

import numpy as np
from numpy import inf

from . import constant as const
from ._abi import SRC_SET, SRC_SUB
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
    """Host generator of one waveform row: value(n) for the half step n."""

    def __init__(self, fn, electric, recorder=None):
        self.fn = fn
        self.electric = electric
        self.recorder = recorder

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
        return [(tuple(int(v) for v in idx), "point", kind, WaveRow(value, electric))]
