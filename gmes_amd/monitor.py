"""Running-DFT flux monitor.

The reference records fields point by point (``Probe``, gmes/file_io.py);
it has no frequency-domain monitor.  This one accumulates, on the device
and in FP64, the DFT of the tangential fields on a plane z = const at a
list of frequencies while the simulation runs, and returns the flux
spectrum  P(f) = Re sum_{x,y} [Ex(f) Hy(f)* - Ey(f) Hx(f)*] dx dy.

Time alignment follows the leapfrog: after step k the E fields are those
of n = k + 1/2 and the H fields those of n = k + 1, so the DFT kernel
exp(-2 pi i f t) is evaluated at t = (k + 1/2) dt for E and (k + 1) dt for
H.  H is averaged over the two planes that straddle the E plane and
co-located with E in x-y: Hy(i+1, j) with Ex(i, j), Hx(i, j+1) with
Ey(i, j) (Yee positions of gmes/geometry.py:360-812).  With decimation M
only steps k = 0, M, 2M, ... are accumulated, each weighted M dt.
"""

import numpy as np

from . import constant as const


def _register(sim, freqs):
    if sim._sim is None:
        raise RuntimeError("init() must be called before adding a monitor")
    freqs = np.atleast_1d(np.asarray(freqs, np.float64))
    if sim._dft_freqs is not None and not np.array_equal(sim._dft_freqs, freqs):
        raise ValueError("all DFT monitors of a simulation share one frequency list")
    sim._dft_freqs = freqs
    return freqs


class DftBox(object):
    """Running DFT of one component over a box [lo, hi) of its storage
    indices: values(f, i, j, k) = sum_k w_k F(i, j, k; t_k)."""

    def __init__(self, sim, comp, lo, hi, freqs, decimation=1):
        self.freqs = _register(sim, freqs)
        self.decimation = int(decimation)
        if self.decimation < 1:
            raise ValueError("decimation must be positive")
        self.sim, self.comp = sim, comp
        self.lo, self.hi = tuple(int(v) for v in lo), tuple(int(v) for v in hi)
        self.box = tuple(h - l for l, h in zip(self.lo, self.hi))
        self._id = sim._lib.add_dft(sim._sim, comp.index, self.lo, self.hi, len(self.freqs),
                                    self.decimation)
        sim._dft_monitors.append(self)
        sim._dft_win = (0, 0)  # weights are uploaded again before the next step

    def values(self):
        n = int(np.prod(self.box))
        v = self.sim._lib.read_dft(self.sim._sim, self._id, len(self.freqs), n)
        return v.reshape((len(self.freqs),) + self.box) * float(self.decimation)


class DftFlux(object):
    def __init__(self, sim, freqs, z, decimation=1):
        if not all(c in sim._active for c in (const.Ex, const.Ey, const.Hx, const.Hy)):
            raise ValueError("the flux plane needs Ex, Ey, Hx and Hy")
        self.sim = sim
        nx, ny, nz = (int(v) for v in sim.space.whole_field_size)
        k0 = int(sim.space.space_to_index(const.Ex, 0, 0, z)[2])
        if not 0 <= k0 < nz:
            raise IndexError("flux plane outside the grid")
        self.k0, self.shape = k0, (nx, ny)
        self.boxes = {
            const.Ex: DftBox(sim, const.Ex, (0, 0, k0), (nx, ny, k0 + 1), freqs, decimation),
            const.Ey: DftBox(sim, const.Ey, (0, 0, k0), (nx, ny, k0 + 1), freqs, decimation),
            const.Hx: DftBox(sim, const.Hx, (0, 1, k0), (nx, ny + 1, k0 + 2), freqs, decimation),
            const.Hy: DftBox(sim, const.Hy, (1, 0, k0), (nx + 1, ny, k0 + 2), freqs, decimation),
        }
        self.freqs = self.boxes[const.Ex].freqs

    def fields(self):
        """DFT fields: Ex, Ey of shape (nfreq, nx, ny); Hx, Hy co-located
        with them, (nfreq, nx, ny)."""
        ex = self.boxes[const.Ex].values()[..., 0]
        ey = self.boxes[const.Ey].values()[..., 0]
        hx = self.boxes[const.Hx].values()
        hy = self.boxes[const.Hy].values()
        return ex, ey, 0.5 * (hx[..., 0] + hx[..., 1]), 0.5 * (hy[..., 0] + hy[..., 1])

    def flux(self):
        """Flux spectrum through the plane, one value per frequency."""
        ex, ey, hx, hy = self.fields()
        dx, dy = self.sim.space.dr[0], self.sim.space.dr[1]
        s = ex * np.conj(hy) - ey * np.conj(hx)
        return np.real(s.sum(axis=(1, 2))) * dx * dy


def dft_weights(freqs, step0, nsteps, dt):
    """w[s][f][0 = E, 1 = H][re, im] = exp(-2 pi i f t) dt at t = (s + 1/2) dt
    and (s + 1) dt for the absolute steps s = step0 .. step0 + nsteps - 1."""
    s = np.arange(step0, step0 + nsteps, dtype=np.float64)[:, None]
    f = np.asarray(freqs, np.float64)[None, :]
    out = np.empty((nsteps, len(freqs), 2, 2), np.float64)
    for e, half in ((0, 0.5), (1, 1.0)):
        w = np.exp(-2j * np.pi * f * ((s + half) * dt)) * dt
        out[:, :, e, 0] = w.real
        out[:, :, e, 1] = w.imag
    return out
