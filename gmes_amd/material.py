"""Material types and their FP64 coefficient setup.

Drop-in counterpart of the reference's Cython module ``gmes.material``
(src/material.pyx).  Each class keeps the reference's constructor
arguments and computes its update coefficients in FP64 with the
reference's formulas, in the same operation order, so the values handed
to the engine are bit-identical to the ones the reference attaches to its
per-cell parameter structs.

Where the reference builds one C++ updater object per cell
(``get_pw_material_<c>``), a material here describes itself once per
component through ``_tile``: an update class (``GMES_CLS_*``), pole
counts, a coefficient list and the permittivity/permeability the update
uses.  The tile builder (``gmes_amd.tiles``) turns these into the
engine's record table.
"""

from cmath import exp as cexp
from collections.abc import Sequence
from math import cos, exp, sin, sqrt, tanh

import numpy as np
from numpy import array, empty

from ._abi import (CLS_CONST, CLS_CPML, CLS_DCP_ADE, CLS_DCP_PLRC, CLS_DIELECTRIC, CLS_DM2,
                   CLS_DRUDE, CLS_DUMMY, CLS_UPML)
from .pygeom import Compound, Material


def _eps_mu(mat, electric, underneath):
    """eps_inf (E) or mu_inf (H) of the update: the underlying medium's when
    there is one (e.g. src/material.pyx:303-306)."""
    src = mat if underneath is None else underneath
    return src.eps_inf if electric else src.mu_inf


class Dummy(Material):
    """No update (src/material.pyx:20-133); also used on ghost planes."""

    def __init__(self, eps_inf=1, mu_inf=1):
        Material.__init__(self, eps_inf, mu_inf)

    def init(self, space, param=None):
        pass

    def display_info(self, indent=0):
        print(" " * indent, "dummy object")

    def _tile(self, electric, underneath, cmplx):
        return dict(cls=CLS_DUMMY, eps_mu=_eps_mu(self, electric, underneath))


class Const(Material):
    """Sets the field to a constant (src/material.pyx:136-276)."""

    def __init__(self, value=0, eps_inf=1, mu_inf=1):
        Material.__init__(self, eps_inf, mu_inf)
        self.value = value if type(value) is complex else float(value)

    def init(self, space, param=None):
        pass

    def display_info(self, indent=0):
        print(" " * indent, "const object", "value:", self.value)

    def _tile(self, electric, underneath, cmplx):
        v = complex(self.value)
        if not cmplx and v.imag != 0:
            raise TypeError("a complex Const value needs complex fields")
        return dict(cls=CLS_CONST, eps_mu=_eps_mu(self, electric, underneath),
                    value=(v.real, v.imag))


class Dielectric(Material):
    """Non-dispersive isotropic medium (src/material.pyx:279-397)."""

    def __init__(self, eps_inf=1, mu_inf=1):
        Material.__init__(self, eps_inf, mu_inf)

    def init(self, space, param=None):
        pass

    def display_info(self, indent=0):
        print(" " * indent, "dielectric", "eps_inf:", self.eps_inf, "mu_inf:", self.mu_inf)

    def _tile(self, electric, underneath, cmplx):
        return dict(cls=CLS_DIELECTRIC, eps_mu=_eps_mu(self, electric, underneath))


class Pml(Material, Compound):
    """Common grading of the PML media (src/material.pyx:400-502)."""

    def __init__(self, eps_inf, mu_inf):
        Material.__init__(self, eps_inf, mu_inf)
        self.initialized = False

    def init(self, space, param):
        self.center = np.array(param[0], np.double)
        self.half_size = array(param[1], np.double)
        self.d = param[2]
        for i in range(3):
            if self.half_size[i] < self.d:
                self.half_size[i] = np.inf
        self.dt = space.dt
        self.dw = array(space.dr, np.double)
        self.sigma_max = self.sigma_max_ratio * self.get_sigma_opt()
        self.initialized = True

    def get_sigma_opt(self):
        eta = sqrt(self.mu_inf / self.eps_inf)
        return 0.8 * (self.m + 1) / (eta * self.dw)

    def _depth(self, w, component):
        """Shifted coordinate -> (inside, depth) of the polynomial grading."""
        w -= self.center[component]
        half_size = self.half_size[component]
        if w <= self.d - half_size:
            return True, min(max((half_size + w) / self.d, 0), 1)
        elif half_size - self.d <= w:
            return True, min(max((half_size - w) / self.d, 0), 1)
        return False, None

    def sigma(self, w, component):
        inside, depth = self._depth(w, component)
        return self.sigma_max[component] * (1 - depth) ** self.m if inside else 0

    def kappa(self, w, component):
        inside, depth = self._depth(w, component)
        return 1 + (self.kappa_max - 1) * (1 - depth) ** self.m if inside else 1

    def profile(self, coords, axis):
        """Rows of per-axis coefficients at each coordinate along ``axis``."""
        return np.array([self._profile_row(w, axis) for w in coords], np.double)


class Upml(Pml):
    """Uniaxial PML (src/material.pyx:505-729)."""

    profile_width = 6

    def __init__(self, eps_inf=1, mu_inf=1, m=3.6, kappa_max=4.6, sigma_max_ratio=.745):
        Pml.__init__(self, eps_inf, mu_inf)
        self.m = float(m)
        self.kappa_max = float(kappa_max)
        self.sigma_max_ratio = float(sigma_max_ratio)

    def display_info(self, indent=0):
        print(" " * indent, "UPML", "m:", self.m, "kappa_max:", self.kappa_max)

    def c1(self, w, component):
        numerator = 2 * self.kappa(w, component) - self.sigma(w, component) * self.dt
        denominator = 2 * self.kappa(w, component) + self.sigma(w, component) * self.dt
        return numerator / denominator

    def c2(self, w, component):
        numerator = 2 * self.dt
        denominator = 2 * self.kappa(w, component) + self.sigma(w, component) * self.dt
        return numerator / denominator

    def c3(self, w, component):
        return self.c1(w, component)

    def c4(self, w, component):
        denominator = 2 * self.kappa(w, component) + self.sigma(w, component) * self.dt
        return 1 / denominator

    def c5(self, w, component):
        return 2 * self.kappa(w, component) + self.sigma(w, component) * self.dt

    def c6(self, w, component):
        return 2 * self.kappa(w, component) - self.sigma(w, component) * self.dt

    def _profile_row(self, w, a):
        return [self.c1(w, a), self.c2(w, a), self.c3(w, a), self.c4(w, a), self.c5(w, a),
                self.c6(w, a)]

    def _tile(self, electric, underneath, cmplx):
        return dict(cls=CLS_UPML, eps_mu=_eps_mu(self, electric, underneath), pml=self)


class Cpml(Pml):
    """Complex-frequency-shifted PML (src/material.pyx:732-954)."""

    profile_width = 3

    def __init__(self, eps_inf=1, mu_inf=1, m=3.4, kappa_max=1, m_a=4.8, a_max=0.8,
                 sigma_max_ratio=0.65):
        Pml.__init__(self, eps_inf, mu_inf)
        self.m = float(m)
        self.kappa_max = float(kappa_max)
        self.m_a = float(m_a)
        self.a_max = float(a_max)
        self.sigma_max_ratio = float(sigma_max_ratio)

    def display_info(self, indent=0):
        print(" " * indent, "CPML", "m:", self.m, "kappa_max:", self.kappa_max,
              "m_a:", self.m_a, "a_max:", self.a_max)

    def a(self, w, component):
        inside, depth = self._depth(w, component)
        return self.a_max * depth ** self.m_a if inside else 0

    def b(self, w, component):
        exponent = -(self.sigma(w, component) / self.kappa(w, component) + self.a(w, component)) * self.dt
        return exp(exponent)

    def c(self, w, component):
        sigma = self.sigma(w, component)
        kappa = self.kappa(w, component)
        numerator = sigma * (self.b(w, component) - 1)
        denominator = (sigma + kappa * self.a(w, component)) * kappa
        if denominator:
            return numerator / denominator
        return 0

    def _profile_row(self, w, a):
        return [self.b(w, a), self.c(w, a), self.kappa(w, a)]

    def _tile(self, electric, underneath, cmplx):
        return dict(cls=CLS_CPML, eps_mu=_eps_mu(self, electric, underneath), pml=self)


class DrudePole(object):
    def __init__(self, omega, gamma):
        self.omega = float(omega)
        self.gamma = float(gamma)


class LorentzPole(object):
    def __init__(self, amp, omega, gamma):
        self.amp = float(amp)
        self.omega = float(omega)
        self.gamma = float(gamma)


class CriticalPoint(object):
    def __init__(self, amp, phi, omega, gamma):
        self.amp = float(amp)
        self.phi = float(phi)
        self.omega = float(omega)
        self.gamma = float(gamma)


class _DispersiveH(object):
    """Dispersive media update H as a dielectric (src/pw_drude.hh:249-298,
    src/pw_lorentz.hh:257-305, src/pw_dcp.hh:347-395,697-745,
    src/pw_dm2.hh:595-677)."""

    def _tile_h(self, underneath):
        return dict(cls=CLS_DIELECTRIC, eps_mu=_eps_mu(self, False, underneath))


class DcpAde(Dielectric, _DispersiveH):
    """Drude-critical-points model, ADE form (src/material.pyx:1029-1224)."""

    def __init__(self, eps_inf=1, mu_inf=1, sigma=0, dps=(), cps=()):
        Dielectric.__init__(self, eps_inf, mu_inf)
        self.sigma = float(sigma)
        self.dps = tuple(dps)
        self.cps = tuple(cps)
        self.initialized = False

    def init(self, space, param=None):
        self.dt = space.dt
        self.a = empty((len(self.dps), 3), np.double)
        for i in range(len(self.dps)):
            pole = self.dps[i]
            denom = float(self.dt * pole.gamma + 2)
            self.a[i, 0] = (self.dt * pole.gamma - 2) / denom
            self.a[i, 1] = 4 / denom
            self.a[i, 2] = 0.5 * (self.dt * pole.omega) ** 2 / denom
        self.b = empty((len(self.cps), 5), np.double)
        for i in range(len(self.cps)):
            pnt = self.cps[i]
            denom = (self.dt * pnt.gamma + 2) ** 2 + (self.dt * pnt.omega) ** 2
            self.b[i, 0] = -((self.dt * pnt.gamma - 2) ** 2 + (self.dt * pnt.omega) ** 2) / denom
            self.b[i, 1] = 2 * (4 - self.dt ** 2 * (pnt.gamma ** 2 + pnt.omega ** 2)) / denom
            self.b[i, 2] = 2 * self.dt * pnt.amp * pnt.omega * (
                self.dt * cos(pnt.phi) * pnt.omega + sin(pnt.phi) * (2 - self.dt * pnt.gamma)) / denom
            self.b[i, 3] = 4 * self.dt ** 2 * pnt.amp * pnt.omega * (
                cos(pnt.phi) * pnt.omega - sin(pnt.phi) * pnt.gamma) / denom
            self.b[i, 4] = 2 * self.dt * pnt.amp * pnt.omega * (
                self.dt * cos(pnt.phi) * pnt.omega - sin(pnt.phi) * (2 + self.dt * pnt.gamma)) / denom
        self.c = empty(4, np.double)
        denom = 0.5 * self.dt * self.sigma + sum(self.a[:, 2]) + sum(self.b[:, 4]) + self.eps_inf
        self.c[0] = self.dt / denom
        self.c[1] = 1 / denom
        self.c[2] = -(sum(self.a[:, 2]) + sum(self.b[:, 2])) / denom
        self.c[3] = -(0.5 * self.dt * self.sigma + 2 * sum(self.a[:, 2]) + sum(self.b[:, 3])
                      - self.eps_inf) / denom
        self.initialized = True

    def _tile(self, electric, underneath, cmplx):
        if not electric:
            return self._tile_h(underneath)
        coef = list(self.a.ravel()) + list(self.b.ravel()) + list(self.c)
        return dict(cls=CLS_DCP_ADE, eps_mu=_eps_mu(self, True, underneath), n1=len(self.dps),
                    n2=len(self.cps), coef=coef)


class DcpPlrc(Dielectric, _DispersiveH):
    """Drude-critical-points model, piecewise-linear recursive convolution
    (src/material.pyx:1227-1471)."""

    def __init__(self, eps_inf=1, mu_inf=1, sigma=0, dps=(), cps=()):
        Dielectric.__init__(self, eps_inf=eps_inf, mu_inf=mu_inf)
        self.sigma = float(sigma)
        self.dps = tuple(dps)
        self.cps = tuple(cps)
        self.initialized = False

    def init(self, space, param=None):
        self.dt = space.dt
        self.a = empty((len(self.dps), 3), np.double)
        for i, pole in enumerate(self.dps):
            self.a[i, 0] = self.delta_chi_dp_0(pole) - self.delta_xi_dp_0(pole)
            self.a[i, 1] = self.delta_xi_dp_0(pole)
            self.a[i, 2] = exp(-pole.gamma * self.dt)
        self.b = empty((len(self.cps), 3), complex)
        for i, pnt in enumerate(self.cps):
            self.b[i, 0] = self.delta_chi_cp_0(pnt) - self.delta_xi_cp_0(pnt)
            self.b[i, 1] = self.delta_xi_cp_0(pnt)
            self.b[i, 2] = cexp(-self.dt * (pnt.gamma + 1j * pnt.omega))
        chi_0 = (sum([self.chi_dp_0(p) for p in self.dps]
                     + [self.chi_cp_0(p) for p in self.cps]) + 0j).real
        xi_0 = (sum([self.xi_dp_0(p) for p in self.dps]
                    + [self.xi_cp_0(p) for p in self.cps]) + 0j).real
        self.c = empty(3, np.double)
        conductivity = 0.5 * self.dt * self.sigma
        denom = self.eps_inf - xi_0 + chi_0 + conductivity
        self.c[0] = self.dt / denom
        self.c[1] = (self.eps_inf - xi_0 - conductivity) / denom
        self.c[2] = 1 / denom
        self.initialized = True

    def chi_dp_0(self, dp):
        gdt = dp.gamma * self.dt
        return (dp.omega / dp.gamma) ** 2 * (exp(-gdt) + gdt - 1)

    def xi_dp_0(self, dp):
        chi_dp_0 = self.chi_dp_0(dp)
        gdt = dp.gamma * self.dt
        return (chi_dp_0 * (1 / (1 - exp(gdt)) + 1 / gdt)
                + (dp.omega / dp.gamma) ** 2 * (gdt / 2 / tanh(gdt / 2) - 1))

    def delta_chi_dp_0(self, dp):
        gdt = dp.gamma * self.dt
        return -(dp.omega / dp.gamma * (1 - exp(-gdt))) ** 2

    def delta_xi_dp_0(self, dp):
        gdt = dp.gamma * self.dt
        return self.delta_chi_dp_0(dp) * (1 / (1 - exp(gdt)) + 1 / gdt)

    def chi_cp_0(self, cp):
        go = cp.gamma + 1j * cp.omega
        return 2j * cp.amp * cp.omega * cexp(1j * cp.phi) * (1 - cexp(-self.dt * go)) / go

    def xi_cp_0(self, cp):
        dtgo = self.dt * (cp.gamma + 1j * cp.omega)
        return self.chi_cp_0(cp) * (1 / (1 - cexp(dtgo)) + 1 / dtgo)

    def delta_chi_cp_0(self, cp):
        go = cp.gamma + 1j * cp.omega
        return 2j * cp.amp * cp.omega * cexp(1j * cp.phi) * (1 - cexp(-self.dt * go)) ** 2 / go

    def delta_xi_cp_0(self, cp):
        dtgo = self.dt * (cp.gamma + 1j * cp.omega)
        return self.delta_chi_cp_0(cp) * (1 / (1 - cexp(dtgo)) + 1 / dtgo)

    def _tile(self, electric, underneath, cmplx):
        if not electric:
            return self._tile_h(underneath)
        coef = list(self.a.ravel())
        for row in self.b:
            for z in row:
                coef += [z.real, z.imag]
        coef += list(self.c)
        return dict(cls=CLS_DCP_PLRC, eps_mu=_eps_mu(self, True, underneath), n1=len(self.dps),
                    n2=len(self.cps), coef=coef)


class DcpRc(DcpPlrc):
    """Recursive-convolution variant: the xi terms vanish
    (src/material.pyx:1474-1497)."""

    def xi_dp_0(self, dp):
        return 0

    def xi_cp_0(self, cp):
        return 0j

    def delta_xi_dp_0(self, dp):
        return 0

    def delta_xi_cp_0(self, cp):
        return 0j


class Drude(Dielectric, _DispersiveH):
    """Multi-pole Drude ADE (src/material.pyx:1500-1681)."""

    def __init__(self, eps_inf=1, mu_inf=1, sigma=0, dps=()):
        Dielectric.__init__(self, eps_inf=eps_inf, mu_inf=mu_inf)
        self.sigma = float(sigma)
        self.dps = tuple(dps)
        self.initialized = False

    def init(self, space, param=None):
        self.dt = space.dt
        self.a = empty((len(self.dps), 3), np.double)
        for i in range(len(self.dps)):
            pole = self.dps[i]
            denom = self.dt * pole.gamma + 2.
            self.a[i, 0] = (self.dt * pole.gamma - 2) / denom
            self.a[i, 1] = 4 / denom
            self.a[i, 2] = 2 * (self.dt * pole.omega) ** 2 / denom
        self.c = empty(3, np.double)
        denom = 2. * self.eps_inf + self.dt * self.sigma
        self.c[0] = 2 * self.dt / denom
        self.c[1] = -2 / denom
        self.c[2] = (2 * self.eps_inf - self.dt * self.sigma) / denom
        self.initialized = True

    def _tile(self, electric, underneath, cmplx):
        if not electric:
            return self._tile_h(underneath)
        return dict(cls=CLS_DRUDE, eps_mu=_eps_mu(self, True, underneath), n1=len(self.dps),
                    coef=list(self.a.ravel()) + list(self.c))


class Lorentz(Dielectric, _DispersiveH):
    """Multi-pole Lorentz ADE (src/material.pyx:1684-1859)."""

    def __init__(self, eps_inf=1, mu_inf=1, sigma=0, lps=()):
        Dielectric.__init__(self, eps_inf=eps_inf, mu_inf=mu_inf)
        self.sigma = float(sigma)
        self.lps = tuple(lps)
        self.initialized = False

    def init(self, space, param=None):
        self.dt = space.dt
        self.a = empty((len(self.lps), 3), np.double)
        for i in range(len(self.lps)):
            pole = self.lps[i]
            denom = self.dt * pole.gamma + 2.
            self.a[i, 0] = (self.dt * pole.gamma - 2) / denom
            self.a[i, 1] = (4 - 2 * (self.dt * pole.omega) ** 2) / denom
            self.a[i, 2] = 2 * pole.amp * (self.dt * pole.omega) ** 2 / denom
        self.c = empty(3, np.double)
        denom = 2. * self.eps_inf + self.dt * self.sigma
        self.c[0] = 2 * self.dt / denom
        self.c[1] = -2 / denom
        self.c[2] = (2 * self.eps_inf - self.dt * self.sigma) / denom
        self.initialized = True

    def _tile(self, electric, underneath, cmplx):
        if not electric:
            return self._tile_h(underneath)
        # n2 = 1 marks a Lorentz medium: the same update as Drude
        # (src/pw_lorentz.hh:96-107, src/pw_drude.hh:90-101); the flag names
        # the reference's updater type for the reference-updater harness
        return dict(cls=CLS_DRUDE, eps_mu=_eps_mu(self, True, underneath), n1=len(self.lps),
                    n2=1, coef=list(self.a.ravel()) + list(self.c))


class Dm2(Dielectric, _DispersiveH):
    """Two-level density-matrix medium, predictor-corrector
    (src/material.pyx:1862-2079).  Real fields only."""

    def __init__(self, eps_inf=1, mu_inf=1, omega=(1,), n_atom=(1,), rho30=-1, gamma=1, t1=1,
                 t2=1, hbar=1, rtol=10e-5):
        Dielectric.__init__(self, eps_inf, mu_inf)
        if isinstance(omega, (Sequence, np.ndarray)):
            self.omega = tuple(map(float, omega))
        else:
            self.omega = (float(omega),)
        if isinstance(n_atom, (Sequence, np.ndarray)):
            self.n_atom = tuple(map(float, n_atom))
        else:
            self.n_atom = (float(n_atom),)
        self.rho30 = float(rho30)
        self.gamma = float(gamma)
        self.t1 = float(t1)
        self.t2 = float(t2)
        self.hbar = float(hbar)
        self.rtol = float(rtol)
        self.initialized = False

    def init(self, space, param=None):
        self.dt = space.dt
        self.initialized = True

    def _tile(self, electric, underneath, cmplx):
        if cmplx:
            raise ValueError("Dm2 class supports real fields only")
        if not electric:
            return self._tile_h(underneath)
        if len(self.omega) != len(self.n_atom):
            raise ValueError("omega and n_atom must have equal lengths")
        coef = list(self.omega) + list(self.n_atom) + [self.rho30, self.gamma, self.t1, self.t2,
                                                       self.hbar, self.rtol]
        return dict(cls=CLS_DM2, eps_mu=_eps_mu(self, True, underneath), n1=len(self.omega),
                    coef=coef)
