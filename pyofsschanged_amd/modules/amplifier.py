"""Gain models of pyofss/modules/amplifier.py.

* ``Amplifier`` (amplifier.py:42-175), saturable gain: the gain enters every
  linear sub-step (linearity.py:248-272); the device multiplies the spectrum
  by exp(G0/(1 + E/E_sat) * profile) with G0 = ln(10^(0.1 G)) h / (2 L) and
  E the energy of the array being propagated.
* ``Amplifier2LevelModel`` (amplifier.py:178-469), two-level Yb rate
  equations: in every linear sub-step the inversion N2 follows from the
  spectral power of the array being propagated, the spectrum is multiplied
  by exp(g_s h 1e3 / 2) and the pump power is advanced by exp(g_p h 1e3).

These classes prepare the constant arrays (set-up, once per domain, as the
reference does); the per-step arithmetic, including the pump state, runs in
the HIP engine.
"""

import os
import warnings

import numpy as np

from ..domain import Domain, lambda_to_nu
from ..field import fftshift, simpson_weights

# Yb cross sections (wavelength m, sigma_12 m^2, sigma_21 m^2), the table
# the reference reads at amplifier.py:247-258
YB_CROSS_SECTIONS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data",
                                 "CrossSectionDataExp.txt")


class Amplifier:
    """amplifier.py:42-175 (simple saturation model, Lorentzian or flat profile)."""

    def __init__(self, name="simple_saturation", gain=None, E_sat=None, P_sat=None, Tr=None, length=1.0, lamb0=None,
                 bandwidth=None, use_Er_profile=False, prg=None, queue=None, ctx=None, dorf="double"):
        if use_Er_profile:
            raise NotImplementedError("the Er gain profile is not supported by the HIP engine")
        self.length = length
        if gain is None:
            warnings.warn("The gain is not defined.")
        if (P_sat is not None) and (E_sat is not None):
            raise Exception("There should be only one parameter of saturation.")
        if bandwidth is not None and lamb0 is not None:
            self.delta = (Domain.vacuum_light_speed * bandwidth) / (lamb0 * lamb0)  # amplifier.py:71-73
            self.lamb = lamb0
        else:
            self.delta = None
            self.lamb = None
        self.name = name
        self.gain = gain
        if E_sat is not None:
            self.E_sat = E_sat
        elif P_sat is not None:
            if Tr is None:
                raise Exception("Repetition rate is not defined.")
            self.E_sat = P_sat * Tr
        else:
            self.E_sat = None
        self.use_Er_profile = use_Er_profile
        self.domain = None
        self.field = None

    def set_domain(self, domain):
        """amplifier.py:171-175"""
        self.domain = domain
        self.shape = domain.nu.shape
        self._profile = None

    @property
    def spectal_filtration_array(self):
        """amplifier.py:109-129: Lorentzian 1/(1 + 4 (nu - nu0)^2 / delta^2), FFT order."""
        if self._profile is None:
            if self.delta is not None and self.lamb is not None:
                f = 1 / (1 + 4 * (((self.domain.nu - lambda_to_nu(self.lamb)) ** 2) / ((self.delta) ** 2)))
            else:
                f = np.ones(len(self.domain.Lambda))
            self._profile = fftshift(f)
        return self._profile

    @property
    def gain_log(self):
        """ln(10^(0.1 G)), the constant of amplifier.py:131-147."""
        return float(np.log(np.power(10, 0.1 * self.gain)))

    def device_arrays(self):
        """(gain_log, length, E_sat, profile, Simpson weights) for gnlse_set_gain."""
        if self.E_sat is None:
            raise ValueError("saturation energy is not defined")
        return (self.gain_log, self.length, self.E_sat, self.spectal_filtration_array,
                simpson_weights(self.domain.t * 1e-3))


class Amplifier2LevelModel:
    """amplifier.py:178-469: two-level Yb gain model (Kirsch et al., Commun.
    Phys. 5, 219 (2022)); same constructor, constants and properties. The
    OpenCL arguments are accepted and ignored."""

    def __init__(self, name="amplifier2LevelModel", Pp=None, N=None, Rr=None, prg=None, queue=None, ctx=None,
                 dorf="double"):
        self.name = name
        self.h_p = 6.62 * 1e-34
        self.T = 850.0 * 1e-6  # s
        self.a = 3.0 * 1e-6  # m
        self.b = 60.0 * 1e-6
        self.NA = 0.13
        self.lamb_p = 976.0
        self.N = N * (1e25) * np.pi * self.a ** 2  # ions per unit length
        self._Pp = Pp
        self.Tr = 1 / Rr
        self.domain = None
        self.Pp_list = []
        self.inversion_factor_list = []
        self.gs_list = []
        self.sigma12_p = 2.5 * 1e-24
        self.sigma21_p = 2.44 * 1e-24
        self.physical_power_factor = None

    def interpolate_sigma_s(self):
        """amplifier.py:247-258: cubic interpolating spline of the table, zero
        outside it. pandas.read_csv(..., ',') with its default header=0 takes
        the table's first row as column names, so the spline starts at the
        second row; the same rows are used here."""
        from scipy import interpolate
        with open(YB_CROSS_SECTIONS) as fh:
            rows = [ln for ln in fh.read().splitlines() if ln.strip()][1:]
        tab = np.array([[float(v) for v in ln.split(",")] for ln in rows])
        IUS = interpolate.InterpolatedUnivariateSpline
        spl12 = IUS(tab[:, 0] * 1e9, tab[:, 1] * 1e27, ext=1)
        spl21 = IUS(tab[:, 0] * 1e9, tab[:, 2] * 1e27, ext=1)
        self.sigma12_s = spl12(self.domain.Lambda) * 1e-27
        self.sigma21_s = spl21(self.domain.Lambda) * 1e-27

    def set_domain(self, domain):
        """amplifier.py:260-266"""
        self.domain = domain
        self.interpolate_sigma_s()
        self.shape = domain.nu.shape
        self.physical_power_factor = np.float64(self.shape[0] * domain.dt / self.Tr)
        for k in ("_rho_s", "_rho_p", "_Psat_s", "_Psat_p", "_eta_s", "_eta_p", "_alpha_s", "_alpha_p", "_ratio_p",
                  "_ratio_s"):
            self.__dict__.pop(k, None)

    def _cached(self, key, fn):
        if key not in self.__dict__:
            with np.errstate(divide="ignore", invalid="ignore"):
                self.__dict__[key] = fn()
        return self.__dict__[key]

    @property
    def rho_s(self):
        """amplifier.py:268-276"""
        def f():
            V = (2.0 * np.pi / self.domain.Lambda) * self.a * self.NA * 1e9
            omega_s = self.a * (0.616 + 1.66 / pow(V, 1.5) + 0.987 / pow(V, 6.0))
            Gamma_s = 1.0 - np.exp(-2.0 * self.a ** 2 / (omega_s * omega_s))
            return Gamma_s / (np.pi * self.a ** 2)
        return self._cached("_rho_s", f)

    @property
    def rho_p(self):
        """amplifier.py:278-284"""
        return self._cached("_rho_p", lambda: ((self.a ** 2) / (self.b ** 2)) / (np.pi * self.a ** 2))

    @property
    def Psat_s(self):
        """amplifier.py:286-296 (inf where both cross sections vanish)."""
        return self._cached("_Psat_s", lambda: (self.h_p * self.domain.nu * 1e12) / (
            self.T * (self.sigma12_s + self.sigma21_s) * self.rho_s))

    @property
    def Psat_p(self):
        """amplifier.py:298-305"""
        return self._cached("_Psat_p", lambda: (self.h_p * lambda_to_nu(self.lamb_p) * 1e12) / (
            self.T * (self.sigma12_p + self.sigma21_p) * self.rho_p))

    @property
    def eta_s(self):
        """amplifier.py:307-315"""
        return self._cached("_eta_s", lambda: self.sigma12_s * self.rho_s * self.N)

    @property
    def eta_p(self):
        """amplifier.py:317-323"""
        return self._cached("_eta_p", lambda: self.sigma12_p * self.rho_p * self.N)

    @property
    def alpha_s(self):
        """amplifier.py:325-333"""
        return self._cached("_alpha_s", lambda: (self.sigma12_s + self.sigma21_s) * self.rho_s)

    @property
    def alpha_p(self):
        """amplifier.py:335-341"""
        return self._cached("_alpha_p", lambda: (self.sigma12_p + self.sigma21_p) * self.rho_p)

    @property
    def Pp(self):
        """Pump power (W): advanced by every linear sub-step of a propagation
        (amplifier.py:386-388) and read back from the device after it."""
        return self._Pp

    @property
    def ratio_p(self):
        """amplifier.py:347-353"""
        return self._cached("_ratio_p", lambda: self.sigma12_p / (self.sigma12_p + self.sigma21_p))

    @property
    def ratio_s(self):
        """amplifier.py:355-366 (NaN where both cross sections vanish)."""
        return self._cached("_ratio_s", lambda: self.sigma12_s / (self.sigma12_s + self.sigma21_s))

    def factor(self, A, h):
        raise NotImplementedError("the Yb gain factor is applied inside the HIP engine's linear step")

    def device_arrays(self):
        """Arguments of gnlse_set_yb_gain: ratio_s/Psat_s, 1/Psat_s, alpha_s,
        eta_s in FFT order with the terms the reference drops as NaN
        (amplifier.py:369-373) set to 0, then len(A) dt / Tr, ratio_p,
        Psat_p, N, alpha_p, eta_p."""
        if self.domain is None:
            raise Exception("Domain is not preset.")
        with np.errstate(divide="ignore", invalid="ignore"):
            rs = self.ratio_s / self.Psat_s
            ip = 1.0 / self.Psat_s

        def clean(a):
            a = np.array(a, dtype=np.float64)
            a[~np.isfinite(a)] = 0.0
            return fftshift(a)
        return (clean(rs), clean(ip), clean(self.alpha_s), clean(self.eta_s), float(self.physical_power_factor),
                float(self.ratio_p), float(self.Psat_p), float(self.N), float(self.alpha_p), float(self.eta_p))
