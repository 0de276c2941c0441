"""Nonlinear operator, as pyofss/modules/nonlinearity.py:29-270.

Set-up (the relative angular frequency grid, the self-steepening factor and
the Raman response spectrum) happens on the host once per domain, as in the
reference; the Raman response is transformed with the device FFT. Every
application of ``non`` / ``exp_non`` runs in the HIP engine.
"""

from math import pi

import numpy as np
from scipy.constants import constants

from .. import _lib, field
from ..domain import Domain
from ..field import fft, fftshift, simpson_weights


class NonlinearityError(Exception):
    pass


class UnknownRamanResponseError(NonlinearityError):
    pass


def calculate_gamma(nonlinear_index, effective_area, centre_omega=2.0 * pi * 193.1):
    """nonlinearity.py:37-50"""
    return 1.0e24 * nonlinear_index * centre_omega / (effective_area * Domain.vacuum_light_speed)


def calculate_raman_term(domain, tau_1=12.2e-3, tau_2=32.0e-3):
    """nonlinearity.py:53-69: Blow-Wood response on t - t_min."""
    t = domain.t - domain.t.min()
    h_R = (tau_2 ** 2 + tau_1 ** 2) / (tau_1 * tau_2 ** 2)
    h_R *= np.exp(-t / tau_2) * np.sin(t / tau_1)
    return h_R


def calculate_raman_term_silica(domain):
    """nonlinearity.py:72-100: Hollenbeck-Cantrell response normalised by the
    old Simpson rule over domain.t."""
    CP = [56.25, 100.00, 231.25, 362.50, 463.00, 497.00, 611.50, 691.67, 793.67, 835.50, 930.00, 1080.00, 1215.00]
    Ai = [1.00, 11.40, 36.67, 67.67, 74.00, 4.50, 6.80, 4.60, 4.20, 4.50, 2.70, 3.10, 3.00]
    Gfwhm = [52.10, 110.42, 175.00, 162.50, 135.33, 24.50, 41.50, 155.00, 59.50, 64.30, 150.00, 91.00, 160.00]
    Lfwhm = [17.37, 38.81, 58.33, 54.17, 45.11, 8.17, 13.83, 51.67, 19.83, 21.43, 50.00, 30.33, 53.33]
    t = domain.t - domain.t.min()
    h_R = np.zeros_like(t)
    for i in range(13):
        h_R += (Ai[i] * np.exp(-np.pi * 1e-10 * constants.c * Lfwhm[i] * t)
                * np.exp(-((np.pi * 1e-10 * constants.c * Gfwhm[i]) ** 2) * (t ** 2) / 4.0)
                * np.sin(2 * np.pi * 1e-10 * constants.c * CP[i] * t))
    norm = float(np.dot(simpson_weights(domain.t), h_R))
    return h_R / norm


# FFTs per call of non / exp_non, as counted by pyofss's fft_counter
_NL_FFTS = {_lib.NL_KERR: 0, _lib.NL_ALL: 4, _lib.NL_SS: 2, _lib.NL_RS: 2, _lib.NL_SS_RS: 4}


class Nonlinearity:
    """nonlinearity.py:103-304 (default and wdm sim_type)."""

    def __init__(self, gamma=None, sim_type=None, self_steepening=False, raman_scattering=False, rs_factor=3e-3,
                 use_all=False, tau_1=12.2e-3, tau_2=32.0e-3, f_R=0.18):
        if sim_type not in (None, "default", "wdm"):
            raise NotImplementedError(f"sim_type={sim_type!r} is not supported by the HIP engine")
        self.wdm = sim_type == "wdm"
        if self.wdm and (use_all or self_steepening or raman_scattering):
            # nonlinearity.py:272-285: the reference's WDM variants with these
            # flags return None (use_all falls back to the single-field operator)
            raise NotImplementedError("WDM fields support the Kerr and cross-phase terms only")
        self.gamma = gamma
        self.self_steepening = self_steepening
        self.raman_scattering = raman_scattering
        self.rs_factor = rs_factor
        self.use_all = use_all
        self.tau_1 = tau_1
        self.tau_2 = tau_2
        self.f_R = f_R
        self.generate_nonlinearity = self.wdm_nonlinearity if self.wdm else self.default_nonlinearity
        # nonlinearity.py:133-152 flag dispatch
        if use_all:
            print("Using general expression for nonlinearity")
            self.kind = _lib.NL_ALL
        elif self_steepening and raman_scattering:
            print("Using self_steepening and raman_scattering")
            self.kind = _lib.NL_SS_RS
        elif self_steepening:
            print("Using self_steepening")
            self.kind = _lib.NL_SS
        elif raman_scattering:
            print("Using raman_scattering")
            self.kind = _lib.NL_RS
        else:
            self.kind = _lib.NL_KERR
        self.non = self.default_f
        self.exp_non = self.default_exp_f
        self.ss_factor = None
        self.h_R = None
        self.omega = None
        self.centre_omega = None
        self.factor = None
        self.domain = None
        self._ctx = None

    def __call__(self, domain):
        """nonlinearity.py:160-185"""
        self.domain = domain
        self.centre_omega = domain.centre_omega
        self.omega = fftshift(domain.omega - domain.centre_omega)
        if self.self_steepening is False:
            self.ss_factor = 0.0
        elif self.self_steepening is True:
            self.ss_factor = 1.0 / self.centre_omega
        else:
            self.ss_factor = float(self.self_steepening)
        if self.use_all is False:
            self.h_R = 0.0
        elif self.use_all is True or self.use_all.lower() == "blowwood":
            self.h_R = fft(calculate_raman_term(domain, self.tau_1, self.tau_2))
        elif self.use_all.lower() == "hollenbeck":
            self.h_R = fft(calculate_raman_term_silica(domain))
        else:
            raise UnknownRamanResponseError('Use False, True or a sting from the list ["blowwood", "hollenbeck"]')
        self.h_R *= domain.window_t
        self.generate_nonlinearity()
        self._upload()

    def default_nonlinearity(self):
        """nonlinearity.py:187-192"""
        self.factor = 0.0 if self.gamma is None else 1j * self.gamma

    def wdm_nonlinearity(self):
        """nonlinearity.py:194-200"""
        self.factor = (0.0, 0.0) if self.gamma is None else (1j * self.gamma[0], 1j * self.gamma[1])

    # -- device binding -----------------------------------------------------------
    def bind(self, ctx):
        self._ctx = ctx

    def _context(self):
        if self._ctx is None:
            self._ctx = _lib.Context(0)
        n = self.domain.total_samples
        if self._ctx.n != n:
            self._ctx.set_grid(n)
        return self._ctx

    def _upload(self):
        ctx = self._context()
        if self.wdm:  # i gamma_c (|A_c|^2 + 2|A_o|^2) A_c, nonlinearity.py:287-304
            g = (0.0, 0.0) if self.gamma is None else (float(self.gamma[0]), float(self.gamma[1]))
            ctx.set_nonlinear_wdm(g[0], g[1])
            return
        gamma = 0.0 if self.gamma is None else float(self.gamma)
        nl_mult = h_R = None
        if self.kind == _lib.NL_ALL:
            nl_mult = self.factor * (1.0 + self.omega * self.ss_factor)  # the multiplier of :207
            h_R = self.h_R
        ctx.set_nonlinear(self.kind, gamma, self.f_R, self.ss_factor, self.rs_factor, self.omega, nl_mult, h_R)

    def default_f(self, A, z):
        """Nonlinearity.non on the GPU (nonlinearity.py:201-207, 223-228, 236-241, 250-256, 266-267)."""
        field.fft_counter += _NL_FFTS[self.kind]
        return self._context().apply_nonlinear(A)

    def default_exp_f(self, A, h, B):
        """Nonlinearity.exp_non on the GPU (nonlinearity.py:209-221, 230-234, 243-248, 258-264, 269-270)."""
        field.fft_counter += _NL_FFTS[self.kind]
        return self._context().apply_exp_nonlinear(A, h, B)
