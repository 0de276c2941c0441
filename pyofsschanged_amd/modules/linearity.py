"""Linear (dispersion and loss) operator, as pyofss/modules/linearity.py:35-272.

``default_linearity`` builds the factor i*sum_n beta_n Domega^n/n! - alpha/2
on the host (set-up, once per domain), as the reference does; it is then held
by the HIP engine, which applies exp(factor*h) in the spectral domain on the
GPU for every call of ``exp_lin``/``lin``.
"""

from math import factorial

import numpy as np
from scipy.constants import constants

from .. import field
from ..field import fftshift


def convert_dispersion_to_physical(D=0.0, S=0.0, Lambda=1550.0):
    """linearity.py:35-55"""
    if (D == 0.0) and (S == 0.0):
        return (0.0, 0.0)
    factor = Lambda ** 2 / (2.0 * np.pi * 1.0e-3 * constants.c)
    return (-factor * D, factor ** 2 * (S + (2.0 * D / Lambda)))


def convert_dispersion_to_engineering(beta_2=0.0, beta_3=0.0, Lambda=1550.0):
    """linearity.py:58-79"""
    if (beta_2 == 0.0) and (beta_3 == 0.0):
        return (0.0, 0.0)
    factor_1 = (2.0 * np.pi * 1.0e-3 * constants.c) / (Lambda ** 2)
    factor_2 = (2.0 * factor_1) / Lambda
    return (-factor_1 * beta_2, factor_1 * factor_1 * beta_3 + factor_2 * beta_2)


def convert_alpha_to_linear(alpha_dB=0.0):
    """linearity.py:82-91"""
    return alpha_dB / (10.0 * np.log10(np.exp(1.0)))


def convert_alpha_to_dB(alpha_linear=0.0):
    """linearity.py:94-103"""
    return alpha_linear * (10.0 * np.log10(np.exp(1.0)))


class Linearity:
    """linearity.py:106-272 (default sim_type)."""

    def __init__(self, alpha=None, beta=None, sim_type=None, use_cache=False, centre_omega=None, phase_lim=False,
                 amplifier=None, use_Er_noise=False):
        if sim_type not in (None, "default", "wdm"):
            raise NotImplementedError(f"sim_type={sim_type!r} is not supported by the HIP engine")
        self.wdm = sim_type == "wdm"
        if self.wdm and amplifier is not None:
            raise NotImplementedError("WDM fields carry no gain model (linearity.py:306-320 ignores the amplifier)")
        if use_Er_noise:
            raise NotImplementedError("Er ASE noise is not supported by the HIP engine")
        self.alpha = alpha
        self.beta = beta
        self.centre_omega = centre_omega
        self.phase_lim = phase_lim  # only the OpenCL path wraps phases; exp() is periodic anyway
        self.amplifier = amplifier
        self.use_Er_noise = use_Er_noise
        self.use_cache = use_cache
        self.generate_linearity = self.wdm_linearity if self.wdm else self.default_linearity
        self.generate_cache = self.default_cache
        self.lin = self.default_f
        self.exp_lin = self.default_exp_f_cached if use_cache else self.default_exp_f
        self.cached_factor = None  # step size whose exp(factor*h) the device caches
        self.factor = None
        self.Domega = None
        self.domain = None
        self._ctx = None

    def __call__(self, domain):
        self.domain = domain
        return self.generate_linearity(domain)

    def default_linearity(self, domain):
        """linearity.py:162-186"""
        if self.amplifier is not None:
            self.amplifier.set_domain(domain)
        if self.beta is None:
            self.factor = 0.0
        else:
            if self.centre_omega is None:
                self.Domega = domain.omega - domain.centre_omega
            else:
                self.Domega = domain.omega - self.centre_omega
            terms = 0.0
            for n, beta in enumerate(self.beta):
                terms += beta * np.power(self.Domega, n) / float(factorial(n))
            self.factor = 1j * fftshift(terms)
        if self.alpha is not None:
            self.factor -= 0.5 * self.alpha
        self.cached_factor = None
        self._upload()
        return self.factor

    def wdm_linearity(self, domain):
        """linearity.py:188-211; the attenuation is subtracted from each
        channel's factor (the reference's in-place update of its tuple
        raises TypeError)."""
        if self.beta is None:
            self.factor = (0.0, 0.0)
        else:
            if self.centre_omega is None:
                self.Domega = (domain.omega - domain.centre_omega, domain.omega - domain.centre_omega)
            else:
                self.Domega = (domain.omega - self.centre_omega[0], domain.omega - self.centre_omega[1])
            terms = [0.0, 0.0]
            for c in range(2):
                for n, beta in enumerate(self.beta[c]):
                    terms[c] += beta * np.power(self.Domega[c], n) / float(factorial(n))
            self.factor = (1j * fftshift(terms[0]), 1j * fftshift(terms[1]))
        if self.alpha is not None:
            self.factor = (self.factor[0] - 0.5 * self.alpha[0], self.factor[1] - 0.5 * self.alpha[1])
        self.cached_factor = None
        self._upload()
        return self.factor

    # -- device binding -----------------------------------------------------------
    def bind(self, ctx):
        """Use engine context ``ctx`` (shared with a Fibre's Nonlinearity)."""
        self._ctx = ctx

    def _context(self):
        if self._ctx is None:
            from .._lib import Context
            self._ctx = Context(0)
        n = self.domain.total_samples
        if self._ctx.n != n:
            self._ctx.set_grid(n)
        return self._ctx

    def _upload(self):
        ctx = self._context()
        n = self.domain.total_samples
        if self.wdm:
            ctx.set_linear_wdm(*[np.broadcast_to(np.asarray(f, dtype=np.complex128), (n,)) for f in self.factor])
            ctx.pump_source = None
            return
        ctx.set_linear(np.broadcast_to(np.asarray(self.factor, dtype=np.complex128), (n,)))
        ctx.pump_source = None
        if self._yb():
            ctx.set_yb_gain(*self.amplifier.device_arrays())
            ctx.pump_source = self.amplifier  # its Pp travels with every launch
        elif self.amplifier is not None:
            ctx.set_gain(True, *self.amplifier.device_arrays())
        else:
            ctx.set_gain(False)

    def _yb(self):
        from .amplifier import Amplifier2LevelModel
        return isinstance(self.amplifier, Amplifier2LevelModel)

    def channels(self):
        return 2 if self.wdm else 1

    def gain_ffts(self):
        """FFTs pyofss counts inside the amplifier per exp_lin call: the Yb
        model's spectral_power (amplifier.py:397)."""
        return 1 if self._yb() else 0

    def default_cache(self, h):
        """linearity.py:220-224: fix the step size of the cached factor."""
        self.cached_factor = h

    def cache(self, h):
        self.generate_cache(h)

    def default_f(self, A, z):
        """linearity.py:238-239 (wdm: :303-304): ifft(factor * fft(A)) on the GPU."""
        field.fft_counter += 2 * self.channels()
        return self._context().apply_linear_factor(A)

    def default_exp_f(self, A, h):
        """linearity.py:241-256: ifft(exp(factor*h) [* exp(gain)] * fft(A)) on the GPU."""
        field.fft_counter += 2 * self.channels() + self.gain_ffts()
        return self._context().apply_linear(A, h)

    def default_exp_f_cached(self, A, h):
        """linearity.py:258-272: as default_exp_f with the step of the first call."""
        if self.cached_factor is None:
            self.cache(h)
        field.fft_counter += 2 * self.channels() + self.gain_ffts()
        return self._context().apply_linear(A, self.cached_factor)
