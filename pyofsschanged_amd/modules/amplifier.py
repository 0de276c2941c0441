"""Saturable gain, as the simple model of pyofss/modules/amplifier.py:42-175.

The gain enters every linear sub-step (linearity.py:248-272): the device
multiplies the spectrum by exp(G0/(1 + E/E_sat) * profile) with G0 =
ln(10^(0.1 G)) h / (2 L) and E the energy of the array being propagated.
This class only prepares the constant arrays; the per-step arithmetic runs in
the HIP engine.
"""

import warnings

import numpy as np

from ..domain import Domain, lambda_to_nu
from ..field import fftshift, simpson_weights


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
