"""Simulation grids and unit conversions, as pyofss/domain.py:39-299.

Units follow pyofss (pyofss/Unit.txt): t in ps, nu in THz, omega in rad/ps,
lambda in nm. The temporal grid is centred, t_n = -T/2 + n T/N
(domain.py:218); the frequency grid is consecutive, centred on centre_nu
(domain.py:223-228).
"""

from math import pi

import numpy as np


class DomainError(Exception):
    pass


class OutOfRangeError(DomainError):
    pass


class NotIntegerError(DomainError):
    pass


def nu_to_omega(nu):
    """domain.py:39-47"""
    return 2.0 * pi * nu


def nu_to_lambda(nu):
    """domain.py:50-58"""
    return Domain.vacuum_light_speed / nu


def omega_to_nu(omega):
    """domain.py:61-69"""
    return omega / (2.0 * pi)


def omega_to_lambda(omega):
    """domain.py:72-80"""
    return 2.0 * pi * Domain.vacuum_light_speed / omega


def lambda_to_nu(Lambda):
    """domain.py:83-91"""
    return Domain.vacuum_light_speed / Lambda


def lambda_to_omega(Lambda):
    """domain.py:94-102"""
    return 2.0 * pi * Domain.vacuum_light_speed / Lambda


def dnu_to_dlambda(dnu, nu=193.1):
    """domain.py:105-115"""
    return Domain.vacuum_light_speed * dnu / (nu ** 2)


def dlambda_to_dnu(dlambda, Lambda=1550.0):
    """domain.py:118-128"""
    return Domain.vacuum_light_speed * dlambda / (Lambda ** 2)


def align_with_nu_grid(nu, domain):
    """domain.py:131-139"""
    return int(nu / domain.dnu) * domain.dnu


class Domain:
    """Temporal and spectral grids (domain.py:142-299)."""

    vacuum_light_speed = 1.0e-3 * 299792458.0  # nm / ps (scipy.constants.c is exact)

    def __init__(self, total_bits=1, samples_per_bit=512, bit_width=100.0, centre_nu=193.1, channels=1):
        if not (0 < total_bits < 4096):
            raise OutOfRangeError("total_bits is out of range. Must be in (0, 4096)")
        if not (0 < samples_per_bit < 262144):
            raise OutOfRangeError("samples_per_bit is out of range. Must be in (0, 262144)")
        if not (0.01 < bit_width < 10000.0):
            raise OutOfRangeError("bit_width is out of range. Must be in (0.01, 10000.0)")
        if not (140.0 < centre_nu < 600.0):
            raise OutOfRangeError("centre_nu is out of range. Must be in (140.0, 600.0)")
        if not (0 < channels < 3):
            raise OutOfRangeError("channels is out of range. Must be in (0, 3)")
        if int(total_bits) != total_bits:
            raise NotIntegerError("total_bits must be an integer")
        if int(samples_per_bit) != samples_per_bit:
            raise NotIntegerError("samples_per_bit must be an integer")
        if int(channels) != channels:
            raise NotIntegerError("channels must be an integer")

        self.total_bits = total_bits
        self.samples_per_bit = samples_per_bit
        self.bit_width = bit_width
        self.centre_nu = centre_nu
        self.total_samples = self.total_bits * self.samples_per_bit
        self.window_t = self.total_bits * self.bit_width
        self.centre_omega = 2.0 * pi * self.centre_nu
        self.centre_lambda = Domain.vacuum_light_speed / self.centre_nu
        self.t, self.dt = np.linspace(-self.window_t / 2.0, self.window_t / 2.0, self.total_samples, False, True)
        self.window_nu = 1.0 / self.dt
        nu_min = self.centre_nu - 0.5 * self.window_nu
        self.nu, self.dnu = np.linspace(nu_min, nu_min + self.window_nu, self.total_samples, False, True)
        self.omega = 2.0 * pi * self.nu
        self.domega = 2.0 * pi * self.dnu
        self.window_omega = 2.0 * pi * self.window_nu
        self.Lambda = Domain.vacuum_light_speed / self.nu
        self.dlambda = Domain.vacuum_light_speed * self.dnu / (self.centre_nu ** 2)
        self.window_lambda = Domain.vacuum_light_speed * self.window_nu / (self.centre_nu ** 2)
        self.channels = channels

    def __eq__(self, obj):
        return (isinstance(obj, Domain) and obj.total_bits == self.total_bits
                and obj.samples_per_bit == self.samples_per_bit and obj.bit_width == self.bit_width
                and obj.centre_nu == self.centre_nu and obj.channels == self.channels)

    def __ne__(self, obj):
        return not self.__eq__(obj)

    __hash__ = object.__hash__

    def __str__(self):
        lines = ["Domain:", "total_bits = {0:d}", "samples_per_bit = {1:d}", "bit_width = {2:.4f} ps",
                 "centre_nu = {3:.4f} THz", "total_samples = {4:d}", "window_t = {5:.4f} ps",
                 "centre_omega = {6:.4f} rad / ps", "centre_lambda = {7:.4f} nm", "dt = {8:.4f} ps",
                 "window_nu = {9:.4f} THz", "dnu = {10:.4f} THz", "window_omega = {11:.4f} rad / ps",
                 "domega = {12:.4f} rad / ps", "window_lambda = {13:.4f} nm", "dlambda = {14:.4f} nm",
                 "channels = {15:d}"]
        return "\n\t".join(lines).format(
            self.total_bits, self.samples_per_bit, self.bit_width, self.centre_nu, self.total_samples,
            self.window_t, self.centre_omega, self.centre_lambda, self.dt, self.window_nu, self.dnu,
            self.window_omega, self.domega, self.window_lambda, self.dlambda, self.channels)
