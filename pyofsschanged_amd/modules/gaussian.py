"""(Super-)Gaussian pulse generator, as pyofss/modules/gaussian.py:42-206."""

from numpy import pi, exp, power
from numpy.lib.scimath import sqrt, log


class GaussianError(Exception):
    pass


class OutOfRangeError(GaussianError):
    pass


class NotIntegerError(GaussianError):
    pass


class Gaussian:
    """sqrt(P0) exp(-tau^(2m)/2) exp(i(phi0 - 2 pi nu_off t - C tau^(2m)/2))
    with tau = (t - position*T)/width; width is the HWIeM (gaussian.py:131-175)."""

    def __init__(self, name="gaussian", position=0.0, width=10.0, peak_power=1e-3, offset_nu=0.0, m=1, C=0.0,
                 initial_phase=0.0, channel=0, using_fwhm=False):
        if not (-0.5 <= position <= 0.5):
            raise OutOfRangeError("position is out of range. Must be in [-0.5, 0.5]")
        if not (1e-3 < width < 1e3):
            raise OutOfRangeError("width is out of range. Must be in (1e-3, 1e3)")
        if not (0.0 <= peak_power < 1e9):
            raise OutOfRangeError("peak_power is out of range. Must be in [0.0, 1e9)")
        if not (-300.0 < offset_nu < 300.0):
            raise OutOfRangeError("offset_nu is out of range. Must be in (-300.0, 300.0)")
        if not (0 < m < 50):
            raise OutOfRangeError("m is out of range. Must be in (0, 50)")
        if not (-1e3 < C < 1e3):
            raise OutOfRangeError("C is out of range. Must be in (-1e3, 1e3)")
        if not (0.0 <= initial_phase < 2.0 * pi):
            raise OutOfRangeError("initial_phase is out of range. Must be in [0.0, 2.0 * pi)")
        if not (0 <= channel < 2):
            raise OutOfRangeError("channel is out of range. Must be in [0, 2)")
        if int(m) != m:
            raise NotIntegerError("m must be an integer")
        if int(channel) != channel:
            raise NotIntegerError("channel must be an integer")
        self.name = name
        self.position = position
        self.width = width
        self.peak_power = peak_power
        self.offset_nu = offset_nu
        self.m = m
        self.C = C
        self.initial_phase = initial_phase
        self.channel = channel
        self.fwhm = None
        if using_fwhm:  # gaussian.py:118-120
            self.fwhm = width
            self.width *= 0.5 / power(log(2.0), 1.0 / (2 * m))
        self.field = None

    def calculate_fwhm(self):
        if self.fwhm is not None:
            return self.fwhm
        return self.width * 2.0 * power(log(2.0), 1.0 / (2 * self.m))

    def __call__(self, domain, field):
        self.field = field
        t_normalised = (domain.t - self.position * domain.window_t) / self.width
        time = power(t_normalised, (2 * self.m))
        phase = self.initial_phase
        phase -= 2.0 * pi * self.offset_nu * domain.t + 0.5 * self.C * time
        magnitude = sqrt(self.peak_power) * exp(-0.5 * time)
        if domain.channels > 1:
            self.field[self.channel] += magnitude * exp(1j * phase)
        else:
            self.field += magnitude * exp(1j * phase)
        return self.field

    def generate(self, t):
        if len(t) < 8:
            raise OutOfRangeError("Require temporal array with at least 8 values")
        t_range = t[-1] - t[0] + (t[1] - t[0])
        t_normalised = (t - self.position * t_range) / self.width
        time = power(t_normalised, (2 * self.m))
        phase = self.initial_phase
        phase -= 2.0 * pi * self.offset_nu * t + 0.5 * self.C * time
        return sqrt(self.peak_power) * exp(-0.5 * time + 1j * phase)

    def __str__(self):
        lines = ["position = {0:f}", "width = {1:f} ps", "fwhm = {2:f} ps", "peak_power = {3:f} W",
                 "offset_nu = {4:f} THz", "m = {5:d}", "C = {6:f}", "initial_phase = {7:f} rad", "channel = {8:d}"]
        return "\n".join(lines).format(self.position, self.width, self.calculate_fwhm(), self.peak_power,
                                       self.offset_nu, self.m, self.C, self.initial_phase, self.channel)
