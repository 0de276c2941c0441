"""Sech pulse generator, as pyofss/modules/sech.py:42-198."""

from numpy import pi, exp, cosh
from numpy.lib.scimath import sqrt, log


class SechError(Exception):
    pass


class OutOfRangeError(SechError):
    pass


class NotIntegerError(SechError):
    pass


class Sech:
    """sqrt(P0) sech(tau) exp(i(phi0 - 2 pi nu_off t - C tau^2 / 2)) with
    tau = (t - position*T)/width; width is the HWIeM (sech.py:124-168)."""

    def __init__(self, name="sech", position=0.0, width=10.0, peak_power=1e-3, offset_nu=0.0, m=0, C=0.0,
                 initial_phase=0.0, channel=0, using_fwhm=False):
        if not (-0.5 <= position <= 0.5):
            raise OutOfRangeError("position is out of range. Must be in [-0.5, 0.5]")
        if not (1e-3 < width < 1e3):
            raise OutOfRangeError("width is out of range. Must be in (1e-3, 1e3)")
        if not (0.0 <= peak_power < 1e9):
            raise OutOfRangeError("peak_power is out of range. Must be in [0.0, 1e9)")
        if not (-200.0 < offset_nu < 200.0):
            raise OutOfRangeError("offset_nu is out of range. Must be in (-200.0, 200.0)")
        if not (-1e3 < C < 1e3):
            raise OutOfRangeError("C is out of range. Must be in (-1e3, 1e3)")
        if not (0.0 <= initial_phase < 2.0 * pi):
            raise OutOfRangeError("initial_phase is out of range. Must be in [0.0, 2.0 * pi)")
        if not (0 <= channel < 2):
            raise OutOfRangeError("channel is out of range. Must be in [0, 2)")
        if int(channel) != channel:
            raise NotIntegerError("channel must be an integer")
        self.name = name
        self.position = position
        self.width = width
        self.peak_power = peak_power
        self.offset_nu = offset_nu
        self.C = C
        self.initial_phase = initial_phase
        self.channel = channel
        self.fwhm = None
        if using_fwhm:  # sech.py:111-113
            self.fwhm = width
            self.width *= 0.5 / log(1.0 + sqrt(2.0))
        self.field = None

    def calculate_fwhm(self):
        if self.fwhm is not None:
            return self.fwhm
        return self.width * 2.0 * log(1.0 + sqrt(2.0))

    def __call__(self, domain, field):
        self.field = field
        t_normalised = (domain.t - self.position * domain.window_t) / self.width
        time = t_normalised * t_normalised
        phase = self.initial_phase
        phase -= 2.0 * pi * self.offset_nu * domain.t + 0.5 * self.C * time
        magnitude = sqrt(self.peak_power) / cosh(t_normalised)
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
        time = t_normalised * t_normalised
        phase = self.initial_phase
        phase -= 2.0 * pi * self.offset_nu * t + 0.5 * self.C * time
        return sqrt(self.peak_power) * exp(1j * phase) / cosh(t_normalised)

    def __str__(self):
        lines = ["position = {0:f}", "width = {1:f} ps", "fwhm = {2:f} ps", "peak_power = {3:f} W",
                 "offset_nu = {4:f} THz", "C = {5:f}", "initial_phase = {6:f} rad", "channel = {7:d}"]
        return "\n".join(lines).format(self.position, self.width, self.calculate_fwhm(), self.peak_power,
                                       self.offset_nu, self.C, self.initial_phase, self.channel)
