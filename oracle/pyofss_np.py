"""NumPy restatement of pyofss's fibre stepping path (TEST INFRASTRUCTURE ONLY).

Restates, in complex128, the arithmetic of these reference files
(KarinaSaraeva/pyofssChanged):

  pyofss/field.py:43-172         fft/ifft convention, shifts, powers, energy
  pyofss/domain.py:39-243        grids and unit conversions
  pyofss/modules/sech.py:62-168, gaussian.py:62-175   input pulses
  pyofss/modules/linearity.py:35-320                  linear operator (incl. WDM)
  pyofss/modules/nonlinearity.py:37-304               nonlinear operator (incl. WDM)
  pyofss/modules/amplifier.py:42-175                  saturable gain
  pyofss/modules/amplifier.py:178-402                 two-level Yb gain (CPU path)
  pyofss/modules/solver.py:24-354                     integrators
  pyofss/modules/stepper.py:33-376                    fixed and adaptive steppers
  pyofss/modules/storage.py:130-192                   trace capture
  pyofss/modules/fibre.py:35-187                      operator wiring

Differences from the reference, all deliberate:
  * scipy.fftpack is replaced by numpy.fft (both pocketfft in the installed
    SciPy 1.15); the sign/scale convention of field.py:109-136 is kept.
  * ``integrate.simps`` no longer exists in SciPy >= 1.14; ``simps_avg``
    restates the old default rule (even='avg').
  * the adaptive stepper returns A_out at the end instead of raising the
    AttributeError of stepper.py:365-371 (``storage.dir_spec`` does not exist).
  * Storage keeps (z, A) and the FFT count of update_characts; the other trace
    characteristics are computed only when ``characts=True``.
  * WDM: Fibre takes no beta_2 from per-channel beta lists (fibre.py:88 fails
    there) and wdm_linearity subtracts the attenuation per channel (the
    reference's item assignment into a tuple fails).

Pins (tests/test_oracle_*.py):
  solver + standard stepper  -> bit-exact scalar-ODE values, test_solver.py:83-187
  rk4ip / ss_* on a soliton   -> peak-power pins, test_solver.py:230-284, and the
                                 closed-form N=2 soliton, soliton.py:30-39
  fft / ifft                  -> O(N^2) long-double DFT, both signs and scalings
  Domain, Sech, Gaussian      -> golden strings and constants of test_domain.py,
                                 test_sech.py, test_gaussian.py
  linear operator             -> closed-form Gaussian dispersion
  Kerr / exp Kerr             -> closed-form SPM phase
  Raman convolution           -> brute-force O(N^2) causal convolution, unit area
  self-steepening factor      -> pure tone eigenvalue
  full nonlinearity           -> reduces to Kerr / Kerr+SS at f_R=0; photon number
  adaptive stepper            -> attempt-by-attempt fixtures written by the
                                 reference's own stepper.py (scalar ODE bit for
                                 bit; Kerr, Raman + self-steepening and gain
                                 fibres), planted controller mistakes failing
                                 them, fifth-order convergence of the local
                                 extrapolation (tests/test_oracle_adaptive.py)
  amplifier                   -> small-signal gain closed form; flat saturated
                                 gain converging to the Lambert-W solution of
                                 dE/dz = g0 E / (1 + E/E_sat)
  two-level Yb gain           -> the df_results values printed by
                                 examples/notebooks/nonlinear_attractor.ipynb
                                 (first five traces, 1e-11; tests/test_yb_oracle.py),
                                 pump-only inversion closed form
  Hollenbeck-Cantrell Raman   -> the same notebook run (its use_all='hollenbeck')
  WDM (two channels, XPM)     -> closed-form cross-phase rotation without dispersion,
                                 closed-form per-channel dispersion without gamma,
                                 per-channel energy conservation (tests/test_wdm_oracle.py)
  simps_avg                   -> exact on cubics, reduces to scipy.simpson on odd N
  OPPM noise, coherence       -> Parseval, |g12| closed forms (build-only parts)
"""

import math
import os

import numpy as np

__all__ = [
    "Domain", "Sech", "Gaussian", "Linearity", "Nonlinearity", "Amplifier", "Amplifier2LevelModel",
    "Solver", "Stepper", "Storage", "Fibre", "System",
    "fft", "ifft", "fftshift", "ifftshift", "temporal_power", "spectral_power",
    "energy", "simps_avg", "get_fft_counter", "reset_fft_counter",
    "calculate_raman_term", "calculate_raman_term_silica", "calculate_gamma",
    "lambda_to_nu", "nu_to_lambda", "nu_to_omega", "omega_to_nu",
    "lambda_to_omega", "omega_to_lambda", "dnu_to_dlambda", "dlambda_to_dnu",
    "SmallStepSizeError", "SuitableStepSizeError", "MaximumStepsAllocatedError",
    "NoDerivativeFunctionError", "UnknownRamanResponseError",
    "OutOfRangeError", "NotIntegerError", "oppm_noise", "coherence_g12",
    "VACUUM_LIGHT_SPEED",
]

# scipy.constants.c is exact (299792458 m/s); domain.py:169 uses 1e-3 * c nm/ps
VACUUM_LIGHT_SPEED = 1.0e-3 * 299792458.0


# --------------------------------------------------------------------------
# field.py
# --------------------------------------------------------------------------
_fft_counter = 0


def get_fft_counter():
    return _fft_counter


def reset_fft_counter():
    """storage.py:174-177"""
    global _fft_counter
    _fft_counter = 0


def fft(A_t):
    """field.py:109-121: physics convention, scipy.fftpack.ifft (sign +, 1/N)."""
    global _fft_counter
    _fft_counter += 1
    return np.fft.ifft(A_t)


def ifft(A_nu):
    """field.py:124-136: scipy.fftpack.fft (sign -, no scaling)."""
    global _fft_counter
    _fft_counter += 1
    return np.fft.fft(A_nu)


def ifftshift(A_nu):
    """field.py:139-147 (named swapped relative to scipy)."""
    return np.fft.fftshift(A_nu)


def fftshift(A_nu):
    """field.py:150-158 (named swapped relative to scipy)."""
    return np.fft.ifftshift(A_nu)


def temporal_power(A_t, normalise=False):
    """field.py:43-57"""
    P = np.abs(A_t) ** 2
    if normalise:
        P /= max(P)
    return P


def spectral_power(A_t, normalise=False):
    """field.py:60-75"""
    P = np.abs(fft(A_t)) ** 2
    if normalise:
        P /= max(P)
    return ifftshift(P)


def _basic_simps(y, start, stop, x):
    """Old scipy.integrate._basic_simps for sampled x (SciPy < 1.11)."""
    h = np.diff(x)
    sl0 = slice(start, stop, 2)
    sl1 = slice(start + 1, stop + 1, 2)
    sl2 = slice(start + 2, stop + 2, 2)
    h0 = h[sl0]
    h1 = h[sl1]
    hsum = h0 + h1
    hprod = h0 * h1
    h0divh1 = h0 / h1
    tmp = hsum / 6.0 * (y[sl0] * (2 - 1.0 / h0divh1) + y[sl1] * hsum * hsum / hprod + y[sl2] * (2 - h0divh1))
    return np.sum(tmp)


def simps_avg(y, x):
    """The integrate.simps(y, x) default (even='avg') used at field.py:170 and
    nonlinearity.py:99, restated because SciPy >= 1.14 removed it."""
    y = np.asarray(y)
    x = np.asarray(x)
    N = y.shape[-1]
    if N % 2 == 0:
        val = 0.0
        last_dx = x[-1] - x[-2]
        val += 0.5 * last_dx * (y[-1] + y[-2])
        result = _basic_simps(y, 0, N - 3, x)
        first_dx = x[1] - x[0]
        val += 0.5 * first_dx * (y[1] + y[0])
        result += _basic_simps(y, 1, N - 2, x)
        val /= 2.0
        result /= 2.0
        return result + val
    return _basic_simps(y, 0, N - 2, x)


def energy(A_t, t):
    """field.py:161-172 (nJ)."""
    return simps_avg(temporal_power(A_t), t * 1e-3)


# --------------------------------------------------------------------------
# domain.py
# --------------------------------------------------------------------------
class DomainError(Exception):
    pass


class OutOfRangeError(DomainError):
    pass


class NotIntegerError(DomainError):
    pass


def nu_to_omega(nu):
    """domain.py:39-47"""
    return 2.0 * math.pi * nu


def nu_to_lambda(nu):
    """domain.py:50-58"""
    return VACUUM_LIGHT_SPEED / nu


def omega_to_nu(omega):
    """domain.py:61-69"""
    return omega / (2.0 * math.pi)


def omega_to_lambda(omega):
    """domain.py:72-80"""
    return 2.0 * math.pi * VACUUM_LIGHT_SPEED / omega


def lambda_to_nu(Lambda):
    """domain.py:83-91"""
    return VACUUM_LIGHT_SPEED / Lambda


def lambda_to_omega(Lambda):
    """domain.py:94-102"""
    return 2.0 * math.pi * VACUUM_LIGHT_SPEED / Lambda


def dnu_to_dlambda(dnu, nu=193.1):
    """domain.py:105-115"""
    return VACUUM_LIGHT_SPEED * dnu / (nu ** 2)


def dlambda_to_dnu(dlambda, Lambda=1550.0):
    """domain.py:118-128"""
    return VACUUM_LIGHT_SPEED * dlambda / (Lambda ** 2)


class Domain:
    """domain.py:142-299"""

    vacuum_light_speed = VACUUM_LIGHT_SPEED

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
        self.centre_omega = 2.0 * math.pi * self.centre_nu
        self.centre_lambda = Domain.vacuum_light_speed / self.centre_nu
        (self.t, self.dt) = np.linspace(-self.window_t / 2.0, self.window_t / 2.0, self.total_samples, False, True)
        self.window_nu = 1.0 / self.dt
        _nu_min = self.centre_nu - 0.5 * self.window_nu
        (self.nu, self.dnu) = np.linspace(_nu_min, _nu_min + self.window_nu, self.total_samples, False, True)
        self.omega = 2.0 * math.pi * self.nu
        self.domega = 2.0 * math.pi * self.dnu
        self.window_omega = 2.0 * math.pi * self.window_nu
        self.Lambda = Domain.vacuum_light_speed / self.nu
        self.dlambda = Domain.vacuum_light_speed * self.dnu / (self.centre_nu ** 2)
        self.window_lambda = Domain.vacuum_light_speed * self.window_nu / (self.centre_nu ** 2)
        self.channels = channels

    def __eq__(self, obj):
        return (isinstance(obj, Domain) and obj.total_bits == self.total_bits
                and obj.samples_per_bit == self.samples_per_bit and obj.bit_width == self.bit_width
                and obj.centre_nu == self.centre_nu and obj.channels == self.channels)

    def __str__(self):
        s = ["Domain:", "total_bits = {0:d}", "samples_per_bit = {1:d}", "bit_width = {2:.4f} ps",
             "centre_nu = {3:.4f} THz", "total_samples = {4:d}", "window_t = {5:.4f} ps",
             "centre_omega = {6:.4f} rad / ps", "centre_lambda = {7:.4f} nm", "dt = {8:.4f} ps",
             "window_nu = {9:.4f} THz", "dnu = {10:.4f} THz", "window_omega = {11:.4f} rad / ps",
             "domega = {12:.4f} rad / ps", "window_lambda = {13:.4f} nm", "dlambda = {14:.4f} nm",
             "channels = {15:d}"]
        return "\n\t".join(s).format(
            self.total_bits, self.samples_per_bit, self.bit_width, self.centre_nu, self.total_samples,
            self.window_t, self.centre_omega, self.centre_lambda, self.dt, self.window_nu, self.dnu,
            self.window_omega, self.domega, self.window_lambda, self.dlambda, self.channels)


# --------------------------------------------------------------------------
# sech.py, gaussian.py
# --------------------------------------------------------------------------
class Sech:
    """sech.py:42-168"""

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
        if not (0.0 <= initial_phase < 2.0 * math.pi):
            raise OutOfRangeError("initial_phase is out of range. Must be in [0.0, 2.0 * pi)")
        if not (0 <= channel < 2):
            raise OutOfRangeError("channel is out of range. Must be in [0, 2)")
        if int(channel) != channel:
            raise NotIntegerError("channel must be an integer")
        self.name, self.position, self.width, self.peak_power = name, position, width, peak_power
        self.offset_nu, self.C, self.initial_phase, self.channel = offset_nu, C, initial_phase, channel
        self.fwhm = None
        if using_fwhm:
            self.fwhm = width
            self.width *= 0.5 / np.log(1.0 + np.sqrt(2.0))
        self.field = None

    def calculate_fwhm(self):
        if self.fwhm is not None:
            return self.fwhm
        return self.width * 2.0 * np.log(1.0 + np.sqrt(2.0))

    def __call__(self, domain, field):
        self.field = field
        t_normalised = (domain.t - self.position * domain.window_t) / self.width
        time = t_normalised * t_normalised
        phase = self.initial_phase
        phase -= 2.0 * np.pi * self.offset_nu * domain.t + 0.5 * self.C * time
        magnitude = np.sqrt(self.peak_power) / np.cosh(t_normalised)
        self.field += magnitude * np.exp(1j * phase)
        return self.field

    def generate(self, t):
        if len(t) < 8:
            raise OutOfRangeError("Require temporal array with at least 8 values")
        t_range = t[-1] - t[0] + (t[1] - t[0])
        t_normalised = (t - self.position * t_range) / self.width
        time = t_normalised * t_normalised
        phase = self.initial_phase
        phase -= 2.0 * np.pi * self.offset_nu * t + 0.5 * self.C * time
        return np.sqrt(self.peak_power) * np.exp(1j * phase) / np.cosh(t_normalised)

    def __str__(self):
        s = ["position = {0:f}", "width = {1:f} ps", "fwhm = {2:f} ps", "peak_power = {3:f} W",
             "offset_nu = {4:f} THz", "C = {5:f}", "initial_phase = {6:f} rad", "channel = {7:d}"]
        return "\n".join(s).format(self.position, self.width, self.calculate_fwhm(), self.peak_power,
                                   self.offset_nu, self.C, self.initial_phase, self.channel)


class Gaussian:
    """gaussian.py:42-175"""

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
        if not (0.0 <= initial_phase < 2.0 * math.pi):
            raise OutOfRangeError("initial_phase is out of range. Must be in [0.0, 2.0 * pi)")
        if not (0 <= channel < 2):
            raise OutOfRangeError("channel is out of range. Must be in [0, 2)")
        if int(m) != m:
            raise NotIntegerError("m must be an integer")
        if int(channel) != channel:
            raise NotIntegerError("channel must be an integer")
        self.name, self.position, self.width, self.peak_power = name, position, width, peak_power
        self.offset_nu, self.m, self.C, self.initial_phase, self.channel = offset_nu, m, C, initial_phase, channel
        self.fwhm = None
        if using_fwhm:
            self.fwhm = width
            self.width *= 0.5 / np.power(np.log(2.0), 1.0 / (2 * m))
        self.field = None

    def calculate_fwhm(self):
        if self.fwhm is not None:
            return self.fwhm
        return self.width * 2.0 * np.power(np.log(2.0), 1.0 / (2 * self.m))

    def __call__(self, domain, field):
        self.field = field
        t_normalised = (domain.t - self.position * domain.window_t) / self.width
        time = np.power(t_normalised, (2 * self.m))
        phase = self.initial_phase
        phase -= 2.0 * np.pi * self.offset_nu * domain.t + 0.5 * self.C * time
        magnitude = np.sqrt(self.peak_power) * np.exp(-0.5 * time)
        self.field += magnitude * np.exp(1j * phase)
        return self.field

    def generate(self, t):
        if len(t) < 8:
            raise OutOfRangeError("Require temporal array with at least 8 values")
        t_range = t[-1] - t[0] + (t[1] - t[0])
        t_normalised = (t - self.position * t_range) / self.width
        time = np.power(t_normalised, (2 * self.m))
        phase = self.initial_phase
        phase -= 2.0 * np.pi * self.offset_nu * t + 0.5 * self.C * time
        return np.sqrt(self.peak_power) * np.exp(-0.5 * time + 1j * phase)

    def __str__(self):
        s = ["position = {0:f}", "width = {1:f} ps", "fwhm = {2:f} ps", "peak_power = {3:f} W",
             "offset_nu = {4:f} THz", "m = {5:d}", "C = {6:f}", "initial_phase = {7:f} rad", "channel = {8:d}"]
        return "\n".join(s).format(self.position, self.width, self.calculate_fwhm(), self.peak_power,
                                   self.offset_nu, self.m, self.C, self.initial_phase, self.channel)


# --------------------------------------------------------------------------
# amplifier.py (simple saturable model only)
# --------------------------------------------------------------------------
class Amplifier:
    """amplifier.py:42-175 without the OpenCL helpers and the Er profile."""

    def __init__(self, name="simple_saturation", gain=None, E_sat=None, P_sat=None, Tr=None, length=1.0,
                 lamb0=None, bandwidth=None):
        self.length = length
        if (P_sat is not None) and (E_sat is not None):
            raise Exception("There should be only one parameter of saturation.")
        if bandwidth is not None and lamb0 is not None:
            self.delta = (VACUUM_LIGHT_SPEED * bandwidth) / (lamb0 * lamb0)
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
        self.domain = None

    def set_domain(self, domain):
        """amplifier.py:171-175"""
        self.domain = domain
        self.shape = domain.nu.shape
        self._profile = None
        self._gain_log = None

    @property
    def spectal_filtration_array(self):
        """amplifier.py:109-129 (Lorentzian or flat)."""
        if self._profile is None:
            if self.delta is not None and self.lamb is not None:
                f = 1 / (1 + 4 * (((self.domain.nu - lambda_to_nu(self.lamb)) ** 2) / ((self.delta) ** 2)))
            else:
                f = np.ones(len(self.domain.Lambda))
            self._profile = fftshift(f)
        return self._profile

    @property
    def gain_times_log(self):
        """amplifier.py:131-147"""
        if self._gain_log is None:
            g = np.ones(self.shape) * np.power(10, 0.1 * self.gain)
            self._gain_log = np.log(g)
        return self._gain_log

    def factor(self, A, h):
        """amplifier.py:149-158: saturation uses the energy of the array passed."""
        G0 = (self.gain_times_log * h) / (2 * self.length)
        E = energy(A, self.domain.t)
        factor = G0 / (1.0 + E / self.E_sat)
        factor *= self.spectal_filtration_array
        return factor


# --------------------------------------------------------------------------
# amplifier.py: two-level Yb rate-equation gain
# --------------------------------------------------------------------------
YB_CROSS_SECTIONS = os.path.join(os.path.dirname(__file__), "..", "pyofsschanged_amd", "data", "CrossSectionDataExp.txt")


def yb_cross_section_table(path=YB_CROSS_SECTIONS):
    """The table read by amplifier.py:247-258 with pandas.read_csv(..., ",")
    and its default header=0: the first data row is consumed as column names,
    so the spline starts from the second row. Columns: wavelength (m),
    sigma_12 (absorption, m^2), sigma_21 (emission, m^2)."""
    rows = []
    with open(path) as fh:
        lines = [ln for ln in fh.read().splitlines() if ln.strip()]
    for ln in lines[1:]:
        rows.append([float(v) for v in ln.split(",")])
    return np.array(rows)


class Amplifier2LevelModel:
    """amplifier.py:178-402, the CPU path (factor, calculate_N2, g_s, g_p,
    update_Pp) without the OpenCL helpers."""

    def __init__(self, name="amplifier2LevelModel", Pp=None, N=None, Rr=None):
        self.h_p = 6.62 * 1e-34
        self.T = 850.0 * 1e-6
        self.a = 3.0 * 1e-6
        self.b = 60.0 * 1e-6
        self.NA = 0.13
        self.lamb_p = 976.0
        self.N = N * (1e25) * np.pi * self.a ** 2
        self._Pp = Pp
        self.Tr = 1 / Rr
        self.domain = None
        self.sigma12_p = 2.5 * 1e-24
        self.sigma21_p = 2.44 * 1e-24

    def interpolate_sigma_s(self):
        """amplifier.py:247-258 (scipy InterpolatedUnivariateSpline, k=3,
        zero outside the table)."""
        from scipy import interpolate
        tab = yb_cross_section_table()
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
        with np.errstate(divide="ignore", invalid="ignore"):
            V = (2.0 * np.pi / domain.Lambda) * self.a * self.NA * 1e9               # :270-276
            omega_s = self.a * (0.616 + 1.66 / pow(V, 1.5) + 0.987 / pow(V, 6.0))
            Gamma_s = 1.0 - np.exp(-2.0 * self.a ** 2 / (omega_s * omega_s))
            self.rho_s = Gamma_s / (np.pi * self.a ** 2)
            Gamma_p = (self.a ** 2) / (self.b ** 2)                                 # :278-284
            self.rho_p = Gamma_p / (np.pi * self.a ** 2)
            self.Psat_s = (self.h_p * domain.nu * 1e12) / (                          # :286-296
                self.T * (self.sigma12_s + self.sigma21_s) * self.rho_s)
            self.Psat_p = (self.h_p * lambda_to_nu(self.lamb_p) * 1e12) / (          # :298-305
                self.T * (self.sigma12_p + self.sigma21_p) * self.rho_p)
            self.eta_s = self.sigma12_s * self.rho_s * self.N                        # :307-315
            self.eta_p = self.sigma12_p * self.rho_p * self.N                        # :317-323
            self.alpha_s = (self.sigma12_s + self.sigma21_s) * self.rho_s            # :325-333
            self.alpha_p = (self.sigma12_p + self.sigma21_p) * self.rho_p            # :335-341
            self.ratio_p = self.sigma12_p / (self.sigma12_p + self.sigma21_p)        # :347-353
            self.ratio_s = self.sigma12_s / (self.sigma12_s + self.sigma21_s)        # :355-366

    @property
    def Pp(self):
        return self._Pp

    def calculate_N2(self, Ps):
        """amplifier.py:368-375: NaN terms (zero cross sections) are dropped."""
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = self.ratio_s * Ps / self.Psat_s
            t2 = Ps / self.Psat_s
        numerator = self.ratio_p * self.Pp / self.Psat_p + np.sum(t1[~np.isnan(t1)])
        denominator = 1 + self.Pp / self.Psat_p + np.sum(t2[~np.isnan(t2)])
        return (numerator / denominator) * self.N

    def calculate_g_s(self, N2):
        return self.alpha_s * N2 - self.eta_s

    def calculate_g_p(self, N2):
        return self.alpha_p * N2 - self.eta_p

    def update_Pp(self, g_p, h):
        """amplifier.py:386-388: the pump evolves with every call of factor."""
        self._Pp = self._Pp * np.exp(g_p * h * 1e3)

    def factor(self, A, h):
        """amplifier.py:390-402: exponent of the gain of one linear sub-step,
        from the spectral power of the array passed (one extra FFT)."""
        N2 = self.calculate_N2(spectral_power(A) * len(A) * self.domain.dt / (self.Tr))
        g_s = self.calculate_g_s(N2)
        g_p = self.calculate_g_p(N2)
        self.update_Pp(g_p, h)
        return fftshift(g_s * h * 1e3 / 2)


# --------------------------------------------------------------------------
# linearity.py
# --------------------------------------------------------------------------
def _factorial(n):
    return float(math.factorial(n))


class Linearity:
    """linearity.py:106-320 (default and wdm sim_type)."""

    def __init__(self, alpha=None, beta=None, sim_type=None, use_cache=False, centre_omega=None, phase_lim=False,
                 amplifier=None):
        self.alpha, self.beta, self.centre_omega, self.phase_lim = alpha, beta, centre_omega, phase_lim
        self.amplifier = amplifier
        self.wdm = sim_type == "wdm"
        if self.wdm:  # linearity.py:136-144 getattr dispatch on sim_type
            self.lin = self.wdm_f
            self.exp_lin = self.wdm_exp_f_cached if use_cache else self.wdm_exp_f
        else:
            self.lin = self.default_f
            self.exp_lin = self.default_exp_f_cached if use_cache else self.default_exp_f
        self.cached_factor = None
        self.factor = None
        self.Domega = None

    def __call__(self, domain):
        self.domain = domain
        return self.wdm_linearity(domain) if self.wdm else self.default_linearity(domain)

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
                terms += beta * np.power(self.Domega, n) / _factorial(n)
            self.factor = 1j * fftshift(terms)
        if self.alpha is None:
            return self.factor
        self.factor -= 0.5 * self.alpha
        return self.factor

    def wdm_linearity(self, domain):
        """linearity.py:188-211. The attenuation is subtracted per channel
        (the reference's item assignment into its tuple raises TypeError)."""
        if self.beta is None:
            self.factor = (0.0, 0.0)
        else:
            if self.centre_omega is None:
                self.Domega = (domain.omega - domain.centre_omega, domain.omega - domain.centre_omega)
            else:
                self.Domega = (domain.omega - self.centre_omega[0], domain.omega - self.centre_omega[1])
            terms = [0.0, 0.0]
            for n, beta in enumerate(self.beta[0]):
                terms[0] += beta * np.power(self.Domega[0], n) / _factorial(n)
            for n, beta in enumerate(self.beta[1]):
                terms[1] += beta * np.power(self.Domega[1], n) / _factorial(n)
            self.factor = (1j * fftshift(terms[0]), 1j * fftshift(terms[1]))
        if self.alpha is not None:
            self.factor = (self.factor[0] - 0.5 * self.alpha[0], self.factor[1] - 0.5 * self.alpha[1])
        return self.factor

    def wdm_cache(self, h):
        """linearity.py:226-231"""
        hf0 = self.factor[0] * h
        hf1 = self.factor[1] * h
        if self.phase_lim:
            hf0 = self._limit_imag_part(hf0)
            hf1 = self._limit_imag_part(hf1)
        self.cached_factor = [np.exp(hf0), np.exp(hf1)]

    def wdm_f(self, As, z):
        """linearity.py:303-304"""
        return np.asarray([ifft(self.factor[0] * fft(As[0])), ifft(self.factor[1] * fft(As[1]))])

    def wdm_exp_f(self, As, h):
        """linearity.py:306-313"""
        hf0 = h * self.factor[0]
        hf1 = h * self.factor[1]
        if self.phase_lim:
            hf0 = self._limit_imag_part(hf0)
            hf1 = self._limit_imag_part(hf1)
        return np.asarray([ifft(np.exp(hf0) * fft(As[0])), ifft(np.exp(hf1) * fft(As[1]))])

    def wdm_exp_f_cached(self, As, h):
        """linearity.py:315-320"""
        if self.cached_factor is None:
            self.wdm_cache(h)
        return np.asarray([ifft(self.cached_factor[0] * fft(As[0])), ifft(self.cached_factor[1] * fft(As[1]))])

    def _limit_imag_part(self, hf):
        """linearity.py:213-218"""
        lhf = np.imag(hf) / (2 * np.pi)
        return np.real(hf) + 1j * 2 * np.pi * np.modf(lhf)[0]

    def default_cache(self, h):
        """linearity.py:220-224"""
        hf = self.factor * h
        if self.phase_lim:
            hf = self._limit_imag_part(hf)
        self.cached_factor = np.exp(hf)

    def cache(self, h):
        self.default_cache(h)

    def default_f(self, A, z):
        """linearity.py:238-239"""
        return ifft(self.factor * fft(A))

    def default_exp_f(self, A, h):
        """linearity.py:241-256"""
        hf = self.factor * h
        if self.phase_lim:
            hf = self._limit_imag_part(hf)
        if self.amplifier is None:
            return ifft(np.exp(hf) * fft(A))
        amp_factor = self.amplifier.factor(A, h)
        return ifft(np.multiply(np.exp(amp_factor), np.exp(hf) * fft(A)))

    def default_exp_f_cached(self, A, h):
        """linearity.py:258-272"""
        if self.cached_factor is None:
            self.cache(h)
        if self.amplifier is None:
            return ifft(self.cached_factor * fft(A))
        amp_factor = self.amplifier.factor(A, h)
        return ifft(np.multiply(np.exp(amp_factor), self.cached_factor * fft(A)))


# --------------------------------------------------------------------------
# nonlinearity.py
# --------------------------------------------------------------------------
class NonlinearityError(Exception):
    pass


class UnknownRamanResponseError(NonlinearityError):
    pass


def calculate_gamma(nonlinear_index, effective_area, centre_omega=2.0 * math.pi * 193.1):
    """nonlinearity.py:37-50"""
    return 1.0e24 * nonlinear_index * centre_omega / (effective_area * VACUUM_LIGHT_SPEED)


def calculate_raman_term(domain, tau_1=12.2e-3, tau_2=32.0e-3):
    """nonlinearity.py:53-69 (Blow-Wood)."""
    t = domain.t - domain.t.min()
    h_R = (tau_2 ** 2 + tau_1 ** 2) / (tau_1 * tau_2 ** 2)
    h_R *= np.exp(-t / tau_2) * np.sin(t / tau_1)
    return h_R


def calculate_raman_term_silica(domain):
    """nonlinearity.py:72-100 (Hollenbeck-Cantrell); the normalisation uses
    the removed simps rule, restated by simps_avg."""
    CP = [56.25, 100.00, 231.25, 362.50, 463.00, 497.00, 611.50, 691.67, 793.67, 835.50, 930.00, 1080.00, 1215.00]
    Ai = [1.00, 11.40, 36.67, 67.67, 74.00, 4.50, 6.80, 4.60, 4.20, 4.50, 2.70, 3.10, 3.00]
    Gfwhm = [52.10, 110.42, 175.00, 162.50, 135.33, 24.50, 41.50, 155.00, 59.50, 64.30, 150.00, 91.00, 160.00]
    Lfwhm = [17.37, 38.81, 58.33, 54.17, 45.11, 8.17, 13.83, 51.67, 19.83, 21.43, 50.00, 30.33, 53.33]
    c = 299792458.0
    t = domain.t - domain.t.min()
    h_R = np.zeros_like(t)
    for i in range(13):
        h_R += (Ai[i] * np.exp(-np.pi * 1e-10 * c * Lfwhm[i] * t)
                * np.exp(-((np.pi * 1e-10 * c * Gfwhm[i]) ** 2) * (t ** 2) / 4.0)
                * np.sin(2 * np.pi * 1e-10 * c * CP[i] * t))
    norm = simps_avg(h_R, domain.t)
    return h_R / norm


class Nonlinearity:
    """nonlinearity.py:103-304 (default and wdm sim_type)."""

    def __init__(self, gamma=None, sim_type=None, self_steepening=False, raman_scattering=False, rs_factor=3e-3,
                 use_all=False, tau_1=12.2e-3, tau_2=32.0e-3, f_R=0.18):
        self.gamma = gamma
        self.self_steepening = self_steepening
        self.raman_scattering = raman_scattering
        self.rs_factor = rs_factor
        self.use_all = use_all
        self.tau_1, self.tau_2, self.f_R = tau_1, tau_2, f_R
        self.wdm = sim_type == "wdm"
        if self.wdm and not (use_all or self_steepening or raman_scattering):
            self.non, self.exp_non = self.wdm_f, self.wdm_exp_f  # nonlinearity.py:148-150 with sim_type wdm
        elif use_all:
            self.non, self.exp_non = self.default_f_all, self.default_exp_f_all
        elif self_steepening and raman_scattering:
            self.non, self.exp_non = self.default_f_with_ss_and_rs, self.default_exp_f_with_ss_and_rs
        elif self_steepening:
            self.non, self.exp_non = self.default_f_with_ss, self.default_exp_f_with_ss
        elif raman_scattering:
            self.non, self.exp_non = self.default_f_with_rs, self.default_exp_f_with_rs
        else:
            self.non, self.exp_non = self.default_f, self.default_exp_f
        self.ss_factor = None
        self.h_R = None
        self.omega = None
        self.centre_omega = None
        self.factor = None

    def __call__(self, domain):
        """nonlinearity.py:160-185"""
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
        if self.wdm:  # nonlinearity.py:194-200
            self.factor = (0.0, 0.0) if self.gamma is None else (1j * self.gamma[0], 1j * self.gamma[1])
        else:
            self.factor = 0.0 if self.gamma is None else 1j * self.gamma

    def default_f_all(self, A, z):
        """nonlinearity.py:201-207"""
        term_spm = np.abs(A) ** 2
        convolution = ifft(self.h_R * fft(term_spm))
        p = fft(A * (1.0 - self.f_R) * term_spm + self.f_R * A * convolution)
        return ifft(self.factor * (1.0 + self.omega * self.ss_factor) * p)

    def default_exp_f_all(self, A, h, B):
        """nonlinearity.py:209-221"""
        term_spm = np.abs(A) ** 2
        convolution = ifft(self.h_R * fft(term_spm))
        term_all = (1.0 - self.f_R) * term_spm + self.f_R * convolution
        with np.errstate(divide="ignore", invalid="ignore"):
            term_ss_all = np.where(np.abs(B) > 1e-15, self.ss_factor / B, 0.0) * ifft(self.omega * fft(term_all * A))
        return np.exp(h * self.factor * (term_all + term_ss_all)) * B

    def default_f_with_ss(self, A, z):
        """nonlinearity.py:223-228"""
        term_spm = np.abs(A) ** 2 * A
        term_ss = self.ss_factor * ifft(self.omega * fft(term_spm))
        return self.factor * (term_spm + term_ss)

    def default_exp_f_with_ss(self, A, h, B):
        """nonlinearity.py:230-234"""
        term_spm = np.abs(A) ** 2
        term_ss = (self.ss_factor / B) * ifft(self.omega * fft(term_spm * A))
        return np.exp(h * self.factor * (term_spm + term_ss)) * B

    def default_f_with_rs(self, A, z):
        """nonlinearity.py:236-241"""
        term_spm = np.abs(A) ** 2
        term_rs = self.rs_factor * ifft(1j * self.omega * fft(term_spm))
        return self.factor * (term_spm + term_rs) * A

    def default_exp_f_with_rs(self, A, h, B):
        """nonlinearity.py:243-248"""
        term_spm = np.abs(A) ** 2
        term_rs = self.rs_factor * ifft(1j * self.omega * fft(term_spm))
        return np.exp(h * self.factor * (term_spm + term_rs)) * B

    def default_f_with_ss_and_rs(self, A, z):
        """nonlinearity.py:250-256 (term_ss is not multiplied by i*gamma)."""
        term_spm = np.abs(A) ** 2
        term_ss = self.ss_factor * ifft(self.omega * fft(term_spm * A))
        term_rs = self.rs_factor * ifft(1j * self.omega * fft(term_spm))
        return self.factor * (term_spm + term_rs) * A + term_ss

    def default_exp_f_with_ss_and_rs(self, A, h, B):
        """nonlinearity.py:258-264"""
        term_spm = np.abs(A) ** 2
        term_ss = (self.ss_factor / B) * ifft(self.omega * fft(term_spm * A))
        term_rs = self.rs_factor * ifft(1j * self.omega * fft(term_spm))
        return np.exp(h * self.factor * (term_spm + term_ss + term_rs)) * B

    def wdm_f(self, As, z):
        """nonlinearity.py:287-293"""
        return np.asarray([
            self.factor[0] * (np.abs(As[0]) ** 2 + 2.0 * np.abs(As[1]) ** 2) * As[0],
            self.factor[1] * (np.abs(As[1]) ** 2 + 2.0 * np.abs(As[0]) ** 2) * As[1],
        ])

    def wdm_exp_f(self, As, h, Bs):
        """nonlinearity.py:295-304"""
        return np.asarray([
            np.exp(h * self.factor[0] * (np.abs(As[0]) ** 2 + 2.0 * np.abs(As[1]) ** 2)) * Bs[0],
            np.exp(h * self.factor[1] * (np.abs(As[1]) ** 2 + 2.0 * np.abs(As[0]) ** 2)) * Bs[1],
        ])

    def default_f(self, A, z):
        """nonlinearity.py:266-267"""
        return self.factor * np.abs(A) ** 2 * A

    def default_exp_f(self, A, h, B):
        """nonlinearity.py:269-270"""
        return np.exp(h * self.factor * np.abs(A) ** 2) * B


# --------------------------------------------------------------------------
# solver.py
# --------------------------------------------------------------------------
class SolverError(Exception):
    pass


class NoDerivativeFunctionError(SolverError):
    pass


class Solver:
    """solver.py:32-354"""

    explicit_solvers = ["euler", "midpoint", "rk4"]
    embedded_solvers = ["bs", "rkf", "ck", "dp"]
    ss_solvers = ["ss_simple", "ss_symmetric", "ss_reduced", "ss_agrawal", "ss_sym_midpoint", "ss_sym_rk4"]
    other_solvers = ["rk4ip", "step_amplifier"]
    errors = {"euler": 2, "midpoint": 3, "rk4": 5, "bs": 2, "rkf": 4, "ck": 4, "dp": 4, "ss_simple": 2,
              "ss_symmetric": 3, "ss_reduced": 2, "ss_agrawal": 3, "ss_sym_midpoint": 3, "ss_sym_rk4": 5,
              "rk4ip": 5}

    def __init__(self, method="rk4", f=None):
        self.embedded = method.lower() in self.embedded_solvers
        self.method = getattr(self, method.lower())
        if f is None:
            raise NoDerivativeFunctionError("Require a function for ODE integration")
        self.f = f

    def __call__(self, A, z, h):
        return self.method(A, z, h, self.f)

    @staticmethod
    def euler(A, z, h, f):
        k0 = h * f(A, z)
        return A + k0

    @staticmethod
    def midpoint(A, z, h, f):
        k0 = h * f(A, z)
        k1 = h * f(A + 0.5 * k0, z + 0.5 * h)
        return A + k1

    @staticmethod
    def rk4(A, z, h, f):
        k0 = h * f(A, z)
        k1 = h * f(A + 0.5 * k0, z + 0.5 * h)
        k2 = h * f(A + 0.5 * k1, z + 0.5 * h)
        k3 = h * f(A + k2, z + h)
        return A + (k0 + 2.0 * (k1 + k2) + k3) / 6.0

    @staticmethod
    def bs(A, z, h, f):
        p = [1.0 / 2.0, 3.0 / 4.0, 2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0]
        q = [1.0 / 2.0, 3.0 / 4.0, 1.0]
        r = [2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0]
        s = [7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0]
        k0 = h * f(A, z)
        k1 = h * f(A + p[0] * k0, z + q[0] * h)
        k2 = h * f(A + p[1] * k1, z + q[1] * h)
        k3 = h * f(A + p[2] * k0 + p[3] * k1 + p[4] * k2, z + q[2] * h)
        A_fine = A + r[0] * k0 + r[1] * k1 + r[2] * k2
        A_coarse = A + s[0] * k0 + s[1] * k1 + s[2] * k2 + s[3] * k3
        return A_fine, A_coarse

    @staticmethod
    def rkf(A, z, h, f):
        p = [1.0 / 4.0, 3.0 / 32.0, 9.0 / 32.0, 1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0,
             439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, -8.0 / 27.0, 2.0, -3544.0 / 2565.0,
             1859.0 / 4104.0, -11.0 / 40.0]
        q = [1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 0.5]
        r = [25.0 / 216.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0]
        s = [16.0 / 135.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0]
        k0 = h * f(A, z)
        k1 = h * f(A + p[0] * k0, z + q[0] * h)
        k2 = h * f(A + p[1] * k0 + p[2] * k1, z + q[1] * h)
        k3 = h * f(A + p[3] * k0 + p[4] * k1 + p[5] * k2, z + q[2] * h)
        k4 = h * f(A + p[6] * k0 + p[7] * k1 + p[8] * k2 + p[9] * k3, z + q[3] * h)
        k5 = h * f(A + p[10] * k0 + p[11] * k1 + p[12] * k2 + p[13] * k3 + p[14] * k4, z + q[4] * h)
        A_coarse = A + r[0] * k0 + r[1] * k2 + r[2] * k3 + r[3] * k4
        A_fine = A + s[0] * k0 + s[1] * k2 + s[2] * k3 + s[3] * k4 + s[4] * k5
        return A_fine, A_coarse

    @staticmethod
    def ck(A, z, h, f):
        p = [1.0 / 5.0, 3.0 / 40.0, 9.0 / 40.0, 3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, -11.0 / 54.0, 5.0 / 2.0,
             -70.0 / 27.0, 35.0 / 27.0, 1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0,
             253.0 / 4096.0]
        q = [1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0]
        r = [37.0 / 378.0, 250.0 / 621.0, 125.0 / 594.0, 512.0 / 1771.0]
        s = [2825.0 / 27648.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0]
        k0 = h * f(A, z)
        k1 = h * f(A + p[0] * k0, z + q[0] * h)
        k2 = h * f(A + p[1] * k0 + p[2] * k1, z + q[1] * h)
        k3 = h * f(A + p[3] * k0 + p[4] * k1 + p[5] * k2, z + q[2] * h)
        k4 = h * f(A + p[6] * k0 + p[7] * k1 + p[8] * k2 + p[9] * k3, z + q[3] * h)
        k5 = h * f(A + p[10] * k0 + p[11] * k1 + p[12] * k2 + p[13] * k3 + p[14] * k4, z + q[4] * h)
        A_coarse = A + r[0] * k0 + r[1] * k2 + r[2] * k3 + r[3] * k5
        A_fine = A + s[0] * k0 + s[1] * k2 + s[2] * k3 + s[3] * k4 + s[4] * k5
        return A_fine, A_coarse

    @staticmethod
    def dp(A, z, h, f):
        p = [1.0 / 5.0, 3.0 / 40.0, 9.0 / 40.0, 44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 19372.0 / 6561.0,
             -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 9017.0 / 3168.0, -355.0 / 33.0,
             46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0,
             -2187.0 / 6784.0, 11.0 / 84.0]
        q = [1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0]
        r = [5179.0 / 57600.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0]
        s = [35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0]
        k0 = h * f(A, z)
        k1 = h * f(A + p[0] * k0, z + q[0] * h)
        k2 = h * f(A + p[1] * k0 + p[2] * k1, z + q[1] * h)
        k3 = h * f(A + p[3] * k0 + p[4] * k1 + p[5] * k2, z + q[2] * h)
        k4 = h * f(A + p[6] * k0 + p[7] * k1 + p[8] * k2 + p[9] * k3, z + q[3] * h)
        k5 = h * f(A + p[10] * k0 + p[11] * k1 + p[12] * k2 + p[13] * k3 + p[14] * k4, z + q[4] * h)
        # solver.py:248 uses k5 twice (p[18] and p[19]); kept as written.
        k6 = h * f(A + p[15] * k0 + p[16] * k2 + p[17] * k3 + p[18] * k5 + p[19] * k5, z + q[5] * h)
        A_coarse = A + r[0] * k0 + r[1] * k2 + r[2] * k3 + r[3] * k4 + r[4] * k5 + r[5] * k6
        A_fine = A + s[0] * k0 + s[1] * k2 + s[2] * k3 + s[3] * k4 + s[4] * k5
        return A_fine, A_coarse

    @staticmethod
    def ss_simple(A, z, h, f):
        A_N = f.nonlinear(A, h, A)
        return f.linear(A_N, h)

    @staticmethod
    def ss_symmetric(A, z, h, f):
        """solver.py:266-275: the nonlinear phase comes from the input A."""
        A_L = f.linear(A, 0.5 * h)
        A_N = f.nonlinear(A, h, A_L)
        return f.linear(A_N, 0.5 * h)

    @staticmethod
    def ss_reduced(A, z, h, f):
        A_L = f.linear(A, 0.5 * h)
        A_N = f.nonlinear(A_L, h, A_L)
        return f.linear(A_N, 0.5 * h)

    @staticmethod
    def ss_agrawal(A, z, h, f):
        Y_0 = A
        A_L = f.linear(A, 0.5 * h)
        with np.errstate(divide="ignore", invalid="ignore"):
            A_N1 = np.exp(0.5 * h * (np.where(abs(Y_0) > 1.0e-5, f.n(Y_0, z) / Y_0, 0.0)
                                     + np.where(abs(A) > 1.0e-5, f.n(A, z) / A, 0.0))) * A_L
            Y_1 = f.linear(A_N1, 0.5 * h)
            A_N2 = np.exp(0.5 * h * (np.where(abs(Y_1) > 1.0e-5, f.n(Y_1, z) / Y_1, 0.0)
                                     + np.where(abs(A) > 1.0e-5, f.n(A, z) / A, 0.0))) * A_L
        return f.linear(A_N2, 0.5 * h)

    @staticmethod
    def step_amplifier(A, z, h, f):
        return f.amplifier_step(A, h)

    def ss_sym_midpoint(self, A, z, h, f):
        A_L = f.linear(A, 0.5 * h)
        A_N = self.midpoint(A_L, z, h, f.n)
        return f.linear(A_N, 0.5 * h)

    def ss_sym_rk4(self, A, z, h, f):
        A_L = f.linear(A, 0.5 * h)
        A_N = self.rk4(A_L, z, h, f.n)
        return f.linear(A_N, 0.5 * h)

    @staticmethod
    def rk4ip(A, z, h, f):
        """solver.py:339-354"""
        hh = 0.5 * h
        A_I = f.linear(A, hh)
        k0 = f.linear(h * f.n(A, z), hh)
        k1 = h * f.n(A_I + 0.5 * k0, z + hh)
        k2 = h * f.n(A_I + 0.5 * k1, z + hh)
        k3 = h * f.n(f.linear(A_I + k2, hh), z + h)
        return (k3 / 6.0) + f.linear(A_I + (k0 + 2.0 * (k1 + k2)) / 6.0, hh)


# --------------------------------------------------------------------------
# storage.py, stepper.py
# --------------------------------------------------------------------------
class Storage:
    """storage.py:130-192: (z, A) trace capture. update_characts' spectral
    FFT is counted (storage.py:203); the rest runs only if characts=True."""

    def __init__(self, f=None, characts=False):
        self.f = f
        self.domain = None
        self.As = []
        self.z = []
        self.step_sizes = []
        self.fft_total = 0
        self.characts = characts
        self.energy_list = []
        self.max_power_list = []

    def __call__(self, domain):
        self.domain = domain

    def reset_array(self):
        self.As = []
        self.z = []

    @staticmethod
    def reset_fft_counter():
        reset_fft_counter()

    def store_current_fft_count(self):
        self.fft_total = _fft_counter

    def append(self, z, A):
        self.z.append(z)
        self.As.append(A)
        if self.domain is not None:
            spectral_power(A)  # storage.py:203 (counted FFT)
            if self.characts:
                self.energy_list.append(energy(A, self.domain.t))
                self.max_power_list.append(np.amax(temporal_power(A)))


class StepperError(Exception):
    pass


class SmallStepSizeError(StepperError):
    pass


class SuitableStepSizeError(StepperError):
    pass


class MaximumStepsAllocatedError(StepperError):
    pass


class Stepper:
    """stepper.py:49-376"""

    def __init__(self, traces=1, local_error=1.0e-6, method="RK4", f=None, length=1.0, total_steps=100,
                 characts=False):
        self.traces = traces
        self.local_error = local_error
        if method.upper().startswith("A"):
            self.adaptive = True
            self.method = method[1:]
        else:
            self.adaptive = False
            self.method = method
        self.solver = Solver(self.method, f)
        self.step = self.solver
        self.length = length
        self.total_steps = total_steps
        self.storage = Storage(characts=characts)
        self.total_attempts = 1000
        self.steps_max = 100000
        self.step_size_min = 1e-37
        self.safety = 0.9
        self.max_factor = 10.0
        self.min_factor = 0.2
        if self.method.lower() != "step_amplifier":
            self.eta = self.solver.errors[self.method.lower()]
        self.A_out = None
        self.attempt_log = []  # (h, delta, accepted) per attempt, adaptive only

    def __call__(self, A, refrence_length=None):
        """stepper.py:147-156"""
        self.storage.reset_fft_counter()
        self.storage.reset_array()
        self.attempt_log = []
        if self.adaptive:
            return self.adaptive_stepper(A, refrence_length)
        return self.standard_stepper(A, refrence_length)

    def standard_stepper(self, A, refrence_length):
        """stepper.py:164-222"""
        self.A_out = A
        h = self.length / self.total_steps
        zs = np.linspace(0.0, self.length, self.total_steps + 1)
        storage_step = int(self.total_steps / self.traces)
        for i, z in enumerate(zs[:-1]):
            if self.solver.embedded:
                self.A_out, A_other = self.step(self.A_out, z, h)
            else:
                self.A_out = self.step(self.A_out, z, h)
            if self.traces != 1:
                if i % storage_step == 0:
                    self.storage.append(z + h, self.A_out)
        self.storage.store_current_fft_count()
        return self.A_out

    @staticmethod
    def relative_local_error(A_fine, A_coarse):
        """stepper.py:230-245"""
        if np.isnan(A_fine).any() or np.isinf(A_fine).any():
            return 1.0
        norm_fine = np.linalg.norm(A_fine)
        if norm_fine != 0.0:
            return np.linalg.norm(A_fine - A_coarse) / norm_fine
        return np.linalg.norm(A_fine - A_coarse)

    def initial_step(self):
        """stepper.py:256-264"""
        if self.traces > self.total_steps:
            h = self.length / self.traces
        else:
            h = self.length / self.total_steps
        if h > 1e-6:
            h = 1e-6
        return h

    def adaptive_stepper(self, A, refrence_length, replay=None, max_attempts=None):
        """stepper.py:247-376. ``replay`` (optional) is a list of step sizes
        to use for successive attempts instead of the controller's;
        ``max_attempts`` (optional) stops the run early and returns the field
        reached, for timing a bounded sample."""
        self.A_out = A
        z = 0.0
        h = self.initial_step()
        step_storage = self.length / self.traces
        z_ignore = step_storage
        f_eta = np.power(2, self.eta - 1.0)
        f_alpha = f_eta / (f_eta - 1.0)
        f_beta = 1.0 / (f_eta - 1.0)
        if self.traces != 1:
            self.storage.append(z, self.A_out)
        n_attempt = 0
        for s in range(1, self.steps_max):
            if (z + h) > self.length:
                h = self.length - z
            for ta in range(0, self.total_attempts):
                if max_attempts is not None and n_attempt >= max_attempts:
                    self.storage.store_current_fft_count()
                    return self.A_out
                if replay is not None:
                    h = replay[n_attempt]
                n_attempt += 1
                h_half = 0.5 * h
                z_half = z + h_half
                if self.solver.embedded:
                    A_fine, A_coarse = self.step(self.A_out, z, h)
                else:
                    A_half = self.step(self.A_out, z, h_half)
                    A_fine = self.step(A_half, z_half, h_half)
                    A_coarse = self.step(self.A_out, z, h)
                delta = self.relative_local_error(A_fine, A_coarse)
                h_temp = h
                if delta > 0.0:
                    error_ratio = self.local_error / delta
                    factor = self.safety * np.power(error_ratio, 1.0 / self.eta)
                    h = h_temp * min(self.max_factor, max(self.min_factor, factor))
                else:
                    h = h_temp * self.max_factor
                accepted = delta < 2.0 * self.local_error
                self.attempt_log.append((h_temp, delta, accepted))
                if accepted:
                    z += h_temp
                    if self.solver.embedded:
                        self.A_out = A_fine
                    else:
                        self.A_out = f_alpha * A_fine - f_beta * A_coarse
                    if z > z_ignore:
                        self.storage.step_sizes.append((z, h_temp))
                        self.storage.append(z, self.A_out)
                        z_ignore += step_storage
                    break
                else:
                    if h < self.step_size_min:
                        raise SmallStepSizeError("Step size is extremely small")
            else:
                raise SuitableStepSizeError("Failed to set suitable step-size")
            if z >= self.length:
                self.storage.store_current_fft_count()
                return self.A_out
        raise MaximumStepsAllocatedError("Failed to complete with maximum steps allocated")


# --------------------------------------------------------------------------
# fibre.py, system.py
# --------------------------------------------------------------------------
class Fibre:
    """fibre.py:35-187 (default sim_type; simple or two-level Yb amplifier)."""

    def __init__(self, name="fibre", length=1.0, alpha=None, beta=None, gamma=0.0, sim_type=None, traces=1,
                 local_error=1.0e-6, method="ss_symmetric", total_steps=100, self_steepening=False,
                 raman_scattering=False, rs_factor=0.003, use_all=False, centre_omega=None, tau_1=12.2e-3,
                 tau_2=32.0e-3, f_R=0.18, small_signal_gain=None, E_sat=None, P_sat=None, Tr=None, lamb0=None,
                 bandwidth=None, use_Yb_model=False, Pp_0=None, N=None, Rr=None, characts=False):
        use_cache = not (method.upper().startswith("A"))
        self.name = name
        self.length = length
        self.gamma = gamma
        # fibre.py:88 fails for WDM's per-channel beta lists; no L_D there
        self.beta_2 = beta[2] if (beta is not None and sim_type != "wdm") else None
        self.domain = None
        self.amplifier = None
        if use_Yb_model:  # fibre.py:97-98
            self.amplifier = Amplifier2LevelModel(Pp=Pp_0, N=N, Rr=Rr)
        elif (small_signal_gain is not None) and (E_sat is not None or (P_sat is not None and Tr is not None)):
            self.amplifier = Amplifier(gain=small_signal_gain, E_sat=E_sat, P_sat=P_sat, Tr=Tr, length=self.length,
                                       lamb0=lamb0, bandwidth=bandwidth)
        self.linearity = Linearity(alpha=alpha, beta=beta, sim_type=sim_type, use_cache=use_cache,
                                   centre_omega=centre_omega, amplifier=self.amplifier)
        self.nonlinearity = Nonlinearity(gamma, sim_type, self_steepening, raman_scattering, rs_factor, use_all,
                                         tau_1, tau_2, f_R)

        fibre = self

        class FunctionStep:
            def __init__(self):
                self.l = fibre.l
                self.n = fibre.n
                self.linear = fibre.linear
                self.nonlinear = fibre.nonlinear

            def __call__(self, A, z):
                return self.l(A, z) + self.n(A, z)

        self.function_step = FunctionStep()
        self.stepper = Stepper(traces, local_error, method, self.function_step, self.length, total_steps,
                               characts=characts)

    def __call__(self, domain, field):
        if self.domain != domain or self.linearity.factor is None or self.nonlinearity.factor is None:
            self.linearity(domain)
            self.nonlinearity(domain)
            self.stepper.storage(domain)
            self.domain = domain
        return self.stepper(field, None)

    def l(self, A, z):
        return self.linearity.lin(A, z)

    def linear(self, A, h):
        return self.linearity.exp_lin(A, h)

    def n(self, A, z):
        return self.nonlinearity.non(A, z)

    def nonlinear(self, A, h, B):
        return self.nonlinearity.exp_non(A, h, B)


class System:
    """system.py:61-243 run loop only."""

    def __init__(self, domain=None, field=None):
        self.domain = domain if domain is not None else Domain()
        self.field = np.zeros([self.domain.total_samples], complex) if field is None else field
        self.fields = {}
        self.modules = []

    def add(self, module):
        self.modules.append(module)

    def __getitem__(self, name):
        for m in self.modules:
            if m.name == name:
                return m

    def run(self):
        for module in self.modules:
            self.field = module(self.domain, self.field)
            self.fields[module.name] = self.field


# --------------------------------------------------------------------------
# Build-only components (no reference counterpart)
# --------------------------------------------------------------------------
PLANCK_W_PS2 = 6.626e-10  # W ps^2, the constant of linearity.py:290


def oppm_noise(domain, seed):
    """One-photon-per-mode noise in the spectral domain (FFT order), returned
    in the temporal domain. |A~_k| = sqrt(h nu_k / (N dt)), phase uniform
    [0, 2pi) from numpy's Philox(key=seed). Seeds are per realisation."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    N = domain.total_samples
    nu = fftshift(domain.nu)
    amp = np.sqrt(PLANCK_W_PS2 * nu / (N * domain.dt))
    phi = rng.uniform(0.0, 2.0 * np.pi, N)
    spec = amp * np.exp(1j * phi)
    return np.fft.fft(spec)  # pyofss ifft (unscaled, sign -)


def coherence_g12(spectra):
    """|g12| over all pairs i != j of the rows of ``spectra`` (M x N):
    |sum_{i!=j} a_i* a_j| / ((M-1) sum_i |a_i|^2)."""
    a = np.asarray(spectra)
    M = a.shape[0]
    s = a.sum(axis=0)
    p = (np.abs(a) ** 2).sum(axis=0)
    cross = np.abs(s) ** 2 - p
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(p > 0, np.abs(cross) / ((M - 1) * p), 0.0)
