"""Field helpers with the pyofss conventions of pyofss/field.py:40-303.

``fft``/``ifft`` run on the GPU through the HIP engine (one context per grid
size); they keep the physics sign convention of field.py:109-136 and the
global ``fft_counter`` that the steppers report as ``fft_total``. The other
helpers are host-side diagnostics, as in the reference.
"""

import threading

import numpy as np
from scipy.interpolate import interp1d
from scipy.signal import find_peaks, peak_widths

from . import _lib

fft_counter = 0
_ctx_lock = threading.Lock()
_contexts = {}


def _grid_context(n):
    with _ctx_lock:
        ctx = _contexts.get(n)
        if ctx is None:
            ctx = _lib.Context(0)
            ctx.set_grid(n)
            _contexts[n] = ctx
        return ctx


def _transform(A, sign):
    A = np.asarray(A)
    return _grid_context(A.shape[-1]).fft(A, sign)


def fft(A_t):
    """field.py:109-121: (1/N) sum_n A_n exp(+2 pi i k n / N), on the GPU."""
    global fft_counter
    fft_counter += 1
    return _transform(A_t, +1)


def ifft(A_nu):
    """field.py:124-136: sum_k A_k exp(-2 pi i k n / N), on the GPU."""
    global fft_counter
    fft_counter += 1
    return _transform(A_nu, -1)


def ifftshift(A_nu):
    """field.py:139-147: FFT order -> consecutive order (name kept as in pyofss)."""
    return np.fft.fftshift(A_nu)


def fftshift(A_nu):
    """field.py:150-158: consecutive order -> FFT order (name kept as in pyofss)."""
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


def phase(A_t, unwrap=True):
    """field.py:78-90"""
    return np.unwrap(np.angle(A_t)) if unwrap else np.angle(A_t)


def chirp(A_t, window_nu, unwrap=True):
    """field.py:93-106"""
    return -np.gradient(phase(A_t, unwrap)) * window_nu


def simpson_weights(x):
    """Weights w with sum_i w_i y_i equal to the old scipy.integrate.simps(y, x)
    (default even='avg'), which field.py:170 and nonlinearity.py:99 call and
    which SciPy >= 1.14 no longer provides. For an even number of samples the
    rule averages Simpson-on-[0, N-2] plus a trapezoid on the last interval
    with a trapezoid on the first interval plus Simpson-on-[1, N-1]."""
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    h = np.diff(x)
    w = np.zeros(N)

    def basic(start, stop, scale):
        h0 = h[start:stop:2]
        h1 = h[start + 1:stop + 1:2]
        hsum = h0 + h1
        hprod = h0 * h1
        h0divh1 = h0 / h1
        i0 = np.arange(start, stop, 2)
        np.add.at(w, i0, scale * hsum / 6.0 * (2 - 1.0 / h0divh1))
        np.add.at(w, i0 + 1, scale * hsum / 6.0 * hsum * hsum / hprod)
        np.add.at(w, i0 + 2, scale * hsum / 6.0 * (2 - h0divh1))

    if N % 2 == 0:
        basic(0, N - 3, 0.5)
        w[-1] += 0.25 * h[-1]
        w[-2] += 0.25 * h[-1]
        basic(1, N - 2, 0.5)
        w[0] += 0.25 * h[0]
        w[1] += 0.25 * h[0]
    else:
        basic(0, N - 2, 1.0)
    return w


def energy(A_t, t):
    """field.py:161-172: pulse energy in nJ (Simpson rule over t * 1e-3)."""
    return float(np.dot(simpson_weights(t * 1e-3), temporal_power(A_t)))


def loss_infrared_db(wl):
    """field.py:187-197 (examples/fopo_cars/FOPO_model_ocl.py:101 builds the
    fibre's attenuation from this and loss_rayleigh_db)"""
    p = np.array([-1.75292532e03, 1.95607318e-02])
    return np.exp(p[1] * (wl + p[0]))


def loss_rayleigh_db(wl):
    """field.py:200-210"""
    factor = 1 - 0.14 / (wl * 1e-3) ** 4
    factor[factor <= 0] = 1e-16
    return -10 * np.log10(factor)


def max_peak_params(P, prominence):
    """field.py:213-234"""
    peaks, properties = find_peaks(P, height=0)
    max_peak_ind = np.argmax(properties["peak_heights"])
    results_fwhm = peak_widths(P, [peaks[max_peak_ind]], rel_height=0.5)
    ind_max = np.argmax(results_fwhm[1])
    return results_fwhm[1][ind_max], results_fwhm[0][ind_max], results_fwhm[2][ind_max], results_fwhm[3][ind_max]


def spectrum_width_params(P, prominence=0.0001):
    """field.py:237-278"""

    def find_x(y_peak, x_peak, is_left, ydata):
        x = np.where(ydata < y_peak)[0]
        if is_left:
            filtered_x = x[np.where(x < x_peak)]
            boundary = np.amax(filtered_x) if len(filtered_x) > 0 else 0
        else:
            filtered_x = x[np.where(x > x_peak)]
            boundary = np.amin(filtered_x) if len(filtered_x) > 0 else 0
        return int(boundary)

    peaks, _ = find_peaks(P, prominence=prominence)
    if len(peaks) > 1:
        heigth_fwhm = np.amax(P) / 10
        peaks.sort()
        left_ind = find_x(heigth_fwhm, peaks[0], True, P)
        right_ind = find_x(heigth_fwhm, peaks[-1], False, P)
        fwhm = right_ind - left_ind
    elif len(peaks) == 0:
        return None, None, None, None
    else:
        results_fwhm = peak_widths(P, peaks, rel_height=0.5)
        heigth_fwhm = np.amax(results_fwhm[1])
        fwhm = results_fwhm[0][0]
        left_ind = results_fwhm[2][0]
        right_ind = results_fwhm[3][0]
    return heigth_fwhm, fwhm, int(left_ind), int(right_ind)


def get_peaks(P):
    """field.py:281-283"""
    peaks, _ = find_peaks(P, height=0, prominence=(np.amax(P) / 10))
    return peaks


def get_duration_spec(P, d_x, prominence=None):
    """field.py:286-290"""
    if prominence is None:
        prominence = np.amax(P) / 100
    fwhm = spectrum_width_params(P, prominence=prominence)[1]
    return abs(fwhm) * d_x if fwhm is not None else None


def get_duration(P, d_x, prominence=None):
    """field.py:293-297"""
    if prominence is None:
        prominence = np.amax(P) / 100
    return max_peak_params(P, prominence=prominence)[1] * d_x


def get_downsampled(P, downsampling):
    """field.py:300-303"""
    f = interp1d(np.arange(len(P)), P)
    return f(np.linspace(0, len(P) - 1, downsampling))
