"""Shared set-ups for the parity tests: the same physics built once with the
oracle (NumPy restatement) and once for the HIP engine."""

import math

import numpy as np

import oracle as O

C_NM_PS = 299792458.0 * 1e-3

# Dispersion of the photonic-crystal fibre of examples/dudley_sc/fig_3.py:35-36 (ps^n / km)
DUDLEY_BETA = [0.0, 0.0, -11.830, 8.1038e-2, -9.5205e-5, 2.0737e-7, -5.3943e-10, 1.3486e-12, -2.5495e-15,
               3.0524e-18, -1.7140e-21]


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def soliton_c1(N=4096):
    """Second-order soliton at 1550 nm: T0 = 1 ps, beta2 = -20 ps^2/km,
    gamma = 1.3 /W/km, P0 = N^2 |beta2| / (gamma T0^2) with N = 2."""
    dom = O.Domain(1, N, 100.0, C_NM_PS / 1550.0)
    A = O.Sech(peak_power=4 * 20.0 / 1.3, width=1.0).generate(dom.t)
    fibre = dict(length=math.pi / 40, beta=[0.0, 0.0, -20.0], gamma=1.3)
    return dom, A, fibre


def supercontinuum_c2(N=8192):
    """Dudley supercontinuum at 835 nm: 50 fs FWHM sech at 10 kW, gamma = 110
    /W/km, self-steepening and Blow-Wood Raman, 15 cm."""
    dom = O.Domain(1, N, 12.5, C_NM_PS / 835.0)
    A = O.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(dom.t)
    fibre = dict(length=1.5e-4, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True)
    return dom, A, fibre


def amplifier_c4(N=16384, C=0.0):
    """Saturable-gain amplifier at 1030 nm: 200 fs FWHM Gaussian of 1 nJ
    (P0 = 4697.2 W) with linear chirp C, 5 m of fibre, beta2 = 0.023 ps^2/m,
    gamma = 5e-3 /W/m, Raman (use_all), 15 dB small-signal gain (3 dB/m),
    E_sat = 5 nJ, Lorentzian gain profile 40 nm wide (amplifier.py:110-129)."""
    dom = O.Domain(1, N, 40.0, C_NM_PS / 1030.0)
    A = O.Gaussian(peak_power=4697.2, width=0.2, using_fwhm=True, C=C).generate(dom.t)
    fibre = dict(length=5e-3, beta=[0.0, 0.0, 23.0], gamma=5.0, use_all=True,
                 amplifier=dict(gain=15.0, E_sat=5.0, lamb0=1030.0, bandwidth=40.0))
    return dom, A, fibre


def oracle_amplifier(dom, length, gain, E_sat, lamb0=None, bandwidth=None):
    amp = O.Amplifier(gain=gain, E_sat=E_sat, length=length, lamb0=lamb0, bandwidth=bandwidth)
    amp.set_domain(dom)
    return amp


def oracle_ops(dom, beta=None, alpha=None, gamma=None, self_steepening=False, raman_scattering=False, use_all=False,
               rs_factor=3e-3, f_R=0.18, tau_1=12.2e-3, tau_2=32e-3, amplifier=None):
    lin = O.Linearity(alpha=alpha, beta=beta, amplifier=amplifier)
    lin(dom)
    nl = O.Nonlinearity(gamma, None, self_steepening, raman_scattering, rs_factor, use_all, tau_1, tau_2, f_R)
    nl(dom)
    return lin, nl


def nl_kind(self_steepening=False, raman_scattering=False, use_all=False):
    if use_all:
        return 1
    if self_steepening and raman_scattering:
        return 4
    if self_steepening:
        return 2
    if raman_scattering:
        return 3
    return 0


def engine_context(dom, lin, nl, kind, amplifier=None):
    """A context loaded with the oracle's operator arrays."""
    from pyofsschanged_amd import _lib
    ctx = _lib.Context(0)
    N = dom.total_samples
    ctx.set_grid(N)
    ctx.set_linear(np.broadcast_to(np.asarray(lin.factor, dtype=complex), (N,)))
    gamma = 0.0 if nl.gamma is None else nl.gamma
    nm = hr = None
    if kind == 1:
        nm = nl.factor * (1.0 + nl.omega * nl.ss_factor)
        hr = nl.h_R
    ctx.set_nonlinear(kind, gamma, nl.f_R, nl.ss_factor, nl.rs_factor, nl.omega, nm, hr)
    if amplifier is not None:
        # energy weights of the old simps(|A|^2, t * 1e-3) rule (field.py:161-172)
        from pyofsschanged_amd.field import simpson_weights
        ctx.set_gain(True, float(np.log(np.power(10, 0.1 * amplifier.gain))), amplifier.length, amplifier.E_sat,
                     amplifier.spectal_filtration_array, simpson_weights(dom.t * 1e-3))
    return ctx
