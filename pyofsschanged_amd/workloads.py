"""Synthetic workloads used by the benchmark and the smoke run, in pyofss units
(length km, beta_n ps^n/km, gamma 1/(W km), power W, time ps).

  soliton       second-order soliton at 1550 nm, 4096 samples over 100 ps,
                T0 = 1 ps, beta2 = -20 ps^2/km, gamma = 1.3 /W/km, one soliton
                period (pi/40 km), fixed steps.
  supercontinuum  50 fs FWHM sech at 10 kW and 835 nm in 15 cm of photonic
                crystal fibre (gamma = 110 /W/km, beta2..beta10 of
                examples/dudley_sc/fig_3.py:35-36), self-steepening and
                Blow-Wood Raman, 8192 samples over 12.5 ps, adaptive RK4IP with
                local error 1e-6.
  amplifier     saturable-gain amplifier at 1030 nm: 200 fs FWHM Gaussian of
                1 nJ (4697.2 W) with linear chirp C, 5 m, beta2 = 0.023 ps^2/m,
                gamma = 5e-3 /W/m, Raman (use_all), 15 dB small-signal gain,
                E_sat = 5 nJ, Lorentzian gain profile 40 nm wide, 16384
                samples over 40 ps, adaptive RK4IP.
  yb_attractor  the two-level Yb amplifier run of pyofss's
                examples/notebooks/nonlinear_attractor.ipynb: 600 W, 2 ps
                FWHM Gaussian at 1028 nm, 16384 samples over 200 ps, 8 m of
                fibre (beta2 = 20 ps^2/km, gamma = 5.1 /W/km, Hollenbeck
                Raman), Yb gain with 10 W pump, N = 4.8, 80 MHz repetition,
                30000 fixed rk4ip steps, 500 traces.
  noise_ensemble  realisations of the soliton fibre (adaptive RK4IP, local
                error 1e-6) on several grids over the same 100 ps window
                (default 128 x 16384, 256 x 8192, 128 x 4096 samples), each
                input with its own one-photon-per-mode noise; one launch,
                |g12| per grid on the device.

Input noise: one photon per mode with a uniform random spectral phase,
|A~_k| = sqrt(h nu_k / (N dt)) (h = 6.626e-10 W ps^2, the constant of
pyofss/modules/linearity.py:290), seeded per realisation with numpy's
Philox(key=seed).
"""

import math

import numpy as np

from .domain import Domain
from .field import fftshift, ifft
from ._lib import Ensemble
from .modules.fibre import Fibre
from .modules.gaussian import Gaussian
from .modules.sech import Sech

C_NM_PS = Domain.vacuum_light_speed
PLANCK_W_PS2 = 6.626e-10

DUDLEY_BETA = [0.0, 0.0, -11.830, 8.1038e-2, -9.5205e-5, 2.0737e-7, -5.3943e-10, 1.3486e-12, -2.5495e-15,
               3.0524e-18, -1.7140e-21]


def oppm_noise(domain, seed):
    """One-photon-per-mode noise field in the temporal domain."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    N = domain.total_samples
    nu = fftshift(domain.nu)
    amp = np.sqrt(PLANCK_W_PS2 * nu / (N * domain.dt))
    spec = amp * np.exp(1j * rng.uniform(0.0, 2.0 * np.pi, N))
    return ifft(spec)


def soliton(samples=4096, method="rk4ip", total_steps=1000, traces=50):
    domain = Domain(1, samples, 100.0, C_NM_PS / 1550.0)
    pulse = Sech(peak_power=4 * 20.0 / 1.3, width=1.0)
    fibre = Fibre(length=math.pi / 40, beta=[0.0, 0.0, -20.0], gamma=1.3, method=method, total_steps=total_steps,
                  traces=traces)
    return domain, pulse, fibre


def supercontinuum(samples=8192, length=1.5e-4, local_error=1e-6, traces=1):
    domain = Domain(1, samples, 12.5, C_NM_PS / 835.0)
    pulse = Sech(peak_power=1e4, width=0.050, using_fwhm=True)
    fibre = Fibre(length=length, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True,
                  method="ark4ip", local_error=local_error, traces=traces)
    return domain, pulse, fibre


def yb_attractor(traces=500, total_steps=30000, length=8e-3, cycle="cycle0"):
    domain = Domain(samples_per_bit=2 ** 14, bit_width=200.0, total_bits=1, centre_nu=C_NM_PS / 1028.0)
    pulse = Gaussian(peak_power=600, width=2, C=0, using_fwhm=True)
    fibre = Fibre(length=length, method="rk4ip", total_steps=total_steps, beta=np.array([0, 0, 20]), gamma=5.1,
                  self_steepening=False, use_all="hollenbeck", cycle=cycle, traces=traces, Pp_0=10, N=4.8,
                  Rr=80 * 1e-6, use_Yb_model=True, downsampling=4000)
    return domain, pulse, fibre


def amplifier(samples=16384, chirps=None, length=5e-3, local_error=1e-6):
    """Domain, input fields (one per chirp value) and Fibre of the amplifier
    batch; chirps default to 64 values evenly spaced in [-2, 2]."""
    chirps = np.linspace(-2.0, 2.0, 64) if chirps is None else np.asarray(chirps, dtype=float)
    domain = Domain(1, samples, 40.0, C_NM_PS / 1030.0)
    fields = np.stack([Gaussian(peak_power=4697.2, width=0.2, using_fwhm=True, C=float(c)).generate(domain.t)
                       for c in chirps])
    fibre = Fibre(length=length, beta=[0.0, 0.0, 23.0], gamma=5.0, use_all=True, small_signal_gain=15.0,
                  E_sat=5.0, lamb0=1030.0, bandwidth=40.0, method="ark4ip", local_error=local_error)
    return domain, fields, fibre


NOISE_ENSEMBLE_GROUPS = ((16384, 128), (8192, 256), (4096, 128))


def noise_ensemble(groups=NOISE_ENSEMBLE_GROUPS, length=math.pi / 40, local_error=1e-6, base_seed=1000,
                   method="ark4ip", inputs=None):
    """Build the noise ensemble: one Fibre (and device context) per grid, the
    inputs (sech + noise; realisation r of the g-th group uses seed
    base_seed + sum of earlier group sizes + r) uploaded, and the Ensemble
    handle. ``inputs`` may give the input arrays per group instead.
    Returns (ensemble, setups) with setups[g] = dict(N, domain, fibre,
    seeds, inputs)."""
    setups = []
    seed = base_seed
    for g, (N, M) in enumerate(groups):
        domain = Domain(1, N, 100.0, C_NM_PS / 1550.0)
        fibre = Fibre(length=length, beta=[0.0, 0.0, -20.0], gamma=1.3, method=method, local_error=local_error)
        fibre.prepare(domain)
        seeds = list(range(seed, seed + M))
        seed += M
        if inputs is None:
            A0 = Sech(peak_power=4 * 20.0 / 1.3, width=1.0).generate(domain.t)
            A = np.stack([A0 + oppm_noise(domain, s) for s in seeds])
        else:
            A = np.asarray(inputs[g], dtype=np.complex128).reshape(M, N)
        setups.append(dict(N=N, domain=domain, fibre=fibre, seeds=seeds, inputs=A))
    stepper = setups[0]["fibre"].stepper
    ens = Ensemble(stepper.solver.method_id, stepper.adaptive_params(),
                   [(s["fibre"]._device_context(), len(s["seeds"])) for s in setups])
    for g, s in enumerate(setups):
        ens.set_input(g, s["inputs"])
    return ens, setups
