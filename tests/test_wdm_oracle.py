"""Pins of the two-channel WDM restatement (oracle Linearity/Nonlinearity
with sim_type="wdm", pyofss/modules/linearity.py:188-211, 303-320 and
nonlinearity.py:194-200, 287-304), CPU only.

* Without dispersion the cross-phase modulation only rotates phases:
  A_c(z) = A_c(0) exp(i gamma_c (|A_c|^2 + 2 |A_o|^2) z); ss_symmetric
  (phase from the un-propagated fields, exact for a pure rotation) to 1e-12,
  rk4ip to 2e-7 at 200 steps with the error falling 16-fold per halving.
* Without nonlinearity each channel disperses with its own beta: a Gaussian
  broadens as T(z) = T0 sqrt(1 + (z/L_D)^2) and a beta_1 delays the pulse by
  beta_1 z.
* With both, each channel's energy is conserved by ss_symmetric (XPM is a
  phase, dispersion unitary); rk4ip converges at fourth order to the
  rotation.
The examples/agrawal_book/fig_7_2.py configuration (two channels, walk-off
beta_1 = 0.125 ps/km on channel 1, gamma = [1.0, 1.2]) is the test case.
"""

import numpy as np
import pytest

import oracle as O


def fig72(length=40.0, steps=400, beta=((0.0, 0.0, 0.0, 0.0), (0.0, 0.125, 0.0, 0.0)), gamma=(1.0, 1.2),
          method="rk4ip", samples=4096):
    dom = O.Domain(bit_width=200.0, samples_per_bit=samples, channels=2)
    A = np.stack([O.Gaussian(width=1.0, peak_power=1.0, channel=0).generate(dom.t),
                  O.Gaussian(width=1.0, peak_power=0.5, channel=1).generate(dom.t)])
    f = O.Fibre("fibre", length=length, gamma=list(gamma), beta=[list(b) for b in beta], sim_type="wdm",
                total_steps=steps, method=method)
    return dom, A, f


def xpm_error(method, steps):
    dom, A, f = fig72(length=2.0, steps=steps, beta=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), method=method)
    out = f(dom, A.copy())
    z = 2.0
    P = np.abs(A) ** 2
    exact = np.stack([A[0] * np.exp(1j * 1.0 * (P[0] + 2 * P[1]) * z), A[1] * np.exp(1j * 1.2 * (P[1] + 2 * P[0]) * z)])
    return np.linalg.norm(out - exact) / np.linalg.norm(exact)


def test_xpm_rotation_closed_form():
    assert xpm_error("ss_symmetric", 200) < 1e-12  # a pure rotation: exact
    e1, e2 = xpm_error("rk4ip", 200), xpm_error("rk4ip", 400)
    assert e1 < 2e-7 and 14.0 < e1 / e2 < 18.0  # fourth order in the step


def test_dispersion_only_closed_form():
    beta2 = (-10.0, 5.0)
    dom, A, f = fig72(length=0.05, steps=50, beta=((0.0, 0.0, beta2[0]), (0.0, 0.3, beta2[1])), gamma=(0.0, 0.0))
    out = f(dom, A.copy())
    for c in range(2):
        T0 = 1.0
        LD = T0 ** 2 / abs(beta2[c])
        Tz = T0 * np.sqrt(1 + (0.05 / LD) ** 2)
        P = np.abs(out[c]) ** 2
        t_mean = np.sum(dom.t * P) / np.sum(P)
        rms = np.sqrt(np.sum((dom.t - t_mean) ** 2 * P) / np.sum(P))
        assert rms == pytest.approx(Tz / np.sqrt(2.0), rel=1e-9)  # Gaussian |A|^2 rms width = T / sqrt(2)
        assert t_mean == pytest.approx(0.0 if c == 0 else 0.3 * 0.05, abs=1e-9)


def test_energy_per_channel_conserved():
    """ss_symmetric is a product of unitary maps per channel (Parseval)."""
    dom, A, f = fig72(length=10.0, steps=200, method="ss_symmetric")
    out = f(dom, A.copy())
    for c in range(2):
        assert np.sum(np.abs(out[c]) ** 2) == pytest.approx(np.sum(np.abs(A[c]) ** 2), rel=1e-12)
    # the walk-off moved channel 1, and XPM broadened channel 0's spectrum
    assert np.linalg.norm(np.abs(out[1]) - np.abs(A[1])) > 1e-3
    assert np.linalg.norm(np.abs(np.fft.fft(out[0])) - np.abs(np.fft.fft(A[0]))) > 1e-3 * np.linalg.norm(np.fft.fft(A[0]))
