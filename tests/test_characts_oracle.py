"""Pins of the trace-characteristics restatement (oracle/characts_np.py)
against the scipy.signal routines the reference calls (field.py:207-287),
CPU only. Signals: random multi-peak sums, quantised signals with flat
peaks, noise, maxima on the border, and temporal/spectral powers of
propagated pulses (soliton and supercontinuum, oracle runs)."""

import numpy as np
import pytest

import oracle as O
from helpers import supercontinuum_c2
from oracle import characts_np as Q


def signals():
    rng = np.random.default_rng(0)
    n = np.arange(512)
    out = []
    for k in range(40):
        x = np.zeros(512)
        for _ in range(rng.integers(1, 6)):
            x += rng.uniform(0.1, 1.0) * np.exp(-0.5 * ((n - rng.uniform(20, 490)) / rng.uniform(2, 60)) ** 2)
        x += rng.uniform(0, 1e-3) * rng.random(512)
        if k % 3 == 1:
            x = np.round(x * 50) / 50 + 1e-9  # flat peaks
        if k % 5 == 2:
            x[0] = x.max() * 1.1  # maximum on the border (not a peak)
        out.append(x)
    out.append(rng.random(512))  # noise: many peaks
    return out


def propagated():
    dom, A, fib = supercontinuum_c2(1024)
    f = O.Fibre(length=3e-6, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True,
                method="ark4ip", local_error=1e-6)
    out = f(dom, A.copy())
    return [O.temporal_power(A), O.spectral_power(A), O.temporal_power(out), O.spectral_power(out)]


@pytest.mark.parametrize("which", ["synthetic", "propagated"])
def test_restated_characteristics_match_scipy(which):
    cases = signals() if which == "synthetic" else propagated()
    for P in cases:
        assert Q.restated_duration(P) == pytest.approx(Q.reference_duration(P), rel=1e-12, abs=1e-12)
        ref = Q.reference_spec_width(P)
        got = Q.restated_spec_width(P)
        assert (got is None) == (ref is None)
        if ref is not None:
            assert got == pytest.approx(ref, rel=1e-12, abs=1e-12)
        assert np.array_equal(Q.restated_peaks(P), Q.reference_peaks(P))


def test_flat_peak_takes_middle_sample():
    x = np.array([0.0, 1.0, 3.0, 3.0, 3.0, 3.0, 1.0, 0.0, 2.0, 0.0])
    assert Q.local_maxima(x) == [3, 8]
    from scipy.signal import find_peaks
    assert list(find_peaks(x)[0]) == [3, 8]
