"""GPU parity of the trace characteristics (Storage.update_characts,
pyofss/modules/storage.py:194-209) against the scipy.signal computations the
reference performs (oracle/characts_np.py reference_*).

The powers the reference analyses are |A|^2 and |fft(A)|^2; for the
comparison the temporal power is formed as Re^2 + Im^2 like the device (numpy's
abs(A)**2 may differ in the last bit), and the spectrum comes from numpy's
FFT (the device's differs by ~1e-16 relative). Bars: energy 1e-13 relative,
maximum power 1e-15, duration and spectral width 1e-9 relative, peak
indices identical.
"""

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, supercontinuum_c2
from oracle import characts_np as Q

pytestmark = pytest.mark.gpu

GRIDS = [256, 1024, 4096, 8192, 16384]


@pytest.fixture(scope="module")
def lib():
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return _lib


def fields(N, seed):
    rng = np.random.default_rng(seed)
    n = np.arange(N)
    out = []
    for k in range(6):
        a = np.zeros(N, complex)
        for _ in range(rng.integers(1, 5)):
            a += rng.uniform(0.3, 1.0) * np.exp(-0.5 * ((n - rng.uniform(0.1, 0.9) * N) / rng.uniform(3, N / 20)) ** 2)
        a = a * np.exp(1j * rng.uniform(-3, 3) * ((n - N / 2) / N) ** 2 * 50) + 1e-4 * rng.standard_normal(N)
        if k == 2:  # real field with flat tops (exact ties in |A|^2)
            a = np.round(np.abs(a) * 20) / 20 + 0j
        out.append(a)
    return np.stack(out)


def expected(A, t):
    P = A.real ** 2 + A.imag ** 2
    S = np.fft.fftshift(np.abs(np.fft.ifft(A)) ** 2)  # pyofss fft = scipy ifft, natural order
    sw = Q.reference_spec_width(S)
    return dict(energy=O.energy(A, t), max_power=P.max(), duration=Q.reference_duration(P),
                spec_width=np.nan if sw is None else sw, peaks=Q.reference_peaks(P))


def check(got, k, exp):
    assert got["energy"][k] == pytest.approx(exp["energy"], rel=1e-13)
    assert got["max_power"][k] == pytest.approx(exp["max_power"], rel=1e-15)
    assert got["duration"][k] == pytest.approx(exp["duration"], rel=1e-9)
    if np.isnan(exp["spec_width"]):
        assert np.isnan(got["spec_width"][k])
    else:
        assert got["spec_width"][k] == pytest.approx(exp["spec_width"], rel=1e-9)
    assert np.array_equal(got["peaks"][k], exp["peaks"])


@pytest.mark.parametrize("N", GRIDS)
def test_characteristics_match_scipy(lib, N):
    from pyofsschanged_amd.field import simpson_weights
    dom = O.Domain(1, N, 20.0, 193.0)
    ctx = lib.Context(0)
    ctx.set_grid(N)
    A = fields(N, N)
    got = ctx.trace_characteristics(A, simpson_weights(dom.t * 1e-3))
    for k in range(len(A)):
        exp = expected(A[k], dom.t)
        print(f"N={N} field {k}: duration {got['duration'][k]:.6f} ({exp['duration']:.6f}), spec width "
              f"{got['spec_width'][k]:.3f} ({exp['spec_width']:.3f}), peaks {len(got['peaks'][k])}", flush=True)
        check(got, k, exp)


def test_storage_characteristics_after_propagation(lib):
    """A System run with traces: the Storage lists filled from the device
    equal the reference computations on the same traces, and fft_total
    counts the per-trace spectral_power FFT as the reference does."""
    import pyofsschanged_amd as P
    from helpers import DUDLEY_BETA, C_NM_PS
    dom = P.Domain(1, 1024, 12.5, C_NM_PS / 835.0)
    sys_ = P.System(dom)
    sys_.add(P.Sech(peak_power=1e4, width=0.050, using_fwhm=True))
    fib = P.Fibre(length=1e-5, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True,
                  method="ark4ip", local_error=1e-6, traces=5)
    sys_.add(fib)
    sys_.run()
    st = fib.stepper.storage
    assert len(st.As) == len(st.energy_list) >= 5
    for k, A in enumerate(st.As):
        exp = expected(np.asarray(A), dom.t)
        assert st.energy_list[k] == pytest.approx(exp["energy"], rel=1e-13)
        assert st.max_power_list[k] == pytest.approx(exp["max_power"], rel=1e-15)
        assert st.duration_list[k] == pytest.approx(exp["duration"] * dom.dt, rel=1e-9)
        assert st.spec_width_list[k] == pytest.approx(exp["spec_width"] * dom.dnu, rel=1e-9)
        assert np.array_equal(st.peaks_list[k], exp["peaks"])
        assert st.l_nl_list[k] == pytest.approx(1.0 / (110.0 * exp["max_power"]), rel=1e-15)
        assert st.l_d_list[k] == pytest.approx((exp["duration"] * dom.dt) ** 2 / 11.830, rel=1e-9)
    # the oracle's run of the same fibre counts the same FFTs
    odom = O.Domain(1, 1024, 12.5, C_NM_PS / 835.0)
    of = O.Fibre(length=1e-5, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True,
                 method="ark4ip", local_error=1e-6, traces=5, characts=True)
    of(odom, O.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(odom.t))
    assert st.fft_total == of.stepper.storage.fft_total
    assert rel_l2(np.asarray(st.As[-1]), of.stepper.storage.As[-1]) < 1e-9
