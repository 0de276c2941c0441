"""GPU runs through the drop-in API (Domain, Sech/Gaussian, Fibre) against
the attempt-by-attempt fixtures that the reference's own stepper.py wrote
(tests/golden/adaptive_fixtures.npz, tests/golden/make_adaptive_fixtures.py):
256-sample adaptive RK4IP fibres with Kerr, strong Kerr, self-steepening +
Blow-Wood Raman (a first step far too long) and Raman + saturable gain.

Bar: identical attempt and acceptance counts, trace positions and stored
step sizes to 1e-9 relative (sums of step sizes that carry the fifth root of
a rounding-level difference), fft_total exactly, output field relative L2
<= 1e-9."""

import os

import numpy as np
import pytest

from helpers import C_NM_PS, DUDLEY_BETA, rel_l2

pytestmark = pytest.mark.gpu
FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "adaptive_fixtures.npz"))
FIELD_TOL = 1e-9


def _case(name):
    return {k.split("/", 1)[1]: FIX[k] for k in FIX.files if k.split("/", 1)[0] == name}


@pytest.fixture(scope="module")
def P():
    import pyofsschanged_amd as P
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return P


def _build(P, name):
    if name in ("kerr", "kerr_strong"):
        dom = P.Domain(1, 256, 20.0, C_NM_PS / 1550.0)
        P0, L, traces, steps = (4 * 20.0 / 1.3, 0.08, 4, 100) if name == "kerr" else (1e3, 0.01, 2, 1)
        A = P.Sech(peak_power=P0, width=1.0).generate(dom.t)
        fib = P.Fibre(length=L, beta=[0.0, 0.0, -20.0], gamma=1.3, method="ark4ip", traces=traces,
                      total_steps=steps, local_error=1e-4)
    elif name == "raman_ss":
        dom = P.Domain(1, 256, 2.0, C_NM_PS / 835.0)
        A = P.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(dom.t)
        fib = P.Fibre(length=5e-6, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True,
                      method="ark4ip", traces=2, total_steps=1, local_error=1e-6)
    else:
        dom = P.Domain(1, 256, 10.0, C_NM_PS / 1030.0)
        A = P.Gaussian(peak_power=1e3, width=0.2, using_fwhm=True).generate(dom.t)
        fib = P.Fibre(length=2e-4, beta=[0.0, 0.0, 23.0], gamma=5.0, use_all=True, method="ark4ip", traces=4,
                      small_signal_gain=15.0, E_sat=5e-3, lamb0=1030.0, bandwidth=40.0, local_error=1e-4)
    return dom, A, fib


@pytest.mark.parametrize("name", ["kerr", "kerr_strong", "raman_ss", "gain"])
def test_engine_reproduces_reference_adaptive_run(P, name):
    ref = _case(name)
    dom, A, fib = _build(P, name)
    assert rel_l2(A, ref["A_in"]) < 1e-15
    out = fib(dom, A.copy())
    st = fib.stepper
    err = rel_l2(out, ref["A_out"])
    print(f"{name}: attempts {st.attempts} (reference {len(ref['h'])}), accepted {st.accepted} "
          f"(reference {int(ref['accepted'].sum())}), rel L2 {err:.2e}", flush=True)
    assert st.attempts == len(ref["h"])
    assert st.accepted == int(ref["accepted"].sum())
    s = st.storage
    assert len(s.z) == len(ref["trace_z"])
    assert np.allclose(s.z, ref["trace_z"], rtol=1e-9, atol=0)
    assert np.allclose([h for _, h in s.step_sizes], ref["step_h"], rtol=1e-9, atol=0)
    assert s.fft_total == int(ref["fft_total"])
    assert err <= FIELD_TOL
