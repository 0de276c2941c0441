"""GPU tests of the pyofss-shaped API (System, Sech, Fibre, Stepper,
Solver, Linearity, Nonlinearity, field) against the oracle's classes on the
same inputs. Field tolerance: relative L2 <= 1e-9 (complex128)."""

import math

import numpy as np
import pytest

import oracle as O
from helpers import C_NM_PS, DUDLEY_BETA, rel_l2

pytestmark = pytest.mark.gpu
FIELD_TOL = 1e-9


@pytest.fixture(scope="module")
def P():
    import pyofsschanged_amd as P
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return P


def run_both(P, domain_args, pulse_kw, fibre_kw, pulse="Sech"):
    sys_p = P.System(P.Domain(*domain_args))
    sys_p.add(getattr(P, pulse)(**pulse_kw))
    fp = P.Fibre(**fibre_kw)
    sys_p.add(fp)
    sys_p.run()
    sys_o = O.System(O.Domain(*domain_args))
    sys_o.add(getattr(O, pulse)(**pulse_kw))
    fo = O.Fibre(**{k: v for k, v in fibre_kw.items() if k not in ("dir", "cycle")})
    sys_o.add(fo)
    sys_o.run()
    return sys_p, fp, sys_o, fo


@pytest.mark.parametrize("method", ["rk4ip", "ss_symmetric"])
def test_system_soliton_fixed(P, method):
    sys_p, fp, sys_o, fo = run_both(
        P, (1, 4096, 100.0, C_NM_PS / 1550.0), dict(peak_power=4 * 20.0 / 1.3, width=1.0),
        dict(length=math.pi / 40, beta=[0.0, 0.0, -20.0], gamma=1.3, method=method, total_steps=1000, traces=50))
    assert rel_l2(sys_p.field, sys_o.field) <= FIELD_TOL
    sp, so = fp.stepper.storage, fo.stepper.storage
    assert np.array_equal(sp.z, so.z) and len(sp.As) == 50
    assert max(rel_l2(a, b) for a, b in zip(sp.As, so.As)) <= FIELD_TOL
    assert sp.fft_total == so.fft_total == (8 if method == "rk4ip" else 4) * 1000 + 50
    # per-trace characteristics and the result frame (storage.py:194-302)
    assert len(sp.energy_list) == 50 and sys_p.df_results is not None and len(sys_p.df_results) == 50
    assert abs(sp.energy_list[0] - O.energy(so.As[0], sys_o.domain.t)) < 1e-12 * abs(sp.energy_list[0])


def test_system_supercontinuum_adaptive_traces(P):
    sys_p, fp, sys_o, fo = run_both(
        P, (1, 2048, 12.5, C_NM_PS / 835.0), dict(peak_power=1e4, width=0.050, using_fwhm=True),
        dict(length=8e-6, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True, method="ark4ip",
             local_error=1e-6, traces=10))
    err = rel_l2(sys_p.field, sys_o.field)
    print(f"adaptive Raman+SS via System: rel L2 {err:.2e}, attempts {fp.stepper.attempts}", flush=True)
    assert err <= FIELD_TOL
    sp, so = fp.stepper.storage, fo.stepper.storage
    assert len(sp.z) == len(so.z) and np.allclose(sp.z, so.z, rtol=1e-8, atol=0)
    assert len(sp.step_sizes) == len(so.step_sizes)
    assert np.allclose(np.array(sp.step_sizes), np.array(so.step_sizes), rtol=1e-8)
    assert max(rel_l2(a, b) for a, b in zip(sp.As, so.As)) <= FIELD_TOL
    assert sp.fft_total == so.fft_total


def test_gaussian_adaptive_ss_symmetric(P):
    sys_p, fp, sys_o, fo = run_both(
        P, (1, 1024, 40.0, C_NM_PS / 1030.0), dict(peak_power=200.0, width=0.2, using_fwhm=True, C=1.0),
        dict(length=2e-3, beta=[0.0, 0.0, 23.0], gamma=5.0, method="ass_symmetric", local_error=1e-6),
        pulse="Gaussian")
    assert rel_l2(sys_p.field, sys_o.field) <= FIELD_TOL
    assert fp.stepper.attempts == len(fo.stepper.attempt_log)


def test_operator_api_and_solver(P):
    """Fibre.linear / n / nonlinear / l and Solver(method, f)(A, z, h)."""
    d = (1, 4096, 12.5, C_NM_PS / 835.0)
    dp, do = P.Domain(*d), O.Domain(*d)
    fp = P.Fibre(length=1e-5, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True, method="rk4ip")
    fo = O.Fibre(length=1e-5, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True, method="rk4ip")
    fp.prepare(dp)
    fo.linearity(do)
    fo.nonlinearity(do)
    A = O.Sech(peak_power=1e4, width=0.05, using_fwhm=True).generate(do.t)
    assert rel_l2(fp.n(A, 0.0), fo.n(A, 0.0)) < 1e-12
    assert rel_l2(fp.linear(A, 1e-7), fo.linear(A, 1e-7)) < 1e-12
    B = fo.linear(A, 1e-7)
    assert rel_l2(fp.nonlinear(A, 1e-7, B), fo.nonlinear(A, 1e-7, B)) < 1e-10
    assert rel_l2(fp.l(A, 0.0), fo.l(A, 0.0)) < 1e-10
    got = P.Solver("rk4ip", fp.function_step)(A, 0.0, 1e-7)
    ref = O.Solver("rk4ip", fo.function_step)(A, 0.0, 1e-7)
    assert rel_l2(got, ref) < 1e-12


def test_field_fft_on_device(P):
    rng = np.random.default_rng(9)
    A = rng.standard_normal(8192) + 1j * rng.standard_normal(8192)
    P.field.fft_counter = 0
    assert rel_l2(P.fft(A), O.fft(A)) < 1e-14
    assert rel_l2(P.ifft(A), O.ifft(A)) < 1e-14
    assert P.field.fft_counter == 2


def test_stepper_raises_small_step_on_nan(P):
    d = P.Domain(1, 1024, 100.0, C_NM_PS / 1550.0)
    f = P.Fibre(length=0.01, beta=[0.0, 0.0, -20.0], gamma=1.3, method="ark4ip")
    f.prepare(d)
    A = P.Sech(peak_power=1.0, width=1.0).generate(d.t)
    A[3] = np.nan
    with pytest.raises(P.SmallStepSizeError):
        f.stepper(A, None)
