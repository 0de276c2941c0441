"""GPU parity on the large grids, 2^15 and 2^16 samples (the two-level
transform and slot-resident fields of csrc/big.hpp), against the oracle
(complex128, relative L2 unless stated).

* FFT (pyofss field.fft / field.ifft) against numpy's pocketfft: <= 1e-15,
  round trip <= 1e-15; impulse and pure-tone known answers to 1e-13
  absolute.
* exp_lin (with and without the saturable gain), non, exp_non for every
  nonlinear variant, one rk4ip / ss_symmetric call: <= 1e-12 (exp_non on a
  field with a weak CW background, as in test_gpu_ops.py); lin within twice
  its conditioning bound.
* Fixed-step rk4ip with the physics of
  examples/notebooks/mamyshev_oscillator_cl.ipynb (2^15 samples over 200 ps
  at 1035 nm, 50 W 3 ps Gaussian with C = 100, Hollenbeck Raman, steps of
  1e-7 km; a passive fibre and an active fibre with the two-level Yb gain)
  through System/Fibre: output and every trace <= 1e-9, pump power <= 1e-9.
* Adaptive RK4IP with the supercontinuum physics over 2 mm: identical
  attempt sequence, step sizes to 1e-6, field <= 1e-9.
* Trace characteristics against the reference's scipy computations, as in
  test_gpu_characts.py.
"""

import math

import numpy as np
import pytest

import oracle as O
from helpers import C_NM_PS, engine_context, nl_kind, oracle_amplifier, oracle_ops, rel_l2, supercontinuum_c2, \
    amplifier_c4

pytestmark = pytest.mark.gpu

GRIDS = [32768, 65536]
OP_TOL = 1e-12
FIELD_TOL = 1e-9


@pytest.fixture(scope="module")
def lib():
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return _lib


def test_large_grids_supported(lib):
    assert lib.grid_supported(32768) and lib.grid_supported(65536)
    assert not lib.grid_supported(131072)


@pytest.mark.parametrize("N", GRIDS)
def test_fft_large(lib, N):
    rng = np.random.default_rng(N)
    A = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    ctx = lib.Context(0)
    ctx.set_grid(N)
    print(f"large fft N={N}", flush=True)
    f = ctx.fft(A, +1)
    b = ctx.fft(A, -1)
    ef, eb = rel_l2(f, O.fft(A)), rel_l2(b, O.ifft(A))
    er = rel_l2(ctx.fft(f, -1), A)
    print(f"N={N} fft err {ef:.2e} ifft err {eb:.2e} round trip {er:.2e}", flush=True)
    assert ef <= 1e-15 and eb <= 1e-15 and er <= 1e-15
    n = np.arange(N, dtype=np.int64)
    for m in (0, 1, 8191, 8192, N // 2 + 3, N - 1):
        tone = np.exp(-2j * np.pi * ((m * n) % N) / N)  # exact phase reduction
        X = ctx.fft(tone, +1)
        expect = np.zeros(N, complex)
        expect[m] = 1.0
        assert np.max(np.abs(X - expect)) < 1e-13
        d = np.zeros(N, complex)
        d[m] = 1.0
        assert np.max(np.abs(ctx.fft(d, -1) - tone)) < 1e-13


@pytest.mark.parametrize("N", GRIDS)
def test_linear_operator_large(lib, N):
    dom, A, fib = supercontinuum_c2(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], alpha=0.05, gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    for h in (1e-6, 3.3e-7):
        err = rel_l2(ctx.apply_linear(A, h), lin.default_exp_f(A, h))
        print(f"N={N} h={h} linear err {err:.2e}", flush=True)
        assert err < OP_TOL
    err = rel_l2(ctx.apply_linear_factor(A), lin.default_f(A, 0.0))
    At = O.fft(A)
    bound = np.finfo(float).eps * np.max(np.abs(lin.factor)) * np.linalg.norm(At) / np.linalg.norm(lin.factor * At)
    print(f"N={N} lin err {err:.2e} (conditioning bound {bound:.2e})", flush=True)
    assert err < max(1e-12, 2 * bound)


def test_linear_operator_with_gain_large(lib):
    N = 32768
    dom, A, fib = amplifier_c4(N, C=1.0)
    amp = oracle_amplifier(dom, fib["length"], **fib["amplifier"])
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], use_all=True, amplifier=amp)
    ctx = engine_context(dom, lin, nl, 1, amplifier=amp)
    for scale in (1.0, 3.0):
        for h in (1e-5, 4e-4):
            err = rel_l2(ctx.apply_linear(scale * A, h), lin.default_exp_f(scale * A, h))
            print(f"N={N} scale={scale} h={h} gain linear err {err:.2e}", flush=True)
            assert err < OP_TOL


KINDS = [dict(), dict(use_all=True), dict(use_all=True, self_steepening=True), dict(self_steepening=True),
         dict(raman_scattering=True), dict(self_steepening=True, raman_scattering=True)]


@pytest.mark.parametrize("flags", KINDS, ids=lambda d: "+".join(sorted(d)) or "kerr")
@pytest.mark.parametrize("N", GRIDS)
def test_nonlinear_operator_large(lib, flags, N):
    dom, A, fib = supercontinuum_c2(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], **flags)
    ctx = engine_context(dom, lin, nl, nl_kind(**flags))
    err = rel_l2(ctx.apply_nonlinear(A), nl.non(A, 0.0))
    print(f"N={N} {flags} non err {err:.2e}", flush=True)
    assert err < OP_TOL
    Ab = A + 1e-2 * np.sqrt(1e4)
    B = lin.default_exp_f(Ab, 2e-7)
    err = rel_l2(ctx.apply_exp_nonlinear(Ab, 2e-7, B), nl.exp_non(Ab, 2e-7, B))
    print(f"N={N} {flags} exp_non err {err:.2e}", flush=True)
    assert err < OP_TOL


@pytest.mark.parametrize("method", ["rk4ip", "ss_symmetric"])
@pytest.mark.parametrize("flags", [dict(), dict(use_all=True, self_steepening=True)], ids=["kerr", "all+ss"])
def test_one_integrator_call_large(lib, method, flags):
    dom, A, fib = supercontinuum_c2(32768)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], **flags)
    ctx = engine_context(dom, lin, nl, nl_kind(**flags))

    class F:
        pass

    f = F()
    f.linear = lin.default_exp_f
    f.n = nl.non
    f.nonlinear = nl.exp_non
    h = 5e-7
    ref = getattr(O.Solver, method)(A, 0.0, h, f)
    err = rel_l2(ctx.step(0 if method == "rk4ip" else 1, A, h), ref)
    print(f"32768 {method} {flags} step err {err:.2e}", flush=True)
    assert err < OP_TOL


@pytest.fixture(scope="module")
def P():
    import pyofsschanged_amd as P
    return P


def mamyshev(mod, active, steps):
    """The notebook's grid, pulse and first passive / active fibre
    (mamyshev_oscillator_cl.ipynb cells 12, 15, 19), shortened to `steps`
    steps of its 1e-7 km step size."""
    lam0 = 1035.0
    dom = mod.Domain(samples_per_bit=2 ** 15, bit_width=200.0, total_bits=1, centre_nu=mod.lambda_to_nu(lam0))
    A = O.Gaussian(peak_power=50.0, width=3.0, C=100.0, using_fwhm=True).generate(dom.t)
    lr = 2 * 2.5e-3 + 4 * 0.8e-3  # round-trip length, km
    Tr = lr * 1e12 / O.VACUUM_LIGHT_SPEED  # ps
    omega0 = O.lambda_to_omega(lam0)
    kw = dict(length=1e-7 * steps, total_steps=steps, traces=10, method="rk4ip", self_steepening=False,
              use_all="hollenbeck")
    if active:
        kw.update(beta=np.array([0, 0, 24.9]), gamma=O.calculate_gamma(2.7e-20, 29, omega0), use_Yb_model=True,
                  Pp_0=10, N=4.2, Rr=1 / Tr)
    else:
        kw.update(beta=np.array([0, 0, 22.2]), gamma=O.calculate_gamma(2.7e-20, 37.8, omega0))
    return dom, A, mod.Fibre(**kw)


@pytest.mark.parametrize("active", [False, True], ids=["passive", "active_yb"])
def test_mamyshev_fixed_steps_32768(lib, P, active):
    steps = 150
    od, A, fo = mamyshev(O, active, steps)
    pd_, A2, fp = mamyshev(P, active, steps)
    assert rel_l2(A2, A) == 0.0
    print(f"mamyshev {'active' if active else 'passive'} 2^15, {steps} steps", flush=True)
    ref = fo(od, A.copy())
    got = fp(pd_, A.copy())
    err = rel_l2(got, ref)
    so, sp = fo.stepper.storage, fp.stepper.storage
    tr_err = max(rel_l2(a, b) for a, b in zip(sp.As, so.As))
    print(f"  rel L2 {err:.2e}, traces {len(sp.As)} (max err {tr_err:.2e})", flush=True)
    assert err <= FIELD_TOL
    assert len(sp.As) == len(so.As) == 10 and tr_err <= FIELD_TOL
    assert np.array_equal(sp.z, so.z)
    if active:
        dp = abs(fp.amplifier.Pp / fo.amplifier.Pp - 1)
        print(f"  pump {fp.amplifier.Pp:.9e} (oracle {fo.amplifier.Pp:.9e}), rel {dp:.2e}", flush=True)
        assert dp <= 1e-9


@pytest.mark.parametrize("N", GRIDS)
def test_adaptive_large(lib, N):
    dom, A, fib = supercontinuum_c2(N)
    fib["length"] = 2.0e-6
    f = O.Fibre(length=fib["length"], beta=fib["beta"], gamma=fib["gamma"], method="ark4ip", local_error=1e-6,
                self_steepening=True, use_all=True)
    ref = f(dom, A.copy())
    log = f.stepper.attempt_log
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    p = lib.AdaptiveParams()
    p.length, p.h0, p.local_error = fib["length"], min(fib["length"] / 100, 1e-6), 1e-6
    p.safety, p.max_factor, p.min_factor, p.step_size_min = 0.9, 10.0, 0.2, 1e-37
    p.traces, p.total_attempts, p.steps_max, p.eta = 1, 1000, 100000, 5
    out = ctx.propagate_adaptive(0, A, p, log_cap=len(log) + 8)
    r = out["results"][0]
    err = rel_l2(out["A"], ref)
    print(f"adaptive N={N}: attempts {r['n_attempts']} (oracle {len(log)}), rel L2 {err:.3e}, "
          f"kernel {ctx.last_kernel_ms():.1f} ms", flush=True)
    assert r["status"] == 0 and r["n_attempts"] == len(log)
    ref_h = np.array([a[0] for a in log])
    assert np.array_equal(out["log_accept"][0, :len(log)], np.array([a[2] for a in log], dtype=np.uint8))
    assert np.max(np.abs(out["log_h"][0, :len(log)] / ref_h - 1)) < 1e-6
    assert err <= FIELD_TOL


@pytest.mark.parametrize("N", GRIDS)
def test_characteristics_large(lib, N):
    from test_gpu_characts import check, expected, fields
    A = fields(N, N + 1)
    dom = O.Domain(1, N, 100.0, C_NM_PS / 1550.0)
    from pyofsschanged_amd.field import simpson_weights
    ctx = lib.Context(0)
    ctx.set_grid(N)
    got = ctx.trace_characteristics(A, simpson_weights(dom.t * 1e-3))
    for k in range(A.shape[0]):
        check(got, k, expected(A[k], dom.t))
