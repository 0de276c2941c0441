"""Pins of the oracle (CPU): the reference's own golden values, closed forms
and invariants, each chosen so that a plausible slip in the restatement (a
dropped term, a wrong sign or index, a transposed operand) fails."""

import math

import numpy as np
import pytest

import oracle as O
from helpers import C_NM_PS, rel_l2

# ---------------------------------------------------------------------------
# pyofss/tests/test_solver.py:63-187 -- bit-exact golden values of the solver
# zoo on dA/dz = 3 exp(-4z) - 2A, A(0) = 1, L = 0.5 (stepper.py:379-390 too).
# ---------------------------------------------------------------------------
GOLDEN = {
    "EULER": (0.72152686293820223, 0.71717199224362427),
    "MIDPOINT": (0.71663952968427402, 0.71669512731047913),
    "RK4": (0.71669567757603803, 0.71669567807363854),
    "BS": (0.71669591977945657, 0.71669567831028291),
    "RKF": (0.71669567807672185, 0.71669567807368773),
    "CK": (0.71669567804234302, 0.71669567807368351),
    "DP": (0.71669567807314427, 0.71669567807368462),
}
A_ANALYTICAL = 0.71669567807368684


def simple(A, z):
    return 3.0 * np.exp(-4.0 * z) - 2.0 * A


@pytest.mark.parametrize("method", sorted(GOLDEN))
def test_solver_golden_values_bit_exact(method):
    for steps, expect in zip((50, 500), GOLDEN[method]):
        st = O.Stepper(traces=1, local_error=1e-6, method=method, f=simple, length=0.5, total_steps=steps)
        assert st(1.0) == expect


def test_solver_defaults_and_lists():
    """test_solver.py:39-60 (other_solvers as in solver.py:54)."""
    with pytest.raises(O.NoDerivativeFunctionError):
        O.Solver()
    s = O.Solver(f=simple)
    assert s.method == s.rk4 and not s.embedded
    assert O.Solver.explicit_solvers == ["euler", "midpoint", "rk4"]
    assert O.Solver.embedded_solvers == ["bs", "rkf", "ck", "dp"]
    assert O.Solver.other_solvers == ["rk4ip", "step_amplifier"]


def test_adaptive_scalar_converges():
    """Adaptive step doubling (stepper.py:247-376) reaches the analytic value."""
    st = O.Stepper(traces=1, local_error=1e-9, method="ARK4", f=simple, length=0.5, total_steps=50)
    assert abs(st(1.0) - A_ANALYTICAL) < 1e-9
    log = st.attempt_log
    assert all(h > 0 for h, _, _ in log)
    # first step is length/total_steps capped at 1e-6 (stepper.py:256-264)
    assert log[0][0] == 1e-6


# ---------------------------------------------------------------------------
# pyofss/tests/test_solver.py:190-284 -- second-order soliton, peak power 4
# after one soliton period, to the decimals the reference asserts.
# ---------------------------------------------------------------------------
class _SolitonF:
    def __init__(self, domain):
        Domega = domain.omega - domain.centre_omega
        self.factor = 1j * np.fft.fftshift(0.5 * -1.0 * Domega ** 2)

    def l(self, A, z):
        return np.fft.fft(self.factor * np.fft.ifft(A))

    def n(self, A, z):
        return 1j * np.abs(A) ** 2 * A

    def linear(self, A, h):
        return np.fft.fft(np.exp(h * self.factor) * np.fft.ifft(A))

    def nonlinear(self, A, h, B):
        return np.exp(h * 1j * np.abs(A) ** 2) * B


@pytest.mark.parametrize("method,decimals", [("SS_SIMPLE", 3), ("SS_SYMMETRIC", 0), ("SS_REDUCED", 3),
                                             ("SS_AGRAWAL", 4), ("SS_SYM_MIDPOINT", 2), ("SS_SYM_RK4", 3),
                                             ("RK4IP", 5)])
def test_soliton_peak_power(method, decimals):
    domain = O.Domain(bit_width=200.0, samples_per_bit=2048)
    A_in = O.Sech(peak_power=4.0, width=1.0).generate(domain.t)
    st = O.Stepper(method=method, f=_SolitonF(domain), length=0.5 * np.pi, total_steps=1000)
    peak = max(np.abs(st(A_in)) ** 2)
    assert abs(peak - 4.0) < 1.5 * 10 ** (-decimals)


def order_2_soliton(xi, tau):
    """examples/second_order_soliton/soliton.py:30-39"""
    num = 4.0 * np.exp(0.5 * 1j * xi) * (np.cosh(3.0 * tau) + 3.0 * np.exp(4.0 * 1j * xi) * np.cosh(tau))
    den = np.cosh(4.0 * tau) + 4.0 * np.cosh(2.0 * tau) + 3.0 * np.cos(4.0 * xi)
    return num / den


def _soliton_run(method, steps):
    domain = O.Domain(bit_width=200.0, samples_per_bit=4096)
    sys = O.System(domain)
    sys.add(O.Sech(peak_power=4.0, width=1.0))
    sys.add(O.Fibre("fibre", length=0.5 * np.pi, beta=[0.0, 0.0, -1.0, 0.0], gamma=1.0, method=method,
                    total_steps=steps))
    sys.run()
    return sys.field, order_2_soliton(0.5 * np.pi, domain.t)


def test_soliton_closed_form_rk4ip():
    """The full Fibre path (Linearity, Nonlinearity, Solver, Stepper) on the
    normalised N=2 soliton against its complex closed form at z = pi/2."""
    field, exact = _soliton_run("rk4ip", 2000)
    assert rel_l2(np.abs(field), np.abs(exact)) < 1e-8
    k = np.argmax(np.abs(exact))  # the closed form fixes the field up to a constant phase
    ph = field[k] / exact[k]
    assert rel_l2(field / (ph / abs(ph)), exact) < 1e-7


def test_soliton_ss_symmetric_is_first_order():
    """solver.py:273-275 takes the nonlinear phase from the un-propagated A,
    which makes the 'symmetric' split step first order: doubling the steps
    halves the error (a phase taken from A_L would quarter it)."""
    e1 = rel_l2(np.abs(_soliton_run("ss_symmetric", 1000)[0]), np.abs(_soliton_run("rk4ip", 2000)[1]))
    e2 = rel_l2(np.abs(_soliton_run("ss_symmetric", 2000)[0]), np.abs(_soliton_run("rk4ip", 2000)[1]))
    assert 1.8 < e1 / e2 < 2.2


def test_fixed_trace_rule():
    """stepper.py:206-209: traces after steps 0, 20, ... stored at z = (i+1) h."""
    dom = O.Domain(1, 256, 100.0, C_NM_PS / 1550.0)
    f = O.Fibre(length=1.0, beta=[0.0, 0.0, -20.0], gamma=1.3, method="rk4ip", total_steps=1000, traces=50)
    sys = O.System(dom)
    sys.add(O.Sech(peak_power=1.0, width=1.0))
    sys.add(f)
    sys.run()
    z = np.array(f.stepper.storage.z)
    zs = np.linspace(0.0, 1.0, 1001)
    assert len(z) == 50
    assert np.array_equal(z, zs[0:1000:20] + 1e-3)
    # fft_total: 8 per rk4ip step + one spectral FFT per stored trace
    assert f.stepper.storage.fft_total == 8 * 1000 + 50


# ---------------------------------------------------------------------------
# Transforms, grids, pulses
# ---------------------------------------------------------------------------
PI_L = np.longdouble("3.141592653589793238462643383279502884")


def dft_longdouble(x, sign):
    N = len(x)
    n = np.arange(N)
    m = (np.outer(n, n) % N).astype(np.longdouble)  # exact reduction of k n mod N
    ang = sign * 2 * PI_L * m / N
    xr = np.asarray(x.real, dtype=np.longdouble)
    xi = np.asarray(x.imag, dtype=np.longdouble)
    re = (np.cos(ang) * xr - np.sin(ang) * xi).sum(axis=1)
    im = (np.sin(ang) * xr + np.cos(ang) * xi).sum(axis=1)
    return (re + 1j * im).astype(np.complex128)


@pytest.mark.parametrize("N", [8, 64, 256])
def test_fft_convention_against_brute_force(N):
    """field.py:109-136: fft has sign + and 1/N, ifft sign - and no scaling."""
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert rel_l2(O.fft(x), dft_longdouble(x, +1) / N) < 1e-15
    assert rel_l2(O.ifft(x), dft_longdouble(x, -1)) < 1e-15


def test_domain_golden():
    """pyofss/tests/test_domain.py:34-151 (the stale t[-1] check of :93 is
    replaced by the centred grid of domain.py:218)."""
    d = O.Domain()
    assert (d.total_bits, d.samples_per_bit, d.bit_width, d.centre_nu, d.channels) == (1, 512, 100.0, 193.1, 1)
    d = O.Domain(2, 32, 200.0, 193.2)
    assert d.total_samples == 64 and d.window_t == 400.0
    assert abs(d.centre_omega - 1213.911401347) < 1.5e-9
    assert abs(d.centre_lambda - 1551.720797) < 1.5e-6
    d = O.Domain(1, 64, 400.0, 193.3)
    assert d.dt == 400.0 / 64
    d = O.Domain(4, 16, 50.0, 190.0)
    assert d.t[0] == -d.window_t / 2 and d.nu[d.total_samples // 2] == d.centre_nu
    expected = ["Domain:", "total_bits = 2", "samples_per_bit = 128", "bit_width = 25.0000 ps",
                "centre_nu = 193.8000 THz", "total_samples = 256", "window_t = 50.0000 ps",
                "centre_omega = 1217.6813 rad / ps", "centre_lambda = 1546.9167 nm", "dt = 0.1953 ps",
                "window_nu = 5.1200 THz", "dnu = 0.0200 THz", "window_omega = 32.1699 rad / ps",
                "domega = 0.1257 rad / ps", "window_lambda = 40.8680 nm", "dlambda = 0.1596 nm", "channels = 2"]
    assert str(O.Domain(2, 128, 25.0, 193.8, 2)) == "\n\t".join(expected)
    for kw in (dict(total_bits=0), dict(samples_per_bit=262144), dict(bit_width=0.01), dict(centre_nu=600.0)):
        with pytest.raises(O.OutOfRangeError):
            O.Domain(**kw)


def test_pulse_golden():
    """test_sech.py:88-108, test_gaussian.py:92-117"""
    assert O.Sech(width=100.0).calculate_fwhm() == 176.2747174039086
    assert O.Gaussian(width=100.0).calculate_fwhm() == 166.51092223153952
    assert O.Gaussian(width=100.0, m=3).calculate_fwhm() == 188.14856305595509
    assert str(O.Sech("sech", 0.2, 5.0, 1.4, 0.3, 0, 0.4, 0.0, 1)) == "\n".join(
        ["position = 0.200000", "width = 5.000000 ps", "fwhm = 8.813736 ps", "peak_power = 1.400000 W",
         "offset_nu = 0.300000 THz", "C = 0.400000", "initial_phase = 0.000000 rad", "channel = 1"])
    assert str(O.Gaussian("gaussian", 0.2, 5.0, 1.4, 0.3, 4, 0.4, 0.0, 1)) == "\n".join(
        ["position = 0.200000", "width = 5.000000 ps", "fwhm = 9.552195 ps", "peak_power = 1.400000 W",
         "offset_nu = 0.300000 THz", "m = 4", "C = 0.400000", "initial_phase = 0.000000 rad", "channel = 1"])
    t = np.arange(0.0, 100.0)
    assert abs(max(abs(O.Sech("sech", 0.4, 5.0, 1.5, 0.3, 0, 1.5).generate(t)) ** 2) - 1.5) < 1.5e-7
    assert abs(max(abs(O.Gaussian("gaussian", 0.4, 5.0, 1.5, 0.3, 4, 1.5, 0).generate(t)) ** 2) - 1.5) < 1.5e-7


# ---------------------------------------------------------------------------
# Operators
# ---------------------------------------------------------------------------
def test_linear_operator_gaussian_dispersion_closed_form():
    """Dispersion only: a Gaussian exp(-t^2/(2 T0^2)) becomes
    T0/sqrt(T0^2 - i b2 z) exp(-t^2 / (2 (T0^2 - i b2 z)))."""
    dom = O.Domain(1, 4096, 200.0, 193.1)
    T0, b2, z = 1.0, -20.0, 0.2
    A = np.exp(-dom.t ** 2 / (2 * T0 ** 2)).astype(complex)
    lin = O.Linearity(beta=[0.0, 0.0, b2])
    lin(dom)
    got = lin.default_exp_f(A, z)
    q = T0 ** 2 - 1j * b2 * z
    exact = T0 / np.sqrt(q) * np.exp(-dom.t ** 2 / (2 * q))
    assert rel_l2(got, exact) < 1e-12
    # energy is conserved to rounding with no loss
    assert abs(np.sum(abs(got) ** 2) / np.sum(abs(A) ** 2) - 1) < 1e-13
    # loss alpha halves the amplitude exponent: |A|^2 scales by exp(-alpha z)
    lin2 = O.Linearity(alpha=0.3, beta=[0.0, 0.0, b2])
    lin2(dom)
    assert abs(np.sum(abs(lin2.default_exp_f(A, z)) ** 2) / np.sum(abs(A) ** 2) - math.exp(-0.3 * z)) < 1e-12


def test_spm_closed_form():
    """SPM only: ss_symmetric is exact, A exp(i gamma |A|^2 z); rk4ip is
    fourth order."""
    dom = O.Domain(1, 512, 100.0, 193.1)
    A = O.Sech(peak_power=2.0, width=3.0).generate(dom.t)
    z = 0.7
    exact = A * np.exp(1j * 1.5 * np.abs(A) ** 2 * z)
    for method, tol in (("ss_symmetric", 1e-13), ("rk4ip", 1e-8)):
        sys = O.System(dom, A.copy())
        sys.add(O.Fibre(length=z, beta=None, gamma=1.5, method=method, total_steps=200))
        sys.run()
        assert rel_l2(sys.field, exact) < tol


def test_raman_convolution_brute_force():
    """ifft(h_R~ * fft(P)) with h_R~ = fft(h_R) * T is the causal circular
    convolution dt * sum_j h_R(t_j - t_0) P[(i - j) mod N]."""
    dom = O.Domain(1, 128, 2.0, 300.0)
    nl = O.Nonlinearity(gamma=1.0, use_all=True)
    nl(dom)
    rng = np.random.default_rng(1)
    P = rng.random(128)
    conv = O.ifft(nl.h_R * O.fft(P))
    h = O.calculate_raman_term(dom)
    brute = np.array([dom.dt * sum(h[j] * P[(i - j) % 128] for j in range(128)) for i in range(128)])
    assert rel_l2(conv, brute) < 1e-13


def test_raman_response_unit_area():
    """Blow-Wood h_R integrates to one. The DC term h~(0) = (1/N) sum h * T =
    sum h(n dt) dt is the rectangle rule on [0, inf) with h(0) = 0, so by
    Euler-Maclaurin it equals 1 - dt^2 h'(0)/12 + dt^4 h'''(0)/720 + O(dt^6).
    With h = A exp(-a t) sin(b t), a = 1/tau2, b = 1/tau1:
    h'(0) = A b and h'''(0) = A (3 a^2 b - b^3)."""
    t1, t2 = 12.2e-3, 32.0e-3
    amp = (t1 ** 2 + t2 ** 2) / (t1 * t2 ** 2)
    a, b = 1 / t2, 1 / t1
    for N in (1 << 13, 1 << 14):
        dom = O.Domain(1, N, 20.0, 300.0)
        nl = O.Nonlinearity(gamma=1.0, use_all=True)
        nl(dom)
        dt = dom.dt
        expect = 1.0 - dt ** 2 * amp * b / 12 + dt ** 4 * amp * (3 * a * a * b - b ** 3) / 720
        assert abs(nl.h_R[0].real - expect) < 1e-8
        assert abs(nl.h_R[0].imag) < 1e-12


def test_self_steepening_pure_tone():
    """ifft(omega * fft(x)) of a grid tone exp(-2 pi i m n / N) is omega_m x,
    with omega_m = 2 pi m dnu (FFT order of nonlinearity.py:162)."""
    dom = O.Domain(1, 256, 10.0, 300.0)
    nl = O.Nonlinearity(gamma=1.0, self_steepening=True)
    nl(dom)
    n = np.arange(256)
    for m in (1, 7, 100, 200):
        x = np.exp(-2j * np.pi * m * n / 256)
        y = O.ifft(nl.omega * O.fft(x))
        om = 2 * np.pi * dom.dnu * (m if m < 128 else m - 256)
        assert rel_l2(y, om * x) < 1e-12


def test_full_nonlinearity_reductions():
    """default_f_all at f_R = 0 equals Kerr + SS (default_f_with_ss) and, at
    ss = 0, plain Kerr; at f_R = 1 on a CW field it equals Kerr with the
    Raman response's unit area."""
    dom = O.Domain(1, 1024, 12.5, 359.0)
    A = O.Sech(peak_power=1e3, width=0.05).generate(dom.t)
    all0 = O.Nonlinearity(gamma=110.0, self_steepening=True, use_all=True, f_R=0.0)
    all0(dom)
    ss = O.Nonlinearity(gamma=110.0, self_steepening=True)
    ss(dom)
    assert rel_l2(all0.non(A, 0), ss.non(A, 0)) < 1e-13
    allk = O.Nonlinearity(gamma=110.0, use_all=True, f_R=0.0)
    allk(dom)
    kerr = O.Nonlinearity(gamma=110.0)
    kerr(dom)
    assert rel_l2(allk.non(A, 0), kerr.non(A, 0)) < 1e-13
    cw = np.full(1024, 3.0 + 1j, complex)
    r1 = O.Nonlinearity(gamma=2.0, use_all=True, f_R=1.0)
    r1(dom)
    assert rel_l2(r1.non(cw, 0), 2j * abs(cw) ** 2 * cw * r1.h_R[0].real) < 1e-12


def test_photon_number_conserved_with_raman_and_ss():
    """With Raman and self-steepening (no loss) the GNLSE conserves the photon
    number sum |A~|^2 / omega_abs while the energy falls (Raman red shift).
    The discrete drift is the integrator's error: small, and shrinking at
    fourth order or better as the step count doubles."""
    dom = O.Domain(1, 4096, 12.5, C_NM_PS / 835.0)
    A0 = O.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(dom.t)
    w = O.fftshift(dom.omega)

    def photons(A):
        return np.sum(np.abs(O.fft(A)) ** 2 / w)

    drift = []
    for steps in (400, 800):
        sys = O.System(dom, A0.copy())
        sys.add(O.Fibre(length=2e-5, beta=[0.0, 0.0, -11.83, 0.081], gamma=110.0, self_steepening=True,
                        use_all=True, method="rk4ip", total_steps=steps))
        sys.run()
        drift.append(abs(photons(sys.field) / photons(A0) - 1))
    assert drift[1] < 1e-6 and drift[0] / drift[1] > 12
    assert np.sum(abs(sys.field) ** 2) / np.sum(abs(A0) ** 2) - 1 < -1e-3


def test_amplifier_small_signal_closed_form():
    """Unsaturated gain with no dispersion or nonlinearity: the spectral power
    grows by 10^(0.1 G profile) over the amplifier length."""
    dom = O.Domain(1, 1024, 40.0, C_NM_PS / 1030.0)
    A = O.Gaussian(peak_power=1.0, width=0.2, using_fwhm=True).generate(dom.t)
    sys = O.System(dom, A.copy())
    fib = O.Fibre(length=5e-3, beta=None, gamma=0.0, small_signal_gain=15.0, E_sat=1e30, lamb0=1030.0,
                  bandwidth=40.0, method="ss_symmetric", total_steps=10)
    sys.add(fib)
    sys.run()
    prof = fib.amplifier.spectal_filtration_array
    ratio = np.abs(O.fft(sys.field)) ** 2 / np.abs(O.fft(A)) ** 2
    sel = np.abs(O.fft(A)) ** 2 > 1e-8 * np.max(np.abs(O.fft(A)) ** 2)
    assert np.allclose(ratio[sel], 10 ** (1.5 * prof[sel]), rtol=1e-9)


def test_amplifier_saturation_uses_energy_of_passed_array():
    """amplifier.py:149-158: the saturation energy is that of the array given."""
    dom = O.Domain(1, 256, 40.0, C_NM_PS / 1030.0)
    amp = O.Amplifier(gain=10.0, E_sat=2.0, length=1.0)
    amp.set_domain(dom)
    A = O.Gaussian(peak_power=100.0, width=0.5).generate(dom.t)
    E = O.energy(A, dom.t)
    f = amp.factor(A, 0.1)
    assert np.allclose(f, np.log(10.0) * 0.1 / 2.0 / (1.0 + E / 2.0))
    assert not np.allclose(amp.factor(2 * A, 0.1), f)


@pytest.mark.parametrize("N", [9, 65])
def test_old_simpson_rule_odd(N):
    """Odd N: the three-point rule is exact for quadratics on any grid and for
    cubics on a uniform grid."""
    x = np.sort(np.random.default_rng(N).random(N)) * 3.0
    x[0], x[-1] = 0.0, 3.0
    q = 1 - 2 * x + 0.5 * x ** 2
    assert abs(O.simps_avg(q, x) - (3 - 9 + 4.5)) < 1e-12
    u = np.linspace(0.0, 3.0, N)
    c = 1 - 2 * u + 0.5 * u ** 2 - 0.25 * u ** 3
    assert abs(O.simps_avg(c, u) - (3 - 9 + 4.5 - 81 / 16)) < 1e-12


@pytest.mark.parametrize("N", [10, 64])
def test_old_simpson_rule_even_average(N):
    """Even N (scipy < 1.11, even='avg'): the mean of [Simpson on 0..N-2 +
    trapezoid on the last interval] and [trapezoid on the first interval +
    Simpson on 1..N-1]; exact for linear data, and for a quadratic off by
    the trapezoid end errors -(h^3/12) y''."""
    u = np.linspace(0.0, 3.0, N)
    h = u[1] - u[0]
    assert abs(O.simps_avg(2 - u, u) - (6 - 4.5)) < 1e-12
    q = u ** 2
    assert abs(O.simps_avg(q, u) - (9.0 + h ** 3 / 6)) < 1e-12


# ---------------------------------------------------------------------------
# Build-only parts: one-photon-per-mode noise, |g12|
# ---------------------------------------------------------------------------
def test_oppm_noise_parseval_and_seed():
    dom = O.Domain(1, 4096, 100.0, C_NM_PS / 1550.0)
    a = O.oppm_noise(dom, 7)
    b = O.oppm_noise(dom, 7)
    c = O.oppm_noise(dom, 8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    spec = O.fft(a)
    nu = O.fftshift(dom.nu)
    # |A~_k|^2 = h nu_k / (N dt) exactly per mode; Parseval in pyofss units
    assert np.allclose(np.abs(spec) ** 2, 6.626e-10 * nu / (4096 * dom.dt), rtol=1e-10)
    assert abs(np.sum(np.abs(a) ** 2) * dom.dt / (4096 * dom.dt * np.sum(np.abs(spec) ** 2)) - 1) < 1e-12


def test_coherence_closed_forms():
    rng = np.random.default_rng(3)
    base = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    same = np.tile(base, (10, 1))
    assert np.allclose(O.coherence_g12(same), 1.0)
    M = 2000
    phases = np.exp(2j * np.pi * rng.random((M, 64)))
    g = O.coherence_g12(phases)
    assert np.mean(g) < 5.0 / M ** 0.5
