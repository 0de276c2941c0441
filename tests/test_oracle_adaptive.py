"""Pins of the oracle's adaptive stepper (stepper.py:247-376) and saturable
gain (amplifier.py:149-158), CPU only.

* Reference fixtures: tests/golden/adaptive_fixtures.npz holds, for seven
  adaptive propagations run by the reference's own stepper.py (written by
  tests/golden/make_adaptive_fixtures.py), every attempt's step size,
  relative local error and acceptance, the trace positions, the stored step
  sizes and the output field. The oracle must reproduce them: bit for bit on
  the scalar ODE (no transforms), to transform rounding on the fibres.
* Planted mistakes: each mutation of the controller a transcription could
  plausibly make is applied to the oracle's source here, on the CPU, and
  must fail the fixture comparison.
* Mathematics: with replayed steps, the locally extrapolated result
  converges at fifth order while the fine result alone converges at fourth;
  flat saturated gain converges to the Lambert-W solution of
  dE/dz = g0 E / (1 + E/E_sat).
"""

import inspect
import os
import textwrap
import types

import numpy as np
import pytest
from scipy.special import lambertw

import oracle as O
import oracle.pyofss_np as ON
from helpers import C_NM_PS, DUDLEY_BETA, rel_l2

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "adaptive_fixtures.npz"))
SCALAR = ("scalar_1e-6", "scalar_1e-9", "scalar_1e-3")
FIBRES = ("kerr", "kerr_strong", "raman_ss", "gain")
TOL = {"scalar_1e-6": 1e-6, "scalar_1e-9": 1e-9, "scalar_1e-3": 1e-3, "kerr": 1e-4, "kerr_strong": 1e-4,
       "raman_ss": 1e-6, "gain": 1e-4}


def simple(A, z):
    return 3.0 * np.exp(-4.0 * z) - 2.0 * A


def _case(name):
    return {k.split("/", 1)[1]: FIX[k] for k in FIX.files if k.split("/", 1)[0] == name}


def _scalar_stepper(name):
    tol = float(name.split("_")[1])
    return O.Stepper(traces=1, local_error=tol, method="ARK4", f=simple, length=0.5, total_steps=50)


def _fibre(name):
    """The oracle's domain, input field and fibre of a fixture case (same
    arguments as make_adaptive_fixtures.py)."""
    if name in ("kerr", "kerr_strong"):
        dom = O.Domain(1, 256, 20.0, C_NM_PS / 1550.0)
        P0, L, traces, steps = (4 * 20.0 / 1.3, 0.08, 4, 100) if name == "kerr" else (1e3, 0.01, 2, 1)
        A = O.Sech(peak_power=P0, width=1.0).generate(dom.t)
        fib = O.Fibre(length=L, beta=[0.0, 0.0, -20.0], gamma=1.3, method="ark4ip", traces=traces,
                      total_steps=steps, local_error=1e-4)
    elif name == "raman_ss":
        dom = O.Domain(1, 256, 2.0, C_NM_PS / 835.0)
        A = O.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(dom.t)
        fib = O.Fibre(length=5e-6, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True,
                      method="ark4ip", traces=2, total_steps=1, local_error=1e-6)
    else:
        dom = O.Domain(1, 256, 10.0, C_NM_PS / 1030.0)
        A = O.Gaussian(peak_power=1e3, width=0.2, using_fwhm=True).generate(dom.t)
        fib = O.Fibre(length=2e-4, beta=[0.0, 0.0, 23.0], gamma=5.0, use_all=True, method="ark4ip", traces=4,
                      small_signal_gain=15.0, E_sat=5e-3, lamb0=1030.0, bandwidth=40.0, local_error=1e-4)
    return dom, A, fib


def _bind(stepper, adaptive_stepper, name):
    """Replace the method of an oracle Stepper instance by a mutated one,
    stopped one attempt after the fixture's count (a longer run cannot match
    and some mutations would take the full 100000 x 1000 attempt budget)."""
    cap = len(_case(name)["h"]) + 1
    stepper.adaptive_stepper = types.MethodType(
        lambda self, A, ref: adaptive_stepper(self, A, ref, max_attempts=cap), stepper)


def _run(name, adaptive_stepper=None):
    """Run the oracle on a fixture case; ``adaptive_stepper`` replaces the
    method of the oracle's Stepper instance (a planted mistake)."""
    if name in SCALAR:
        st = _scalar_stepper(name)
        if adaptive_stepper is not None:
            _bind(st, adaptive_stepper, name)
        try:
            A = st(1.0)
        except ON.StepperError:
            return None
        return {"A_out": np.array([A]), "log": st.attempt_log, "storage": st.storage}
    dom, A, fib = _fibre(name)
    if adaptive_stepper is not None:
        _bind(fib.stepper, adaptive_stepper, name)
    try:
        out = fib(dom, A.copy())
    except ON.StepperError:
        return None
    return {"A_in": A, "A_out": out, "log": fib.stepper.attempt_log, "storage": fib.stepper.storage}


def _mismatch(name, got):
    """Empty string when the oracle run ``got`` reproduces the fixture, else
    the first difference. Scalar cases: bit for bit. Fibre cases: identical
    acceptance sequence and fft count, step sizes to 1e-9 relative (they
    carry the fifth root of a rounding-level difference of two fields),
    errors to 1e-6, trace positions (sums of step sizes) to 1e-9, output
    field to 1e-10
    (numpy.fft here, scipy.fftpack in the reference: their rounding differs
    and 177 strongly nonlinear attempts carry it to 1.4e-12)."""
    if got is None:
        return "the run raised"
    ref = _case(name)
    h = np.array([a[0] for a in got["log"]])
    d = np.array([a[1] for a in got["log"]])
    acc = np.array([a[2] for a in got["log"]])
    if len(h) != len(ref["h"]):
        return f"{len(h)} attempts, reference {len(ref['h'])}"
    if not np.array_equal(acc, ref["accepted"]):
        return "acceptance sequence"
    exact = name in SCALAR
    if exact:
        if not (np.array_equal(h, ref["h"]) and np.array_equal(d, ref["delta"])):
            return "step sizes / errors not bit-identical"
        if not np.array_equal(got["A_out"], ref["A_out"]):
            return "output"
        return ""
    if not np.allclose(h, ref["h"], rtol=1e-9, atol=0):
        return f"step sizes (max rel {np.max(np.abs(h / ref['h'] - 1)):.2e})"
    if not np.allclose(d, ref["delta"], rtol=1e-6, atol=0):
        return "errors"
    st = got["storage"]
    if not np.allclose(np.array(st.z), ref["trace_z"], rtol=1e-9, atol=0):
        return "trace positions"
    if not np.allclose([hh for _, hh in st.step_sizes], ref["step_h"], rtol=1e-9, atol=0):
        return "stored step sizes"
    if st.fft_total != int(ref["fft_total"]):
        return "fft count"
    if rel_l2(got["A_out"], ref["A_out"]) > 1e-10:
        return f"output field ({rel_l2(got['A_out'], ref['A_out']):.2e})"
    return ""


@pytest.mark.parametrize("name", SCALAR + FIBRES)
def test_oracle_reproduces_reference_adaptive_run(name):
    got = _run(name)
    if name not in SCALAR:
        # the input pulse of the fixture is the reference generator's
        assert rel_l2(got["A_in"], _case(name)["A_in"]) < 1e-15
    assert _mismatch(name, got) == ""


def test_fixtures_discriminate():
    """The fixtures hold what the mutations below need: rejected attempts and
    accepted attempts with tol <= delta < 2 tol."""
    rejected = sum(int((~_case(n)["accepted"]).sum()) for n in SCALAR + FIBRES)
    band = clipped = 0
    for n in SCALAR + FIBRES:
        c = _case(n)
        band += int(np.sum(c["accepted"] & (c["delta"] >= TOL[n])))
        # factor 0.9 (tol/delta)^(1/5) below the clip 0.2
        clipped += int(np.sum(c["delta"] > TOL[n] * (0.9 / 0.2) ** 5))
    assert rejected >= 3 and band >= 10 and clipped >= 1


# source-level mutations of Stepper.adaptive_stepper: (description, old, new)
MUTATIONS = [
    ("accept at delta < tol", "accepted = delta < 2.0 * self.local_error", "accepted = delta < self.local_error"),
    ("no local extrapolation", "self.A_out = f_alpha * A_fine - f_beta * A_coarse", "self.A_out = A_fine"),
    ("safety factor 0.8", "factor = self.safety * np.power", "factor = 0.8 * np.power"),
    ("step exponent 1/4", "np.power(error_ratio, 1.0 / self.eta)", "np.power(error_ratio, 1.0 / (self.eta - 1))"),
    ("extrapolation weights of eta - 1", "f_eta = np.power(2, self.eta - 1.0)", "f_eta = np.power(2, self.eta - 2.0)"),
    ("no lower clip of the factor", "max(self.min_factor, factor)", "factor"),
    ("advance z by the next step", "z += h_temp", "z += h"),
    ("coarse step from the half step", "    A_coarse = self.step(self.A_out, z, h)", "    A_coarse = self.step(A_half, z, h)"),
    ("fine step from the accepted field", "A_fine = self.step(A_half, z_half, h_half)",
     "A_fine = self.step(self.A_out, z_half, h_half)"),
    ("no clip of the last step", "h = self.length - z", "pass"),
]


def _mutated(old, new):
    src = textwrap.dedent(inspect.getsource(ON.Stepper.adaptive_stepper))
    assert src.count(old) == 1, old
    ns = {}
    exec(compile(src.replace(old, new), "<mutated adaptive_stepper>", "exec"), vars(ON), ns)
    return ns["adaptive_stepper"]


@pytest.mark.parametrize("desc,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_controller_mistake_fails_fixtures(desc, old, new):
    fn = _mutated(old, new)
    failures = [n for n in SCALAR + FIBRES if _mismatch(n, _run(n, fn))]
    assert failures, f"mutation '{desc}' reproduces every fixture"


def test_planted_local_error_mistake_fails_fixtures(monkeypatch):
    """relative_local_error (stepper.py:230-245) normalised by the coarse
    field instead of the fine one."""

    def rle(A_fine, A_coarse):
        return np.linalg.norm(A_fine - A_coarse) / np.linalg.norm(A_coarse)

    monkeypatch.setattr(ON.Stepper, "relative_local_error", staticmethod(rle))
    assert any(_mismatch(n, _run(n)) for n in SCALAR)


def _replayed_error(n, adaptive_stepper=None):
    """Global error at z = 0.5 of the scalar ODE after n replayed steps of
    0.5/n, every attempt accepted (local error 1: delta < 2 always)."""
    st = O.Stepper(traces=1, local_error=1.0, method="ARK4", f=simple, length=0.5, total_steps=50)
    if adaptive_stepper is not None:
        st.adaptive_stepper = types.MethodType(adaptive_stepper, st)
    A = st.adaptive_stepper(1.0, None, replay=[0.5 / n] * n)
    assert len(st.attempt_log) == n and all(a for _, _, a in st.attempt_log)
    exact = (5.0 * np.exp(-1.0) - 3.0 * np.exp(-2.0)) / 2.0
    return abs(A - exact)


def test_local_extrapolation_is_fifth_order():
    """Step doubling of rk4 with local extrapolation (16 A_f - A_c)/15
    (stepper.py:272-274, 338) cancels the h^5 term of the local error, so the
    global error falls as h^5; the fine result alone falls as h^4."""
    ns = (8, 16, 32)
    e = [_replayed_error(n) for n in ns]
    order = np.log2(e[0] / e[1]), np.log2(e[1] / e[2])
    assert all(4.7 < o < 5.3 for o in order), order
    fine = _mutated("self.A_out = f_alpha * A_fine - f_beta * A_coarse", "self.A_out = A_fine")
    ef = [_replayed_error(n, fine) for n in ns]
    order_f = np.log2(ef[0] / ef[1]), np.log2(ef[1] / ef[2])
    assert all(3.8 < o < 4.2 for o in order_f), order_f


def _gain_energy(n, G_dB=20.0, L=1e-3, E_sat=0.05):
    """Energy after a lossless, dispersionless, nonlinearity-free fibre with
    a flat saturable gain (no lamb0/bandwidth: profile 1), n fixed rk4ip steps
    (each two linear half steps whose gain follows the energy of their
    input, amplifier.py:149-158 via linearity.py:264-272)."""
    dom = O.Domain(1, 512, 20.0, C_NM_PS / 1030.0)
    A = O.Gaussian(peak_power=100.0, width=1.0, using_fwhm=True).generate(dom.t)
    fib = O.Fibre(length=L, beta=[0.0, 0.0, 0.0], gamma=0.0, method="rk4ip", total_steps=n, traces=1,
                  small_signal_gain=G_dB, E_sat=E_sat)
    out = fib(dom, A.copy())
    return O.energy(A, dom.t), O.energy(out, dom.t)


def test_saturated_gain_lambert_w():
    """dE/dz = g0 E / (1 + E/E_sat), g0 = ln(10^(G/10)) / L, has the closed
    form E(L) = E_sat W((E0/E_sat) exp(E0/E_sat + g0 L)). Each linear half
    step multiplies the energy by exp(g0 (h/2) / (1 + E/E_sat)), an
    exponential Euler step: first order in h, so the Richardson combination
    2 E(2n) - E(n) matches the closed form far more tightly than either."""
    G_dB, L, E_sat = 20.0, 1e-3, 0.05
    g0 = np.log(10 ** (0.1 * G_dB)) / L
    E0, E1 = _gain_energy(100, G_dB, L, E_sat)
    _, E2 = _gain_energy(200, G_dB, L, E_sat)
    exact = E_sat * lambertw((E0 / E_sat) * np.exp(E0 / E_sat + g0 * L)).real
    assert E0 / E_sat > 0.5  # saturated regime: the gain is far below 20 dB
    e1, e2 = abs(E1 / exact - 1), abs(E2 / exact - 1)
    assert 1.9 < e1 / e2 < 2.1  # first order
    assert abs((2 * E2 - E1) / exact - 1) < 1e-2 * e2
    assert e2 < 2e-3
