"""Attempt-by-attempt fixtures of the reference's adaptive stepper.

The adaptive controller of pyofss/modules/stepper.py:247-376 has no golden
value in the reference's own tests (pyofss/tests/test_stepper.py:61-68 is a
stub), so its restatement in oracle/pyofss_np.py is pinned here against the
reference itself: this script loads the unmodified reference modules from
/root/reference (build container only; the reference never travels), runs
seven adaptive propagations and writes every attempt's step size, relative
local error and acceptance, the trace positions, the stored step sizes and
the output field to adaptive_fixtures.npz.

Loading without running pyofss/__init__.py (it imports the OpenCL fibre):
the packages `pyofss` and `pyofss.modules` are created as empty shells whose
search path is the reference directory, pyopencl is an in-memory stub (its
names are only touched by the OpenCL helpers, never on these paths), and the
names SciPy >= 1.14 removed are aliased: scipy.power/sqrt/log10/exp/log to
numpy's, integrate.simps to the old even='avg' rule as restated in
oracle.simps_avg (pinned by tests/test_oracle_pins.py; it is used only for
the amplifier's energy and the trace characteristics, whose one transform
per stored trace counts in fft_total). After a run, `storage.dir_spec = None` lets stepper.py:365
return A_out instead of raising on the attribute that does not exist.

Cases (all "ark4ip" with local error 1e-4 except the scalar ones and raman_ss):
  scalar_*     dA/dz = 3 exp(-4z) - 2A, A(0) = 1, L = 0.5, method "ARK4"
               (pyofss/tests/test_solver.py:63-80), local errors 1e-6, 1e-9, 1e-3
  kerr         256 samples / 20 ps at 1550 nm, N = 2 sech T0 1 ps, beta2 -20,
               gamma 1.3, one soliton period (8 cm), 4 traces
  kerr_strong  the same fibre, 1 kW sech, 1 cm, 2 traces, total_steps 1
  raman_ss     256 samples / 2 ps at 835 nm, 50 fs FWHM sech 10 kW, Dudley beta
               list, gamma 110, self-steepening + Blow-Wood Raman, 5 mm, 2 traces,
               total_steps 1, local error 1e-6 (a first step so far too long
               that the controller's lower factor clip 0.2 acts)
  gain         256 samples / 10 ps at 1030 nm, 200 fs FWHM Gaussian 1 kW,
               beta2 23, gamma 5, Raman, 15 dB small-signal gain over 0.2 m,
               E_sat 5e-3 nJ, 40 nm Lorentzian at 1030 nm, 4 traces
Between them they hold rejected attempts and accepted attempts with
tol <= delta < 2 tol, so a mis-stated acceptance rule or step update shows.

Run from the repository root (build container only):

    python tests/golden/make_adaptive_fixtures.py
"""

import collections
import collections.abc
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/pyofss"
OUT = os.path.join(HERE, "adaptive_fixtures.npz")

sys.path.insert(0, REPO)


def _load_reference():
    import scipy
    import scipy.integrate

    import oracle as O

    for name in ("power", "sqrt", "log10", "exp", "log"):
        if not hasattr(scipy, name):
            setattr(scipy, name, getattr(np, name))
    if not hasattr(scipy.integrate, "simps"):
        scipy.integrate.simps = lambda y, x=None, **kw: O.simps_avg(np.asarray(y), np.asarray(x))
    if not hasattr(collections, "Iterable"):
        collections.Iterable = collections.abc.Iterable
    cl = types.ModuleType("pyopencl")
    cl_array = types.ModuleType("pyopencl.array")
    cl.array = cl_array
    sys.modules.setdefault("pyopencl", cl)
    sys.modules.setdefault("pyopencl.array", cl_array)
    pkg = types.ModuleType("pyofss")
    pkg.__path__ = [REF]
    sys.modules["pyofss"] = pkg
    mods = types.ModuleType("pyofss.modules")
    mods.__path__ = [os.path.join(REF, "modules")]
    sys.modules["pyofss.modules"] = mods
    pkg.modules = mods
    import importlib

    R = types.SimpleNamespace()
    R.field = importlib.import_module("pyofss.field")
    R.domain = importlib.import_module("pyofss.domain")
    R.stepper = importlib.import_module("pyofss.modules.stepper")
    R.fibre = importlib.import_module("pyofss.modules.fibre")
    R.sech = importlib.import_module("pyofss.modules.sech")
    R.gaussian = importlib.import_module("pyofss.modules.gaussian")
    return R


class Recorder:
    """Wraps a reference Stepper instance: every integrator call and every
    relative local error is recorded; attempts are read back from them."""

    def __init__(self, stepper):
        self.calls = []  # (id of input, h)
        self.deltas = []
        self.st = stepper
        step = stepper.step
        rle = stepper.relative_local_error

        def step_rec(A, z, h):
            self.calls.append((id(A), h))
            return step(A, z, h)

        def rle_rec(A_fine, A_coarse):
            d = rle(A_fine, A_coarse)
            self.deltas.append(d)
            return d

        stepper.step = step_rec
        stepper.relative_local_error = rle_rec
        stepper.storage.dir_spec = None

    def attempts(self):
        """(h, delta, accepted) per attempt: three calls (h/2, h/2, h) per
        attempt; an attempt was accepted when the next one starts from
        another field (the last one ends the run, so it was accepted)."""
        n = len(self.deltas)
        assert len(self.calls) == 3 * n
        h = np.array([self.calls[3 * k + 2][1] for k in range(n)])
        src = [self.calls[3 * k][0] for k in range(n)]
        acc = np.array([src[k + 1] != src[k] for k in range(n - 1)] + [True])
        return h, np.array(self.deltas, dtype=float), acc


BETA_DUDLEY = [0.0, 0.0, -11.830, 8.1038e-2, -9.5205e-5, 2.0737e-7, -5.3943e-10, 1.3486e-12, -2.5495e-15,
               3.0524e-18, -1.7140e-21]  # examples/dudley_sc/fig_3.py:35-36


def scalar_case(R, tol):
    def simple(A, z):
        return 3.0 * np.exp(-4.0 * z) - 2.0 * A

    st = R.stepper.Stepper(traces=1, local_error=tol, method="ARK4", f=simple, length=0.5, total_steps=50)
    rec = Recorder(st)
    A = st(1.0, None)
    h, d, acc = rec.attempts()
    return {"A_out": np.array([A], dtype=float), "h": h, "delta": d, "accepted": acc}


def fibre_case(R, domain, field, fibre):
    rec = Recorder(fibre.stepper)
    A = fibre(domain, field.copy())
    h, d, acc = rec.attempts()
    s = fibre.stepper.storage
    return {"A_in": field, "A_out": np.asarray(A), "h": h, "delta": d, "accepted": acc,
            "trace_z": np.array(s.z, dtype=float), "step_z": np.array([z for z, _ in s.step_sizes], dtype=float),
            "step_h": np.array([hh for _, hh in s.step_sizes], dtype=float), "fft_total": np.array(s.fft_total)}


def main():
    R = _load_reference()
    c = R.domain.Domain.vacuum_light_speed
    out = {}
    for tol, tag in ((1e-6, "1e-6"), (1e-9, "1e-9"), (1e-3, "1e-3")):
        for k, v in scalar_case(R, tol).items():
            out[f"scalar_{tag}/{k}"] = v

    def put(name, dom, A0, fib):
        for k, v in fibre_case(R, dom, A0, fib).items():
            out[f"{name}/{k}"] = v

    for name, P0, L, traces, steps in (("kerr", 4 * 20.0 / 1.3, 0.08, 4, 100), ("kerr_strong", 1e3, 0.01, 2, 1)):
        dom = R.domain.Domain(total_bits=1, samples_per_bit=256, bit_width=20.0, centre_nu=c / 1550.0)
        A0 = R.sech.Sech(peak_power=P0, width=1.0).generate(dom.t)
        fib = R.fibre.Fibre(length=L, beta=[0.0, 0.0, -20.0], gamma=1.3, method="ark4ip", traces=traces,
                            total_steps=steps, local_error=1e-4)
        put(name, dom, A0, fib)

    dom = R.domain.Domain(total_bits=1, samples_per_bit=256, bit_width=2.0, centre_nu=c / 835.0)
    A0 = R.sech.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(dom.t)
    fib = R.fibre.Fibre(length=5e-6, beta=BETA_DUDLEY, gamma=110.0, self_steepening=True, use_all=True,
                        method="ark4ip", traces=2, total_steps=1, local_error=1e-6)
    put("raman_ss", dom, A0, fib)

    dom = R.domain.Domain(total_bits=1, samples_per_bit=256, bit_width=10.0, centre_nu=c / 1030.0)
    A0 = R.gaussian.Gaussian(peak_power=1e3, width=0.2, using_fwhm=True).generate(dom.t)
    fib = R.fibre.Fibre(length=2e-4, beta=[0.0, 0.0, 23.0], gamma=5.0, use_all=True, method="ark4ip", traces=4,
                        small_signal_gain=15.0, E_sat=5e-3, lamb0=1030.0, bandwidth=40.0, local_error=1e-4)
    put("gain", dom, A0, fib)

    np.savez_compressed(OUT, **out)
    for key in sorted({k.split("/")[0] for k in out}):
        acc = out[f"{key}/accepted"]
        d = out[f"{key}/delta"]
        tol = float(key.split("_")[1]) if key.startswith("scalar") else (1e-6 if key == "raman_ss" else 1e-4)
        band = int(np.sum(acc & (d >= tol)))
        print(f"{key}: {len(acc)} attempts, {int((~acc).sum())} rejected, {band} accepted with tol <= delta < 2 tol")


if __name__ == "__main__":
    main()
