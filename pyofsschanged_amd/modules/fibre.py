"""Fibre module, as pyofss/modules/fibre.py:31-187.

Wires Linearity, Nonlinearity and Stepper exactly as the reference does; the
three share one HIP engine context, created at the first call, that holds
the operator arrays on the GPU for the whole propagation.
"""

import numpy as np
from numpy import abs, amax

from .. import _lib, field
from ..field import get_duration, temporal_power
from .amplifier import Amplifier, Amplifier2LevelModel
from .linearity import Linearity
from .nonlinearity import Nonlinearity, _NL_FFTS
from .stepper import Stepper


class FiberInitError(Exception):
    pass


class Fibre:
    """fibre.py:35-187"""

    def __init__(self, name="fibre", length=1.0, alpha=None, beta=None, gamma=0.0, sim_type=None, traces=1,
                 local_error=1.0e-6, method="ss_symmetric", total_steps=100, self_steepening=False,
                 raman_scattering=False, rs_factor=0.003, use_all=False, centre_omega=None, tau_1=12.2e-3,
                 tau_2=32.0e-3, f_R=0.18, small_signal_gain=None, E_sat=None, P_sat=None, Tr=None, lamb0=None,
                 bandwidth=None, use_Er_profile=False, use_Er_noise=False, use_Yb_model=False, Pp_0=None, N=None,
                 Rr=None, dir=None, save_represent="power", cycle=None, downsampling=None):
        use_cache = not (method.upper().startswith("A"))
        self.name = name
        self.length = length
        self.small_signal_gain = small_signal_gain
        self.wdm = sim_type == "wdm"
        # fibre.py:88; with WDM's per-channel beta lists the reference fails
        # there (beta[2] of a two-element list): no dispersion length then
        self.beta_2 = beta[2] if (beta is not None and not self.wdm) else None
        self.gamma = gamma
        self.L_D = None
        self.L_NL = None
        self.reference_length = None
        self.cycle = cycle
        self.domain = None
        self.amplifier = None
        if use_Yb_model:  # fibre.py:97-98
            self.amplifier = Amplifier2LevelModel(Pp=Pp_0, N=N, Rr=Rr)
        elif (small_signal_gain is not None) and (E_sat is not None or (P_sat is not None and Tr is not None)):
            self.amplifier = Amplifier(gain=self.small_signal_gain, E_sat=E_sat, P_sat=P_sat, Tr=Tr,
                                       length=self.length, lamb0=lamb0, bandwidth=bandwidth,
                                       use_Er_profile=use_Er_profile)
        self.linearity = Linearity(alpha=alpha, beta=beta, sim_type=sim_type, use_cache=use_cache,
                                   centre_omega=centre_omega, amplifier=self.amplifier, use_Er_noise=use_Er_noise)
        self.nonlinearity = Nonlinearity(gamma, sim_type, self_steepening, raman_scattering, rs_factor, use_all,
                                         tau_1, tau_2, f_R)
        self._ctx = None
        fibre = self

        class FunctionStep:
            """fibre.py:110-121; ``fibre`` lets the engine find the operators."""

            def __init__(self, l, n, linear, nonlinear, amplifier_step):
                self.l = l
                self.n = n
                self.linear = linear
                self.nonlinear = nonlinear
                self.amplifier_step = amplifier_step
                self.fibre = fibre

            def __call__(self, A, z):
                return self.l(A, z) + self.n(A, z)

        class FunctionCharacts:
            def __init__(self, l_d, l_nl, lengths):
                self.l_d = l_d
                self.l_nl = l_nl
                self.lengths = lengths

        self.function_step = FunctionStep(self.l, self.n, self.linear, self.nonlinear, self.amplifier_step)
        self.function_characts = FunctionCharacts(self.get_dispersion_length, self.get_nonlinear_length,
                                                  self.lengths_from)
        self.stepper = Stepper(traces, local_error, method, self.function_step, self.function_characts, self.length,
                               total_steps, dir, save_represent, cycle=self.cycle, fibre_name=self.name,
                               downsampling=downsampling)

    def _device_context(self):
        if self._ctx is None:
            self._ctx = _lib.Context(0)
            self.linearity.bind(self._ctx)
            self.nonlinearity.bind(self._ctx)
        return self._ctx

    def linear_step(self, h):
        """Step of exp(factor*h) used by a linear sub-step of size h: with a
        caching Linearity (fixed-step methods) the step of its first call,
        as linearity.py:258-262 caches exp(factor*h) once."""
        lin = self.linearity
        if not lin.use_cache:
            return h
        if lin.cached_factor is None:
            lin.cache(h)
        return lin.cached_factor

    def _count_ffts(self, method, calls):
        """Advance field.fft_counter as pyofss counts FFTs (field.py:118-134)."""
        nl = _NL_FFTS[self.nonlinearity.kind]
        lin = 2 * self.linearity.channels() + self.linearity.gain_ffts()  # per linear call
        if method == _lib.METHOD_RK4IP:
            per_call = 4 * lin + 4 * nl
        elif method == _lib.METHOD_SS_SYM_RK4:  # two linear half steps, four nonlinear evaluations
            per_call = 2 * lin + 4 * nl
        else:
            per_call = 2 * lin + nl
        field.fft_counter += per_call * calls

    def prepare(self, domain):
        """Operator set-up of fibre.py:137-143: build the linear and nonlinear
        operators for ``domain`` and load them into the device context."""
        if self.domain != domain or self.linearity.factor is None or self.nonlinearity.factor is None:
            self._device_context()
            self.linearity(domain)
            self.nonlinearity(domain)
            self.stepper.storage(domain)
            self.domain = domain
        return self._device_context()

    def __call__(self, domain, field_in):
        """fibre.py:136-148"""
        self.prepare(domain)
        self.calculate_reference_length(field_in)
        return self.stepper(field_in, self.reference_length)

    def l(self, A, z):
        return self.linearity.lin(A, z)

    def linear(self, A, h):
        return self.linearity.exp_lin(A, h)

    def n(self, A, z):
        return self.nonlinearity.non(A, z)

    def nonlinear(self, A, h, B):
        return self.nonlinearity.exp_non(A, h, B)

    def amplifier_step(self, A, h):
        raise NotImplementedError("step_amplifier is not supported by the HIP engine")

    def calculate_reference_length(self, field_in):
        """fibre.py:169-177 (no reference length for WDM fields, whose
        per-channel gamma and beta the reference cannot evaluate here)."""
        if self.wdm:
            self.L_D = self.L_NL = self.reference_length = None
            return
        self.L_D = self.get_dispersion_length(field_in)
        self.L_NL = self.get_nonlinear_length(field_in)
        if self.L_NL and self.L_D is not None:
            self.reference_length = min(self.L_NL, self.L_D)
        elif self.L_NL or self.L_D is None:
            self.reference_length = None
        else:
            self.reference_length = self.L_NL if self.L_D is None else self.L_D

    def lengths_from(self, T_0, P_0):
        """(L_D, L_NL) of fibre.py:179-187 from a trace's duration T_0 and
        maximum power P_0."""
        with np.errstate(divide="ignore"):
            L_NL = 1 / (self.gamma * np.float64(P_0))
            L_D = None if self.beta_2 is None else T_0 ** 2 / abs(np.float64(self.beta_2))
        return L_D, L_NL

    def get_nonlinear_length(self, field_in):
        """fibre.py:179-182"""
        P_0 = amax(temporal_power(field_in))
        with np.errstate(divide="ignore"):
            return 1 / (self.gamma * P_0)

    def get_dispersion_length(self, field_in):
        """fibre.py:184-187; returns None without dispersion, where the
        reference fails on abs(None)."""
        if self.beta_2 is None:
            return None
        T_0 = get_duration(temporal_power(field_in), self.domain.dt)
        with np.errstate(divide="ignore"):
            return T_0 ** 2 / abs(np.float64(self.beta_2))
