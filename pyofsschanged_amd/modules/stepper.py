"""Fixed and adaptive steppers, as pyofss/modules/stepper.py:33-376.

This is the drop-in boundary: ``Stepper.__call__`` (stepper.py:147-156)
hands the whole propagation of a Fibre to the HIP engine, which runs every
step -- and, for the adaptive stepper, the step-doubling controller with its
error reductions, accept/reject decisions and trace selection -- inside one
persistent kernel. Only the traces come back. Status codes map onto the
reference's exceptions.
"""

import warnings

import numpy as np

from .. import _lib, field
from .solver import Solver
from .storage import Storage


class StepperError(Exception):
    pass


class SmallStepSizeError(StepperError):
    pass


class SuitableStepSizeError(StepperError):
    pass


class MaximumStepsAllocatedError(StepperError):
    pass


class EngineStatusError(StepperError):
    pass


def raise_for_status(status):
    """Map a device status code to the reference's exception (stepper.py:33-46)."""
    if status == _lib.STATUS_OK:
        return
    if status == _lib.STATUS_SMALL_STEP:
        raise SmallStepSizeError("Step size is extremely small")
    if status == _lib.STATUS_SUITABLE_STEP:
        raise SuitableStepSizeError("Failed to set suitable step-size")
    if status == _lib.STATUS_MAX_STEPS:
        raise MaximumStepsAllocatedError("Failed to complete with maximum steps allocated")
    raise EngineStatusError(f"engine status {status}")


def fixed_trace_steps(total_steps, traces):
    """Step indices i after which stepper.py:206-209 stores a trace."""
    if traces == 1:
        return []
    storage_step = int(total_steps / traces)
    if storage_step == 0:
        # stepper.py:208 evaluates i % 0
        raise ZeroDivisionError("integer division or modulo by zero")
    return list(range(0, total_steps, storage_step))


def adaptive_initial_step(length, traces, total_steps):
    """stepper.py:256-264"""
    h = length / traces if traces > total_steps else length / total_steps
    return 1e-6 if h > 1e-6 else h


class Stepper:
    """stepper.py:49-376"""

    def __init__(self, traces=1, local_error=1.0e-6, method="RK4", f=None, f_characts=None, length=1.0,
                 total_steps=100, dir=None, save_represent="power", cycle=None, fibre_name=None, downsampling=None):
        self.traces = traces
        self.local_error = local_error
        self.save_represent = save_represent
        try:
            self.save_df = getattr(self, "save_df_" + save_represent)
        except AttributeError:
            print("No such method: save_represent should be either 'complex' or 'power'")
        self.cycle = cycle
        self.fibre_name = fibre_name
        if method.upper().startswith("A"):
            self.adaptive = True
            self.method = method[1:]
        else:
            self.adaptive = False
            self.method = method
        self.solver = Solver(self.method, f)
        self.step = self.solver
        self.f = f
        self.length = length
        self.total_steps = total_steps
        self.storage = Storage(dir, cycle=self.cycle, fibre_name=self.fibre_name, f=f_characts,
                               downsampling=downsampling)
        self.total_attempts = 1000
        self.steps_max = 100000
        self.step_size_min = 1e-37
        self.safety = 0.9
        self.max_factor = 10.0
        self.min_factor = 0.2
        self.eta = self.solver.errors[self.method.lower()]
        self.A_out = None
        self.attempts = 0
        self.accepted = 0

    def __call__(self, A, refrence_length=None):
        """stepper.py:147-156"""
        self.storage.reset_fft_counter()
        self.storage.reset_array()
        if self.adaptive:
            return self.adaptive_stepper(A, refrence_length)
        return self.standard_stepper(A, refrence_length)

    def save_df_power(self):
        self.storage.save_all_storage_to_dir_as_df(save_power=True)

    def save_df_complex(self):
        self.storage.save_all_storage_to_dir_as_df(save_power=False)

    def _fibre(self):
        fibre = getattr(self.f, "fibre", None)
        if fibre is None:
            raise TypeError("the HIP engine integrates a Fibre's FunctionStep; arbitrary callables are not supported")
        return fibre

    def standard_stepper(self, A, refrence_length):
        """stepper.py:164-222 as one persistent device launch."""
        fibre = self._fibre()
        h = self.length / self.total_steps
        if refrence_length is not None and h > refrence_length * (10 ** (-2)):
            warnings.warn(
                f"{self.cycle}: {self.fibre_name}: h must be much less than dispersion length (L_D) and the "
                f"nonlinear length (L_NL)\n now the minimum of the characteristic distances is equal to "
                f"{refrence_length:.6f}*km* \n step is equal to {h}*km*")
        zs = np.linspace(0.0, self.length, self.total_steps + 1)
        steps = fixed_trace_steps(self.total_steps, self.traces)
        trace_every = int(self.total_steps / self.traces) if steps else 0
        ctx = fibre._device_context()
        A_out, tr, ntr = ctx.propagate_fixed(self.solver.method_id, np.asarray(A, dtype=np.complex128), h,
                                             self.total_steps, trace_every, len(steps),
                                             h_exp=fibre.linear_step(0.5 * h))
        fibre._count_ffts(self.solver.method_id, self.total_steps)
        self.storage.append_traces([zs[i] + h for i in steps], [tr[0, k].copy() for k in range(len(steps))], ctx)
        self.storage.store_current_fft_count()
        self.A_out = A_out
        self.save_df()
        return self.A_out

    @staticmethod
    def relative_local_error(A_fine, A_coarse):
        """stepper.py:230-245, on the host for callers that use it directly."""
        if np.isnan(A_fine).any() or np.isinf(A_fine).any():
            return 1.0
        norm_fine = np.linalg.norm(A_fine)
        if norm_fine != 0.0:
            return np.linalg.norm(A_fine - A_coarse) / norm_fine
        return np.linalg.norm(A_fine - A_coarse)

    def adaptive_params(self):
        """The controller constants of stepper.py:133-139 and 256-274."""
        if self.traces == 0:
            raise ZeroDivisionError("float division by zero")  # stepper.py:268 length / traces
        p = _lib.AdaptiveParams()
        p.length = self.length
        p.h0 = adaptive_initial_step(self.length, self.traces, self.total_steps)
        p.local_error = self.local_error
        p.safety = self.safety
        p.max_factor = self.max_factor
        p.min_factor = self.min_factor
        p.step_size_min = self.step_size_min
        p.traces = self.traces
        p.total_attempts = self.total_attempts
        p.steps_max = self.steps_max
        p.eta = self.eta
        return p

    def adaptive_stepper(self, A, refrence_length):
        """stepper.py:247-376 as one persistent device launch. The reference
        raises AttributeError at stepper.py:365 after finishing; this returns
        A_out as intended."""
        print("Starting ODE integration with adaptive step-size... ")
        fibre = self._fibre()
        p = self.adaptive_params()
        print(f"initial step size equals {p.h0}")
        max_traces = self.traces + 2 if self.traces != 1 else 0
        ctx = fibre._device_context()
        out = ctx.propagate_adaptive(self.solver.method_id, np.asarray(A, dtype=np.complex128), p,
                                     max_traces=max_traces)
        res = out["results"][0]
        self.attempts = res["n_attempts"]
        self.accepted = res["n_accepted"]
        fibre._count_ffts(self.solver.method_id, 3 * res["n_attempts"])
        zs, traces = [], []
        for k in range(res["n_traces"]):
            z = float(out["trace_z"][0, k])
            if k > 0:  # stepper.py:343: step sizes only with the traces after z = 0
                self.storage.step_sizes.append((z, float(out["step_h"][0, k])))
            zs.append(z)
            traces.append(out["traces"][0, k].copy())
        self.storage.append_traces(zs, traces, ctx)
        raise_for_status(res["status"])
        self.storage.store_current_fft_count()
        self.A_out = out["A"]
        return self.A_out
