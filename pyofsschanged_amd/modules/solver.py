"""Integrators, as pyofss/modules/solver.py:24-354.

The HIP engine implements the fibre integrators of the stepping path,
``rk4ip`` (solver.py:339-354), ``ss_symmetric`` (solver.py:266-275) and
``ss_sym_rk4`` (solver.py:330-337, the symmetric split step with the
classical rk4 on the nonlinear operator). A
call ``Solver(method, f)(A, z, h)`` with a Fibre's FunctionStep runs one
integrator call on the GPU. The generic ODE methods of the reference (euler,
rk4, embedded pairs, other split-step variants) take arbitrary Python
callables and are not part of the device path; asking for them raises.
"""

from .. import _lib


class SolverError(Exception):
    pass


class NoDerivativeFunctionError(SolverError):
    pass


class UnsupportedMethodError(SolverError, NotImplementedError):
    pass


_DEVICE_METHODS = {"rk4ip": _lib.METHOD_RK4IP, "ss_symmetric": _lib.METHOD_SS_SYMMETRIC,
                   "ss_sym_rk4": _lib.METHOD_SS_SYM_RK4}


def device_method(name):
    """Engine method id for a (case-insensitive) solver name, or raise."""
    m = _DEVICE_METHODS.get(name.lower())
    if m is None:
        raise UnsupportedMethodError(
            f"method {name!r} is not implemented by the HIP engine; use one of "
            f"{sorted(_DEVICE_METHODS)} (prefix 'A' for the adaptive stepper)")
    return m


def _engine_of(f):
    fibre = getattr(f, "fibre", None)
    if fibre is None or not hasattr(fibre, "_device_context"):
        raise TypeError("the HIP engine integrates a Fibre's FunctionStep; arbitrary callables are not supported")
    return fibre


class Solver:
    """solver.py:32-91"""

    explicit_solvers = ["euler", "midpoint", "rk4"]
    embedded_solvers = ["bs", "rkf", "ck", "dp"]
    ss_solvers = ["ss_simple", "ss_symmetric", "ss_reduced", "ss_agrawal", "ss_sym_midpoint", "ss_sym_rk4"]
    other_solvers = ["rk4ip", "step_amplifier"]
    errors = {"euler": 2, "midpoint": 3, "rk4": 5, "bs": 2, "rkf": 4, "ck": 4, "dp": 4, "ss_simple": 2,
              "ss_symmetric": 3, "ss_reduced": 2, "ss_agrawal": 3, "ss_sym_midpoint": 3, "ss_sym_rk4": 5,
              "rk4ip": 5}

    def __init__(self, method="rk4", f=None):
        self.embedded = method.lower() in self.embedded_solvers
        self.method_id = device_method(method)
        self.method = getattr(self, method.lower())
        if f is None:
            raise NoDerivativeFunctionError("Require a function for ODE integration")
        self.f = f

    def __call__(self, A, z, h):
        return self.method(A, z, h, self.f)

    @staticmethod
    def rk4ip(A, z, h, f):
        """solver.py:339-354 on the GPU."""
        fibre = _engine_of(f)
        fibre._count_ffts(_lib.METHOD_RK4IP, 1)
        return fibre._device_context().step(_lib.METHOD_RK4IP, A, h, fibre.linear_step(0.5 * h))

    @staticmethod
    def ss_symmetric(A, z, h, f):
        """solver.py:266-275 on the GPU."""
        fibre = _engine_of(f)
        fibre._count_ffts(_lib.METHOD_SS_SYMMETRIC, 1)
        return fibre._device_context().step(_lib.METHOD_SS_SYMMETRIC, A, h, fibre.linear_step(0.5 * h))

    @staticmethod
    def ss_sym_rk4(A, z, h, f):
        """solver.py:330-337 on the GPU."""
        fibre = _engine_of(f)
        fibre._count_ffts(_lib.METHOD_SS_SYM_RK4, 1)
        return fibre._device_context().step(_lib.METHOD_SS_SYM_RK4, A, h, fibre.linear_step(0.5 * h))
