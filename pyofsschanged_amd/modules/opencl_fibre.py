"""OpenclProgramm and OpenclFibre, as pyofss/modules/opencl_fibre.py:287-1003,
running on the HIP engine.

Notebook code written for the OpenCL fibre (``OpenclFibre(cl_programm=...,
...)`` with an ``OpenclProgramm(fast_math=..., use_all=..., dorf=...,
ctx=..., downsampling=...)``, e.g. examples/notebooks/mamyshev_oscillator_cl.
ipynb and nonlinear_attractor.ipynb) runs unchanged: the constructors take
the same arguments, the fixed-step propagation (cl_rk4ip, cl_ss_symmetric, cl_ss_sym_rk4)
runs in the engine's persistent kernel, and every stored trace is reduced on
the device the way the OpenCL fibre does it: cl_power / cl_fftshift /
cl_interpolate down-sampled powers, the maximum power, and the energy of
cl_simpson (plain composite Simpson over the sample triples (0,1,2), (2,3,4),
..., (N-4, N-3, N-2), opencl_fibre.py:226-236, which differs from the
Fibre's scipy rule when the field reaches the window edges). The OpenCL
context, queue and program arguments are accepted and ignored.
"""

import warnings

import numpy as np
import pandas as pd

from .. import _lib
from ..field import get_duration, temporal_power
from .amplifier import Amplifier, Amplifier2LevelModel
from .linearity import Linearity
from .nonlinearity import Nonlinearity


class FiberInitError(Exception):
    pass


_METHODS = {"cl_rk4ip": _lib.METHOD_RK4IP, "cl_ss_symmetric": _lib.METHOD_SS_SYMMETRIC,
            "cl_ss_sym_rk4": _lib.METHOD_SS_SYM_RK4}


def opencl_simpson_weights(n, dt):
    """Weights w with sum_i w_i P_i = cl_simpson + cl.array.sum of P
    (opencl_fibre.py:226-236, 795-796; h = dt * 1e-3)."""
    w = np.zeros(n)
    g = np.arange(0, n - 3, 2)
    np.add.at(w, g, 1.0)
    np.add.at(w, g + 1, 4.0)
    np.add.at(w, g + 2, 1.0)
    return w * (dt * 1e-3) / 3.0


class OpenclProgramm:
    """opencl_fibre.py:287-409: holds the settings shared by OpenCL fibres."""

    def __init__(self, name="cl_programm", dorf="double", ctx=None, fast_math=False, use_all=False, downsampling=500):
        if dorf != "double":
            raise NotImplementedError("the HIP engine computes in complex128 (dorf='double')")
        self.name = name
        self.cached_factor = False
        self.use_all = use_all
        self.dorf = dorf
        self.downsampling = downsampling
        self.domain = None
        self.fast_math = fast_math
        self.ctx = ctx
        self.queue = None
        self.np_float = np.float64
        self.np_complex = np.complex128
        self.prg = None
        self.compiler_options = None
        self.plan = None
        self.shape = None

    def cl_initialise(self):
        """opencl_fibre.py:324-349: nothing to build for the HIP engine."""

    def set_domain(self, domain):
        """opencl_fibre.py:351-364"""
        if self.domain is None:
            self.domain = domain


class OpenclFibre:
    """opencl_fibre.py:422-1003 (methods cl_rk4ip, cl_ss_symmetric and cl_ss_sym_rk4)."""

    def __init__(self, cl_programm, name="cl_fibre", length=1.0, alpha=None, beta=None, gamma=0.0,
                 method="cl_ss_symmetric", total_steps=100, self_steepening=False, centre_omega=None, tau_1=12.2e-3,
                 tau_2=32.0e-3, f_R=0.18, small_signal_gain=None, E_sat=None, P_sat=None, Tr=None, lamb0=None,
                 bandwidth=None, use_Er_profile=False, use_Er_noise=False, use_Yb_model=False, Pp_0=None, N=None,
                 Rr=None, save_represent="power", cycle="cycle0", traces=None, dir=None, amplifier=None):
        self.cl_programm = cl_programm
        self.dir = dir
        self.name = name
        self.cycle = cycle
        self.domain = None
        self.gamma = gamma
        self.beta_2 = beta[2] if beta is not None else None
        self.f_R = f_R
        self.f_R_inv = 1.0 - f_R
        self.length = length
        self.total_steps = total_steps
        self.traces = traces if traces is not None else total_steps
        self.stepsize = self.length / self.total_steps
        self.zs = np.linspace(0.0, self.length, self.total_steps + 1)
        key = method.lower()
        if key not in _METHODS:
            raise AttributeError(f"OpenclFibre has no method {method!r}")
        self.method = key
        if not self.use_all and self_steepening:
            raise NotImplementedError("Self-steepening without general nonlinearity is not implemented")
        if amplifier is not None:
            self.amplifier = amplifier
        elif use_Yb_model:
            print("Use Yb model")
            self.amplifier = Amplifier2LevelModel(Pp=Pp_0, N=N, Rr=Rr)
        elif (small_signal_gain is not None) and (E_sat is not None or (P_sat is not None and Tr is not None)):
            print("Use simple saturation model")
            self.amplifier = Amplifier(gain=small_signal_gain, E_sat=E_sat, P_sat=P_sat, Tr=Tr, length=self.length,
                                       lamb0=lamb0, bandwidth=bandwidth, use_Er_profile=use_Er_profile)
        else:
            print("Passive fibre modulation")
            self.amplifier = None
        self.linearity = Linearity(alpha, beta, sim_type="default", use_cache=True, centre_omega=centre_omega,
                                   phase_lim=True, use_Er_noise=use_Er_noise, amplifier=self.amplifier)
        self.nonlinearity = Nonlinearity(gamma, None, self_steepening, False, 0, self.use_all, tau_1, tau_2, f_R)
        self.factor = None
        self.energy_list = []
        self.max_power_list = []
        self.peaks_list = []
        self.l_nl_list = []
        self.l_d_list = []
        self.z_list = []
        self.temp_field_list = []
        self.spec_field_list = []
        self._ctx = None

    # settings shared through the program object (opencl_fibre.py:522-617)
    @property
    def dorf(self):
        return self.cl_programm.dorf

    @property
    def use_all(self):
        return self.cl_programm.use_all

    @property
    def queue(self):
        return self.cl_programm.queue

    @property
    def np_float(self):
        return self.cl_programm.np_float

    @property
    def np_complex(self):
        return self.cl_programm.np_complex

    @property
    def prg(self):
        return self.cl_programm.prg

    @property
    def fast_math(self):
        return self.cl_programm.fast_math

    @property
    def ctx(self):
        return self.cl_programm.ctx

    @property
    def shape(self):
        return None if self.domain is None else (self.domain.total_samples,)

    def _device_context(self):
        if self._ctx is None:
            self._ctx = _lib.Context(0)
            self.linearity.bind(self._ctx)
            self.nonlinearity.bind(self._ctx)
        return self._ctx

    def __call__(self, domain, field):
        """opencl_fibre.py:618-671: total_steps fixed steps; after step i with
        i % int(total_steps / traces) == 0 the trace at zs[i + 1] is reduced
        (down-sampled powers, maximum power, energy, L_D, L_NL)."""
        ctx = self._device_context()
        if self.domain != domain:
            self.domain = domain
            self.nonlinearity(domain)
            self.calculate_reference_length(field)
            if self.reference_length is not None and self.stepsize > self.reference_length * (10 ** (-2)):
                warnings.warn(f"{self.cycle}: {self.name}: h must be much less than dispersion length (L_D) and the "
                              f"nonlinear length (L_NL)\n now the minimum of the characteristic distances is equal "
                              f"to {self.reference_length:.6f}*km* \n step is equal to {self.stepsize}*km*")
        if self.factor is None:
            self.factor = self.linearity(domain)
        if ctx.n != domain.total_samples:
            ctx.set_grid(domain.total_samples)
        storage_step = int(self.total_steps / self.traces)
        n_tr = len(range(0, self.total_steps, storage_step)) if storage_step > 0 else 0
        A = np.asarray(field, dtype=np.complex128)
        out, tr, ntr = ctx.propagate_fixed(_METHODS[self.method], A, self.stepsize, self.total_steps,
                                           trace_every=max(storage_step, 0), max_traces=n_tr,
                                           h_exp=0.5 * self.stepsize)
        k = int(ntr[0]) if n_tr > 0 else 0
        if k > 0:
            traces = tr[0, :k]
            self.z_list.extend(self.zs[1:][0::storage_step][:k].tolist())
            t_pow, s_pow = ctx.trace_powers(traces, self.cl_programm.downsampling, _lib.DOWNSAMPLE_OPENCL)
            self.temp_field_list.extend(list(t_pow))
            self.spec_field_list.extend(list(s_pow))
            ch = ctx.trace_characteristics(traces, opencl_simpson_weights(domain.total_samples, domain.dt))
            for j in range(k):
                P0 = float(ch["max_power"][j])
                T0 = float(ch["duration"][j]) * domain.dt
                self.energy_list.append(float(ch["energy"][j]))
                self.max_power_list.append(P0)
                with np.errstate(divide="ignore"):
                    self.l_d_list.append(abs(T0 ** 2 / self.beta_2))
                    self.l_nl_list.append(1 / (self.gamma * P0))
        return out

    def calculate_reference_length(self, field):
        """opencl_fibre.py:673-681"""
        self.L_D = self.get_dispersion_length(field)
        self.L_NL = self.get_nonlinear_length(field)
        if self.L_NL and self.L_D is not None:
            self.reference_length = min(self.L_NL, self.L_D)
        elif self.L_NL or self.L_D is None:
            self.reference_length = None
        else:
            self.reference_length = self.L_NL if self.L_D is None else self.L_D

    def get_nonlinear_length(self, field):
        """opencl_fibre.py:683-686"""
        with np.errstate(divide="ignore"):
            return 1 / (self.gamma * np.amax(temporal_power(field)))

    def get_dispersion_length(self, field):
        """opencl_fibre.py:688-692"""
        T_0 = get_duration(temporal_power(field), self.domain.dt)
        with np.errstate(divide="ignore"):
            return abs(T_0 ** 2 / (self.beta_2))

    def _index(self, z_curr):
        arr_z = np.array(self.z_list) * 10 ** 6 + z_curr  # mm
        if self.cycle and self.name is not None:
            return pd.MultiIndex.from_product([[self.cycle], [self.name], arr_z], names=["cycle", "fibre", "z [mm]"])
        return pd.MultiIndex.from_product([arr_z], names=["z [mm]"])

    def get_df(self, type="complex", z_curr=0, channel=None):
        """opencl_fibre.py:694-713: down-sampled powers; the reference has no
        complex frame (it warns, then fails on an unbound name): an empty
        frame is returned after the warning."""
        if type == "temp":
            y = self.temp_field_list
        elif type == "spec":
            y = self.spec_field_list
        elif type == "complex":
            warnings.warn("Saving downsampled complex power has not been made yet!")
            return pd.DataFrame()
        else:
            raise ValueError()
        return pd.DataFrame(y, index=self._index(z_curr))

    def get_df_result(self, z_curr=0):
        """opencl_fibre.py:715-740: peak power, energy, L_D and L_NL; the
        width, spectral width and peak columns stay empty as in the reference."""
        cols = ["Peak Power [W]", "Energy [nJ]", "Temp width [ps]", "Spec width [THz]", "Dispersion length [km]",
                "Nonlinear length [km]", "Peaks [idx]"]
        df = pd.DataFrame(index=self._index(z_curr), columns=cols)
        df["Peak Power [W]"] = self.max_power_list
        df["Energy [nJ]"] = self.energy_list
        df["Dispersion length [km]"] = self.l_d_list
        df["Nonlinear length [km]"] = self.l_nl_list
        return df

    def clear_arrays_on_device(self):
        """opencl_fibre.py:773-779: device buffers belong to the engine context."""
