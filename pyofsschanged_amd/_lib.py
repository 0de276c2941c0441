"""ctypes binding of libgnlse_hip.so (the C ABI of include/gnlse.h).

The shared library is built in-tree by ``python -m pyofsschanged_amd.build``.
Every numerical entry point runs on the GPU; if the library or a HIP device is
missing, the calls raise ``EngineUnavailableError`` instead of computing
anything on the host.
"""

import ctypes
import os
import threading

import numpy as np

LIB_PATH = os.environ.get("GNLSE_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "libgnlse_hip.so")

METHOD_RK4IP = 0
METHOD_SS_SYMMETRIC = 1
METHOD_SS_SYM_RK4 = 2

NL_KERR = 0
NL_ALL = 1
NL_SS = 2
NL_RS = 3
NL_SS_RS = 4

DOWNSAMPLE_PYOFSS = 0
DOWNSAMPLE_OPENCL = 1

STATUS_OK = 0
STATUS_SMALL_STEP = 1
STATUS_SUITABLE_STEP = 2
STATUS_MAX_STEPS = 3
STATUS_TRACE_OVERFLOW = 4
STATUS_REPLAY_EXHAUSTED = 5
STATUS_PAIR_TIMEOUT = 6


class EngineUnavailableError(RuntimeError):
    """The HIP engine cannot run: library not built or no GPU visible."""


class GNLSEError(RuntimeError):
    """A call into the HIP engine failed."""


class AdaptiveParams(ctypes.Structure):
    _fields_ = [
        ("length", ctypes.c_double),
        ("h0", ctypes.c_double),
        ("local_error", ctypes.c_double),
        ("safety", ctypes.c_double),
        ("max_factor", ctypes.c_double),
        ("min_factor", ctypes.c_double),
        ("step_size_min", ctypes.c_double),
        ("traces", ctypes.c_int64),
        ("total_attempts", ctypes.c_int64),
        ("steps_max", ctypes.c_int64),
        ("eta", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class InstanceResult(ctypes.Structure):
    _fields_ = [
        ("z", ctypes.c_double),
        ("h", ctypes.c_double),
        ("n_attempts", ctypes.c_int64),
        ("n_accepted", ctypes.c_int64),
        ("n_traces", ctypes.c_int64),
        ("status", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int

# name -> argtypes; every function returns int
SIGNATURES = {
    "gnlse_version": [],
    "gnlse_grid_supported": [_I64],
    "gnlse_device_count": [ctypes.POINTER(ctypes.c_int)],
    "gnlse_ctx_create": [_INT, ctypes.POINTER(_P)],
    "gnlse_ctx_destroy": [_P],
    "gnlse_set_grid": [_P, _I64],
    "gnlse_set_linear": [_P, _P],
    "gnlse_set_nonlinear": [_P, _INT, _D, _D, _D, _D, _P, _P, _P],
    "gnlse_set_gain": [_P, _INT, _D, _D, _D, _P, _P],
    "gnlse_set_linear_wdm": [_P, _P, _P],
    "gnlse_set_nonlinear_wdm": [_P, _D, _D],
    "gnlse_set_yb_gain": [_P, _P, _P, _P, _P, _D, _D, _D, _D, _D, _D],
    "gnlse_set_pump": [_P, _I64, _P],
    "gnlse_get_pump": [_P, _I64, _P],
    "gnlse_fft": [_P, _INT, _I64, _P, _P],
    "gnlse_apply_linear": [_P, _D, _I64, _P, _P],
    "gnlse_apply_linear_factor": [_P, _I64, _P, _P],
    "gnlse_apply_nonlinear": [_P, _I64, _P, _P],
    "gnlse_apply_exp_nonlinear": [_P, _D, _I64, _P, _P, _P],
    "gnlse_step": [_P, _INT, _D, _D, _I64, _P, _P],
    "gnlse_propagate_fixed": [_P, _INT, _I64, _D, _D, _I64, _I64, _I64, _P, _P, _P, _P],
    "gnlse_propagate_adaptive": [_P, _INT, _I64, ctypes.POINTER(AdaptiveParams), _P, _P, _I64, _P, _P, _I64, _P,
                                 _P, _P, _P, _I64, _P, _P, _P, _P],
    "gnlse_malloc": [_P, _I64, ctypes.POINTER(_P)],
    "gnlse_free": [_P, _P],
    "gnlse_memcpy_htod": [_P, _P, _P, _I64],
    "gnlse_memcpy_dtoh": [_P, _P, _P, _I64],
    "gnlse_memcpy_dtod": [_P, _P, _P, _I64],
    "gnlse_propagate_adaptive_device": [_P, _INT, _I64, ctypes.POINTER(AdaptiveParams), _P, _P, _P],
    "gnlse_copy_rows": [_P, _I64, _I64, _P, _P, _P, _P],
    "gnlse_propagate_adaptive_device_noalloc": [_P, _INT, _I64, ctypes.POINTER(AdaptiveParams), _P, _P, _P],
    "gnlse_last_kernel_ms": [_P, ctypes.POINTER(ctypes.c_double)],
    "gnlse_schedule_count": [_I64, ctypes.POINTER(ctypes.c_int)],
    "gnlse_schedule_info": [_I64, _INT, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                            _INT],
    "gnlse_set_schedule": [_P, _INT],
    "gnlse_get_schedule": [_P, ctypes.POINTER(ctypes.c_int)],
    "gnlse_set_pairing": [_P, _INT],
    "gnlse_tune_adaptive": [_P, _I64, ctypes.POINTER(AdaptiveParams), _P, _I64, _INT, ctypes.POINTER(ctypes.c_int),
                            _P],
    "gnlse_trace_characteristics": [_P, _I64, _P, _P, _P, _P, _P, _INT],
    "gnlse_trace_powers": [_P, _I64, _P, _I64, _INT, _P, _P],
    "gnlse_ensemble_create": [_INT, ctypes.POINTER(AdaptiveParams), _INT, _P, _P, ctypes.POINTER(_P)],
    "gnlse_ensemble_tune": [_P, _INT, _INT, ctypes.POINTER(ctypes.c_int)],
    "gnlse_ensemble_threads": [_P, ctypes.POINTER(ctypes.c_int)],
    "gnlse_ensemble_group_threads": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "gnlse_ensemble_set_group_threads": [_P, ctypes.c_int, ctypes.c_int],
    "gnlse_ensemble_device_ptrs": [_P, _INT, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)],
    "gnlse_ensemble_set_input": [_P, _INT, _P],
    "gnlse_ensemble_run": [_P, _INT],
    "gnlse_ensemble_output": [_P, _INT, _P, _P, _P],
    "gnlse_ensemble_last_ms": [_P, ctypes.POINTER(ctypes.c_double)],
    "gnlse_ensemble_destroy": [_P],
    "gnlse_sno_create": [_INT, _I64, _INT, _INT, _INT, _INT, _INT, _I64, ctypes.POINTER(_P)],
    "gnlse_sno_param_count": [_P, ctypes.POINTER(ctypes.c_int64)],
    "gnlse_sno_set_params": [_P, _P],
    "gnlse_sno_get_params": [_P, _P],
    "gnlse_sno_forward": [_P, _I64, _P, _P],
    "gnlse_sno_loss_grad": [_P, _I64, _P, _P, ctypes.POINTER(ctypes.c_double), _P, _P],
    "gnlse_sno_adam": [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float],
    "gnlse_sno_last_ms": [_P, ctypes.POINTER(ctypes.c_double)],
    "gnlse_sno_destroy": [_P],
    "gnlse_sno_encode": [_P, _P, _I64, _P, _D, _P, _INT, _INT, _P],
    "gnlse_sno_decode": [_P, _P, _I64, _P, _D, _P, _INT, _P],
    "gnlse_sno_energy_defect": [_P, _I64, _P, _P, _P],
    "gnlse_sno_combine": [_P, _I64, _P, _P, _P, _P],
    "gnlse_check_reductions": [_INT, _INT, _INT, _I64, _P, _P, _P],
    "gnlse_gemm_bf16": [_INT, _INT, _INT, _I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _I64, _I64, _I64, _I64, _INT],
}

_lib = None
_lib_lock = threading.Lock()


def load_library(path=LIB_PATH):
    """Load the engine library (no GPU needed to load it)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise EngineUnavailableError(
                f"{path} is not built; run `python -m pyofsschanged_amd.build`")
        lib = ctypes.CDLL(path)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        lib.gnlse_last_error.argtypes = []
        lib.gnlse_last_error.restype = ctypes.c_char_p
        _lib = lib
        return lib


def _check(lib, rc, what):
    if rc != 0:
        msg = lib.gnlse_last_error().decode(errors="replace")
        raise GNLSEError(f"{what} failed ({rc}): {msg}")


def device_count():
    lib = load_library()
    n = ctypes.c_int(0)
    rc = lib.gnlse_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def grid_supported(n):
    return bool(load_library().gnlse_grid_supported(int(n)))


def schedules(n):
    """The compiled schedules of the n-sample grid's spectral-domain adaptive
    RK4IP kernels: list of dicts (index, threads, cfg, desc); index 0 is the
    grid's default. No GPU needed."""
    lib = load_library()
    c = ctypes.c_int()
    _check(lib, lib.gnlse_schedule_count(int(n), ctypes.byref(c)), "gnlse_schedule_count")
    out = []
    for i in range(c.value):
        t, cfg = ctypes.c_int(), ctypes.c_int()
        buf = ctypes.create_string_buffer(128)
        _check(lib, lib.gnlse_schedule_info(int(n), i, ctypes.byref(t), ctypes.byref(cfg), buf, 128),
               "gnlse_schedule_info")
        out.append({"index": i, "threads": t.value, "cfg": cfg.value, "desc": buf.value.decode()})
    return out


def _c128(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if n is not None and a.shape[-1] != n:
        raise ValueError(f"field has {a.shape[-1]} samples, the context grid has {n}")
    return a


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.shape[-1] != n:
        raise ValueError(f"array has {a.shape[-1]} samples, the context grid has {n}")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One engine context: a grid, operator arrays and device buffers on one
    GPU. Calls are serialised by a lock (the C context is not thread-safe)."""

    def __init__(self, device=0):
        self.lib = load_library()
        if device_count() < 1:
            raise EngineUnavailableError("no HIP device is visible to the GNLSE engine")
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.gnlse_ctx_create(int(device), ctypes.byref(h)), "gnlse_ctx_create")
        self._h = h
        self.device = device
        self.n = None
        self.lock = threading.RLock()
        # Object with a pump power `_Pp` (an Amplifier2LevelModel) whose state
        # every launch with the Yb gain starts from and writes back
        self.pump_source = None
        self.fields = 1  # fields per instance: 2 after set_nonlinear_wdm

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gnlse_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        with self.lock:
            _check(self.lib, getattr(self.lib, name)(self._h, *args), name)

    # -- set-up -------------------------------------------------------------
    def set_grid(self, n):
        self._call("gnlse_set_grid", int(n))
        if self.n != int(n):
            self.fields = 1  # a new grid starts with single-channel fields
        self.n = int(n)

    def set_linear(self, factor):
        f = _c128(np.broadcast_to(np.asarray(factor, dtype=np.complex128), (self.n,)), self.n)
        self._call("gnlse_set_linear", _ptr(f))

    def set_nonlinear(self, kind, gamma, f_R=0.0, ss=0.0, rs=0.0, omega=None, nl_mult=None, h_R=None):
        om = None if omega is None else _f64(omega, self.n)
        nm = None if nl_mult is None else _c128(nl_mult, self.n)
        hr = None if h_R is None else _c128(h_R, self.n)
        self._call("gnlse_set_nonlinear", int(kind), float(gamma), float(f_R), float(ss), float(rs), _ptr(om),
                   _ptr(nm), _ptr(hr))
        self.fields = 1

    def set_gain(self, enabled, gain_log=0.0, length=1.0, E_sat=1.0, profile=None, simpson_w=None):
        pr = None if profile is None else _f64(profile, self.n)
        sw = None if simpson_w is None else _f64(simpson_w, self.n)
        self._call("gnlse_set_gain", int(bool(enabled)), float(gain_log), float(length), float(E_sat), _ptr(pr),
                   _ptr(sw))

    def set_linear_wdm(self, factor0, factor1):
        """Linear factors of the two WDM channels (linearity.py:188-211)."""
        f0 = _c128(np.broadcast_to(factor0, (self.n,)), self.n)
        f1 = _c128(np.broadcast_to(factor1, (self.n,)), self.n)
        self._call("gnlse_set_linear_wdm", _ptr(f0), _ptr(f1))

    def set_nonlinear_wdm(self, gamma0, gamma1):
        """WDM coupling i gamma_c (|A_c|^2 + 2 |A_o|^2) A_c; from here on every
        field argument holds two channels per instance."""
        self._call("gnlse_set_nonlinear_wdm", float(gamma0), float(gamma1))
        self.fields = 2

    def _instances(self, rows):
        if rows % self.fields:
            raise ValueError(f"WDM fields need {self.fields} channels per instance, got {rows} rows")
        return rows // self.fields

    def set_yb_gain(self, ratio_over_psat, inv_psat, alpha_s, eta_s, power_factor, ratio_p, psat_p, n_total,
                    alpha_p, eta_p):
        """Two-level Yb gain (arrays in FFT order, finite)."""
        arrs = [_f64(a, self.n) for a in (ratio_over_psat, inv_psat, alpha_s, eta_s)]
        self._call("gnlse_set_yb_gain", *[_ptr(a) for a in arrs], float(power_factor), float(ratio_p), float(psat_p),
                   float(n_total), float(alpha_p), float(eta_p))

    def set_pump(self, pump):
        """Pump power per instance for launches with the Yb gain."""
        p = np.ascontiguousarray(np.atleast_1d(np.asarray(pump, dtype=np.float64)))
        self._call("gnlse_set_pump", p.size, _ptr(p))

    def get_pump(self, count):
        p = np.empty(int(count), dtype=np.float64)
        self._call("gnlse_get_pump", p.size, _ptr(p))
        return p

    def _pump_in(self, batch):
        if self.pump_source is not None:
            self.set_pump(np.full(int(batch), float(self.pump_source._Pp)))

    def _pump_out(self, batch):
        """After a launch: the pump powers per instance; a single instance's
        becomes the amplifier's state, as the reference's object carries it."""
        if self.pump_source is not None:
            p = self.get_pump(batch)
            self.pump_source.pump_out = p
            if int(batch) == 1:
                self.pump_source._Pp = float(p[0])

    # -- operators ------------------------------------------------------------
    def _batched(self, A):
        A = _c128(A, self.n)
        shape = A.shape
        return A.reshape(-1, self.n), shape

    def fft(self, A, sign):
        a, shape = self._batched(A)
        out = np.empty_like(a)
        self._call("gnlse_fft", int(sign), a.shape[0], _ptr(a), _ptr(out))
        return out.reshape(shape)

    def apply_linear(self, A, h):
        a, shape = self._batched(A)
        out = np.empty_like(a)
        B = self._instances(a.shape[0])
        self._pump_in(B)
        self._call("gnlse_apply_linear", float(h), B, _ptr(a), _ptr(out))
        self._pump_out(B)
        return out.reshape(shape)

    def apply_linear_factor(self, A):
        a, shape = self._batched(A)
        out = np.empty_like(a)
        self._call("gnlse_apply_linear_factor", self._instances(a.shape[0]), _ptr(a), _ptr(out))
        return out.reshape(shape)

    def apply_nonlinear(self, A):
        a, shape = self._batched(A)
        out = np.empty_like(a)
        self._call("gnlse_apply_nonlinear", self._instances(a.shape[0]), _ptr(a), _ptr(out))
        return out.reshape(shape)

    def apply_exp_nonlinear(self, A, h, B):
        a, shape = self._batched(A)
        b, _ = self._batched(B)
        if b.shape != a.shape:
            raise ValueError("A and B must have the same shape")
        out = np.empty_like(a)
        self._call("gnlse_apply_exp_nonlinear", float(h), self._instances(a.shape[0]), _ptr(a), _ptr(b), _ptr(out))
        return out.reshape(shape)

    def step(self, method, A, h, h_exp=None):
        """One integrator call; linear sub-steps use exp(factor*h_exp)
        (default h/2)."""
        a, shape = self._batched(A)
        out = np.empty_like(a)
        B = self._instances(a.shape[0])
        self._pump_in(B)
        self._call("gnlse_step", int(method), float(h), float(h_exp or 0.0), B, _ptr(a), _ptr(out))
        self._pump_out(B)
        return out.reshape(shape)

    # -- propagation ------------------------------------------------------------
    def propagate_fixed(self, method, A, h, nsteps, trace_every=0, max_traces=0, h_exp=None):
        """Traces come back as [B, max_traces, N] ([B, max_traces, 2, N] for
        WDM fields)."""
        a, shape = self._batched(A)
        B = self._instances(a.shape[0])
        out = np.empty_like(a)
        traces = np.empty((B, max(max_traces, 0)) + ((self.fields,) if self.fields > 1 else ()) + (self.n,),
                          np.complex128)
        ntr = np.zeros(B, np.int64)
        self._pump_in(B)
        self._call("gnlse_propagate_fixed", int(method), B, float(h), float(h_exp or 0.0), int(nsteps), int(trace_every),
                   int(max_traces), _ptr(a), _ptr(out), _ptr(traces) if max_traces > 0 else None,
                   _ptr(ntr) if max_traces > 0 else None)
        self._pump_out(B)
        return out.reshape(shape), traces, ntr

    def propagate_adaptive(self, method, A, params, max_traces=0, log_cap=0, replay=None):
        a, shape = self._batched(A)
        B = self._instances(a.shape[0])
        out = np.empty_like(a)
        tA = np.empty((B, max_traces) + ((self.fields,) if self.fields > 1 else ()) + (self.n,), np.complex128)
        tz = np.zeros((B, max_traces))
        sz = np.zeros((B, max_traces))
        sh = np.zeros((B, max_traces))
        lh = np.zeros((B, log_cap))
        ld = np.zeros((B, log_cap))
        la = np.zeros((B, log_cap), np.uint8)
        res = (InstanceResult * B)()
        rh = ra = None
        rlen = 0
        if replay is not None:
            rh = _f64(replay[0])
            ra = np.ascontiguousarray(replay[1], dtype=np.uint8)
            rlen = len(rh)
        mt = max_traces > 0
        lg = log_cap > 0
        self._pump_in(B)
        self._call("gnlse_propagate_adaptive", int(method), B, ctypes.byref(params), _ptr(rh), _ptr(ra), int(rlen),
                   _ptr(a), _ptr(out), int(max_traces), _ptr(tA) if mt else None, _ptr(tz) if mt else None,
                   _ptr(sz) if mt else None, _ptr(sh) if mt else None, int(log_cap), _ptr(lh) if lg else None,
                   _ptr(ld) if lg else None, _ptr(la) if lg else None, res)
        self._pump_out(B)
        results = [{"z": r.z, "h": r.h, "n_attempts": r.n_attempts, "n_accepted": r.n_accepted,
                    "n_traces": r.n_traces, "status": r.status} for r in res]
        return {"A": out.reshape(shape), "traces": tA, "trace_z": tz, "step_z": sz, "step_h": sh,
                "log_h": lh, "log_delta": ld, "log_accept": la, "results": results}

    # -- device-resident -----------------------------------------------------------
    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self._call("gnlse_malloc", int(nbytes), ctypes.byref(p))
        return p

    def free(self, p):
        self._call("gnlse_free", p)

    def htod(self, dptr, host):
        host = np.ascontiguousarray(host)
        self._call("gnlse_memcpy_htod", dptr, _ptr(host), host.nbytes)

    def dtoh(self, host, dptr):
        self._call("gnlse_memcpy_dtoh", _ptr(host), dptr, host.nbytes)

    def dtod(self, dst, src, nbytes):
        self._call("gnlse_memcpy_dtod", dst, src, int(nbytes))

    def propagate_adaptive_device(self, method, batch, params, d_in, d_out):
        res = (InstanceResult * int(batch))()
        self._pump_in(batch)
        self._call("gnlse_propagate_adaptive_device", int(method), int(batch), ctypes.byref(params), d_in, d_out, res)
        self._pump_out(batch)
        return [{"z": r.z, "h": r.h, "n_attempts": r.n_attempts, "n_accepted": r.n_accepted,
                 "n_traces": r.n_traces, "status": r.status} for r in res]

    def copy_rows(self, row_bytes, src_rows, dst_rows, d_src, d_dst):
        """Row src_rows[r] of d_src to row dst_rows[r] of d_dst, one launch."""
        s = np.ascontiguousarray(src_rows, dtype=np.int64)
        d = np.ascontiguousarray(dst_rows, dtype=np.int64)
        if s.size:
            self._call("gnlse_copy_rows", s.size, int(row_bytes), _ptr(s), _ptr(d), d_src, d_dst)

    def propagate_adaptive_device_noalloc(self, method, batch, params, d_in, d_out, d_res):
        """Enqueue the no-allocation propagation; d_res receives batch
        InstanceResult records in device memory (read them with dtoh)."""
        self._call("gnlse_propagate_adaptive_device_noalloc", int(method), int(batch), ctypes.byref(params), d_in,
                   d_out, d_res)

    def set_schedule(self, index):
        """Select the schedule of the spectral-domain RK4IP launches (0 is
        the grid's default; see schedules())."""
        self._call("gnlse_set_schedule", int(index))

    def set_pairing(self, enable):
        """Two workgroups per instance for adaptive runs whose batch leaves
        room on the device (on by default; results are identical)."""
        self._call("gnlse_set_pairing", 1 if enable else 0)

    def get_schedule(self):
        v = ctypes.c_int()
        self._call("gnlse_get_schedule", ctypes.byref(v))
        return v.value

    def tune_adaptive(self, A, params, trial_attempts=40, repeats=3):
        """Autotuner: time every schedule of this grid on trial_attempts
        replayed steps of params.h0 over the rows of A (repeats runs each, the
        first untimed), keep the fastest for this context. Returns (index,
        list of ms per schedule)."""
        a, _ = self._batched(A)
        c = ctypes.c_int()
        ms = np.zeros(len(schedules(self.n)), dtype=np.float64)
        self._call("gnlse_tune_adaptive", a.shape[0], ctypes.byref(params), _ptr(a), int(trial_attempts), int(repeats),
                   ctypes.byref(c), _ptr(ms))
        return c.value, ms.tolist()

    def last_kernel_ms(self):
        v = ctypes.c_double()
        self._call("gnlse_last_kernel_ms", ctypes.byref(v))
        return v.value

    def trace_characteristics(self, A, simpson_w, max_peaks=256):
        """Storage.update_characts quantities of every row of A on the device:
        dict of energy, max_power, duration (samples), spec_width (samples,
        NaN without a qualifying spectral peak), peaks (list of index arrays)."""
        a, _ = self._batched(A)
        w = _f64(simpson_w, self.n)
        k = a.shape[0]
        out = np.empty((k, 4), dtype=np.float64)
        npk = np.empty(k, dtype=np.int32)
        pk = np.empty((k, max_peaks), dtype=np.int32)
        self._call("gnlse_trace_characteristics", k, _ptr(a), _ptr(w), _ptr(out), _ptr(npk), _ptr(pk), int(max_peaks))
        if np.any(npk > max_peaks):
            raise GNLSEError(f"more than {max_peaks} temporal peaks in a trace")
        return {"energy": out[:, 0], "max_power": out[:, 1], "duration": out[:, 2], "spec_width": out[:, 3],
                "peaks": [pk[i, :npk[i]].astype(np.int64) for i in range(k)]}

    def trace_powers(self, A, n_out, mode=DOWNSAMPLE_PYOFSS):
        """Down-sampled temporal and spectral powers (natural order) of every
        row of A on the device: (temporal[k, n_out], spectral[k, n_out])."""
        a, _ = self._batched(A)
        k = a.shape[0]
        t = np.empty((k, int(n_out)), dtype=np.float64)
        sp = np.empty((k, int(n_out)), dtype=np.float64)
        self._call("gnlse_trace_powers", k, _ptr(a), int(n_out), int(mode), _ptr(t), _ptr(sp))
        return t, sp


def _results(res):
    return [{"z": r.z, "h": r.h, "n_attempts": r.n_attempts, "n_accepted": r.n_accepted, "n_traces": r.n_traces,
             "status": r.status} for r in res]


class Ensemble:
    """Realisations on up to four grids (4096/8192/16384) advanced in one
    launch, each with its own adaptive step sizes, plus the spectral
    coherence |g12| of each group on the device.

    groups: list of (context, batch); every context holds its grid's
    operators and must outlive the ensemble."""

    def __init__(self, method, params, groups):
        self.lib = load_library()
        self.groups = [(ctx, int(b)) for ctx, b in groups]
        self.contexts = [ctx for ctx, _ in self.groups]  # keep them alive
        handles = (ctypes.c_void_p * len(self.groups))(*[ctx._h for ctx, _ in self.groups])
        batch = (ctypes.c_int64 * len(self.groups))(*[b for _, b in self.groups])
        h = ctypes.c_void_p()
        self.params = params
        _check(self.lib, self.lib.gnlse_ensemble_create(int(method), ctypes.byref(params), len(self.groups),
                                                        ctypes.cast(handles, ctypes.c_void_p),
                                                        ctypes.cast(batch, ctypes.c_void_p), ctypes.byref(h)),
               "gnlse_ensemble_create")
        self._h = h
        self.lock = threading.RLock()

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gnlse_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        with self.lock:
            _check(self.lib, getattr(self.lib, name)(self._h, *args), name)

    def tune(self, threads=0, trial_attempts=32):
        """threads 512 or 1024 fixes the launch shape of every group; 0 times
        each group alone at both on a short trial and keeps the faster for
        that group. Returns the thread count of the largest grid's group
        (group_threads gives each group's)."""
        c = ctypes.c_int()
        self._call("gnlse_ensemble_tune", int(threads), int(trial_attempts), ctypes.byref(c))
        return c.value

    @property
    def threads(self):
        c = ctypes.c_int()
        self._call("gnlse_ensemble_threads", ctypes.byref(c))
        return c.value

    def group_threads(self, group):
        c = ctypes.c_int()
        self._call("gnlse_ensemble_group_threads", int(group), ctypes.byref(c))
        return c.value

    def set_group_threads(self, group, threads):
        self._call("gnlse_ensemble_set_group_threads", int(group), int(threads))

    def device_ptrs(self, group):
        a, b, g = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._call("gnlse_ensemble_device_ptrs", int(group), ctypes.byref(a), ctypes.byref(b), ctypes.byref(g))
        return a.value, b.value, g.value

    def set_input(self, group, A):
        ctx, b = self.groups[group]
        a = _c128(A, ctx.n).reshape(b, ctx.n)
        self._call("gnlse_ensemble_set_input", int(group), _ptr(a))

    def run(self, graph=False):
        self._call("gnlse_ensemble_run", 1 if graph else 0)
        return self.last_ms()

    def output(self, group):
        ctx, b = self.groups[group]
        A = np.empty((b, ctx.n), dtype=np.complex128)
        g12 = np.empty(ctx.n, dtype=np.float64)
        res = (InstanceResult * b)()
        self._call("gnlse_ensemble_output", int(group), _ptr(A), _ptr(g12), ctypes.cast(res, ctypes.c_void_p))
        return {"A": A, "g12": g12, "results": _results(res)}

    def last_ms(self):
        v = ctypes.c_double()
        self._call("gnlse_ensemble_last_ms", ctypes.byref(v))
        return v.value


_default = {}


def default_context(device=0):
    """A process-wide context for standalone operator calls (field.fft, ...)."""
    ctx = _default.get(device)
    if ctx is None:
        ctx = Context(device)
        _default[device] = ctx
    return ctx
