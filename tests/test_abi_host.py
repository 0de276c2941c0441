"""CPU tests: the C-ABI library loads and exports every function that
include/gnlse.h declares, and the host-side logic (grids, pulses, set-up
arrays, trace schedules, status mapping) matches the oracle. No GPU calls."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import C_NM_PS, DUDLEY_BETA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "gnlse.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gnlse_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    from pyofsschanged_amd import _lib
    lib = _lib.load_library()
    names = header_functions()
    assert len(names) >= 20
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for name in names:
        assert hasattr(raw, name), name
    # the binding declares exactly the header's functions
    assert set(_lib.SIGNATURES) | {"gnlse_last_error"} == set(names)
    assert lib.gnlse_version() == 1
    for n in (256, 512, 1024, 2048, 4096, 8192, 16384):
        assert _lib.grid_supported(n)
    assert not _lib.grid_supported(1000)
    assert _lib.grid_supported(32768) and _lib.grid_supported(65536)
    assert not _lib.grid_supported(131072)


def _radices(plan):
    return [(int(r.rstrip("L")), r.endswith("L")) for r in plan.split(".")]


def test_schedules_describe_valid_plans():
    """Every compiled schedule of the spectral-domain RK4IP kernels: its
    N-point and N/2-point pass plans factor the transform length, every
    in-thread radix divides the points per thread P and every lane-pair pass
    has radix 2P (two lanes of P points each); schedule 0 is the grid's
    default launch shape."""
    from pyofsschanged_amd import _lib
    for n in (4096, 8192, 16384):
        sc = _lib.schedules(n)
        assert len(sc) >= 2 and [s["index"] for s in sc] == list(range(len(sc)))
        descs = set()
        for s in sc:
            parts = s["desc"].split(" ")
            t_part, n_part, r_part, tile = parts[:4]
            assert parts[4:] in ([], ["unrolled"], ["passes", "unrolled"], ["raman", "unrolled"]), s["desc"]
            T = int(t_part[1:])
            assert T == s["threads"] and T % 64 == 0 and tile.startswith("tile")
            for length, plan in ((n, n_part[2:]), (n // 2, r_part[2:])):
                P = length // T
                rad = _radices(plan)
                assert int(np.prod([r for r, _ in rad])) == length, s["desc"]
                for r, lane in rad:
                    assert (r == 2 * P) if lane else (P % r == 0), s["desc"]
            descs.add(s["desc"])
        assert len(descs) == len(sc)  # no duplicate schedule
    # defaults: a lane-pair pass in the 8192- and 16384-point transforms and
    # unrolled passes; schedule 1 keeps the passes in their loops
    assert _lib.schedules(8192)[0]["desc"] == "T512 N:32L.16.16 R:8.8.8.8 tile8 unrolled"
    assert _lib.schedules(8192)[1]["desc"] == "T512 N:32L.16.16 R:8.8.8.8 tile4"
    assert _lib.schedules(16384)[0]["desc"] == "T1024 N:32L.32L.16 R:8.8.8.2.8 tile2 unrolled"
    assert _lib.schedules(16384)[1]["desc"] == "T1024 N:32L.32L.16 R:8.8.8.2.8 tile4"
    assert _lib.schedules(4096)[0]["desc"] == "T512 N:8.8.8.8 R:4.4.4.4.2.4 tile4"

    for n in (256, 1024):
        assert len(_lib.schedules(n)) == 1
    assert _lib.schedules(32768)[0]["desc"] == "T1024 two-level"
    with pytest.raises(_lib.GNLSEError):
        _lib.schedules(1000)


def test_struct_layouts_match_header():
    from pyofsschanged_amd import _lib
    assert ctypes.sizeof(_lib.AdaptiveParams) == 7 * 8 + 3 * 8 + 2 * 4
    assert ctypes.sizeof(_lib.InstanceResult) == 2 * 8 + 3 * 8 + 2 * 4


def test_context_fails_loudly_without_gpu():
    from pyofsschanged_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.EngineUnavailableError):
        _lib.Context(0)


def test_domain_and_pulses_match_oracle():
    import pyofsschanged_amd as P
    for args in ((1, 4096, 100.0, C_NM_PS / 1550.0), (1, 8192, 12.5, C_NM_PS / 835.0), (2, 128, 25.0, 193.8, 2)):
        a, b = P.Domain(*args), O.Domain(*args)
        for attr in ("t", "nu", "omega", "Lambda"):
            assert np.array_equal(getattr(a, attr), getattr(b, attr))
        assert str(a) == str(b)
    d = P.Domain(1, 8192, 12.5, C_NM_PS / 835.0)
    od = O.Domain(1, 8192, 12.5, C_NM_PS / 835.0)
    assert np.array_equal(P.Sech(peak_power=1e4, width=0.05, using_fwhm=True).generate(d.t),
                          O.Sech(peak_power=1e4, width=0.05, using_fwhm=True).generate(od.t))
    assert np.array_equal(P.Gaussian(peak_power=4697.2, width=0.2, using_fwhm=True, C=1.5).generate(d.t),
                          O.Gaussian(peak_power=4697.2, width=0.2, using_fwhm=True, C=1.5).generate(od.t))
    assert str(P.Sech("sech", 0.2, 5.0, 1.4, 0.3, 0, 0.4, 0.0, 1)) == str(O.Sech("sech", 0.2, 5.0, 1.4, 0.3, 0, 0.4,
                                                                                  0.0, 1))


def test_linear_factor_matches_oracle():
    """The host set-up of linearity.py:162-186 (no device upload)."""
    import pyofsschanged_amd as P
    d = P.Domain(1, 8192, 12.5, C_NM_PS / 835.0)
    lin = P.Linearity(alpha=0.1, beta=DUDLEY_BETA)
    lin._upload = lambda: None
    lin(d)
    ref = O.Linearity(alpha=0.1, beta=DUDLEY_BETA)
    ref(O.Domain(1, 8192, 12.5, C_NM_PS / 835.0))
    assert np.array_equal(lin.factor, ref.factor)


def test_simpson_weights_match_old_rule():
    from pyofsschanged_amd.field import simpson_weights
    rng = np.random.default_rng(5)
    for N in (8, 9, 256, 257):
        x = np.cumsum(rng.random(N))
        y = rng.random(N)
        assert abs(simpson_weights(x) @ y - O.simps_avg(y, x)) < 1e-12 * abs(O.simps_avg(y, x))


def test_trace_schedule_and_initial_step():
    from pyofsschanged_amd.modules.stepper import adaptive_initial_step, fixed_trace_steps
    assert fixed_trace_steps(1000, 1) == []
    assert fixed_trace_steps(1000, 50) == list(range(0, 1000, 20))
    assert fixed_trace_steps(1000, 3) == [0, 333, 666, 999]
    with pytest.raises(ZeroDivisionError):
        fixed_trace_steps(10, 20)
    assert adaptive_initial_step(1.5e-4, 1, 100) == 1e-6  # L/total_steps = 1.5e-6, capped at 1e-6
    assert adaptive_initial_step(1e-5, 300, 100) == 1e-5 / 300  # L/traces when traces > total_steps


def test_status_maps_to_reference_exceptions():
    from pyofsschanged_amd import _lib
    from pyofsschanged_amd.modules import stepper as S
    S.raise_for_status(_lib.STATUS_OK)
    for code, exc in ((_lib.STATUS_SMALL_STEP, S.SmallStepSizeError),
                      (_lib.STATUS_SUITABLE_STEP, S.SuitableStepSizeError),
                      (_lib.STATUS_MAX_STEPS, S.MaximumStepsAllocatedError),
                      (_lib.STATUS_TRACE_OVERFLOW, S.EngineStatusError)):
        with pytest.raises(exc):
            S.raise_for_status(code)


def test_unsupported_methods_fail_loudly():
    import pyofsschanged_amd as P
    from pyofsschanged_amd.modules.solver import UnsupportedMethodError
    with pytest.raises(UnsupportedMethodError):
        P.Fibre(method="rk4")
    with pytest.raises(UnsupportedMethodError):
        P.Fibre(method="ass_agrawal")
    f = P.Fibre(method="ARK4IP")
    assert f.stepper.adaptive and f.stepper.eta == 5 and not f.linearity.use_cache
    f = P.Fibre(method="ss_symmetric")
    assert not f.stepper.adaptive and f.stepper.eta == 3 and f.linearity.use_cache


def test_fft_count_per_call():
    """Counts of pyofss's fft_counter per integrator call (field.py:118-134)."""
    import pyofsschanged_amd as P
    from pyofsschanged_amd import _lib, field
    cases = [(dict(), _lib.METHOD_RK4IP, 8), (dict(use_all=True), _lib.METHOD_RK4IP, 24),
             (dict(self_steepening=True), _lib.METHOD_RK4IP, 16), (dict(), _lib.METHOD_SS_SYMMETRIC, 4),
             (dict(use_all=True), _lib.METHOD_SS_SYMMETRIC, 8)]
    for kw, method, per_call in cases:
        f = P.Fibre(**kw)
        field.fft_counter = 0
        f._count_ffts(method, 10)
        assert field.fft_counter == 10 * per_call
