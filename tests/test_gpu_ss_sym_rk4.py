"""The symmetric split step with the classical rk4 on the nonlinear operator,
Solver.ss_sym_rk4 (pyofss/modules/solver.py:330-337) and the OpenCL fibre's
cl_ss_sym_rk4 (opencl_fibre.py:929-970, the same scheme), on the GPU against
the oracle's ss_sym_rk4 (complex128, relative L2).

Bars: one integrator call <= 1e-12 (Kerr; Raman + self-steepening; 4096
and 32768 samples); the soliton over one period with 1000 fixed steps and 50
traces through System/Fibre <= 1e-9 with fft_total equal; an adaptive run
("ass_sym_rk4", local error order 5) with the identical attempt count and
the field <= 1e-9; the OpenCL front end's cl_ss_sym_rk4 against the oracle's
fixed-step ss_sym_rk4 <= 1e-9."""

import math

import numpy as np
import pytest

import oracle as O
from helpers import C_NM_PS, DUDLEY_BETA, engine_context, nl_kind, oracle_ops, rel_l2, supercontinuum_c2

pytestmark = pytest.mark.gpu
FIELD_TOL = 1e-9


@pytest.fixture(scope="module")
def P():
    import pyofsschanged_amd as P
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return P


@pytest.mark.parametrize("N", [4096, 32768])
@pytest.mark.parametrize("flags", [dict(), dict(use_all=True, self_steepening=True)], ids=["kerr", "all+ss"])
def test_one_call(P, N, flags):
    from pyofsschanged_amd import _lib
    dom, A, fib = supercontinuum_c2(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], **flags)
    ctx = engine_context(dom, lin, nl, nl_kind(**flags))

    class F:
        pass

    f = F()
    f.linear = lin.default_exp_f
    f.n = nl.non
    h = 5e-7
    ref = O.Solver("ss_sym_rk4", f)(A, 0.0, h)
    err = rel_l2(ctx.step(_lib.METHOD_SS_SYM_RK4, A, h), ref)
    print(f"N={N} {flags} ss_sym_rk4 call err {err:.2e}", flush=True)
    assert err < 1e-12


def test_system_soliton_fixed(P):
    d = (1, 4096, 100.0, C_NM_PS / 1550.0)
    kw = dict(length=math.pi / 40, beta=[0.0, 0.0, -20.0], gamma=1.3, method="ss_sym_rk4", total_steps=1000,
              traces=50)
    sp_ = P.System(P.Domain(*d))
    sp_.add(P.Sech(peak_power=4 * 20.0 / 1.3, width=1.0))
    fp = P.Fibre(**kw)
    sp_.add(fp)
    sp_.run()
    so = O.System(O.Domain(*d))
    so.add(O.Sech(peak_power=4 * 20.0 / 1.3, width=1.0))
    fo = O.Fibre(**kw)
    so.add(fo)
    so.run()
    err = rel_l2(sp_.field, so.field)
    tr = max(rel_l2(a, b) for a, b in zip(fp.stepper.storage.As, fo.stepper.storage.As))
    print(f"soliton ss_sym_rk4: rel L2 {err:.2e}, traces max {tr:.2e}", flush=True)
    assert err <= FIELD_TOL and tr <= FIELD_TOL
    assert len(fp.stepper.storage.As) == 50
    assert fp.stepper.storage.fft_total == fo.stepper.storage.fft_total


def test_adaptive(P):
    d = (1, 2048, 12.5, C_NM_PS / 835.0)
    kw = dict(length=4e-6, beta=DUDLEY_BETA, gamma=110.0, self_steepening=True, use_all=True, method="ass_sym_rk4",
              local_error=1e-6)
    sp_ = P.System(P.Domain(*d))
    sp_.add(P.Sech(peak_power=1e4, width=0.050, using_fwhm=True))
    fp = P.Fibre(**kw)
    sp_.add(fp)
    sp_.run()
    so = O.System(O.Domain(*d))
    so.add(O.Sech(peak_power=1e4, width=0.050, using_fwhm=True))
    fo = O.Fibre(**kw)
    so.add(fo)
    so.run()
    err = rel_l2(sp_.field, so.field)
    print(f"adaptive ss_sym_rk4: attempts {fp.stepper.attempts} (oracle {len(fo.stepper.attempt_log)}), "
          f"rel L2 {err:.2e}", flush=True)
    assert fp.stepper.attempts == len(fo.stepper.attempt_log)
    assert err <= FIELD_TOL


def test_opencl_front_end(P):
    d = (1, 4096, 100.0, C_NM_PS / 1550.0)
    dom = P.Domain(*d)
    prog = P.OpenclProgramm(fast_math=True, use_all=False, dorf="double", ctx=None)
    prog.set_domain(dom)
    sp_ = P.System(dom)
    sp_.add(P.Sech(peak_power=4 * 20.0 / 1.3, width=1.0))
    sp_.add(P.OpenclFibre(cl_programm=prog, length=math.pi / 80, total_steps=300, beta=np.array([0.0, 0.0, -20.0]),
                          gamma=1.3, method="cl_ss_sym_rk4", traces=10))
    sp_.run()
    so = O.System(O.Domain(*d))
    so.add(O.Sech(peak_power=4 * 20.0 / 1.3, width=1.0))
    so.add(O.Fibre(length=math.pi / 80, beta=[0.0, 0.0, -20.0], gamma=1.3, method="ss_sym_rk4", total_steps=300))
    so.run()
    err = rel_l2(sp_.field, so.field)
    print(f"OpenCL front end cl_ss_sym_rk4: rel L2 {err:.2e}", flush=True)
    assert err <= FIELD_TOL
