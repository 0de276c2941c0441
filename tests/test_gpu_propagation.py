"""GPU parity of whole propagations against the oracle (complex128).

Tolerance: relative L2 field error <= 1e-9 against the oracle's NumPy
complex128 path on identical inputs, for every returned field. Adaptive runs
must also reproduce the oracle's attempt sequence: identical accept/reject
decisions and attempt count, step sizes to 1e-6 relative. The controller's
error estimate delta is the norm of A_fine - A_coarse, a difference of two
fields that agree to ~delta ~ 1e-6, so two FP64 implementations whose fields
differ by e (~1e-12 relative) give deltas that differ by ~e/delta (~1e-6);
h depends on delta through its fifth root.
"""

import math

import numpy as np
import pytest

import oracle as O
from helpers import (C_NM_PS, amplifier_c4, engine_context, nl_kind, oracle_amplifier, oracle_ops, rel_l2, soliton_c1,
                     supercontinuum_c2)

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-9


@pytest.fixture(scope="module")
def lib():
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return _lib


def oracle_fixed(dom, A, fib, method, steps, traces):
    f = O.Fibre(length=fib["length"], beta=fib["beta"], gamma=fib["gamma"], method=method, total_steps=steps,
                traces=traces)
    out = f(dom, A.copy())
    st = f.stepper.storage
    return out, st.z, st.As, st.fft_total


@pytest.mark.parametrize("method,mid", [("rk4ip", 0), ("ss_symmetric", 1)])
def test_c1_soliton_fixed(lib, method, mid):
    """C1: N=2 soliton over one soliton period, 1000 fixed steps, 50 traces."""
    dom, A, fib = soliton_c1(4096)
    ref, ref_z, ref_As, _ = oracle_fixed(dom, A, fib, method, 1000, 50)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    h = fib["length"] / 1000
    got, tr, ntr = ctx.propagate_fixed(mid, A, h, 1000, trace_every=20, max_traces=50)
    err = rel_l2(got, ref)
    print(f"C1 {method}: rel L2 {err:.3e}, traces {ntr[0]}, kernel {ctx.last_kernel_ms():.2f} ms", flush=True)
    assert err <= FIELD_TOL
    assert ntr[0] == len(ref_As) == 50
    for k in range(50):
        assert rel_l2(tr[0, k], ref_As[k]) <= FIELD_TOL


def test_c1_batch_is_bitwise_consistent(lib):
    """Every instance of a batch computes exactly what a single run computes."""
    dom, A, fib = soliton_c1(4096)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    h = fib["length"] / 200
    rng = np.random.default_rng(0)
    batch = np.stack([A * (1 + 0.01 * k) + 1e-3 * rng.standard_normal(A.size) for k in range(5)])
    got, _, _ = ctx.propagate_fixed(0, batch, h, 200)
    for k in range(5):
        one, _, _ = ctx.propagate_fixed(0, batch[k], h, 200)
        assert np.array_equal(one, got[k])


def adaptive_oracle(dom, A, fib, method, tol, traces=1):
    amp = fib.get("amplifier") or {}
    f = O.Fibre(length=fib["length"], beta=fib["beta"], gamma=fib["gamma"], method=method, local_error=tol,
                traces=traces, self_steepening=fib.get("self_steepening", False), use_all=fib.get("use_all", False),
                small_signal_gain=amp.get("gain"), E_sat=amp.get("E_sat"), lamb0=amp.get("lamb0"),
                bandwidth=amp.get("bandwidth"))
    out = f(dom, A.copy())
    return out, f


def adaptive_params(lib, fib, tol, traces=1, eta=5, total_steps=100):
    p = lib.AdaptiveParams()
    p.length = fib["length"]
    h = fib["length"] / traces if traces > total_steps else fib["length"] / total_steps
    p.h0 = min(h, 1e-6)
    p.local_error = tol
    p.safety, p.max_factor, p.min_factor, p.step_size_min = 0.9, 10.0, 0.2, 1e-37
    p.traces, p.total_attempts, p.steps_max, p.eta = traces, 1000, 100000, eta
    return p


@pytest.mark.parametrize("N,length", [(1024, 1.0e-5), (8192, 1.0e-5), (256, 1.0e-5), (512, 1.0e-5), (2048, 1.0e-5),
                                      (4096, 1.0e-5), (16384, 2.0e-6)])
def test_c2_adaptive_short(lib, N, length):
    """C2 physics (Raman + self-steepening, tol 1e-6) over the first part of
    the fibre: same attempt sequence and final field as the oracle."""
    dom, A, fib = supercontinuum_c2(N)
    fib["length"] = length
    ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
    log = f.stepper.attempt_log
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    out = ctx.propagate_adaptive(0, A, adaptive_params(lib, fib, 1e-6), log_cap=len(log) + 8)
    r = out["results"][0]
    err = rel_l2(out["A"], ref)
    print(f"C2 N={N} L={length}: attempts {r['n_attempts']} (oracle {len(log)}), rel L2 {err:.3e}, "
          f"kernel {ctx.last_kernel_ms():.1f} ms", flush=True)
    assert r["status"] == 0
    assert r["n_attempts"] == len(log)
    ref_h = np.array([a[0] for a in log])
    ref_acc = np.array([a[2] for a in log], dtype=np.uint8)
    dh = np.max(np.abs(out["log_h"][0, :len(log)] / ref_h - 1))
    print(f"  max rel step-size difference {dh:.2e}", flush=True)
    assert np.array_equal(out["log_accept"][0, :len(log)], ref_acc)
    assert dh < 1e-6
    assert err <= FIELD_TOL


@pytest.mark.parametrize("N", [512, 8192])
def test_c2_physics_fixed_steps(lib, N):
    """Raman + self-steepening with fixed rk4ip steps (the spectral-domain
    fixed-step launch): every trace and the output field against the
    oracle's Fibre run."""
    dom, A, fib = supercontinuum_c2(N)
    steps, traces, length = 200, 10, 2.0e-6
    f = O.Fibre(length=length, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True,
                method="rk4ip", total_steps=steps, traces=traces)
    ref = f(dom, A.copy())
    ref_As = f.stepper.storage.As
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    got, tr, ntr = ctx.propagate_fixed(0, A, length / steps, steps, trace_every=steps // traces, max_traces=traces)
    err = rel_l2(got, ref)
    print(f"C2 physics fixed N={N}: rel L2 {err:.3e}, traces {ntr[0]}", flush=True)
    assert err <= FIELD_TOL
    assert ntr[0] == len(ref_As) == traces
    for k in range(traces):
        assert rel_l2(tr[0, k], ref_As[k]) <= FIELD_TOL


def test_c2_replay_matches_oracle(lib):
    """Replay the oracle's step sequence: isolates arithmetic from control."""
    dom, A, fib = supercontinuum_c2(8192)
    fib["length"] = 1.0e-5
    ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
    log = f.stepper.attempt_log
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    rh = np.array([a[0] for a in log])
    ra = np.array([a[2] for a in log], dtype=np.uint8)
    out = ctx.propagate_adaptive(0, A, adaptive_params(lib, fib, 1e-6), log_cap=len(log), replay=(rh, ra))
    d_dev = out["log_delta"][0]
    d_ref = np.array([a[1] for a in log])
    err = rel_l2(out["A"], ref)
    print(f"replay: rel L2 {err:.3e}, max rel delta diff {np.max(np.abs(d_dev - d_ref) / d_ref):.2e}", flush=True)
    assert out["results"][0]["status"] == 0
    assert np.allclose(d_dev, d_ref, rtol=1e-6)
    assert err <= FIELD_TOL


def amplifier_batch(lib, N, length, chirps):
    """C4 runs for several chirps: oracle runs, and the engine context loaded
    with the (chirp-independent) operators."""
    refs, logs, inputs = [], [], []
    for C in chirps:
        dom, A, fib = amplifier_c4(N, C)
        fib["length"] = length
        ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
        refs.append(ref)
        logs.append(f.stepper.attempt_log)
        inputs.append(A)
    amp = oracle_amplifier(dom, length, **fib["amplifier"])
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], use_all=True, amplifier=amp)
    ctx = engine_context(dom, lin, nl, 1, amplifier=amp)
    return ctx, fib, np.stack(inputs), refs, logs


@pytest.mark.parametrize("N,length,chirps", [(2048, 5e-3, (-2.0, 0.5, 2.0)), (16384, 2e-4, (-2.0, 0.5, 2.0)),
                                             (4096, 1e-3, (-2.0, 0.5)), (8192, 5e-4, (-1.0, 2.0)),
                                              (16384, 5e-3, (1.0,))], ids=["2048-full", "16384-short", "4096-short", "8192-short", "16384-full"])
def test_c4_amplifier_batch(lib, N, length, chirps):
    """C4 physics (Lorentzian saturable gain inside every linear sub-step,
    Raman, adaptive RK4IP): chirps in one launch, each instance with its own
    step sequence, each equal to the oracle run of its chirp. The 16384-full
    case is the whole 5 m amplifier (about 750 attempted steps)."""
    ctx, fib, A, refs, logs = amplifier_batch(lib, N, length, chirps)
    out = ctx.propagate_adaptive(0, A, adaptive_params(lib, fib, 1e-6), log_cap=max(len(l) for l in logs) + 8)
    print(f"C4 N={N} L={length}: kernel {ctx.last_kernel_ms():.1f} ms", flush=True)
    for i, C in enumerate(chirps):
        r, log = out["results"][i], logs[i]
        err = rel_l2(out["A"][i], refs[i])
        ref_h = np.array([a[0] for a in log])
        ref_acc = np.array([a[2] for a in log], dtype=np.uint8)
        dh = np.max(np.abs(out["log_h"][i, :len(log)] / ref_h - 1))
        print(f"  C={C}: attempts {r['n_attempts']} (oracle {len(log)}), rel L2 {err:.3e}, dh {dh:.1e}", flush=True)
        assert r["status"] == 0
        assert r["n_attempts"] == len(log)
        assert np.array_equal(out["log_accept"][i, :len(log)], ref_acc)
        assert dh < 1e-6
        assert err <= FIELD_TOL


@pytest.mark.parametrize("N", [2048, 16384])
def test_c4_physics_fixed_steps(lib, N):
    """The saturable-gain amplifier with fixed rk4ip steps (the spectral-domain
    fixed-step launch with the gain): every trace and the output field."""
    dom, A, fib = amplifier_c4(N, 0.5)
    amp_args = fib["amplifier"]
    steps, traces, length = 120, 4, 2e-4
    f = O.Fibre(length=length, beta=fib["beta"], gamma=fib["gamma"], method="rk4ip", total_steps=steps,
                traces=traces, use_all=True, small_signal_gain=amp_args["gain"], E_sat=amp_args["E_sat"],
                lamb0=amp_args["lamb0"], bandwidth=amp_args["bandwidth"])
    ref = f(dom, A.copy())
    ref_As = f.stepper.storage.As
    amp = oracle_amplifier(dom, length, **amp_args)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], use_all=True, amplifier=amp)
    ctx = engine_context(dom, lin, nl, 1, amplifier=amp)
    got, tr, ntr = ctx.propagate_fixed(0, A, length / steps, steps, trace_every=steps // traces, max_traces=traces)
    err = rel_l2(got, ref)
    print(f"C4 physics fixed N={N}: rel L2 {err:.3e}, traces {ntr[0]}", flush=True)
    assert err <= FIELD_TOL
    assert ntr[0] == len(ref_As) == traces
    for k in range(traces):
        assert rel_l2(tr[0, k], ref_As[k]) <= FIELD_TOL


@pytest.mark.parametrize("N,length", [(4096, 1e-3), (16384, 2e-4)])
def test_kerr_amplifier_adaptive_and_fixed(lib, N, length):
    """The saturable gain with the Kerr operator alone (no Raman): the
    time-domain Kerr kernels that keep the gain branch (the gain-free Kerr
    launches compile without it). Adaptive rk4ip of two chirps in one launch
    (the oracle's attempt sequences), then fixed rk4ip steps."""
    refs, logs, inputs = [], [], []
    for C in (-1.0, 1.5):
        dom, A, fib = amplifier_c4(N, C)
        fib = dict(fib, length=length, use_all=False)
        ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
        refs.append(ref)
        logs.append(f.stepper.attempt_log)
        inputs.append(A)
    amp = oracle_amplifier(dom, length, **fib["amplifier"])
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], amplifier=amp)
    ctx = engine_context(dom, lin, nl, 0, amplifier=amp)
    out = ctx.propagate_adaptive(0, np.stack(inputs), adaptive_params(lib, fib, 1e-6),
                                 log_cap=max(len(l) for l in logs) + 8)
    for i, (ref, log) in enumerate(zip(refs, logs)):
        r = out["results"][i]
        err = rel_l2(out["A"][i], ref)
        print(f"Kerr amplifier N={N} chirp {i}: attempts {r['n_attempts']} (oracle {len(log)}), rel L2 {err:.3e}",
              flush=True)
        assert r["status"] == 0
        assert r["n_attempts"] == len(log)
        assert np.array_equal(out["log_accept"][i, :len(log)], np.array([a[2] for a in log], dtype=np.uint8))
        assert err <= FIELD_TOL
    steps = 40
    f = O.Fibre(length=length, beta=fib["beta"], gamma=fib["gamma"], method="rk4ip", total_steps=steps,
                small_signal_gain=fib["amplifier"]["gain"], E_sat=fib["amplifier"]["E_sat"],
                lamb0=fib["amplifier"]["lamb0"], bandwidth=fib["amplifier"]["bandwidth"])
    ref = f(dom, inputs[0].copy())
    got, _, _ = ctx.propagate_fixed(0, inputs[0], length / steps, steps)
    err = rel_l2(got, ref)
    print(f"Kerr amplifier N={N} fixed: rel L2 {err:.3e}", flush=True)
    assert err <= FIELD_TOL


@pytest.mark.parametrize("N", [8192, 16384])
def test_kerr_soliton_large_grids(lib, N):
    """The gain-free Kerr kernels (written-out rk4ip, no gain branch) on the
    large grids through the single-grid launches: adaptive rk4ip of the N = 2
    soliton over 1/16 of its period with two workgroups per instance and with
    one, and 64 fixed rk4ip steps, each against the oracle."""
    dom, A, fib = soliton_c1(N)  # the whole period: 155 attempted steps
    ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
    log = f.stepper.attempt_log
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    for pair in (True, False):
        ctx.set_pairing(pair)
        out = ctx.propagate_adaptive(0, A, adaptive_params(lib, fib, 1e-6), log_cap=len(log) + 8)
        r = out["results"][0]
        err = rel_l2(out["A"], ref)
        print(f"Kerr soliton N={N} pairing {pair}: attempts {r['n_attempts']} (oracle {len(log)}), "
              f"rel L2 {err:.3e}", flush=True)
        assert r["status"] == 0 and r["n_attempts"] == len(log)
        assert np.array_equal(out["log_accept"][0, :len(log)], np.array([a[2] for a in log], dtype=np.uint8))
        assert err <= FIELD_TOL
    steps = 64
    fo = O.Fibre(length=fib["length"], beta=fib["beta"], gamma=fib["gamma"], method="rk4ip", total_steps=steps)
    ref_f = fo(dom, A.copy())
    got, _, _ = ctx.propagate_fixed(0, A, fib["length"] / steps, steps)
    err = rel_l2(got, ref_f)
    print(f"Kerr soliton N={N} fixed: rel L2 {err:.3e}", flush=True)
    assert err <= FIELD_TOL


GENERIC_KINDS = [dict(self_steepening=True), dict(raman_scattering=True),
                 dict(self_steepening=True, raman_scattering=True)]


@pytest.mark.parametrize("flags", GENERIC_KINDS, ids=lambda d: "+".join(sorted(d)))
@pytest.mark.parametrize("N", [4096, 16384])
def test_generic_kinds_adaptive_and_fixed(lib, flags, N):
    """The self-steepening and T_R Raman operators (nonlinearity.py:223-256)
    through the gain-free propagating kernels: adaptive rk4ip over the first
    1e-5 km (4096 samples) or 4e-6 km (16384) of the C2 fibre (the oracle's
    attempt sequence) and 50 fixed rk4ip steps, against the oracle."""
    dom, A, fib = supercontinuum_c2(N)
    length = 1e-5 if N == 4096 else 4e-6
    f = O.Fibre(length=length, beta=fib["beta"], gamma=fib["gamma"], method="ark4ip", local_error=1e-6, **flags)
    ref = f(dom, A.copy())
    log = f.stepper.attempt_log
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], **flags)
    ctx = engine_context(dom, lin, nl, nl_kind(**flags))
    out = ctx.propagate_adaptive(0, A, adaptive_params(lib, dict(fib, length=length), 1e-6), log_cap=len(log) + 8)
    r = out["results"][0]
    err = rel_l2(out["A"], ref)
    print(f"{flags} N={N}: attempts {r['n_attempts']} (oracle {len(log)}), rel L2 {err:.3e}", flush=True)
    assert r["status"] == 0 and r["n_attempts"] == len(log)
    assert np.array_equal(out["log_accept"][0, :len(log)], np.array([a[2] for a in log], dtype=np.uint8))
    assert err <= FIELD_TOL
    steps = 50
    fo = O.Fibre(length=length, beta=fib["beta"], gamma=fib["gamma"], method="rk4ip", total_steps=steps, **flags)
    ref_f = fo(dom, A.copy())
    got, _, _ = ctx.propagate_fixed(0, A, length / steps, steps)
    err = rel_l2(got, ref_f)
    print(f"{flags} N={N} fixed: rel L2 {err:.3e}", flush=True)
    assert err <= FIELD_TOL


def test_adaptive_nan_input_raises_small_step(lib):
    """stepper.py:234-237, 349-351: NaN fields are rejected until h < 1e-37."""
    dom, A, fib = soliton_c1(1024)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    A = A.copy()
    A[10] = np.nan
    out = ctx.propagate_adaptive(0, A, adaptive_params(lib, fib, 1e-6))
    assert out["results"][0]["status"] == lib.STATUS_SMALL_STEP


def test_spectral_adaptive_nan_and_large_batch(lib):
    """The spectral-domain adaptive kernel (Raman + self-steepening): a NaN
    field ends in SMALL_STEP (stepper.py:234-237, 349-351) beside regular
    instances, and a batch larger than the CU count (several workgroups per
    CU over the launch) gives every instance exactly its single-run result."""
    dom, A, fib = supercontinuum_c2(1024)
    fib["length"] = 2e-6
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    rng = np.random.default_rng(3)
    B = 300
    batch = np.stack([A * (1 + 0.002 * (k % 7)) + 1e-2 * rng.standard_normal(A.size) for k in range(B)])
    batch[17, 5] = np.nan
    params = adaptive_params(lib, fib, 1e-6)
    out = ctx.propagate_adaptive(0, batch, params)
    st = [r["status"] for r in out["results"]]
    assert st[17] == lib.STATUS_SMALL_STEP
    assert all(s == 0 for k, s in enumerate(st) if k != 17)
    for k in (0, 1, 150, 299):
        one = ctx.propagate_adaptive(0, batch[k], params)
        assert one["results"][0]["n_attempts"] == out["results"][k]["n_attempts"]
        assert np.array_equal(one["A"][0] if one["A"].ndim == 2 else one["A"], out["A"][k])
    ref, f = adaptive_oracle(dom, batch[150], fib, "ark4ip", 1e-6)
    assert out["results"][150]["n_attempts"] == len(f.stepper.attempt_log)
    assert rel_l2(out["A"][150], ref) <= FIELD_TOL


def test_adaptive_steps_max(lib):
    dom, A, fib = soliton_c1(1024)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    p = adaptive_params(lib, fib, 1e-9)
    p.steps_max = 3
    out = ctx.propagate_adaptive(0, A, p)
    assert out["results"][0]["status"] == lib.STATUS_MAX_STEPS
    assert out["results"][0]["n_accepted"] == 2
