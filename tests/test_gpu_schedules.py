"""GPU parity of every compiled schedule of the spectral-domain adaptive
RK4IP kernels (threads per instance, pass plans of the N-point and N/2-point
transforms including lane-pair passes, block of the stage passes) against
the oracle, and the autotuner that picks among them.

Bar (as tests/test_gpu_propagation.py): the oracle's attempt sequence
(identical accept/reject decisions and count, step sizes to 1e-6 relative)
and output field relative L2 <= 1e-9. Workloads: the C2 supercontinuum
physics (Raman + self-steepening, no gain) over the first 1e-5 km (2e-6 km
at 16384 samples), and the C4 amplifier (saturable gain) over 5e-4 km at
8192 and 2e-4 km at 16384 samples.
"""

import functools

import numpy as np
import pytest

from helpers import amplifier_c4, engine_context, oracle_amplifier, oracle_ops, rel_l2, supercontinuum_c2
from test_gpu_propagation import adaptive_oracle, adaptive_params

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-9
GRIDS = (4096, 8192, 16384)


def _schedule_cases():
    try:
        from pyofsschanged_amd import _lib
        return [(n, s["index"]) for n in GRIDS for s in _lib.schedules(n)]
    except Exception:  # library not built: one case, which fails loudly when run
        return [(n, 0) for n in GRIDS]


@pytest.fixture(scope="module")
def lib():
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return _lib


@functools.lru_cache(maxsize=None)
def c2_reference(N):
    dom, A, fib = supercontinuum_c2(N)
    fib["length"] = 2e-6 if N == 16384 else 1e-5
    ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
    return dom, A, fib, ref, f.stepper.attempt_log


@functools.lru_cache(maxsize=None)
def c4_reference(N):
    dom, A, fib = amplifier_c4(N, 0.5)
    fib["length"] = 2e-4 if N == 16384 else 5e-4
    ref, f = adaptive_oracle(dom, A, fib, "ark4ip", 1e-6)
    return dom, A, fib, ref, f.stepper.attempt_log


def check_against_oracle(lib, ctx, A, fib, ref, log, tag):
    out = ctx.propagate_adaptive(0, A, adaptive_params(lib, fib, 1e-6), log_cap=len(log) + 8)
    r = out["results"][0]
    err = rel_l2(out["A"], ref)
    ref_h = np.array([a[0] for a in log])
    ref_acc = np.array([a[2] for a in log], dtype=np.uint8)
    n = min(len(log), int(r["n_attempts"]))
    dh = np.max(np.abs(out["log_h"][0, :n] / ref_h[:n] - 1))
    print(f"{tag}: attempts {r['n_attempts']} (oracle {len(log)}), rel L2 {err:.3e}, dh {dh:.1e}, "
          f"kernel {ctx.last_kernel_ms():.1f} ms", flush=True)
    assert r["status"] == 0
    assert r["n_attempts"] == len(log)
    assert np.array_equal(out["log_accept"][0, :len(log)], ref_acc)
    assert dh < 1e-6
    assert err <= FIELD_TOL
    return out


@pytest.mark.parametrize("N,index", _schedule_cases())
def test_schedule_c2_against_oracle(lib, N, index):
    """Raman + self-steepening without gain (the headline's kernel)."""
    dom, A, fib, ref, log = c2_reference(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    desc = lib.schedules(N)[index]["desc"]
    ctx.set_schedule(index)
    assert ctx.get_schedule() == index
    check_against_oracle(lib, ctx, A, fib, ref, log, f"C2 N={N} schedule {index} ({desc})")


@pytest.mark.parametrize("N,index", [c for c in _schedule_cases() if c[0] in (8192, 16384)])
def test_schedule_c4_against_oracle(lib, N, index):
    """The saturable-gain amplifier (gain energies reduced over the
    workgroup inside the stage passes)."""
    dom, A, fib, ref, log = c4_reference(N)
    amp = oracle_amplifier(dom, fib["length"], **fib["amplifier"])
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], use_all=True, amplifier=amp)
    ctx = engine_context(dom, lin, nl, 1, amplifier=amp)
    ctx.set_schedule(index)
    check_against_oracle(lib, ctx, A, fib, ref, log, f"C4 N={N} schedule {index}")


def test_autotuner_picks_a_schedule_and_keeps_it(lib):
    """gnlse_tune_adaptive times every schedule, selects the fastest for the
    context and leaves it selected: the next propagation is bit for bit the
    one of that schedule set by hand, and still the oracle's."""
    N = 8192
    dom, A, fib, ref, log = c2_reference(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    params = adaptive_params(lib, fib, 1e-6)
    batch = np.stack([A, A * (1 + 1e-3), A * (1 - 1e-3), A])
    chosen, ms = ctx.tune_adaptive(batch, params, trial_attempts=20, repeats=2)
    count = len(lib.schedules(N))
    print(f"tuner: chose {chosen} of {count}; trial ms {[round(m, 3) for m in ms]}", flush=True)
    assert 0 <= chosen < count
    assert len(ms) == count and all(m > 0 for m in ms)
    assert ms[chosen] == min(ms)
    assert ctx.get_schedule() == chosen
    tuned = check_against_oracle(lib, ctx, A, fib, ref, log, f"C2 N={N} tuned schedule {chosen}")
    ctx2 = engine_context(dom, lin, nl, 1)
    ctx2.set_schedule(chosen)
    by_hand = ctx2.propagate_adaptive(0, A, params, log_cap=len(log) + 8)
    assert np.array_equal(tuned["A"], by_hand["A"])


def test_schedule_selection_errors(lib):
    dom, A, fib, ref, log = c2_reference(4096)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"])
    ctx = engine_context(dom, lin, nl, 0)  # Kerr only: no schedules cover it
    with pytest.raises(lib.GNLSEError):
        ctx.set_schedule(len(lib.schedules(4096)))
    with pytest.raises(lib.GNLSEError, match="no schedules"):
        ctx.tune_adaptive(A, adaptive_params(lib, fib, 1e-6))
    # a new grid starts from its default schedule
    ctx.set_schedule(1)
    ctx.set_grid(8192)
    assert ctx.get_schedule() == 0


def test_pairing_equals_one_workgroup(lib):
    """Two workgroups per instance (the coarse step beside the half and fine
    steps; on by default when the batch leaves room) compute the operations
    of one workgroup in its order. Kerr soliton at 4096 (rk4ip and
    ss_symmetric, one field: a plain Fibre run): outputs, attempt logs,
    traces and results equal bit for bit. Saturable-gain amplifier at 16384
    samples (3 chirps): the unrolled spectral-domain passes of the two
    kernels are scheduled apart and may round differently (fused
    multiply-adds where the compiler forms them), so the attempt sequences
    are equal and the fields and traces agree to 1e-10, the outputs each
    within the oracle bar."""
    N = 16384
    ctx, fib, A, refs, logs = amplifier_batch_ctx(lib, N)
    params = adaptive_params(lib, fib, 1e-6, traces=4)
    cap = max(len(l) for l in logs) + 8
    ctx.set_pairing(False)
    one = ctx.propagate_adaptive(0, A, params, max_traces=8, log_cap=cap)
    t_one = ctx.last_kernel_ms()
    ctx.set_pairing(True)
    two = ctx.propagate_adaptive(0, A, params, max_traces=8, log_cap=cap)
    print(f"amplifier {N}: one workgroup {t_one:.1f} ms, two {ctx.last_kernel_ms():.1f} ms", flush=True)
    for i, ref in enumerate(refs):
        r1, r2 = one["results"][i], two["results"][i]
        na = int(r2["n_attempts"])
        assert r2["status"] == 0 and na == len(logs[i]) == r1["n_attempts"]
        assert r1["n_accepted"] == r2["n_accepted"] and r1["n_traces"] == r2["n_traces"]
        assert np.array_equal(one["log_accept"][i, :na], two["log_accept"][i, :na])
        # (the last step is the remainder length - z: relative rounding grows there)
        assert np.max(np.abs(two["log_h"][i, :na] / one["log_h"][i, :na] - 1)) < 1e-6
        assert rel_l2(two["A"][i], one["A"][i]) <= 1e-10
        for k in range(int(r2["n_traces"])):
            assert rel_l2(two["traces"][i, k], one["traces"][i, k]) <= 1e-10
        assert rel_l2(two["A"][i], ref) <= FIELD_TOL
        assert rel_l2(one["A"][i], ref) <= FIELD_TOL
    from helpers import soliton_c1
    dom, A1, fib1 = soliton_c1(4096)
    lin, nl = oracle_ops(dom, beta=fib1["beta"], gamma=fib1["gamma"])
    ctx = engine_context(dom, lin, nl, 0)
    fib1 = dict(fib1, length=fib1["length"] / 8)
    for method, eta in ((0, 5), (1, 3)):
        p = adaptive_params(lib, fib1, 1e-6, traces=3, eta=eta)
        ctx.set_pairing(False)
        one = ctx.propagate_adaptive(method, A1, p, max_traces=4, log_cap=4000)
        t_one = ctx.last_kernel_ms()
        ctx.set_pairing(True)
        two = ctx.propagate_adaptive(method, A1, p, max_traces=4, log_cap=4000)
        print(f"Kerr 4096 method {method}: {two['results'][0]['n_attempts']} attempts, one workgroup {t_one:.1f} ms, "
              f"two {ctx.last_kernel_ms():.1f} ms", flush=True)
        assert two["results"][0]["status"] == 0
        _same(one, two)


def _same(one, two):
    """Outputs, results, and the written parts of the attempt logs and traces."""
    assert np.array_equal(one["A"], two["A"])
    for i, (r1, r2) in enumerate(zip(one["results"], two["results"])):
        assert r1 == r2
        na, nt = int(r1["n_attempts"]), int(r1["n_traces"])
        for key in ("log_h", "log_delta", "log_accept"):
            assert np.array_equal(one[key][i, :na], two[key][i, :na]), key
        for key in ("traces", "trace_z", "step_z", "step_h"):
            assert np.array_equal(one[key][i, :nt], two[key][i, :nt]), key


def test_pairing_falls_back_when_the_batch_does_not_fit(lib):
    """A batch whose workgroups cannot all be resident (two per instance)
    runs one workgroup per instance: the same results as pairing off."""
    N = 8192
    dom, A, fib, ref, log = c2_reference(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    fib = dict(fib, length=2e-6)
    params = adaptive_params(lib, fib, 1e-6)
    batch = np.stack([A * (1 + 1e-4 * k) for k in range(300)])
    ctx.set_pairing(False)
    one = ctx.propagate_adaptive(0, batch, params)
    ctx.set_pairing(True)
    two = ctx.propagate_adaptive(0, batch, params)
    assert np.array_equal(one["A"], two["A"])
    assert all(r["status"] == 0 for r in two["results"])


def amplifier_batch_ctx(lib, N):
    from test_gpu_propagation import amplifier_batch
    return amplifier_batch(lib, N, 2e-4, (-2.0, 0.5, 2.0))
