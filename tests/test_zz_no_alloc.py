"""The no-allocation entry point gnlse_propagate_adaptive_device_noalloc
against gnlse_propagate_adaptive_device, through the C ABI.

Both run the same adaptive kernel on the same device inputs; the only
difference is where the per-instance workspace lives (the library's static
device array instead of context buffers) and where the per-instance results
go (device memory instead of a host array). So the bar is bit-identity of
the output fields and of every result record, on two calls in a row, the
second with a smaller batch (the array keeps the first call's contents).
Workload: the supercontinuum physics (Raman + self-steepening, adaptive
RK4IP, local error 1e-6) on 1024 samples over the first 5 mm, each instance
with its own seeded one-photon-per-mode input noise.
"""

import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import engine_context, oracle_ops, supercontinuum_c2

pytestmark = pytest.mark.gpu

N = 1024
LENGTH = 5e-6


@pytest.fixture(scope="module")
def setup():
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    dom, A, fib = supercontinuum_c2(N)
    lin, nl = oracle_ops(dom, beta=fib["beta"], gamma=fib["gamma"], self_steepening=True, use_all=True)
    ctx = engine_context(dom, lin, nl, 1)
    p = _lib.AdaptiveParams()
    p.length = LENGTH
    p.h0 = min(LENGTH / 100, 1e-6)
    p.local_error = 1e-6
    p.safety, p.max_factor, p.min_factor, p.step_size_min = 0.9, 10.0, 0.2, 1e-37
    p.traces, p.total_attempts, p.steps_max, p.eta = 1, 1000, 100000, 5
    yield _lib, dom, A, ctx, p
    ctx.close()


def _inputs(dom, A, batch, seed0):
    return np.stack([A + O.oppm_noise(dom, seed0 + r) for r in range(batch)]).astype(np.complex128)


def _run_both(lib, ctx, p, fields):
    B = fields.shape[0]
    d_in = ctx.malloc(fields.nbytes)
    d_ref = ctx.malloc(fields.nbytes)
    d_out = ctx.malloc(fields.nbytes)
    d_res = ctx.malloc(B * ctypes.sizeof(lib.InstanceResult))
    try:
        ctx.htod(d_in, fields)
        ref_res = ctx.propagate_adaptive_device(0, B, p, d_in, d_ref)
        ref = np.empty_like(fields)
        ctx.dtoh(ref, d_ref)
        print(f"no-alloc check: batch {B}, N {N}, seeds from {fields[0, 0]!r}", flush=True)
        ctx.propagate_adaptive_device_noalloc(0, B, p, d_in, d_out, d_res)
        got = np.empty_like(fields)
        ctx.dtoh(got, d_out)
        res = (lib.InstanceResult * B)()
        ctx.dtoh(np.frombuffer(res, dtype=np.uint8), d_res)
    finally:
        for d in (d_in, d_ref, d_out, d_res):
            ctx.free(d)
    return ref, ref_res, got, res


def _check(ref, ref_res, got, res):
    assert np.array_equal(got, ref)
    for r, q in zip(ref_res, res):
        assert q.status == r["status"] == 0
        assert (q.z, q.h, q.n_attempts, q.n_accepted, q.n_traces) == (r["z"], r["h"], r["n_attempts"],
                                                                      r["n_accepted"], r["n_traces"])


def test_noalloc_matches_device_entry_on_two_calls(setup):
    lib, dom, A, ctx, p = setup
    first = _inputs(dom, A, 8, 100)
    ref, ref_res, got, res = _run_both(lib, ctx, p, first)
    _check(ref, ref_res, got, res)
    assert not np.array_equal(ref[0], ref[1])  # the instances really differ
    second = _inputs(dom, A, 3, 900)  # smaller, after the first call filled the array
    ref, ref_res, got, res = _run_both(lib, ctx, p, second)
    _check(ref, ref_res, got, res)


def test_noalloc_rejects_batch_beyond_the_array(setup):
    lib, dom, A, ctx, p = setup
    per = 11 * N * 16  # NSLOT field slots of complex128 per instance
    too_many = (64 * 1024 * 1024) // per + 1
    with pytest.raises(lib.GNLSEError, match="no-allocation entry point holds"):
        # rejected on the host before any launch: the pointers are never read
        ctx.propagate_adaptive_device_noalloc(0, too_many, p, ctypes.c_void_p(16), ctypes.c_void_p(16 + (1 << 44)),
                                              ctypes.c_void_p(16))


@pytest.mark.parametrize("offset", [0, N * 16 - 16, -(N * 16 - 16)], ids=["same", "tail", "head"])
def test_overlapping_input_and_output_rejected(setup, offset):
    # the kernels take input and output as restrict pointers; both device
    # entry points refuse ranges that share a byte, before any launch
    lib, dom, A, ctx, p = setup
    base = ctx.malloc(4 * N * 16)
    try:
        d_in = ctypes.c_void_p(base.value + N * 16)
        d_out = ctypes.c_void_p(base.value + N * 16 + offset)
        with pytest.raises(lib.GNLSEError, match="overlap"):
            ctx.propagate_adaptive_device(0, 1, p, d_in, d_out)
        with pytest.raises(lib.GNLSEError, match="overlap"):
            ctx.propagate_adaptive_device_noalloc(0, 1, p, d_in, d_out, ctypes.c_void_p(base.value + 3 * N * 16))
    finally:
        ctx.free(base)
