"""The workgroup reductions of the kernels (csrc/ops.hpp wg_sum, wg_sum_n)
through gnlse_check_reductions: m sums at once (the routine of the gain
stages and of the adaptive error norm) against m single sums, bit for bit in
every thread, at every workgroup size the propagation kernels use (64-1024
threads) and every m they reduce at once (2 and 4: Yb gain; 9 and 18:
saturable gain from spectra; 3: error norm). Inputs span 16 decades with
random signs, so any change of summation order would show in the low bits.
The sums also match numpy's to 1e-12 of the sum of magnitudes."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return _lib


@pytest.mark.parametrize("m", [2, 3, 4, 9, 18])
@pytest.mark.parametrize("threads", [64, 128, 256, 512, 1024])
def test_sum_n_equals_single_sums(lib, threads, m):
    import ctypes
    blocks = 16
    rng = np.random.default_rng(threads * 100 + m)
    x = rng.choice([-1.0, 1.0], size=blocks * m * threads) * 10.0 ** rng.uniform(-8, 8, blocks * m * threads)
    on = np.empty_like(x)
    o1 = np.empty_like(x)
    L = lib.load_library()
    rc = L.gnlse_check_reductions(0, threads, m, blocks, x.ctypes.data_as(ctypes.c_void_p),
                                  on.ctypes.data_as(ctypes.c_void_p), o1.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, L.gnlse_last_error()
    assert np.array_equal(on.view(np.uint64), o1.view(np.uint64))
    v = on.reshape(blocks, m, threads)
    assert np.all(v == v[:, :, :1])  # every thread holds the same sum
    xs = x.reshape(blocks, m, threads)
    ref = xs.sum(axis=2)
    assert np.all(np.abs(v[:, :, 0] - ref) <= 1e-12 * np.abs(xs).sum(axis=2))
