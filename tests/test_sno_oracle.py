"""Pins of the surrogate's FP64 oracle (oracle/sno_np.py), CPU only.

* The truncated DFT matrices equal numpy's rfft (first K modes) and irfft
  (modes >= K zeroed) to rounding.
* The hand-written backward pass equals central finite differences of the
  loss for every parameter group.
* The layout and a random initialisation have the expected sizes and scales.
"""

import numpy as np

from oracle.sno_np import SNO, dft_mats, init_params, layout, unpack


def test_dft_matrices_match_numpy_fft():
    N, K = 64, 8
    Ft, G = dft_mats(N, K)
    rng = np.random.default_rng(0)
    h = rng.standard_normal((3, N))
    V = h @ Ft.T
    R = np.fft.rfft(h, axis=1)[:, :K]
    assert np.allclose(V[:, 0::2], R.real, atol=1e-12)
    assert np.allclose(V[:, 1::2], R.imag, atol=1e-12)
    W = rng.standard_normal((3, K)) + 1j * rng.standard_normal((3, K))
    W[:, 0] = W[:, 0].real  # irfft ignores the imaginary part of the DC mode
    Ws = np.empty((3, 2 * K))
    Ws[:, 0::2], Ws[:, 1::2] = W.real, W.imag
    full = np.zeros((3, N // 2 + 1), complex)
    full[:, :K] = W
    assert np.allclose(Ws @ G.T, np.fft.irfft(full, n=N, axis=1), atol=1e-12)


def test_backward_matches_finite_differences():
    net = SNO(N=64, C=4, K=8, L=2, CI=3, CH=8)
    rng = np.random.default_rng(1)
    w = init_params(net, seed=2) + 0.05 * rng.standard_normal(net.size)
    x = rng.standard_normal((2, 64, 3))
    t = rng.standard_normal((2, 64, 2))
    loss, rel, g = net.loss_grad(w, x, t)
    assert np.isclose(loss, np.mean(rel ** 2))
    eps = 1e-6
    for name, v in net.lay.items():
        if name == "_total":
            continue
        o, s = v
        n = int(np.prod(s))
        idx = o + rng.choice(n, size=min(n, 6), replace=False)
        for i in idx:
            wp, wm = w.copy(), w.copy()
            wp[i] += eps
            wm[i] -= eps
            fd = (net.loss_grad(wp, x, t)[0] - net.loss_grad(wm, x, t)[0]) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[i])


def test_layout_and_init():
    lay = layout(4096, 64, 128, 4, 3, 128)
    # 4 layers x (2 x 128 x 64 x 64 + 64 x 64 + 64) + lift + projection
    assert lay["_total"] == 4 * (2 * 128 * 64 * 64 + 64 * 64 + 64) + 3 * 64 + 64 + 64 * 128 + 128 + 128 * 2 + 2
    net = SNO(N=256, C=64, K=32, L=2, CI=3, CH=128)
    p = unpack(init_params(net, seed=0), net.lay)
    assert np.all(p["bl"] == 0) and np.all(p["bq2"] == 0)
    assert 0 <= p["Rr0"].min() and p["Rr0"].max() <= 1.0 / 64 ** 2
    assert np.abs(p["Wp1"]).max() <= 1.0 / 8
