"""GPU tests of the surrogate's bf16 matrix-core kernels against FP64
references.

* GEMM: every operand layout, batching, split-K and ragged edges against
  numpy on the same bf16-rounded operands (FP32 accumulation: <= 1e-5
  relative).
* Network forward and every parameter group's gradient at the production
  shape (4096 samples, width 64, 128 modes, 4 layers) against the FP64
  oracle (oracle/sno_np.py, itself pinned by finite differences). The GPU
  rounds every GEMM operand to bf16 (8-bit mantissa, relative 2^-9 per
  element), so the bars are 2e-2 for the network output and 5e-2 for each
  gradient group (relative L2).
* Training on solver-generated data lowers the loss, and the trained network
  is more accurate than the method without it (linear propagation only,
  which is what the untrained residual network computes).
* The certified judge: the two-network combine kernel against numpy (mean
  to FP32 rounding, disagreement 1e-6 relative), and on the held-out fields
  it was calibrated on, no delivered field exceeds the tolerance (every
  field above it is re-solved by construction of the gate).
"""

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from pyofsschanged_amd import _lib, surrogate
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    ctx = _lib.Context(0)
    return _lib, surrogate, ctx


def to_bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(h):
    return (h.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("at,bt", [(0, 0), (1, 0), (0, 1)])
@pytest.mark.parametrize("M,N,K,batch,split", [(128, 128, 64, 1, 1), (200, 72, 96, 2, 1), (64, 256, 4096, 3, 4),
                                               (296, 136, 256, 1, 2)])
def test_gemm_bf16(env, at, bt, M, N, K, batch, split):
    lib, S, ctx = env
    rng = np.random.default_rng(M + N + K)
    A = to_bf16(rng.standard_normal((batch, K, M) if at else (batch, M, K)))
    B = to_bf16(rng.standard_normal((batch, N, K) if bt else (batch, K, N)))
    dA = S.DeviceArray(ctx, A.shape, np.uint16)
    dB = S.DeviceArray(ctx, B.shape, np.uint16)
    dC = S.DeviceArray(ctx, (batch * split, M, N), np.float32)
    dA.upload(A)
    dB.upload(B)
    lda, ldb = A.shape[2], B.shape[2]
    lib._check(lib.load_library(), lib.load_library().gnlse_gemm_bf16(
        0, at, bt, M, N, K, dA.ptr, lda, dB.ptr, ldb, dC.ptr, N, batch, A[0].size, B[0].size, M * N, split), "gemm")
    C = dC.download().reshape(batch, split, M, N).sum(axis=1)
    Af = from_bf16(A).astype(np.float64)
    Bf = from_bf16(B).astype(np.float64)
    opA = np.transpose(Af, (0, 2, 1)) if at else Af
    opB = np.transpose(Bf, (0, 2, 1)) if bt else Bf
    ref = opA @ opB
    err = rel_l2(C, ref)
    print(f"gemm at={at} bt={bt} {M}x{N}x{K} batch {batch} split {split}: rel {err:.1e}", flush=True)
    assert err < 1e-5


@pytest.fixture(scope="module")
def net_and_oracle(env):
    _, S, ctx = env
    from oracle.sno_np import SNO, init_params
    net = S.SpectralOperator(n=4096, max_batch=32)
    ora = SNO(N=4096)
    assert net.size == ora.size
    rng = np.random.default_rng(3)
    w = init_params(ora, seed=4)
    # larger spectral weights than the initialisation, so the Fourier path
    # carries a share of the output comparable to the pointwise path
    for l in range(4):
        for part in ("Rr", "Ri"):
            o, s = ora.lay[f"{part}{l}"]
            w[o:o + int(np.prod(s))] = rng.uniform(-1, 1, int(np.prod(s))) / 64.0
    net.set_params(w.astype(np.float32))
    w32 = net.get_params().astype(np.float64)
    return net, ora, w32


def smooth_inputs(rng, B, N):
    n = np.arange(N)
    x = np.zeros((B, N, 3))
    for b in range(B):
        w = rng.uniform(20, 200)
        c = rng.uniform(-200, 200)
        env_ = np.exp(-0.5 * ((n - N / 2 - c) / w) ** 2)
        ph = rng.uniform(-0.01, 0.01) * (n - N / 2) ** 2 / N
        x[b, :, 0] = env_ * np.cos(ph)
        x[b, :, 1] = env_ * np.sin(ph)
        x[b, :, 2] = (n - N / 2) / (N / 2)
    return x


def test_forward_matches_oracle(env, net_and_oracle):
    _, S, ctx = env
    net, ora, w = net_and_oracle
    rng = np.random.default_rng(5)
    x = smooth_inputs(rng, 4, 4096)
    dx = S.DeviceArray(ctx, x.shape, np.float32)
    dy = S.DeviceArray(ctx, (4, 4096, 2), np.float32)
    dx.upload(x)
    net.forward(4, dx.ptr, dy.ptr)
    y = dy.download().astype(np.float64)
    ref = ora.forward(w, dx.download().astype(np.float64))
    err = rel_l2(y - x[..., :2], ref - x[..., :2])
    print(f"forward: network output rel L2 {err:.2e} ({net.last_ms():.2f} ms)", flush=True)
    assert err < 2e-2


def test_gradients_match_oracle(env, net_and_oracle):
    _, S, ctx = env
    net, ora, w = net_and_oracle
    rng = np.random.default_rng(6)
    B = 32
    x = smooth_inputs(rng, B, 4096)
    t = x[..., :2] * (1.0 + 0.1 * rng.standard_normal((B, 1, 1))) + 0.02 * rng.standard_normal((B, 4096, 2))
    dx = S.DeviceArray(ctx, x.shape, np.float32)
    dt = S.DeviceArray(ctx, t.shape, np.float32)
    dx.upload(x)
    dt.upload(t)
    loss, rel, g = net.loss_grad(B, dx.ptr, dt.ptr, want_grad=True)
    x32 = dx.download().astype(np.float64)
    t32 = dt.download().astype(np.float64)
    rloss, rrel, rg = ora.loss_grad(w, x32, t32)
    print(f"loss {loss:.6e} oracle {rloss:.6e} ({net.last_ms():.2f} ms per step)", flush=True)
    assert abs(loss - rloss) <= 2e-2 * rloss
    assert np.allclose(rel, rrel, rtol=2e-2)
    worst = 0.0
    for name, v in ora.lay.items():
        if name == "_total":
            continue
        o, s = v
        n = int(np.prod(s))
        e = rel_l2(g[o:o + n], rg[o:o + n])
        worst = max(worst, e)
        print(f"  grad {name:5s}: rel L2 {e:.2e}", flush=True)
        assert e < 5e-2, name
    print(f"worst gradient group {worst:.2e}", flush=True)


def test_training_on_solver_data(env):
    _, S, ctx = env
    prob = S.SegmentProblem()
    net = S.SpectralOperator(n=prob.n, max_batch=64)
    net.init_params(seed=0)
    train = S.Dataset(prob, net, 1024, seed=11)
    test = S.Dataset(prob, net, 64, seed=12, scale=train.scale, keep_fields=True)
    # the untrained residual network is the identity: linear propagation only
    l0, rel0, _ = net.loss_grad(64, test.x.ptr, test.t.ptr)
    hist = S.train(net, train, steps=400, batch=32, lr=2e-3, log_every=100)
    l1, rel1, _ = net.loss_grad(64, test.x.ptr, test.t.ptr)
    print(f"held-out relative L2: linear only median {np.median(rel0):.3e}, trained median {np.median(rel1):.3e}; "
          f"train loss {hist[0]:.3e} -> {hist[-10:].mean():.3e}", flush=True)
    assert hist[-10:].mean() < 0.2 * hist[:10].mean()
    assert np.median(rel1) < 0.5 * np.median(rel0)


def test_certified_judge(env):
    _, S, ctx = env
    prob = S.SegmentProblem()
    nets = []
    for k in range(2):
        net = S.SpectralOperator(n=prob.n, max_batch=256)
        net.init_params(seed=3 + k)
        nets.append(net)
    train = S.Dataset(prob, nets[0], 1024, seed=11)
    for k, net in enumerate(nets):
        S.train(net, train, steps=300, batch=32, lr=2e-3, seed=k)
    # combine kernel against numpy on the two networks' raw outputs
    held = S.Dataset(prob, nets[0], 256, seed=12, scale=train.scale, keep_fields=True)
    y1 = S.DeviceArray(ctx, (256, prob.n, 2), np.float32)
    y2 = S.DeviceArray(ctx, (256, prob.n, 2), np.float32)
    ym = S.DeviceArray(ctx, (256, prob.n, 2), np.float32)
    nets[0].forward(256, held.x.ptr, y1.ptr)
    nets[1].forward(256, held.x.ptr, y2.ptr)
    dist = nets[0].combine(256, y1.ptr, y2.ptr, ym.ptr)
    a, b, m = y1.download().astype(np.float64), y2.download().astype(np.float64), ym.download().astype(np.float64)
    assert np.allclose(m, 0.5 * (a + b), rtol=1e-6, atol=1e-7 * np.abs(m).max())
    ref_d = np.linalg.norm((a - b).reshape(256, -1), axis=1) / np.linalg.norm(m.reshape(256, -1), axis=1)
    assert np.allclose(dist, ref_d, rtol=1e-6)
    # calibration, then judging the same held-out fields: nothing above tol passes
    sur = S.FibreSurrogate(prob, nets[0], train.scale, partner=nets[1], chunk=256)
    d_in = S.DeviceArray(ctx, held.fields_in.shape, np.complex128)
    d_out = S.DeviceArray(ctx, held.fields_in.shape, np.complex128)
    d_in.upload(held.fields_in)
    cal = sur.calibrate(d_in.ptr, d_out.ptr, 256, held.fields_out)
    rep = sur.propagate(d_in.ptr, d_out.ptr, 256, reference=held.fields_out)
    print(f"certified judge: calibration {cal}, within tol {rep['within_tol']:.3f}, resolved {rep['resolved']}",
          flush=True)
    assert rep["within_tol"] == 1.0
    assert rep["err_max"] <= sur.tol
