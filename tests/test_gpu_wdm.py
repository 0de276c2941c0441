"""GPU parity of the two-channel WDM path (sim_type="wdm": linearity.py:
188-211, 303-320; nonlinearity.py:194-200, 287-304) against the oracle.

Case: examples/agrawal_book/fig_7_2.py (4096 samples over 200 ps, Gaussian
pulses of 1 W and 0.5 W, gamma = [1.0, 1.2] /W/km, walk-off beta_1 = 0.125
ps/km on channel 1), plus a dispersive variant (beta_2 = -5 and +3 ps^2/km)
so that both channels' linear factors matter.
* operators (exp_lin, lin, non, exp_non on 2 x N fields): <= 1e-13 (lin:
  twice its conditioning bound eps max|D| ||A~|| / ||D A~||);
* fixed-step rk4ip (the example's 400 steps over 40 km, 10 traces) and
  ss_symmetric through System/Fibre as the example writes them: every trace
  and the output <= 1e-10;
* adaptive ark4ip: the same attempt count as the oracle, output <= 1e-9.
"""

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu

BETAS = {"fig72": [[0.0, 0.0, 0.0, 0.0], [0.0, 0.125, 0.0, 0.0]],
         "dispersive": [[0.0, 0.0, -5.0], [0.0, 0.125, 3.0]]}


@pytest.fixture(scope="module")
def P():
    import pyofsschanged_amd as P
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return P


def fields(dom):
    return np.stack([O.Gaussian(width=1.0, peak_power=1.0, channel=0).generate(dom.t),
                     O.Gaussian(width=1.0, peak_power=0.5, channel=1).generate(dom.t)])


@pytest.mark.parametrize("case", sorted(BETAS))
def test_wdm_operators(P, case):
    od = O.Domain(bit_width=200.0, samples_per_bit=4096, channels=2)
    pd_ = P.Domain(bit_width=200.0, samples_per_bit=4096, channels=2)
    kw = dict(length=40.0, gamma=[1.0, 1.2], beta=BETAS[case], sim_type="wdm", total_steps=400, method="ark4ip")
    fo, fp = O.Fibre("fibre", **kw), P.Fibre("fibre", **kw)
    fo.linearity(od)
    fo.nonlinearity(od)
    fp.prepare(pd_)
    A = fields(od) * np.exp(1j * np.linspace(0, 3, 4096))
    B = fo.linear(A, 0.37)
    pairs = [("exp_lin", fp.linear(A, 0.37), B), ("lin", fp.l(A, 0.0), fo.l(A, 0.0)),
             ("non", fp.n(A, 0.0), fo.n(A, 0.0)), ("exp_non", fp.nonlinear(A, 0.37, B), fo.nonlinear(A, 0.37, B))]
    # lin = ifft(D fft(A)) amplifies the transform's roundoff by |D| (relative
    # to the result): twice that conditioning bound, as for one channel
    At = np.stack([O.fft(A[c]) for c in range(2)])
    D = np.stack([np.broadcast_to(fo.linearity.factor[c], (4096,)) for c in range(2)])
    lin_bound = np.finfo(float).eps * np.max(np.abs(D)) * np.linalg.norm(At) / np.linalg.norm(D * At)
    for name, got, ref in pairs:
        e = rel_l2(got, ref)
        bar = max(1e-13, 2 * lin_bound) if name == "lin" else 1e-13
        print(f"{case} {name}: rel L2 {e:.2e} (bar {bar:.1e})", flush=True)
        assert got.shape == (2, 4096) and e <= bar


@pytest.mark.parametrize("method,case", [("rk4ip", "fig72"), ("rk4ip", "dispersive"), ("ss_symmetric", "dispersive")])
def test_wdm_fixed_steps_via_system(P, method, case):
    od = O.Domain(bit_width=200.0, samples_per_bit=4096, channels=2)
    kw = dict(length=40.0, gamma=[1.0, 1.2], beta=BETAS[case], sim_type="wdm", total_steps=400, method=method,
              traces=10)
    fo = O.Fibre("fibre", **kw)
    out_o = fo(od, fields(od))
    system = P.System(P.Domain(bit_width=200.0, samples_per_bit=4096, channels=2))
    system.add(P.Gaussian(width=1.0, peak_power=1.0, channel=0))
    system.add(P.Gaussian(width=1.0, peak_power=0.5, channel=1))
    system.add(P.Fibre("fibre", **kw))
    system.run()
    out_p = system.fields["fibre"]
    st = system["fibre"].stepper.storage
    assert len(st.As) == len(fo.stepper.storage.As) == 10
    for k in range(10):
        assert st.As[k].shape == (2, 4096)
        assert st.z[k] == pytest.approx(fo.stepper.storage.z[k], rel=1e-15)
        assert rel_l2(st.As[k], fo.stepper.storage.As[k]) <= 1e-10
    e = rel_l2(out_p, out_o)
    print(f"{method} {case}: rel L2 {e:.2e}", flush=True)
    assert e <= 1e-10
    assert st.fft_total == fo.stepper.storage.fft_total


def test_wdm_adaptive(P):
    od = O.Domain(bit_width=200.0, samples_per_bit=4096, channels=2)
    pd_ = P.Domain(bit_width=200.0, samples_per_bit=4096, channels=2)
    kw = dict(length=5.0, gamma=[1.0, 1.2], beta=BETAS["dispersive"], sim_type="wdm", method="ark4ip",
              local_error=1e-7, traces=2)
    fo, fp = O.Fibre("fibre", **kw), P.Fibre("fibre", **kw)
    out_o = fo(od, fields(od))
    out_p = fp(pd_, fields(od))
    n_o = len(fo.stepper.attempt_log)
    e = rel_l2(out_p, out_o)
    print(f"ark4ip: rel L2 {e:.2e}, attempts {fp.stepper.attempts} (oracle {n_o}, "
          f"{sum(1 for a in fo.stepper.attempt_log if not a[2])} rejected)", flush=True)
    assert fp.stepper.attempts == n_o
    assert e <= 1e-9
