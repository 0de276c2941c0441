"""GPU tests of the OpenCL-fibre front end (OpenclProgramm / OpenclFibre of
pyofss/modules/opencl_fibre.py:287-1003 on the HIP engine) and of the
device down-sampling of traces (gnlse_trace_powers).

* The second system of examples/notebooks/nonlinear_attractor.ipynb cell 16,
  written as the notebook writes it (OpenclProgramm(fast_math=True,
  use_all='hollenbeck', dorf='double', ctx=..., downsampling=4000),
  Amplifier2LevelModel(..., prg=, queue=, ctx=), OpenclFibre(cl_programm=...,
  method='cl_rk4ip', traces=500, amplifier=...)): its df_results against the
  notebook's OpenCL table (tests/golden/yb_attractor.json). Peak power and
  energy (cl_simpson rule) of the first five traces within 1e-11, of the last
  five within 5e-9 (the CPU and OpenCL runs of the reference differ by
  6.7e-10 in peak power there); L_D and L_NL to the printed six decimals.
* Down-sampling: GNLSE_DOWNSAMPLE_PYOFSS against field.get_downsampled of
  temporal_power / spectral_power (scipy interp1d), and
  GNLSE_DOWNSAMPLE_OPENCL against a NumPy restatement of cl_interpolate
  (float32 coordinate t = i/n_out, mix(y[j], y[j+1], t(N-1) - j)); both to
  one rounding (2.3e-16 of the trace's maximum power) for the same powers:
  |A|^2 formed as re^2 + im^2 (within 1e-15 of numpy's abs(A)**2) and the
  spectra of the device's own FFT (within 4e-15 of numpy's).
"""

import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "yb_attractor.json")))


@pytest.fixture(scope="module")
def P():
    import pyofsschanged_amd as P
    from pyofsschanged_amd import _lib
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible")
    return P


def cl_interpolate(P, n_out):
    """opencl_fibre.py:207-216 restated with NumPy (float32 t as written)."""
    n = len(P)
    t = np.arange(n_out, dtype=np.float32) / np.float32(n_out)
    tt = t.astype(np.float64) * float(n - 1)
    j = tt.astype(np.int64)
    u = tt - j
    return (P[j + 1] - P[j]) * u + P[j]


def test_cl_interpolate_restatement_invariants():
    """Pins of the restatement: a linear ramp is reproduced at the float32
    coordinates, and the output starts at the first sample."""
    ramp = np.arange(16384, dtype=np.float64) * 0.5
    t = np.arange(4000, dtype=np.float32) / np.float32(4000)
    assert np.allclose(cl_interpolate(ramp, 4000), 0.5 * t.astype(np.float64) * 16383, rtol=1e-15, atol=0)
    assert cl_interpolate(ramp, 4000)[0] == 0.0


@pytest.mark.parametrize("n_out", [4000, 500, 16384])
def test_trace_powers(P, n_out):
    from pyofsschanged_amd import _lib
    dom = O.Domain(1, 16384, 40.0, 291.0)
    rng = np.random.default_rng(n_out)
    A = np.stack([O.Sech(peak_power=p, width=w).generate(dom.t) * np.exp(1j * rng.uniform(-2, 2, dom.total_samples))
                  for p, w in ((100.0, 0.3), (5.0, 2.0), (1e3, 0.05))])
    ctx = _lib.Context(0)
    ctx.set_grid(dom.total_samples)
    S = ctx.fft(A, +1)  # the device's pyofss spectra: the down-sampling is checked on its own
    for mode in (_lib.DOWNSAMPLE_PYOFSS, _lib.DOWNSAMPLE_OPENCL):
        t, s = ctx.trace_powers(A, n_out, mode)
        for k in range(len(A)):
            # the device's power formula (re^2 + im^2) on the device's spectra
            Pt = A[k].real ** 2 + A[k].imag ** 2
            Ps = O.ifftshift(S[k].real ** 2 + S[k].imag ** 2)
            assert np.max(np.abs(Pt - O.temporal_power(A[k]))) <= 1e-15 * Pt.max()
            assert np.max(np.abs(Ps - O.spectral_power(A[k]))) <= 4e-15 * Ps.max()
            if mode == _lib.DOWNSAMPLE_PYOFSS:
                et, es = P.get_downsampled(Pt, n_out), P.get_downsampled(Ps, n_out)
            else:
                et, es = cl_interpolate(Pt, n_out), cl_interpolate(Ps, n_out)
            dt = np.max(np.abs(t[k] - et)) / Pt.max()
            ds = np.max(np.abs(s[k] - es)) / Ps.max()
            print(f"mode {mode} n_out {n_out} field {k}: temporal {dt:.1e}, spectral {ds:.1e}", flush=True)
            assert dt <= 2.3e-16 and ds <= 2.3e-16


def gold(col, which):
    rows = sorted(GOLD[which][col].items(), key=lambda kv: float(kv[0]))
    return [(float(z), v) for z, v in rows]


def test_published_attractor_run_opencl_fibre(P):
    """examples/notebooks/nonlinear_attractor.ipynb cell 16, second system."""
    Pp_0, N, Rr = 10, 4.8, 80 * 1e-6
    Tr = 1 / Rr
    domain = P.Domain(samples_per_bit=2 ** 14, bit_width=200.0, total_bits=1, centre_nu=P.lambda_to_nu(1028))
    gaussian = P.Gaussian(peak_power=600, width=2, C=0, using_fwhm=True)
    cl_program = P.OpenclProgramm(fast_math=True, use_all="hollenbeck", dorf="double", ctx=None, downsampling=4000)
    cl_program.set_domain(domain)
    sys1 = P.System(domain)
    sys1.add(gaussian)
    amplifier = P.Amplifier2LevelModel(Pp=Pp_0, N=N, Rr=1 / Tr, prg=cl_program.prg, queue=cl_program.queue,
                                       ctx=cl_program.ctx)
    sys1.add(P.OpenclFibre(cl_programm=cl_program, length=8 * 1e-3, total_steps=30000, beta=np.array([0, 0, 20]),
                           gamma=5.1, self_steepening=False, method="cl_rk4ip", cycle="cycle0", traces=500,
                           amplifier=amplifier))
    sys1.run()
    df = sys1.df_results
    assert len(df) == 500
    assert list(df.index.get_level_values("fibre"))[0] == "cl_fibre"
    z = df.index.get_level_values("z [mm]").values
    for j, r in enumerate(list(range(5)) + list(range(495, 500))):
        assert z[r] == pytest.approx(gold("Peak Power [W]", "opencl")[j][0], abs=1e-6)
        rec = df.iloc[r]
        bar = 1e-11 if j < 5 else 5e-9
        p_cl, e_cl = float(gold("Peak Power [W]", "opencl")[j][1]), float(gold("Energy [nJ]", "opencl")[j][1])
        print(f"z={z[r]:.6f} mm: peak {rec['Peak Power [W]']:.13g} (OpenCL {p_cl:.13g}), energy "
              f"{rec['Energy [nJ]']:.13g} (OpenCL {e_cl:.13g})", flush=True)
        assert rec["Peak Power [W]"] == pytest.approx(p_cl, rel=bar)
        assert rec["Energy [nJ]"] == pytest.approx(e_cl, rel=bar)
        for col in ("Dispersion length [km]", "Nonlinear length [km]"):
            ref = float(gold(col, "opencl")[j][1])
            assert abs(rec[col] - ref) <= 6e-7 + 5e-9 * abs(ref), col
        assert np.isnan(rec["Temp width [ps]"]) and np.isnan(rec["Spec width [THz]"])
    assert sys1.df_temp.shape == (500, 4000) and sys1.df_spec.shape == (500, 4000)
    assert amplifier.Pp < Pp_0
