"""Pins of the two-level Yb gain restatement (oracle Amplifier2LevelModel,
pyofss/modules/amplifier.py:178-402), CPU only.

Golden vectors: the df_results tables that examples/notebooks/
nonlinear_attractor.ipynb prints for a 2^14-sample rk4ip run with Hollenbeck
Raman and the Yb model (tests/golden/yb_attractor.json, extracted from the
notebook by tests/golden/make_yb_attractor.py). The first five traces (steps
1, 61, 121, 181, 241) are reproduced here by the oracle with the same step
size; the OpenCL column carries full digits, the CPU column six decimals and
the characteristics (duration, spectral width, L_D, L_NL, peaks) of
Storage.update_characts. These also pin the Hollenbeck-Cantrell response
(nonlinearity.py:72-100) and its simps normalisation.

Closed forms: without signal the inversion is the pump-only value
N2 = N (r_p P_p/P_sat,p) / (1 + P_p/P_sat,p), the gain exponent is
(alpha_s N2 - eta_s) h 1e3 / 2 and the pump evolves by exp(g_p h 1e3).
"""

import json
import os

import numpy as np
import pytest

import oracle as O
from oracle import characts_np as Q
from oracle.pyofss_np import yb_cross_section_table

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "yb_attractor.json")))


def attractor_run(steps, traces):
    """The notebook's system over the first `steps` of its 30000 steps."""
    cfg = GOLD["config"]
    dom = O.Domain(samples_per_bit=cfg["samples"], bit_width=cfg["bit_width_ps"], total_bits=1,
                   centre_nu=O.lambda_to_nu(cfg["lambda0_nm"]))
    A0 = O.Gaussian(peak_power=cfg["peak_power_W"], width=cfg["fwhm_ps"], C=0, using_fwhm=True).generate(dom.t)
    h0 = cfg["length_km"] / cfg["total_steps"]
    f = O.Fibre(length=steps * h0, method="rk4ip", total_steps=steps, beta=np.array(cfg["beta"]),
                gamma=cfg["gamma"], self_steepening=False, use_all="hollenbeck", traces=traces, Pp_0=cfg["Pp_0"],
                N=cfg["N"], Rr=cfg["Rr"], use_Yb_model=True)
    f(dom, A0.copy())
    return dom, f


@pytest.fixture(scope="module")
def first_traces():
    return attractor_run(241, 4)  # int(241 / 4) = 60: traces after steps 1, 61, ..., 241


def golden_rows(col, which):
    rows = sorted(GOLD[which][col].items(), key=lambda kv: float(kv[0]))
    return [(float(z), v) for z, v in rows]


def test_first_traces_match_published_values(first_traces):
    dom, f = first_traces
    st = f.stepper.storage
    zs = np.array(st.z) * 1e6
    gold_p = golden_rows("Peak Power [W]", "opencl")[:5]
    gold_e = golden_rows("Energy [nJ]", "opencl")[:5]
    assert np.allclose(zs, [z for z, _ in gold_p], atol=1e-6)
    for k, A in enumerate(st.As):
        P = O.temporal_power(A)
        assert P.max() == pytest.approx(float(gold_p[k][1]), rel=1e-11)
        assert O.energy(A, dom.t) == pytest.approx(float(gold_e[k][1]), rel=1e-11)
        # CPU column: six printed decimals
        assert abs(P.max() - float(golden_rows("Peak Power [W]", "cpu")[k][1])) <= 6e-7
        assert abs(O.energy(A, dom.t) - float(golden_rows("Energy [nJ]", "cpu")[k][1])) <= 6e-7


def test_first_traces_characteristics_match_published_values(first_traces):
    """Storage.update_characts columns (storage.py:194-209) of the CPU run."""
    dom, f = first_traces
    st = f.stepper.storage
    for k, A in enumerate(st.As):
        P = O.temporal_power(A)
        S = O.spectral_power(A)
        dur = Q.reference_duration(P) * dom.dt
        sw = Q.reference_spec_width(S) * dom.dnu
        assert abs(dur - float(golden_rows("Temp width [ps]", "cpu")[k][1])) <= 6e-7
        assert abs(sw - float(golden_rows("Spec width [THz]", "cpu")[k][1])) <= 6e-7
        assert abs(dur ** 2 / 20.0 - float(golden_rows("Dispersion length [km]", "cpu")[k][1])) <= 6e-7
        assert abs(1 / (5.1 * P.max()) - float(golden_rows("Nonlinear length [km]", "cpu")[k][1])) <= 6e-7
        assert [int(i) for i in Q.reference_peaks(P)] == json.loads(golden_rows("Peaks [idx]", "cpu")[k][1])
        assert list(Q.restated_peaks(P)) == list(Q.reference_peaks(P))


def test_fft_accounting_counts_gain_spectra(first_traces):
    """24 stepping FFTs (Raman) + 4 gain-spectrum FFTs per rk4ip step."""
    _, f = first_traces
    assert f.stepper.storage.fft_total - 241 * 28 in range(0, 8)


def test_pump_only_inversion_closed_form():
    dom = O.Domain(samples_per_bit=4096, bit_width=100.0, total_bits=1, centre_nu=O.lambda_to_nu(1030))
    amp = O.Amplifier2LevelModel(Pp=3.0, N=4.8, Rr=80e-6)
    amp.set_domain(dom)
    tp = 3.0 / amp.Psat_p
    N2 = amp.N * (amp.ratio_p * tp) / (1.0 + tp)
    h = 1e-6
    fac = amp.factor(np.zeros(dom.total_samples, complex), h)
    g_s = amp.alpha_s * N2 - amp.eta_s
    scale = np.max(amp.alpha_s * N2 + amp.eta_s) * h * 1e3 / 2  # g_s is a difference: bound by its terms
    assert np.allclose(O.ifftshift(fac), g_s * h * 1e3 / 2, rtol=0, atol=1e-14 * scale)
    g_p = amp.alpha_p * N2 - amp.eta_p
    assert amp.Pp == pytest.approx(3.0 * np.exp(g_p * h * 1e3), rel=1e-13)
    # with a weak pump the fibre absorbs at 976 nm and in-band: g_p < 0
    assert g_p < 0 and amp.Pp < 3.0


def test_cross_sections_interpolate_the_table():
    """The spline passes through the table's points (the first row is the
    pandas header, so it is not among them) and is zero outside the range."""
    tab = yb_cross_section_table()
    assert tab.shape == (290, 3) and tab[0, 0] == pytest.approx(851e-9)
    dom = O.Domain(samples_per_bit=4096, bit_width=100.0, total_bits=1, centre_nu=O.lambda_to_nu(1030))
    amp = O.Amplifier2LevelModel(Pp=3.0, N=4.8, Rr=80e-6)
    amp.domain = dom
    amp.interpolate_sigma_s()
    from scipy import interpolate
    spl = interpolate.InterpolatedUnivariateSpline(tab[:, 0] * 1e9, tab[:, 2] * 1e27, ext=1)
    assert np.allclose(spl(tab[:, 0] * 1e9) * 1e-27, tab[:, 2], rtol=1e-12, atol=1e-40)
    outside = (dom.Lambda < 851.0) | (dom.Lambda > 1140.0)
    assert np.all(amp.sigma12_s[outside] == 0) and np.all(amp.sigma21_s[outside] == 0)
