"""Full-length oracle runs of the supercontinuum workload (8192 samples over
12.5 ps at 835 nm, 50 fs FWHM sech at 10 kW, 15 cm of fibre with gamma 110
/W/km and the Dudley beta2..beta10, self-steepening + Blow-Wood Raman,
adaptive RK4IP at local error 1e-6) for two inputs: the noiseless pulse and
the pulse plus the one-photon-per-mode noise of seed 0 (the benchmark's
first realisation). For each: the output field, every attempt's (h, delta,
accepted), the photon number sum |A~_k|^2 / omega_k at both ends, and the
oracle's own sensitivity: the relative L2 distance between the output and a
second oracle run whose input differs by one rounding unit per sample
(relative 2^-53, seeded sign). A device run that differs from the oracle by
rounding can be expected to land within a few times that distance, not
within 1e-9, when the propagation amplifies rounding.

About 40 s per run on one core (four runs). Run from the repository root:

    python tests/golden/make_c2_full_oracle.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

C_NM_PS = 299792458.0 * 1e-3
BETA = [0.0, 0.0, -11.830, 8.1038e-2, -9.5205e-5, 2.0737e-7, -5.3943e-10, 1.3486e-12, -2.5495e-15, 3.0524e-18,
        -1.7140e-21]
LENGTH = 1.5e-4


def photon_number(A, dom):
    """sum_k |fft(A)_k|^2 / omega_k with absolute angular frequencies (FFT order)."""
    omega = O.fftshift(dom.omega)
    return float(np.sum(np.abs(O.fft(A)) ** 2 / omega))


def run(dom, A):
    f = O.Fibre(length=LENGTH, beta=BETA, gamma=110.0, self_steepening=True, use_all=True, method="ark4ip",
                local_error=1e-6)
    out = f(dom, A.copy())
    log = f.stepper.attempt_log
    return out, np.array([a[0] for a in log]), np.array([a[1] for a in log]), np.array([a[2] for a in log])


def main():
    dom = O.Domain(1, 8192, 12.5, C_NM_PS / 835.0)
    A0 = O.Sech(peak_power=1e4, width=0.050, using_fwhm=True).generate(dom.t)
    res = {}
    for tag, A in (("noiseless", A0), ("seed0", A0 + O.oppm_noise(dom, 0))):
        t0 = time.time()
        out, h, d, acc = run(dom, A)
        rng = np.random.default_rng(1)
        ulp = 2.0 ** -53
        Ap = A * (1.0 + ulp * rng.choice([-1.0, 1.0], size=A.size))
        out_p, h_p, _, acc_p = run(dom, Ap)
        sens = float(np.linalg.norm(out_p - out) / np.linalg.norm(out))
        res.update({f"{tag}/A_in": A, f"{tag}/A_out": out, f"{tag}/h": h, f"{tag}/delta": d,
                    f"{tag}/accepted": acc, f"{tag}/photons_in": photon_number(A, dom),
                    f"{tag}/photons_out": photon_number(out, dom), f"{tag}/sensitivity": sens,
                    f"{tag}/sensitivity_same_steps": bool(len(h_p) == len(h) and np.array_equal(acc_p, acc))})
        print(f"{tag}: {len(h)} attempts, photon number change "
              f"{res[f'{tag}/photons_out'] / res[f'{tag}/photons_in'] - 1:.2e}, oracle sensitivity {sens:.2e} "
              f"(same attempt sequence: {res[f'{tag}/sensitivity_same_steps']}), {time.time() - t0:.0f} s",
              flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_full_oracle.npz"), **res)


if __name__ == "__main__":
    main()
