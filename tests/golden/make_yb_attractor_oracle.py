"""Full-length oracle run of the published Yb attractor system
(examples/notebooks/nonlinear_attractor.ipynb cell 16: 2^14 samples, 30000
rk4ip steps over 8 m, Hollenbeck Raman, Yb gain, 500 traces) -> peak power
and energy of the first and last five traces and the final pump power, in
yb_attractor_oracle.json. About 5 minutes on one core; the GPU test compares
the device run with these values (the full run is too long for the CPU
suite). Run from the repository root:

    python tests/golden/make_yb_attractor_oracle.py
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def main():
    od = O.Domain(samples_per_bit=2 ** 14, bit_width=200.0, total_bits=1, centre_nu=O.lambda_to_nu(1028))
    A = O.Gaussian(peak_power=600, width=2, C=0, using_fwhm=True).generate(od.t)
    t0 = time.time()
    f = O.Fibre(length=8e-3, method="rk4ip", total_steps=30000, beta=np.array([0, 0, 20]), gamma=5.1,
                self_steepening=False, use_all="hollenbeck", traces=500, Pp_0=10, N=4.8, Rr=80e-6,
                use_Yb_model=True)
    f(od, A.copy())
    st = f.stepper.storage
    rows = []
    for k in list(range(5)) + list(range(len(st.As) - 5, len(st.As))):
        a = st.As[k]
        rows.append({"z_mm": st.z[k] * 1e6, "peak_W": float(np.max(O.temporal_power(a))),
                     "energy_nJ": float(O.energy(a, od.t))})
    out = {"source": "oracle run of examples/notebooks/nonlinear_attractor.ipynb cell 16 (first system)",
           "traces": len(st.As), "rows": rows, "final_pump_W": float(f.amplifier.Pp),
           "fft_total": int(st.fft_total), "seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "yb_attractor_oracle.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
