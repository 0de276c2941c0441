"""Extract the published trace characteristics of the two-level Yb amplifier
run in examples/notebooks/nonlinear_attractor.ipynb (cells 24 and 26: the
df_results tables of the CPU Fibre and of the OpenCL fibre, 2^14 samples,
rk4ip, 30000 steps over 8 m, Hollenbeck Raman, Yb gain with Pp_0 = 10 W,
N = 4.8, Rr = 80e-6 THz, 500 traces) into yb_attractor.json.

The notebook is read as JSON text; nothing of the reference is run. Run here
(where the reference tree is present):

    python tests/golden/make_yb_attractor.py /root/reference
"""

import json
import os
import sys


COLUMNS = ["Peak Power [W]", "Energy [nJ]", "Temp width [ps]", "Spec width [THz]", "Dispersion length [km]",
           "Nonlinear length [km]", "Peaks [idx]"]  # storage.py:276-284


def parse_table(text):
    """pandas' wrapped text rendering -> {column: {z: value}} (the first and
    last five rows of each column block; the '...' row is skipped)."""
    cols = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i]
        if not line.strip() or line.startswith("["):
            i += 1
            continue
        header = sorted((line.index(c), c) for c in COLUMNS if c in line)
        header = [c for _, c in header]
        i += 2  # column names, index names
        block = []
        while i < len(lines) and lines[i].strip():
            row = lines[i].split()
            if row[0] != "...":
                z = float(row[2] if row[0].startswith("cycle") else row[0])
                vals = row[3:] if row[0].startswith("cycle") else row[1:]
                block.append((z, vals))
            i += 1
        for k, name in enumerate(header):
            cols.setdefault(name, {})
            for z, vals in block:
                cols[name][repr(z)] = vals[k]
    return cols


def main(ref):
    nb = json.load(open(os.path.join(ref, "examples", "notebooks", "nonlinear_attractor.ipynb")))
    out = {"source": "examples/notebooks/nonlinear_attractor.ipynb cells 24 (CPU Fibre) and 26 (OpenCL fibre)",
           "config": {"samples": 16384, "bit_width_ps": 200.0, "lambda0_nm": 1028, "peak_power_W": 600,
                      "fwhm_ps": 2, "length_km": 8e-3, "total_steps": 30000, "traces": 500, "beta": [0, 0, 20],
                      "gamma": 5.1, "use_all": "hollenbeck", "self_steepening": False, "method": "rk4ip",
                      "Pp_0": 10, "N": 4.8, "Rr": 80e-6}}
    for cell, key in ((24, "cpu"), (26, "opencl")):
        text = "".join(nb["cells"][cell]["outputs"][0]["data"]["text/plain"])
        out[key] = parse_table(text)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "yb_attractor.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
