"""Write the golden fixtures for the seeded (H2O)8 workload and a few small
molecules: bohr coordinates (the exact input bits) and the oracle's energy
(with its D4 two- and three-body parts, and the energy with dispersion
switched off), charges and iteration counts.  The fixture is data; the tests compare both
the oracle (regression) and the HIP engine against it.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tblite-cuda_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
import tblite_amd  # noqa: E402
import workloads  # noqa: E402


def main():
    orc = oracle.Oracle()
    nodisp = tblite_amd.default_param()
    nodisp.d4.enabled = 0
    orc0 = oracle.Oracle(nodisp)
    rng = np.random.default_rng(2024)
    cases = {
        "water8_seed0": workloads.config0()[0],
        "water1": workloads.water_cluster(1, 0.5, rng),
        "water4": workloads.water_cluster(4, workloads.liquid_radius(4), rng),
        "organic_c5": workloads.organic_molecule(5, np.random.default_rng(11)),
        "organic_c12": workloads.organic_molecule(12, np.random.default_rng(12)),
    }
    out = {}
    for name, (num, xyz) in cases.items():
        r = orc.singlepoint(num, xyz)
        out[name] = {
            "num": [int(z) for z in num],
            "xyz_bohr": [[float(v) for v in row] for row in xyz],
            "energy": r["energy"],
            "erep": r["erep"],
            "edisp2": r["edisp2"],
            "edisp3": r["edisp3"],
            "energy_nodisp": orc0.singlepoint(num, xyz)["energy"],
            "iterations": int(r["iterations"]),
            "qsh": [float(v) for v in r["qsh"]],
            "qat": [float(v) for v in r["qat"]],
        }
    with open(os.path.join(HERE, "singlepoints.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
