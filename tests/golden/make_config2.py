"""Write the full-size fixture of the single-large-molecule workload: one
seeded (H2O)700 cluster at liquid density (2100 atoms, 4200 AOs), run by the
CPU oracle at the default settings (accuracy 1, 300 K, damping 0.4, 250
iterations).  The fixture holds a hash of the bohr coordinates (the GPU test
regenerates them from the seed and checks the hash), the energy and its
parts, the iteration count, the shell and atomic charges and the orbital
energies.  The oracle needs a few minutes per SCF iteration on 8 cores at
this size, so the fixture is computed once and committed.

    python tests/golden/make_config2.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tblite-cuda_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
import workloads  # noqa: E402


def geometry_hash(num, xyz):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(num, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(xyz, dtype=np.float64).tobytes())
    return h.hexdigest()


def main():
    num, xyz = workloads.config2()[0]
    orc = oracle.Oracle()
    t = time.time()
    r = orc.singlepoint(num, xyz)
    dt = time.time() - t
    out = {
        "workload": "one seeded (H2O)700 cluster at liquid density, workloads.config2()",
        "nat": int(len(num)), "nao": int(orc.dims(num)[2]),
        "geometry_sha256": geometry_hash(num, xyz),
        "energy": r["energy"], "erep": r["erep"], "eelec": r["eelec"],
        "edisp2": r["edisp2"], "edisp3": r["edisp3"],
        "iterations": int(r["iterations"]), "status": int(r["status"]),
        "qsh": [float(x) for x in r["qsh"]], "qat": [float(x) for x in r["qat"]],
        "emo": [float(x) for x in r["emo"]], "focc_sum": float(np.sum(r["focc"])),
        "oracle_seconds": round(dt, 1),
    }
    with open(os.path.join(HERE, "config2.json"), "w") as f:
        json.dump(out, f)
    print(f"config2: E = {r['energy']:.10f} Eh, {r['iterations']} iterations, {dt:.0f} s")


if __name__ == "__main__":
    main()
