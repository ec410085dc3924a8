"""File -> engine -> JSON on the GPU: a structure read from an xyz file (in
angstrom) and from a Turbomole coord file (bohr, with $eht charge and
unpaired electrons) runs through tblite_amd_xtb_singlepoint, and its result
written by tblite_amd_write_json reads back equal to the oracle's singlepoint
on the same coordinates (|dE| <= 1e-6 Eh, charges and atomic dipoles 1e-5,
iterations +-1; the JSON dipole is sum_a (q_a r_a + mu_a) of the written
arrays)."""
import json

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu
AUTOAA = 0.52917721090380
SYM = {1: "H", 6: "C", 7: "N", 8: "O", 16: "S"}


def _roundtrip(engine, orc, path, charge_ref, uhf_ref, tmp_path, tag):
    num, xyz, charge, uhf = tblite_amd.read_structure(path)
    assert charge == charge_ref and uhf == uhf_ref
    print(f"[gpu] file -> engine -> JSON: {tag}, nat={len(num)}", flush=True)
    r = engine.singlepoint(num, xyz, charge=charge, uhf=uhf)
    assert r["rc"] == 0, engine.error()
    out = tmp_path / f"{tag}.json"
    tblite_amd.write_json(out, num, xyz, r["energy"], r["iterations"], r["qat"],
                          r["dpat"].reshape(-1, 3), r["qpat"].reshape(-1, 6), r["emo"], r["focc"])
    d = json.loads(out.read_text())
    ref = orc.singlepoint(num, xyz, charge=charge, uhf=uhf)
    assert abs(d["energy"] - ref["energy"]) < 1e-6, (tag, d["energy"], ref["energy"])
    assert abs(d["scf-iterations"] - ref["iterations"]) <= 1
    assert d["numbers"] == [int(z) for z in num]
    assert np.max(np.abs(np.array(d["charges"]) - ref["qat"])) < 1e-5
    assert abs(sum(d["charges"]) - charge) < 1e-6
    dip = (ref["qat"][:, None] * xyz + ref["dpat"]).sum(0)
    assert np.max(np.abs(np.array(d["dipole"]) - dip)) < 1e-4
    assert np.allclose(np.array(d["orbital-occupations"]), ref["focc"], rtol=0, atol=1e-6)


def test_xyz_file_to_json(engine, orc, tmp_path):
    rng = np.random.default_rng(61)
    num, xyz = workloads.organic_molecule(10, rng)
    p = tmp_path / "mol.xyz"
    lines = [str(len(num)), "organic"] + [
        f"{SYM[int(z)]} {x * AUTOAA:.14f} {y * AUTOAA:.14f} {w * AUTOAA:.14f}"
        for z, (x, y, w) in zip(num, xyz)]
    p.write_text("\n".join(lines) + "\n")
    _roundtrip(engine, orc, p, 0.0, 0, tmp_path, "xyz")


def test_coord_file_with_charge_to_json(engine, orc, tmp_path):
    rng = np.random.default_rng(62)
    num, xyz = workloads.water_cluster(6, workloads.liquid_radius(6), rng)
    p = tmp_path / "coord"
    lines = ["$coord"] + [f" {x:.14f} {y:.14f} {w:.14f} {SYM[int(z)].lower()}"
                          for z, (x, y, w) in zip(num, xyz)]
    lines += ["$eht charge=-1 unpaired=1", "$end"]
    p.write_text("\n".join(lines) + "\n")
    _roundtrip(engine, orc, p, -1.0, 1, tmp_path, "coord")
