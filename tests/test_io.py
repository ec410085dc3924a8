"""Structure readers and the JSON result writer of the C ABI (host code, no
GPU).  Pins: unit conversion (xyz in Angstrom -> bohr, CODATA 2018), the
Turbomole $coord/$eht groups, element symbols in either case, round trips,
and the molecular dipole sum_a (q_a r_a + mu_a) written to JSON."""
import json

import numpy as np
import pytest

import tblite_amd
import workloads

AUTOAA = 0.52917721090380


def test_xyz_angstrom_to_bohr(tmp_path):
    p = tmp_path / "w.xyz"
    p.write_text("3\nwater\nO 0.0 0.0 0.1173\nh 0.0 0.7572 -0.4692\n1 0.0 -0.7572 -0.4692\n")
    num, xyz, charge, uhf = tblite_amd.read_structure(p)
    assert list(num) == [8, 1, 1]
    assert np.allclose(xyz[1], np.array([0.0, 0.7572, -0.4692]) / AUTOAA, rtol=0, atol=1e-12)
    assert charge == 0.0 and uhf == 0


def test_coord_bohr_with_charge_and_spin(tmp_path):
    p = tmp_path / "coord"
    p.write_text("$coord\n 0.0 0.0 0.0 c\n 0.0 0.0 2.1 o\n$eht charge=1 unpaired=1\n$end\n")
    num, xyz, charge, uhf = tblite_amd.read_structure(p)
    assert list(num) == [6, 8]
    assert np.array_equal(xyz, np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.1]]))
    assert charge == 1.0 and uhf == 1


def test_round_trip_cluster(tmp_path):
    rng = np.random.default_rng(5)
    num, xyz = workloads.water_cluster(6, workloads.liquid_radius(6), rng)
    p = tmp_path / "c.xyz"
    sym = {1: "H", 8: "O"}
    lines = [str(len(num)), "cluster"] + [
        f"{sym[int(z)]} {x * AUTOAA:.12f} {y * AUTOAA:.12f} {w * AUTOAA:.12f}" for z, (x, y, w) in zip(num, xyz)]
    p.write_text("\n".join(lines) + "\n")
    n2, x2, _, _ = tblite_amd.read_structure(p)
    assert np.array_equal(n2, num)
    assert np.abs(x2 - xyz).max() < 1e-10


@pytest.mark.parametrize("text", ["2\nx\nXx 0 0 0\nH 0 0 1\n", "3\nshort\nH 0 0 0\n"])
def test_bad_xyz_raises(tmp_path, text):
    p = tmp_path / "bad.xyz"
    p.write_text(text)
    with pytest.raises(ValueError):
        tblite_amd.read_structure(p)


def test_missing_file_raises(tmp_path):
    with pytest.raises(ValueError):
        tblite_amd.read_structure(tmp_path / "none.xyz")


def test_json_fields_and_dipole(tmp_path):
    num = np.array([8, 1, 1])
    xyz = np.array([[0.0, 0.0, 0.2], [0.0, 1.4, -0.9], [0.0, -1.4, -0.9]])
    qat = np.array([-0.6, 0.3, 0.3])
    dpat = np.array([[0.0, 0.0, 0.1], [0.0, 0.02, 0.0], [0.0, -0.02, 0.0]])
    qpat = np.zeros((3, 6))
    emo = np.array([-1.0, -0.5, 0.2])
    focc = np.array([2.0, 2.0, 0.0])
    p = tmp_path / "r.json"
    tblite_amd.write_json(p, num, xyz, -5.07, 9, qat, dpat, qpat, emo, focc)
    d = json.loads(p.read_text())
    assert d["energy"] == -5.07 and d["scf-iterations"] == 9
    assert d["numbers"] == [8, 1, 1]
    assert np.allclose(d["charges"], qat) and np.allclose(d["orbital-occupations"], focc)
    want = (qat[:, None] * xyz + dpat).sum(0)
    assert np.allclose(d["dipole"], want, rtol=0, atol=1e-15)
