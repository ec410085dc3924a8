"""Charged and open-shell molecules on the GPU against the oracle, on all
three SCF paths: cations and anions (total charge +-1, +2) with the spin the
electron count implies, and explicit unpaired electrons (uhf), as the
reference's structure_type carries them (charge, uhf: src/structure.h:5-15;
electron split src/xtb.cu:300-313).  These exercise the alpha/beta Fermi
filling and the D4 charge scale away from neutral atoms.

Tolerances as elsewhere: |dE| <= 1e-6 Eh, charges 1e-5 e, SCF iterations
+-1; the atomic charges sum to the total charge (to the Fermi solver's
tolerance).  Single atoms and diatomics are the smallest inputs.
"""
import os

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cases():
    rng = np.random.default_rng(808)
    w4 = workloads.water_cluster(4, workloads.liquid_radius(4), rng)
    w9 = workloads.water_cluster(9, workloads.liquid_radius(9), rng)
    o7 = workloads.organic_molecule(7, rng)
    o22 = workloads.organic_molecule(22, rng)   # above the LDS-resident size
    # (molecule, charge, uhf)
    return [(w4, 1.0, 1), (w4, -1.0, 1), (w9, 2.0, 0), (o7, 1.0, 1), (o7, 0.0, 2),
            (o22, -1.0, 1), (o22, 1.0, 1)]


def _run(eng, orc, cases, tag):
    mols = [c[0] for c in cases]
    mol_nat, num, xyz = workloads.flatten(mols)
    charge = np.array([c[1] for c in cases])
    uhf = np.array([c[2] for c in cases], dtype=np.int32)
    print(f"[gpu] charged {tag}: nat={list(mol_nat)} charge={list(charge)} uhf={list(uhf)}", flush=True)
    e, it, st = eng.singlepoint_flat(mol_nat, num, xyz, charge=charge, uhf=uhf)
    for m, (mol, ch, u) in enumerate(cases):
        r = orc.singlepoint(mol[0], mol[1], charge=ch, uhf=u)
        assert r["rc"] == 0 and st[m] == 0, (tag, m, r["rc"], st[m])
        assert abs(e[m] - r["energy"]) < 1e-6, (tag, m, e[m], r["energy"])
        assert abs(int(it[m]) - r["iterations"]) <= 1, (tag, m, it[m], r["iterations"])
        q = eng.charges(m)
        assert np.max(np.abs(q["qsh"] - r["qsh"])) < 1e-5, (tag, m)
        # the Fermi level is solved to ~1e-8 electrons (oracle and engine alike)
        assert abs(q["qat"].sum() - ch) < 1e-6
        assert abs(q["focc"].sum() - (r["focc"].sum())) < 1e-8


def test_charged_one_workgroup_paths(engine, orc, cases):
    _run(engine, orc, cases, "one-workgroup paths")


def test_charged_multi_workgroup_path(orc, cases):
    os.environ["TBLITE_AMD_GRID_MIN_LD"] = "64"
    try:
        eng = tblite_amd.Engine(tuning={"auto_grid": 0})
    finally:
        del os.environ["TBLITE_AMD_GRID_MIN_LD"]
    try:
        big = [c for c in cases if orc.dims(c[0][0])[2] > 64]
        assert big
        _run(eng, orc, big, "multi-workgroup path")
    finally:
        eng.close()


def test_atoms_and_diatomics(engine, orc):
    """Single atoms (one AO for H, nat = 1) with their ground-state spin, and
    diatomics: the smallest inputs the batch layout and the LDS-resident
    solver see (padded dimension 4)."""
    mols = [(np.array([1]), np.zeros((1, 3)), 1), (np.array([8]), np.zeros((1, 3)), 2),
            (np.array([16]), np.zeros((1, 3)), 2), (np.array([7]), np.zeros((1, 3)), 3),
            (np.array([1, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]), 0),
            (np.array([6, 8]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.13]]), 0)]
    _run(engine, orc, [((m[0], m[1]), 0.0, m[2]) for m in mols], "atoms and diatomics")
