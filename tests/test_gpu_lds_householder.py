"""The alternative eigensolver of the LDS-resident SCF kernels (padded AO
dimension <= 96): Householder tridiagonalisation with the reflectors applied
to the basis rows, bisection, twisted-factorisation eigenvectors
(TBLITE_AMD_LDS_EIG=householder), against the oracle on all three LDS kernel
sizes, a cation and single atoms; and against the default Jacobi solver.

Tolerances: |dE| <= 1e-6 Eh, charges 1e-5 e, iterations +-1; Householder vs
Jacobi |dE| <= 1e-8 Eh."""
import os

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hh_engine():
    os.environ["TBLITE_AMD_LDS_EIG"] = "householder"
    try:
        eng = tblite_amd.Engine(tuning={"auto_grid": 0})
    finally:
        del os.environ["TBLITE_AMD_LDS_EIG"]
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def cases():
    rng = np.random.default_rng(303)
    mols = [workloads.water_cluster(n, workloads.liquid_radius(n), rng) for n in (2, 5, 9, 14, 16)]
    mols += [workloads.organic_molecule(h, rng) for h in (4, 9, 14)]
    mols += [(np.array([8]), np.zeros((1, 3))), (np.array([1, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]))]
    charge = np.zeros(len(mols))
    uhf = np.zeros(len(mols), dtype=np.int32)
    charge[2], uhf[2] = 1.0, 1        # (H2O)9 cation
    uhf[-2] = 2                        # O atom triplet
    return mols, charge, uhf


def test_householder_vs_oracle(hh_engine, orc, cases):
    mols, charge, uhf = cases
    mol_nat, num, xyz = workloads.flatten(mols)
    print(f"[gpu] LDS Householder eigensolver: nat={list(mol_nat)}", flush=True)
    e, it, st = hh_engine.singlepoint_flat(mol_nat, num, xyz, charge=charge, uhf=uhf)
    for m, (n_, x_) in enumerate(mols):
        assert hh_engine.molecule_path(m) == 0
        r = orc.singlepoint(n_, x_, charge=charge[m], uhf=int(uhf[m]))
        assert st[m] == 0, (m, st[m])
        assert abs(e[m] - r["energy"]) < 1e-6, (m, e[m], r["energy"])
        assert abs(int(it[m]) - r["iterations"]) <= 1, (m, it[m], r["iterations"])
        q = hh_engine.charges(m)
        assert np.max(np.abs(q["qsh"] - r["qsh"])) < 1e-5, m
        assert np.max(np.abs(q["emo"] - r["emo"])) < 1e-6, m


def test_householder_vs_jacobi(hh_engine, engine, cases):
    mols, charge, uhf = cases
    mol_nat, num, xyz = workloads.flatten(mols)
    print("[gpu] LDS Householder vs Jacobi", flush=True)
    e1, it1, st1 = hh_engine.singlepoint_flat(mol_nat, num, xyz, charge=charge, uhf=uhf)
    e2, it2, st2 = engine.singlepoint_flat(mol_nat, num, xyz, charge=charge, uhf=uhf)
    assert np.all(st1 == 0) and np.all(st2 == 0)
    assert np.max(np.abs(e1 - e2)) < 1e-8
