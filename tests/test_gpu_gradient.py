"""Analytic gradients on the GPU (tblite_amd_singlepoint_flat_gradient and
tblite_amd_xtb_singlepoint_gradient) against the oracle's analytic gradient,
which tests/test_oracle_gradient.py pins to central differences of the
energy.  Covers the three SCF paths (LDS-resident, one workgroup in global
memory, multi-workgroup with its size threshold lowered), d shells (S), a
charged open-shell cluster with fractional occupations and single atoms.

Tolerances: |g_gpu - g_oracle| <= 1e-8 Eh/bohr (measured 1.7e-10 on the
one-workgroup paths, 1.2e-12 on the multi-workgroup path) and
|sigma_gpu - sigma_oracle| <= 1e-6 Eh at the default SCF accuracy, energies
1e-6 Eh; sum_a g_a = 0 to 1e-9.  At that accuracy both SCFs stop at a charge residual of 2e-5 e, so
the density is stationary only to that order and sigma is symmetric only to
about 1e-7 Eh, in the oracle as on the GPU (the symmetry itself is pinned at
tight accuracy on the CPU, tests/test_oracle_gradient.py)."""
import os

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu
GTOL = 1e-8


@pytest.fixture(scope="module")
def cases():
    rng = np.random.default_rng(811)
    mols = [workloads.water_cluster(n, workloads.liquid_radius(n), rng) for n in (2, 6, 13)]
    mols += [workloads.organic_molecule(h, rng) for h in (5, 9, 14)]
    mols += [workloads.water_cluster(20, workloads.liquid_radius(20), rng),   # 120 AOs: global memory
             workloads.organic_molecule(30, rng)]
    mols += [(np.array([8]), np.zeros((1, 3))),
             (np.array([1, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.3, 1.4]]))]
    charge = np.zeros(len(mols))
    uhf = np.zeros(len(mols), dtype=np.int32)
    charge[1], uhf[1] = 1.0, 1          # (H2O)6 cation
    uhf[-2] = 2                          # O atom triplet
    return mols, charge, uhf


def _check(e, g, sg, st, mols, charge, uhf, orc, tag):
    off = 0
    worst = 0.0
    for m, (num, xyz) in enumerate(mols):
        nat = len(num)
        r = orc.singlepoint(num, xyz, charge=charge[m], uhf=int(uhf[m]), gradient=True)
        gm = g[off:off + nat]
        off += nat
        assert st[m] == 0, (tag, m, st[m])
        assert abs(e[m] - r["energy"]) < 1e-6, (tag, m, e[m], r["energy"])
        d = np.max(np.abs(gm - r["gradient"]))
        worst = max(worst, d)
        assert d < GTOL, (tag, m, d)
        assert np.max(np.abs(gm.sum(0))) < 1e-9, (tag, m)
        assert np.max(np.abs(sg[m] - r["sigma"])) < 1e-6, (tag, m, np.max(np.abs(sg[m] - r["sigma"])))
        assert np.allclose(sg[m], np.einsum("ai,aj->ij", gm, xyz), atol=1e-10)
    print(f"[gpu] {tag}: max |g_gpu - g_oracle| = {worst:.2e} Eh/bohr", flush=True)
    return worst


def test_gradient_vs_oracle(engine, orc, cases):
    mols, charge, uhf = cases
    mol_nat, num, xyz = workloads.flatten(mols)
    print(f"[gpu] gradient batch: nat={list(mol_nat)}", flush=True)
    e, g, sg, it, st = engine.singlepoint_flat_gradient(mol_nat, num, xyz, charge=charge, uhf=uhf)
    paths = {engine.molecule_path(m) for m in range(len(mols))}
    assert paths == {0, 1}
    _check(e, g, sg, st, mols, charge, uhf, orc, "LDS + global-memory paths")


def test_gradient_energy_equals_energy_only_call(engine, cases):
    mols, charge, uhf = cases
    mol_nat, num, xyz = workloads.flatten(mols)
    e1, it1, st1 = engine.singlepoint_flat(mol_nat, num, xyz, charge=charge, uhf=uhf)
    e2, g, sg, it2, st2 = engine.singlepoint_flat_gradient(mol_nat, num, xyz, charge=charge, uhf=uhf)
    assert np.array_equal(it1, it2)
    assert np.max(np.abs(e1 - e2)) < 1e-12


def test_gradient_multi_workgroup_path(orc, cases):
    mols, charge, uhf = cases
    sel = [2, 5, 6]
    sub = [mols[k] for k in sel]
    os.environ["TBLITE_AMD_GRID_MIN_LD"] = "64"
    try:
        eng = tblite_amd.Engine()
    finally:
        del os.environ["TBLITE_AMD_GRID_MIN_LD"]
    try:
        mol_nat, num, xyz = workloads.flatten(sub)
        print(f"[gpu] gradient, multi-workgroup path: nat={list(mol_nat)}", flush=True)
        e, g, sg, it, st = eng.singlepoint_flat_gradient(mol_nat, num, xyz)
        assert all(eng.molecule_path(m) == 2 for m in range(len(sub)))
        _check(e, g, sg, st, sub, np.zeros(len(sub)), np.zeros(len(sub), dtype=np.int32), orc,
               "multi-workgroup path")
    finally:
        eng.close()


def test_per_molecule_gradient_entry(engine, orc, cases):
    mols, charge, uhf = cases
    num, xyz = mols[5]
    print(f"[gpu] xtb_singlepoint_gradient: nat={len(num)}", flush=True)
    r = engine.singlepoint(num, xyz, gradient=True)
    assert r["rc"] == 0
    ref = orc.singlepoint(num, xyz, gradient=True)
    assert np.max(np.abs(r["gradient"] - ref["gradient"])) < GTOL
    assert np.max(np.abs(r["sigma"] - ref["sigma"])) < 1e-5
    g, sg = engine.gradient(0)
    assert np.array_equal(g, r["gradient"])


def test_device_resident_gradient_batch(orc, cases):
    """tblite_amd_batch_prepare_gradient / run_gradient / gradients on two
    geometries of one composition (a displaced copy second): each run's
    gradients equal the oracle's at its own geometry."""
    mols, charge, uhf = cases
    sub = [mols[0], mols[3], mols[6]]
    mol_nat, num, xyz = workloads.flatten(sub)
    eng = tblite_amd.Engine(tuning={"auto_grid": 0})
    try:
        eng.prepare(mol_nat, num, gradient=True)
        rng = np.random.default_rng(5)
        for step, x in enumerate((xyz, xyz + 0.05 * rng.standard_normal(xyz.shape))):
            print(f"[gpu] resident gradient batch, geometry {step}", flush=True)
            eng.set_geometry(x)
            eng.run_gradient()
            g, sg = eng.gradients()
            e, it, st = eng.results()
            assert np.all(st == 0)
            off = 0
            for m, (n_, _) in enumerate(sub):
                xm = x[off:off + len(n_)]
                r = orc.singlepoint(n_, xm, gradient=True)
                assert abs(e[m] - r["energy"]) < 1e-6
                assert np.max(np.abs(g[off:off + len(n_)] - r["gradient"])) < GTOL
                off += len(n_)
    finally:
        eng.close()
