"""D4 dispersion on the GPU against the oracle, on all three SCF paths.

* two-body energy at the final charges and three-body energy per molecule
  (|dE| <= 1e-8 Eh: both are smooth functions of the charges, which agree to
  1e-5 e, and of the geometry), total energies within 1e-6 Eh, iterations +-1;
* molecules with more than 116 atoms split their three-body triples over
  several work items (the chunking of k_d4_atm);
* a parameter set with dispersion switched off gives the oracle's
  dispersion-free energy and zero dispersion parts;
* reference charges that differ between an element's references take the
  general (not factorised) weighting path;
* the multi-workgroup path (threshold lowered) carries the same terms.
"""
import os

import numpy as np
import pytest

import oracle
import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mols():
    rng = np.random.default_rng(404)
    return [workloads.water_cluster(5, workloads.liquid_radius(5), rng),       # LDS path
            workloads.organic_molecule(9, rng),                                 # LDS path, S d shells possible
            workloads.organic_molecule(30, rng),                                # global-memory path
            workloads.water_cluster(45, workloads.liquid_radius(45), rng)]      # 135 atoms: chunked triples


def _check(eng, orc, mols, tag):
    mol_nat, num, xyz = workloads.flatten(mols)
    print(f"[gpu] d4 {tag}: nat={list(mol_nat)}", flush=True)
    e, it, st = eng.singlepoint_flat(mol_nat, num, xyz)
    for m, (n_, x_) in enumerate(mols):
        r = orc.singlepoint(n_, x_)
        assert st[m] == 0, (tag, m, st[m])
        e2, e3 = eng.dispersion(m)
        assert abs(e3 - r["edisp3"]) < 1e-8, (tag, m, e3, r["edisp3"])
        assert abs(e2 - r["edisp2"]) < 1e-8, (tag, m, e2, r["edisp2"])
        assert abs(e[m] - r["energy"]) < 1e-6, (tag, m, e[m], r["energy"])
        assert abs(int(it[m]) - r["iterations"]) <= 1
        q = eng.charges(m)
        assert np.max(np.abs(q["qat"] - r["qat"])) < 1e-5


def test_d4_vs_oracle(engine, orc, mols):
    assert max(len(m[0]) for m in mols) > 116
    _check(engine, orc, mols, "one-workgroup paths")


def test_d4_multi_workgroup_path(orc, mols):
    os.environ["TBLITE_AMD_GRID_MIN_LD"] = "64"
    try:
        eng = tblite_amd.Engine(tuning={"auto_grid": 0})
    finally:
        del os.environ["TBLITE_AMD_GRID_MIN_LD"]
    try:
        _check(eng, orc, mols[1:3], "multi-workgroup path")
    finally:
        eng.close()


def test_d4_reference_charges_per_reference(mols):
    """Reference charges that differ between the references of an element
    (as in dftd4's tables) take the general weighting path of the SCF kernels
    instead of the factorised charge scale."""
    p = tblite_amd.default_param()
    p.d4.elem[0].q[1] = 0.05   # H, second reference
    p.d4.elem[1].q[3] = -0.04  # C
    p.d4.elem[3].q[2] = -0.10  # O
    orc1 = oracle.Oracle(p)
    eng = tblite_amd.Engine(param=p, tuning={"auto_grid": 0})
    try:
        _check(eng, orc1, mols[:3], "per-reference charges")
    finally:
        eng.close()


def test_d4_disabled(mols):
    p = tblite_amd.default_param()
    p.d4.enabled = 0
    orc0 = oracle.Oracle(p)
    eng = tblite_amd.Engine(param=p, tuning={"auto_grid": 0})
    try:
        sub = mols[:3]
        mol_nat, num, xyz = workloads.flatten(sub)
        print("[gpu] d4 disabled", flush=True)
        e, it, st = eng.singlepoint_flat(mol_nat, num, xyz)
        for m, (n_, x_) in enumerate(sub):
            r = orc0.singlepoint(n_, x_)
            assert st[m] == 0
            assert abs(e[m] - r["energy"]) < 1e-6
            assert eng.dispersion(m) == (0.0, 0.0)
    finally:
        eng.close()
