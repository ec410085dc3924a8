"""HIP engine vs the fixtures and vs the oracle on organic molecules with
sulfur d shells, including molecules above the LDS-resident size (nao > 96)
that run on the global-memory path."""
import json
import os

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "singlepoints.json")))


def test_engine_vs_fixtures(engine):
    names = sorted(GOLD)
    mols = [(np.array(GOLD[k]["num"]), np.array(GOLD[k]["xyz_bohr"])) for k in names]
    mol_nat, num, xyz = workloads.flatten(mols)
    print(f"[gpu] fixtures: {names}", flush=True)
    e, it, st = engine.singlepoint_flat(mol_nat, num, xyz)
    for m, k in enumerate(names):
        g = GOLD[k]
        assert st[m] == 0
        assert abs(e[m] - g["energy"]) < 1e-6, (k, e[m], g["energy"])
        assert abs(int(it[m]) - g["iterations"]) <= 1
        q = engine.charges(m)
        assert np.max(np.abs(q["qsh"] - np.array(g["qsh"]))) < 1e-5


@pytest.fixture(scope="module")
def organics():
    rng = np.random.default_rng(33)
    mols = [workloads.organic_molecule(int(nh), rng) for nh in (4, 8, 16, 24, 30, 40, 55)]
    return mols


def test_organics_parity(engine, orc, organics):
    mol_nat, num, xyz = workloads.flatten(organics)
    naos = [orc.dims(m[0])[2] for m in organics]
    print(f"[gpu] organics nao={naos}", flush=True)
    assert max(naos) > 96  # exercises the global-memory SCF path
    e, it, st = engine.singlepoint_flat(mol_nat, num, xyz)
    for m, (n_, x_) in enumerate(organics):
        r = orc.singlepoint(n_, x_)
        assert st[m] == 0, (m, st[m])
        assert abs(e[m] - r["energy"]) < 1e-6, (m, naos[m], e[m], r["energy"])
        assert abs(int(it[m]) - r["iterations"]) <= 1
        q = engine.charges(m)
        assert np.max(np.abs(q["qsh"] - r["qsh"])) < 1e-5
        assert np.max(np.abs(q["dpat"] - r["dpat"])) < 1e-5
        assert np.max(np.abs(q["qpat"] - r["qpat"])) < 1e-5


def test_organic_integrals(engine, orc, organics):
    # the batch of the previous test is still resident
    for m in (0, 3):
        n_, x_ = organics[m]
        S, D, Q, H = engine.integrals(m)
        So, Do, Qo, Ho = orc.hamiltonian(n_, x_)
        assert np.max(np.abs(S - So)) < 1e-12
        assert np.max(np.abs(D - Do)) < 1e-11
        assert np.max(np.abs(Q - Qo)) < 1e-11
        assert np.max(np.abs(H - Ho)) < 1e-12
