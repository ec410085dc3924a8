"""Parity of the HIP engine with the CPU oracle, stage by stage and end to end
(energy within 1e-6 Eh, Mulliken charges within 1e-5 e, SCF iterations within
+-1; the neighbour list bit-exact).  Runs on a GPU through the C ABI."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def _mols_small():
    mols = workloads.config0()
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 5, 8, 16):
        mols.append(workloads.water_cluster(n, workloads.liquid_radius(n), rng))
    return mols


@pytest.fixture(scope="module")
def batch(engine):
    mols = _mols_small()
    mol_nat, num, xyz = workloads.flatten(mols)
    print(f"[gpu] batched singlepoint over {len(mols)} water clusters", flush=True)
    e, it, st = engine.singlepoint_flat(mol_nat, num, xyz)
    return mols, e, it, st


def test_status(batch):
    _, e, it, st = batch
    assert np.all(st == 0), st
    assert np.all(np.isfinite(e))


def test_energy_parity(batch, orc):
    mols, e, it, st = batch
    for m, (num, xyz) in enumerate(mols):
        r = orc.singlepoint(num, xyz)
        assert r["converged"]
        assert abs(e[m] - r["energy"]) < 1e-6, (m, e[m], r["energy"])
        assert abs(int(it[m]) - r["iterations"]) <= 1, (m, it[m], r["iterations"])


def test_charges_parity(batch, orc, engine):
    mols, *_ = batch
    for m, (num, xyz) in enumerate(mols):
        r = orc.singlepoint(num, xyz)
        g = engine.charges(m)
        assert np.max(np.abs(g["qsh"] - r["qsh"])) < 1e-5
        assert np.max(np.abs(g["qat"] - r["qat"])) < 1e-5
        assert np.max(np.abs(g["dpat"] - r["dpat"])) < 1e-5
        assert np.max(np.abs(g["qpat"] - r["qpat"])) < 1e-5


def test_adjlist_bitexact(batch, orc, engine):
    mols, *_ = batch
    for m, (num, xyz) in enumerate(mols):
        inl, nnl, nlat = engine.adjlist(m)
        cut = orc.cutoff(num)
        oi, on, ol = orc.adjlist(xyz, cut)
        assert np.array_equal(inl, oi) and np.array_equal(nnl, on) and np.array_equal(nlat, ol)


def test_integrals_parity(batch, orc, engine):
    mols, *_ = batch
    for m, (num, xyz) in enumerate(mols[:4]):
        S, D, Q, H = engine.integrals(m)
        So, Do, Qo, Ho = orc.hamiltonian(num, xyz)
        assert np.max(np.abs(S - So)) < 1e-12
        assert np.max(np.abs(D - Do)) < 1e-11
        assert np.max(np.abs(Q - Qo)) < 1e-11
        assert np.max(np.abs(H - Ho)) < 1e-12


def test_cn_selfenergy(batch, orc, engine):
    mols, *_ = batch
    for m, (num, xyz) in enumerate(mols):
        cn = engine.cn(m)
        assert np.max(np.abs(cn - orc.cn(num, xyz))) < 1e-12
        se = engine.selfenergy(m)
        assert np.max(np.abs(se - orc.selfenergy(num, orc.cn(num, xyz)))) < 1e-12
