"""Multi-workgroup SCF path (molecules above the single-workgroup size, e.g.
one (H2O)700 cluster).  The path's size threshold is lowered for this test
(TBLITE_AMD_GRID_MIN_LD) so molecules the oracle finishes in seconds run
through it; results are compared with the oracle and with the single-workgroup
path on the same molecules.

Tolerances: oracle |dE| <= 1e-6 Eh, charges 1e-5 e, iterations +-1;
single-workgroup path vs grid path |dE| <= 1e-8 Eh."""
import os

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid_engine():
    os.environ["TBLITE_AMD_GRID_MIN_LD"] = "64"
    try:
        eng = tblite_amd.Engine()
    finally:
        del os.environ["TBLITE_AMD_GRID_MIN_LD"]
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def mols():
    rng = np.random.default_rng(55)
    return [workloads.water_cluster(14, workloads.liquid_radius(14), rng),
            workloads.organic_molecule(18, rng),
            workloads.water_cluster(40, workloads.liquid_radius(40), rng)]


def test_grid_vs_oracle(grid_engine, orc, mols):
    small = mols[:2]
    mol_nat, num, xyz = workloads.flatten(small)
    print(f"[gpu] grid path vs oracle, nao={[orc.dims(m[0])[2] for m in small]}", flush=True)
    e, it, st = grid_engine.singlepoint_flat(mol_nat, num, xyz)
    for m, (n_, x_) in enumerate(small):
        r = orc.singlepoint(n_, x_)
        assert st[m] == 0, (m, st[m])
        assert abs(e[m] - r["energy"]) < 1e-6, (m, e[m], r["energy"])
        assert abs(int(it[m]) - r["iterations"]) <= 1
        q = grid_engine.charges(m)
        assert np.max(np.abs(q["qsh"] - r["qsh"])) < 1e-5
        assert np.max(np.abs(q["dpat"] - r["dpat"])) < 1e-5


def test_grid_vs_single_workgroup(grid_engine, engine, mols):
    mol_nat, num, xyz = workloads.flatten(mols)
    print(f"[gpu] grid path vs single-workgroup path, {len(mols)} molecules", flush=True)
    e1, it1, st1 = grid_engine.singlepoint_flat(mol_nat, num, xyz)
    e2, it2, st2 = engine.singlepoint_flat(mol_nat, num, xyz)
    assert np.all(st1 == 0) and np.all(st2 == 0)
    assert np.max(np.abs(e1 - e2)) < 1e-8, (e1, e2)
    assert np.max(np.abs(it1.astype(int) - it2.astype(int))) <= 1


def test_grid_graph_replay_equals_eager(mols):
    """The eigenvalue phase of SCF iterations 2.. replays a captured HIP graph;
    between replays the eager eigenvector, density and mixing kernels and host
    synchronisations run.  Every replay feeds the next iteration, so equal
    energies and iteration counts with and without the graph check them all."""
    out = {}
    for flag in ("1", "0"):
        os.environ["TBLITE_AMD_GRID_MIN_LD"] = "64"
        os.environ["TBLITE_AMD_GRID_GRAPH"] = flag
        try:
            eng = tblite_amd.Engine()
        finally:
            del os.environ["TBLITE_AMD_GRID_MIN_LD"]
            del os.environ["TBLITE_AMD_GRID_GRAPH"]
        res = []
        for num, xyz in mols[1:]:
            print(f"[gpu] grid graph={flag} nat={len(num)}", flush=True)
            e, it, st = eng.singlepoint_flat(np.array([len(num)]), num, xyz)
            assert st[0] == 0
            res.append((float(e[0]), int(it[0])))
        # the same engine twice: the second run reuses nothing from the first graph
        e2, it2, st2 = eng.singlepoint_flat(np.array([len(mols[1][0])]), *mols[1])
        assert abs(float(e2[0]) - res[0][0]) <= 1e-12 and int(it2[0]) == res[0][1]
        eng.close()
        out[flag] = res
    for (eg, ig), (ee, ie) in zip(out["1"], out["0"]):
        assert ig == ie
        assert abs(eg - ee) <= 1e-10, (eg, ee)
