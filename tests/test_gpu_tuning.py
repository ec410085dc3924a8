"""SCF path routing and its autotuner (tblite_amd_tuning).

A molecule of the global-memory size range that is alone in its batch runs
faster across the device (multi-workgroup path) than on one workgroup; in a
batch that fills the device it stays on one workgroup.  Both paths give the
same energy (|dE| <= 1e-8 Eh, the tolerance of the grid-vs-one-workgroup
test).  The autotuner's fit must be sane and applicable."""
import os
import sys

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_small_batch_goes_multi_workgroup(engine):
    rng = np.random.default_rng(71)
    num, xyz = workloads.water_cluster(60, workloads.liquid_radius(60), rng)
    auto = tblite_amd.Engine()
    try:
        assert auto.tuning()["auto_grid"] == 1
        print("[gpu] routing: one (H2O)60 alone", flush=True)
        e, it, st = auto.singlepoint_flat(np.array([len(num)]), num, xyz)
        assert st[0] == 0
        assert auto.molecule_path(0) == 2
        e1, it1, st1 = engine.singlepoint_flat(np.array([len(num)]), num, xyz)
        assert engine.molecule_path(0) == 1
        assert abs(float(e[0]) - float(e1[0])) <= 1e-8
        assert abs(int(it[0]) - int(it1[0])) <= 1
        # a batch that fills the device keeps one workgroup per molecule
        small = workloads.water_cluster(20, workloads.liquid_radius(20), rng)
        mols = [small] * 300
        mol_nat, nums, xyzs = workloads.flatten(mols)
        print("[gpu] routing: 300 x (H2O)20", flush=True)
        e, it, st = auto.singlepoint_flat(mol_nat, nums, xyzs)
        assert (st == 0).all()
        assert all(auto.molecule_path(m) == 1 for m in range(len(mols)))
        assert np.ptp(e) <= 1e-9
    finally:
        auto.close()


def test_set_tuning_rejects_bad_values(engine):
    with pytest.raises(tblite_amd.EngineError):
        engine.set_tuning(lds_max_ld=100)
    with pytest.raises(tblite_amd.EngineError):
        engine.set_tuning(wg_a=-1.0)
    assert engine.tuning()["lds_max_ld"] == 96


def test_autotune_quick_fit():
    sys.path.insert(0, os.path.join(ROOT, "tblite-cuda_amd"))
    import autotune

    eng = tblite_amd.Engine()
    try:
        print("[gpu] autotune (quick)", flush=True)
        t = autotune.autotune(eng, quick=True, log=lambda s: print(s, flush=True))
        assert t["wg_a"] > 0 and 1.0 < t["wg_p"] < 4.0
        assert t["grid_b"] >= 0
        eng.set_tuning(**t)
        got = eng.tuning()
        assert abs(got["wg_p"] - t["wg_p"]) < 1e-12 and got["lds_max_ld"] == t["lds_max_ld"]
    finally:
        eng.close()
