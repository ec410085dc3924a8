"""SCF path routing, work tiles and their autotuner (tblite_amd_tuning).

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


def test_work_tiles_do_not_change_results(engine):
    """Every work-tile setting (integral work-item size, pair-plane and
    three-body chunks, fine buckets, one molecule per CU on the
    global-memory path) gives the default's energies within 1e-9 Eh (only
    summation orders and bucket shapes change) and iteration counts."""
    rng = np.random.default_rng(72)
    mols = [workloads.water_cluster(n, workloads.liquid_radius(n), rng) for n in (3, 8, 15, 22)]
    mols += [workloads.organic_molecule(h, rng) for h in (6, 17, 33)]
    mol_nat, num, xyz = workloads.flatten(mols)
    e0, it0, st0 = engine.singlepoint_flat(mol_nat, num, xyz)
    assert (st0 == 0).all()
    eng = tblite_amd.Engine(tuning={"auto_grid": 0})
    try:
        for tiles in ({"integral_chunk": 32}, {"integral_chunk": 64}, {"pair_chunk_log2": 8},
                      {"atm_chunk_log2": 10}, {"fine_buckets": 1}, {"big_lds_min_kb": 96}):
            eng.set_tuning(**tiles)
            print(f"[gpu] work tiles {tiles}", flush=True)
            e, it, st = eng.singlepoint_flat(mol_nat, num, xyz)
            assert (st == 0).all()
            assert np.max(np.abs(e - e0)) <= 1e-9, tiles
            assert np.array_equal(it, it0), tiles
            eng.set_tuning(integral_chunk=128, pair_chunk_log2=15, atm_chunk_log2=18,
                           fine_buckets=0, big_lds_min_kb=0)
        with pytest.raises(tblite_amd.EngineError):
            eng.set_tuning(integral_chunk=256)
        with pytest.raises(tblite_amd.EngineError):
            eng.set_tuning(pair_chunk_log2=30)
    finally:
        eng.close()


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
        for k in autotune.TILE_CANDIDATES:
            assert got[k] == t[k] and t[k] in autotune.TILE_CANDIDATES[k]
    finally:
        eng.close()
