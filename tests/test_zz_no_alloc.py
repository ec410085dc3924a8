"""tblite_amd_singlepoint_flat_noalloc (the entry point that allocates no
device memory: its workspace is a static device array) against
tblite_amd_singlepoint_flat on the same inputs, on two calls in a row, the
second with a smaller batch; and its size limits, which are checked on the
host before any kernel runs.

Bar: energies within 1e-10 Eh (the same kernels run on the same inputs),
equal iteration counts and status words, shell charges within 1e-10 e."""
import ctypes as C

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engines():
    a = tblite_amd.Engine(tuning={"auto_grid": 0})
    b = tblite_amd.Engine(tuning={"auto_grid": 0})
    yield a, b
    a.close()
    b.close()


def _noalloc(eng, mols):
    mol_nat, num, xyz = workloads.flatten(mols)
    mol_nat = np.ascontiguousarray(mol_nat, dtype=np.int32)
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = len(mols)
    e = np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    rc = eng.lib.tblite_amd_singlepoint_flat_noalloc(
        eng.ctx, eng.calc, n, tblite_amd._ip(mol_nat), tblite_amd._ip(num), tblite_amd._dp(xyz),
        None, None, 1.0, tblite_amd._dp(e), tblite_amd._ip(it), tblite_amd._ip(st))
    eng._nmol = n
    return rc, e, it, st


def _compare(fixed, ref, mols, tag):
    rc, e, it, st = _noalloc(fixed, mols)
    assert rc == 0, (tag, rc, fixed.error())
    mol_nat, num, xyz = workloads.flatten(mols)
    e0, it0, st0 = ref.singlepoint_flat(mol_nat, num, xyz)
    assert np.all(st == 0) and np.all(st0 == 0), tag
    assert np.array_equal(it, it0), (tag, it, it0)
    assert np.max(np.abs(e - e0)) < 1e-10, (tag, np.max(np.abs(e - e0)))
    for m in range(len(mols)):
        q, q0 = fixed.charges(m), ref.charges(m)
        assert np.max(np.abs(q["qsh"] - q0["qsh"])) < 1e-10, (tag, m)


def test_noalloc_matches_flat_two_calls(engines):
    fixed, ref = engines
    rng = np.random.default_rng(123)
    first = [workloads.water_cluster(int(n), workloads.liquid_radius(int(n)), rng)
             for n in (3, 8, 14)]
    first += [workloads.organic_molecule(int(h), rng) for h in (6, 12, 24)]  # one above 96 AOs
    second = [workloads.water_cluster(5, workloads.liquid_radius(5), rng),
              workloads.organic_molecule(9, rng)]
    print(f"[gpu] no-alloc entry, call 1: {len(first)} molecules", flush=True)
    _compare(fixed, ref, first, "call 1")
    print(f"[gpu] no-alloc entry, call 2: {len(second)} molecules", flush=True)
    _compare(fixed, ref, second, "call 2")


def test_noalloc_limits(engines):
    fixed, _ = engines
    rng = np.random.default_rng(124)
    # workspace above the fixed array: rejected before any kernel runs
    mols = [workloads.water_cluster(16, workloads.liquid_radius(16), rng) for _ in range(40)]
    lib = tblite_amd.load()
    mol_nat, num, _ = workloads.flatten(mols)
    need, admit = C.c_double(), C.c_double()
    assert lib.tblite_amd_workspace_estimate(
        None, len(mols), tblite_amd._ip(np.ascontiguousarray(mol_nat, dtype=np.int32)),
        tblite_amd._ip(np.ascontiguousarray(num, dtype=np.int32)), 1024, C.byref(need),
        C.byref(admit)) == 0
    assert need.value > (64 << 20)
    print("[gpu] no-alloc entry, oversize batch (host-side rejection)", flush=True)
    rc, *_ = _noalloc(fixed, mols)
    assert rc == -8 and "fixed workspace" in fixed.error()
    # a molecule routed to the multi-workgroup path: rejected on the host
    eng = tblite_amd.Engine(tuning={"auto_grid": 0, "grid_min_ld": 64})
    try:
        print("[gpu] no-alloc entry, multi-workgroup molecule (host-side rejection)", flush=True)
        rc, *_ = _noalloc(eng, [workloads.organic_molecule(24, rng)])
        assert rc == -8 and "multi-workgroup" in eng.error()
    finally:
        eng.close()
