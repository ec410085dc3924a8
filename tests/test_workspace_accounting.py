"""Serving admission accounting, on the host (no GPU calls): the footprint the
server counts for a batch of requests covers the device workspace the engine
allocates for that batch, on all three SCF paths (the multi-workgroup path by
lowering its size threshold), so a slot's reserved workspace never has to be
freed and reallocated."""
import ctypes as C

import numpy as np
import pytest

import tblite_amd
import workloads


def _estimate(mols, grid_min_ld):
    lib = tblite_amd.load()
    mol_nat = np.array([len(m[0]) for m in mols], dtype=np.int32)
    num = np.concatenate([m[0] for m in mols]).astype(np.int32)
    need, admit = C.c_double(), C.c_double()
    rc = lib.tblite_amd_workspace_estimate(None, len(mols), tblite_amd._ip(mol_nat),
                                           tblite_amd._ip(num), grid_min_ld, C.byref(need),
                                           C.byref(admit))
    assert rc == 0
    return need.value, admit.value


@pytest.mark.parametrize("grid_min_ld", [1024, 128, 64])
def test_admission_covers_workspace(grid_min_ld):
    rng = np.random.default_rng(91)
    for trial in range(6):
        mols = []
        for _ in range(int(rng.integers(1, 12))):
            if rng.random() < 0.5:
                n = int(rng.integers(2, 40))
                mols.append(workloads.water_cluster(n, workloads.liquid_radius(n), rng))
            else:
                mols.append(workloads.organic_molecule(int(rng.integers(4, 40)), rng))
        need, admit = _estimate(mols, grid_min_ld)
        assert need > 0
        assert admit >= need, (trial, grid_min_ld, need, admit)
        # and not wildly above it (the reserve is a few hundred KiB per batch)
        assert admit <= 1.6 * need + 2 ** 20, (trial, need, admit)


def test_grid_molecule_brings_pair_planes():
    """One multi-workgroup molecule makes the engine allocate the AES2 pair
    planes of every molecule in the batch; the admission counts them."""
    rng = np.random.default_rng(92)
    # the planes (18 nat^2 doubles per molecule, 12 MB here) outweigh the
    # difference between the large molecule's own path-2 and path-3 scratch
    small = [workloads.water_cluster(8, workloads.liquid_radius(8), rng) for _ in range(150)]
    big = workloads.water_cluster(30, workloads.liquid_radius(30), rng)  # 180 AOs
    need_lo, _ = _estimate(small + [big], 1024)
    need_hi, admit_hi = _estimate(small + [big], 128)
    assert need_hi > need_lo
    assert admit_hi >= need_hi


def test_unknown_element_is_an_error():
    lib = tblite_amd.load()
    mol_nat = np.array([2], dtype=np.int32)
    num = np.array([1, 92], dtype=np.int32)
    need, admit = C.c_double(), C.c_double()
    assert lib.tblite_amd_workspace_estimate(None, 1, tblite_amd._ip(mol_nat), tblite_amd._ip(num),
                                             1024, C.byref(need), C.byref(admit)) < 0
