"""Molecules of the global-memory single-workgroup path that share a size
bucket with a larger one: each result must not depend on what else is in the
bucket.  The bucket's LDS layout is sized by its largest padded AO dimension;
a (H2O)117 cluster (702 AOs) next to a (H2O)130 cluster (780 AOs) exercises
the layout of a bucket whose largest member is above 768 AOs, where the
tridiagonalisation keeps fewer column-sum partials in LDS.

Tolerances: alone vs together |dE| <= 1e-9 Eh (the same kernels; the partial
count of the tridiagonalisation may differ, i.e. the summation order), the
same iteration count, status 0."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def test_bucket_neighbour_does_not_change_result(engine):
    rng = np.random.default_rng(61)
    mols = [workloads.water_cluster(117, workloads.liquid_radius(117), rng),
            workloads.water_cluster(130, workloads.liquid_radius(130), rng)]
    alone = []
    for num, xyz in mols:
        print(f"[gpu] bucket mix: alone nat={len(num)}", flush=True)
        e, it, st = engine.singlepoint_flat(np.array([len(num)]), num, xyz)
        assert st[0] == 0
        alone.append((float(e[0]), int(it[0])))
    mol_nat, num, xyz = workloads.flatten(mols)
    print("[gpu] bucket mix: together", flush=True)
    e, it, st = engine.singlepoint_flat(mol_nat, num, xyz)
    for m in range(2):
        assert st[m] == 0, (m, st[m])
        assert int(it[m]) == alone[m][1], (m, it[m], alone[m][1])
        assert abs(float(e[m]) - alone[m][0]) <= 1e-9, (m, e[m], alone[m][0])
