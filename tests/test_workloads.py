"""Seeded workload generators: determinism and the geometric rules of the
synthetic inputs (rigid waters, minimum distances, valences, sizes)."""
import numpy as np

import workloads

A = workloads.AUTOAA


def test_deterministic():
    a = workloads.config0()[0]
    b = workloads.config0()[0]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_rigid_water_geometry():
    num, xyz = workloads.config0()[0]
    assert len(num) == 24 and list(num[:3]) == [8, 1, 1]
    x = xyz * A
    for w in range(8):
        o, h1, h2 = x[3 * w], x[3 * w + 1], x[3 * w + 2]
        assert abs(np.linalg.norm(h1 - o) - 0.9572) < 1e-12
        assert abs(np.linalg.norm(h2 - o) - 0.9572) < 1e-12
        ang = np.degrees(np.arccos((h1 - o) @ (h2 - o) / 0.9572**2))
        assert abs(ang - 104.52) < 1e-9
        assert np.linalg.norm(o) <= 4.5 + 1e-12
    oxy = x[::3]
    d = np.linalg.norm(oxy[:, None] - oxy[None], axis=-1) + np.eye(8) * 10
    assert d.min() >= 2.6


def test_config1_sizes():
    mols = workloads.config1(count=200)
    n = np.array([len(m[0]) // 3 for m in mols])
    assert n.min() >= 2 and n.max() <= 16


def test_organic_valences_and_closed_shell():
    rng = np.random.default_rng(1)
    for _ in range(20):
        num, xyz = workloads.organic_molecule(int(rng.integers(4, 20)), rng)
        x = xyz * A
        d = np.linalg.norm(x[:, None] - x[None], axis=-1) + np.eye(len(num)) * 10
        # bonds are 1.50 (heavy-heavy) or 1.09 (X-H); every other pair is >= 1.2
        bonded = np.isclose(d, 1.5) | np.isclose(d, 1.09)
        assert d[~bonded].min() >= 1.2 - 1e-12
        nel = sum({1: 1, 6: 4, 7: 5, 8: 6, 16: 6}[int(z)] for z in num)
        assert nel % 2 == 0
        heavy = np.where(num != 1)[0]
        # every H sits at 1.09 A from one heavy atom
        for h in np.where(num == 1)[0]:
            assert np.isclose(d[h, heavy].min(), 1.09)
