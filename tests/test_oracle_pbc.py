"""Pins of the oracle's periodic (Gamma-point) terms by what the mathematics
fixes, independent of the formulas the oracle uses:

- the Ewald splitting parameter drops out of every Coulomb matrix element
  (ES2, AES2) and of the SCF energy;
- a 2 x 1 x 1 supercell reproduces the unit cell: per-atom CN, repulsion, D4
  CN and three-body energies are equal for both copies, and the supercell's
  interaction matrices summed over the two copies of an atom equal the unit
  cell's (this holds for the reciprocal sum only with the right 4 pi / V
  normalisation and self terms);
- the SCF energy is invariant under a rigid translation, a rotation of atoms
  and lattice together, moving one atom by a lattice vector, and a change of
  lattice basis (a_1 -> a_1 + a_2);
- a molecule in a cubic box of edge L tends to the molecular result, with
  the difference falling like 1/L^3 (dipole lattice sums);
- the neighbour list holds exactly the (atom, image) pairs that a brute-force
  search over a larger set of translations finds within the cutoff.

The reference carries the periodic branches only as comments
(src/adjlist.cu:26-30, src/h0.cu:359, src/coulomb.cu:94-96, :212-213,
:554-557); tblite itself is not available here, so the absolute periodic
energies are parity-unpinned against upstream (the Ewald split, the switch of
the ES2 kernel to 1/r and the image cutoffs are this oracle's definitions).
"""
import numpy as np
import pytest

import workloads


@pytest.fixture(scope="module")
def cell():
    return workloads.water_box(4, np.random.default_rng(11))


def _supercell(num, xyz, lat):
    num2 = np.concatenate([num, num])
    xyz2 = np.concatenate([xyz, xyz + lat[0]])
    lat2 = lat.copy()
    lat2[0] = 2.0 * lat[0]
    return num2, xyz2, lat2


def test_ewald_alpha_independence_of_matrices(orc, cell):
    num, xyz, lat = cell
    a = orc.pbc_terms(num, xyz, lat)
    try:
        for s in (0.7, 1.4):
            orc.set_ewald_alpha_scale(s)
            b = orc.pbc_terms(num, xyz, lat)
            for k in ("amat", "sd", "dd", "sq"):
                assert np.abs(a[k] - b[k]).max() < 1e-12 * max(1.0, np.abs(a[k]).max()), (k, s)
    finally:
        orc.set_ewald_alpha_scale(1.0)
    # the ES2 matrix is symmetric, dd symmetric in (atom, component) pairs,
    # sd odd under exchange of the atoms
    assert np.abs(a["amat"] - a["amat"].T).max() < 1e-12
    nat = len(num)
    dd = a["dd"].reshape(3 * nat, 3 * nat)
    assert np.abs(dd - dd.T).max() < 1e-12
    assert np.abs(a["sd"] + a["sd"].transpose(1, 0, 2)).max() < 1e-12


def test_supercell_reproduces_unit_cell(orc, cell):
    num, xyz, lat = cell
    nat = len(num)
    nsh = orc.dims(num)[1]
    a = orc.pbc_terms(num, xyz, lat)
    b = orc.pbc_terms(*_supercell(num, xyz, lat))
    for k in ("cn", "erep", "cn4", "e3", "mrad"):
        assert np.abs(b[k][:nat] - a[k]).max() < 1e-12, k
        assert np.abs(b[k][nat:] - a[k]).max() < 1e-12, k
    # matrices: the two copies of atom j together carry the unit cell's images
    amat2 = b["amat"][:nsh, :nsh] + b["amat"][:nsh, nsh:]
    assert np.abs(amat2 - a["amat"]).max() < 1e-11
    for k in ("sd", "sq"):
        assert np.abs(b[k][:nat, :nat] + b[k][:nat, nat:] - a[k]).max() < 1e-11, k
    dd2 = b["dd"][:nat, :, :nat, :] + b["dd"][:nat, :, nat:, :]
    assert np.abs(dd2 - a["dd"]).max() < 1e-11
    dm2 = b["dmat"][:nat, :nat] + b["dmat"][:nat, nat:]
    assert np.abs(dm2 - a["dmat"]).max() < 1e-12


def test_planted_self_term_fails_supercell(orc, cell):
    """What the supercell test catches: a unit-cell ES2 matrix without the
    -2 alpha / sqrt(pi) self term of the Ewald split misses the copy sum by
    far more than its tolerance."""
    num, xyz, lat = cell
    nsh = orc.dims(num)[1]
    a = orc.pbc_terms(num, xyz, lat)
    b = orc.pbc_terms(*_supercell(num, xyz, lat))
    amat2 = b["amat"][:nsh, :nsh] + b["amat"][:nsh, nsh:]
    assert np.abs(amat2 - a["amat"]).max() < 1e-11
    bad = a["amat"] - np.diag(np.full(nsh, 2 * orc.pbc_tables(lat, xyz)["alpha"] / np.sqrt(np.pi)))
    assert np.abs(amat2 - bad).max() > 0.05


def test_adjlist_bruteforce(orc, cell):
    num, xyz, lat = cell
    cutoff = orc.cutoff(num, 1.0)
    inl, nnl, nlat, nltr = orc.adjlist_pbc(xyz, lat, cutoff)
    tr = orc.pbc_tables(lat, xyz)["trans"]
    got = set()
    for i in range(len(num)):
        for k in range(inl[i], inl[i] + nnl[i]):
            got.add((i, int(nlat[k]), tuple(np.round(tr[nltr[k]], 9))))
    # brute force over a wider integer range
    want = set()
    n = int(np.ceil(cutoff / np.min(np.abs(np.diag(lat))))) + 3
    rng = np.arange(-n, n + 1)
    T = np.array([[a, b, c] for a in rng for b in rng for c in rng], dtype=float) @ lat
    for i in range(len(num)):
        for j in range(i + 1):
            d = xyz[i] - xyz[j] - T
            r2 = (d * d).sum(1)
            for t in np.nonzero((r2 >= 2.220446049250313e-16) & (r2 <= cutoff * cutoff))[0]:
                want.add((i, j, tuple(np.round(T[t], 9))))
    assert got == want
    # every image appears with both signs for the atom's own images
    assert len(got) > len(num) * (len(num) + 1) // 2


def _energy(orc, num, xyz, lat, acc=0.01):
    r = orc.singlepoint(num, xyz, lattice=lat, accuracy=acc)
    assert r["rc"] == 0
    return r


def test_scf_invariances(orc, cell):
    num, xyz, lat = cell
    e0 = _energy(orc, num, xyz, lat)
    assert abs(e0["qat"].sum()) < 1e-9
    # rigid translation
    assert abs(_energy(orc, num, xyz + np.array([0.37, -1.1, 2.3]), lat)["energy"] - e0["energy"]) < 1e-8
    # rotation of atoms and lattice together
    rng = np.random.default_rng(3)
    q = rng.standard_normal((3, 3))
    R, _ = np.linalg.qr(q)
    assert abs(_energy(orc, num, xyz @ R.T, lat @ R.T)["energy"] - e0["energy"]) < 1e-8
    # one atom moved by a lattice vector
    x2 = xyz.copy()
    x2[4] += lat[1]
    assert abs(_energy(orc, num, x2, lat)["energy"] - e0["energy"]) < 1e-8
    # another basis of the same lattice
    lat2 = lat.copy()
    lat2[0] = lat[0] + lat[1]
    assert abs(_energy(orc, num, xyz, lat2)["energy"] - e0["energy"]) < 1e-8
    # Ewald splitting parameter
    try:
        orc.set_ewald_alpha_scale(1.3)
        assert abs(_energy(orc, num, xyz, lat)["energy"] - e0["energy"]) < 1e-9
    finally:
        orc.set_ewald_alpha_scale(1.0)


def test_large_box_tends_to_molecule(orc):
    num, xyz = workloads.config0()[0]
    num, xyz = num[:6], xyz[:6]  # a water dimer: dipolar, neutral
    mol = orc.singlepoint(num, xyz, accuracy=0.01)["energy"]
    d = []
    for L in (60.0, 120.0):
        d.append(_energy(orc, num, xyz, np.eye(3) * L)["energy"] - mol)
    assert abs(d[1]) < 2e-6
    # dipole lattice sums fall like 1/L^3: a factor of 8 between the boxes
    assert 5.0 < d[0] / d[1] < 11.0, d


def test_cell_gradient_not_built(orc, cell):
    num, xyz, lat = cell
    r = orc.singlepoint(num, xyz, lattice=lat, gradient=True)
    assert r["rc"] == -1
