"""Pins of the oracle's SCF pieces and of the whole singlepoint.

* generalized eigensolver vs scipy.linalg.eigh (LAPACK dsygvd) and the
  closed-form A = 2B case of the reference's own demo (src/lapack/sygvd.cu:183-186);
* Fermi filling invariants (electron count, kT -> 0 aufbau, entropy sign,
  Fermi-Dirac form);
* electrostatics: potential = dE/dq by finite differences (ES2 + AES2 + ES3),
  gamma limits, far-field charge-dipole law, symmetry of the interaction
  matrices;
* neighbour list vs a brute-force loop (bit for bit), CN and repulsion
  special cases;
* singlepoint: charge conservation, occupation sum, SCF fixed point
  (independent re-diagonalisation with scipy from the converged charges is
  not needed: the mixer residual is checked), size consistency, rotation,
  translation and permutation invariance.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle
import workloads

KT300 = 300 * oracle.KB_EH


# ------------------------------------------------------------ eigensolver
def test_sygvd_closed_form_reference_demo():
    A = np.array([[4.0, 1.0], [1.0, 3.0]])
    B = np.array([[2.0, 0.5], [0.5, 1.5]])
    rc, w, Cm = oracle.sygvd(A, B)
    assert rc == 0
    assert np.allclose(w, [2.0, 2.0], atol=1e-14)
    assert np.allclose(Cm.T @ B @ Cm, np.eye(2), atol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 5, 17, 48, 96])
def test_sygvd_vs_scipy_random(n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    S = X @ X.T / n + np.eye(n)
    F = rng.standard_normal((n, n))
    F = 0.5 * (F + F.T)
    rc, w, Cm = oracle.sygvd(F, S)
    assert rc == 0
    wr = sl.eigh(F, S, eigvals_only=True)
    assert np.allclose(w, wr, atol=1e-11 * max(1.0, np.abs(wr).max()))
    assert np.allclose(Cm.T @ S @ Cm, np.eye(n), atol=1e-10)
    assert np.allclose(F @ Cm, S @ Cm * w, atol=1e-10)


def test_sygvd_vs_scipy_molecule(orc):
    num, xyz = workloads.config0()[0]
    S, D, Q, H = orc.hamiltonian(num, xyz)
    rc, w, Cm = oracle.sygvd(H, S)
    wr, Cr = sl.eigh(H, S)
    assert np.allclose(w, wr, atol=1e-12)
    nocc = 16
    P = Cm[:, :nocc] @ Cm[:, :nocc].T
    Pr = Cr[:, :nocc] @ Cr[:, :nocc].T
    assert np.allclose(P, Pr, atol=1e-10)


def test_sygvd_rejects_indefinite_overlap():
    S = np.array([[1.0, 2.0], [2.0, 1.0]])
    rc, _, _ = oracle.sygvd(np.eye(2), S)
    assert rc == -4  # TBLITE_AMD_ERR_CHOLESKY


# ------------------------------------------------------------ occupation
def test_fermi_invariants():
    rng = np.random.default_rng(3)
    emo = np.sort(rng.uniform(-1.0, 0.5, 40))
    for nel in (7.0, 10.0, 10.5, 19.0):
        homo, occ, ef = oracle.fermi_filling(nel, KT300, emo)
        assert abs(occ.sum() - np.floor(nel) - (1 if nel % 1 > 0.5 else 0)) < 2e-8 or abs(occ.sum() - nel) < 2e-8
        assert np.all((occ >= 0) & (occ <= 1))
        assert oracle.entropy(occ, KT300) <= 0.0
    # kT -> 0: aufbau
    homo, occ, ef = oracle.fermi_filling(10.0, 1e-9, emo)
    assert np.allclose(occ, np.r_[np.ones(10), np.zeros(30)], atol=1e-12)
    # Fermi-Dirac form with a near degeneracy at the Fermi level
    e2 = np.array([-1.0, -0.5, -0.30, -0.2999, 0.1])
    homo, occ, ef2 = oracle.fermi_filling(3.0, 0.01, e2)
    ef = 2 * ef2  # get_fermi_filling returns half the Fermi energy (upstream convention)
    fd = 1.0 / (np.exp((e2 - ef) / 0.01) + 1.0)
    assert abs(occ.sum() - 3.0) < 2e-8
    assert np.allclose(occ, fd, atol=1e-6)


# --------------------------------------------------------- electrostatics
def test_coulomb_potential_is_energy_gradient(orc):
    rng = np.random.default_rng(5)
    num, xyz = workloads.water_cluster(3, 3.0, rng)
    nat, nsh, _ = orc.dims(num)
    qsh = rng.normal(0, 0.3, nsh)
    dp = rng.normal(0, 0.1, (nat, 3))
    qp = rng.normal(0, 0.1, (nat, 6))
    e0, vsh, vat, vdp, vqp = orc.coulomb_eval(num, xyz, qsh, dp, qp)
    sh2at = np.repeat(np.arange(nat), [2 if z == 8 else 1 for z in num])
    h = 1e-6

    def E(a, b, c):
        return orc.coulomb_eval(num, xyz, a, b, c)[0].sum()

    g = np.array([(E(qsh + h * np.eye(nsh)[i], dp, qp) - E(qsh - h * np.eye(nsh)[i], dp, qp)) / (2 * h)
                  for i in range(nsh)])
    assert np.allclose(g, vsh + vat[sh2at], atol=1e-8)
    for i in range(nat):
        for k in range(3):
            d1 = dp.copy(); d1[i, k] += h
            d2 = dp.copy(); d2[i, k] -= h
            assert abs((E(qsh, d1, qp) - E(qsh, d2, qp)) / (2 * h) - vdp[i, k]) < 1e-8
        for k in range(6):
            q1 = qp.copy(); q1[i, k] += h
            q2 = qp.copy(); q2[i, k] -= h
            # off-diagonal quadrupole components enter twice (xy and yx)
            assert abs((E(qsh, dp, q1) - E(qsh, dp, q2)) / (2 * h) - vqp[i, k]) < 1e-8


def test_gamma_limits_and_symmetry(orc):
    num = np.array([8, 1])
    far = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 500.0]])
    amat, sd, dd, sq, mrad = orc.coulomb_matrices(num, far)
    assert np.allclose(amat, amat.T, atol=1e-15)
    # far apart: gamma -> 1/R, with a relative correction of order 1/(2 R^2 eta^2)
    assert abs(amat[0, 2] * 500.0 - 1.0) < 1e-4 and abs(amat[1, 2] * 500.0 - 1.0) < 1e-4
    assert amat[0, 2] < 1 / 500.0
    # on-site: the shell hardness (diagonal > off-diagonal between atoms)
    assert amat[0, 0] > amat[0, 2] and amat[2, 2] > 0.3
    # antisymmetric charge-dipole, symmetric dipole-dipole and charge-quadrupole kernels
    assert np.allclose(sd, -sd.transpose(1, 0, 2), atol=1e-15)
    assert np.allclose(sq, sq.transpose(1, 0, 2), atol=1e-15)
    assert np.allclose(dd, dd.transpose(2, 3, 0, 1), atol=1e-15)


def test_charge_dipole_far_field(orc):
    """At large separation the damping vanishes and the AES2 charge-dipole
    energy is the classical q mu.(R_q - R_mu)/R^3 (note tblite's dipole sign:
    dpat is the electronic moment with the electron charge folded in)."""
    num = np.array([1, 1])
    R = 600.0  # damping 1/(1 + 6 (rbar/R)^3) ~ 1 - 1e-6 here
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, R]])
    nat, nsh, _ = orc.dims(num)
    qsh = np.array([0.5, 0.0])
    dp = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.3]])
    qp = np.zeros((2, 6))
    e_all = orc.coulomb_eval(num, xyz, qsh, dp, qp)[0].sum()
    e_q = orc.coulomb_eval(num, xyz, qsh, 0 * dp, qp)[0].sum()
    kernel = 5.563889e-2 * 0.3**2  # H on-site dipole kernel
    e_qd = e_all - e_q - kernel
    classical = 0.5 * 0.3 * (0.0 - R) / R**3
    assert abs(e_qd / classical - 1.0) < 1e-5


# -------------------------------------------------- list / CN / repulsion
def test_adjlist_bruteforce(orc):
    num, xyz = workloads.config0()[0]
    cut = orc.cutoff(num)
    inl, nnl, nlat = orc.adjlist(xyz, cut)
    ref = []
    for i in range(len(num)):
        for j in range(i + 1):
            d = xyz[i] - xyz[j]
            r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
            if not (r2 < np.finfo(float).eps or r2 > cut * cut):
                ref.append((i, j))
    got = [(i, int(nlat[inl[i] + k])) for i in range(len(num)) for k in range(nnl[i])]
    assert got == ref
    # a tight cutoff drops pairs, monotonically
    _, nnl2, _ = orc.adjlist(xyz, 3.0)
    assert nnl2.sum() < nnl.sum()


def test_cutoff_closed_form(orc):
    # water: smallest primitive exponent is the most diffuse H 1s primitive
    a_h = orc.cgto(1, 0)[1].min()
    a_o = min(orc.cgto(8, 0)[1].min(), orc.cgto(8, 1)[1].min())
    assert abs(orc.cutoff([8, 1, 1]) - min(np.sqrt(2 * 25.0 / min(a_h, a_o)), 40.0)) < 1e-14


def test_cn_special_cases(orc):
    # two H at exactly the sum of their covalent radii: first counting function = 1/2
    rcov = 0.32 * 4 / 3 / workloads.AUTOAA
    r = 2 * rcov
    cn = orc.cn([1, 1], [[0, 0, 0], [0, 0, r]])
    second = 1.0 / (1.0 + np.exp(-20.0 * ((2 * rcov + 2.0) / r - 1.0)))
    assert np.allclose(cn, 0.5 * second, atol=1e-15)
    assert np.all(orc.cn([1, 1], [[0, 0, 0], [0, 0, 30.0]]) == 0.0)  # beyond the 25 bohr cutoff


def test_repulsion_pair_exponents(orc):
    # H-H uses distance exponent 1, O-H 1.5 (GFN2 light-pair rule)
    r = 2.1
    e_hh = orc.repulsion([1, 1], [[0, 0, 0], [0, 0, r]]).sum()
    z, a = 1.105388, 2.213717
    assert abs(e_hh - z * z * np.exp(-a * r) / r) < 1e-15
    e_oh = orc.repulsion([8, 1], [[0, 0, 0], [0, 0, r]]).sum()
    ref = 5.784415 * 1.105388 * np.exp(-np.sqrt(2.165712 * 2.213717) * r**1.5) / r
    assert abs(e_oh - ref) < 1e-15


# ------------------------------------------------------------ singlepoint
def _rot(xyz, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return xyz @ R.T


def test_singlepoint_invariants(orc):
    num, xyz = workloads.config0()[0]
    r = orc.singlepoint(num, xyz)
    assert r["converged"] and r["rc"] == 0
    assert abs(r["qsh"].sum()) < 1e-10
    assert abs(r["qat"].sum()) < 1e-10
    assert abs(r["focc"].sum() - 64.0) < 1e-8
    assert abs(r["energy"] - (r["erep"] + r["edisp3"] + r["eelec"])) < 1e-12
    assert np.all(np.diff(r["emo"]) >= 0)
    # rotation / translation / permutation invariance
    r2 = orc.singlepoint(num, _rot(xyz, 1) + 2.0)
    assert abs(r2["energy"] - r["energy"]) < 1e-7
    perm = np.random.default_rng(2).permutation(len(num))
    r3 = orc.singlepoint(num[perm], xyz[perm])
    assert abs(r3["energy"] - r["energy"]) < 1e-7
    assert np.allclose(np.sort(r3["qat"]), np.sort(r["qat"]), atol=1e-5)


def test_size_consistency(orc):
    rng = np.random.default_rng(9)
    a = workloads.water_cluster(1, 0.5, rng)
    b = workloads.organic_molecule(3, rng)
    ea = orc.singlepoint(*a)["energy"]
    eb = orc.singlepoint(*b)["energy"]
    num = np.r_[a[0], b[0]]
    xyz = np.r_[a[1], b[1] + np.array([0.0, 0.0, 2000.0])]
    eab = orc.singlepoint(num, xyz, accuracy=0.01)
    ea1 = orc.singlepoint(*a, accuracy=0.01)["energy"]
    eb1 = orc.singlepoint(*b, accuracy=0.01)["energy"]
    assert eab["converged"]
    assert abs(eab["energy"] - (ea1 + eb1)) < 1e-7
    assert abs(ea1 - ea) < 1e-5 and abs(eb1 - eb) < 1e-5


def test_charged_and_open_shell(orc):
    num, xyz = workloads.water_cluster(1, 0.5, np.random.default_rng(4))
    cation = orc.singlepoint(num, xyz, charge=1.0, uhf=1)
    assert cation["converged"]
    assert abs(cation["qat"].sum() - 1.0) < 1e-9
    assert abs(cation["focc"].sum() - 7.0) < 1e-8
    neutral = orc.singlepoint(num, xyz)
    assert cation["energy"] > neutral["energy"]


def test_scf_fixed_point(orc):
    """At convergence the mixer residual is below the threshold: charges
    going in and coming out of the last diagonalisation agree."""
    num, xyz = workloads.water_cluster(4, 3.2, np.random.default_rng(6))
    r = orc.singlepoint(num, xyz)
    assert r["converged"]
    r_tight = orc.singlepoint(num, xyz, accuracy=0.01)
    assert abs(r_tight["energy"] - r["energy"]) < 2e-6
    assert np.max(np.abs(r_tight["qsh"] - r["qsh"])) < 1e-4
