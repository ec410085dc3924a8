"""Pins of the oracle's analytic gradient by what the mathematics fixes: each
derivative equals the central finite difference of the function the oracle
already computes (and which the other oracle tests pin), and the total
gradient of a molecule is translation and rotation invariant.

* integral blocks: dS, dD, dQ of every shell-pair kind (s/p/d on H, C, N, O,
  S) against central differences of the contracted multipole block by the
  separation vector (h = 1e-5 bohr, |err| <= 1e-8);
* classical terms at fixed moments: repulsion, the electrostatics (ES2,
  AES2 with the CN-dependent multipole radii, ES3), D4 two-body and the
  three-body term, against central differences of their energies;
* the integral terms at a fixed random symmetric P, W and random potentials
  (H0 with its CN-dependent self energies, overlap, Mulliken charges,
  dipole and quadrupole moments with the operator shift);
* the full SCF gradient (accuracy 1e-4) against central differences of the
  converged energy (h = 1e-4 bohr, |err| <= 5e-8 Eh/bohr): this also fixes
  the energy-weighted density term and the stationarity argument behind it,
  on a neutral closed-shell organic with d shells and on a charged
  open-shell cluster with fractional occupations;
* sum_a g_a = 0 and sum_a R_a x g_a = 0; sigma = sum_a g_a (x) R_a is
  symmetric.

Planted in a copy of the oracle, each of these mistakes moved the full
gradient away from the finite differences by (organic, cation) Eh/bohr:
W term dropped 0.27, 0.27; Mulliken term dropped 0.025, 0.055; unshifted
dipole operator on the second atom 0.006, 0.010; self-energy CN chain
dropped 9e-4, 3e-4; multipole-radius CN chain dropped 3e-3, 1.2e-4; D4 CN
chain dropped 3.4e-5, 2.9e-6 -- all far above the 5e-8 tolerance."""
import numpy as np
import pytest

import oracle
import workloads


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


PAIRS = [(1, 0, 1, 0), (8, 1, 1, 0), (6, 1, 8, 1), (16, 2, 6, 1), (16, 2, 16, 2),
         (7, 0, 16, 2), (1, 0, 16, 2), (6, 0, 7, 1), (16, 1, 16, 0)]


@pytest.mark.parametrize("pair", PAIRS)
def test_integral_block_derivatives(orc, pair):
    zj, jsh, zi, ish = pair
    rng = np.random.default_rng(sum(pair))
    h = 1e-5
    for _ in range(3):
        vec = rng.standard_normal(3) * 1.5
        S, D, Q, dS, dD, dQ = orc.multipole_grad_cgto(zj, jsh, zi, ish, vec)
        S0, D0, Q0 = orc.multipole_cgto(zj, jsh, zi, ish, vec)
        assert np.allclose(S, S0, atol=1e-14) and np.allclose(D, D0, atol=1e-14)
        assert np.allclose(Q, Q0, atol=1e-14)
        for m in range(3):
            e = np.zeros(3)
            e[m] = h
            p = orc.multipole_cgto(zj, jsh, zi, ish, vec + e)
            q = orc.multipole_cgto(zj, jsh, zi, ish, vec - e)
            for ana, a, b in ((dS[m], p[0], q[0]), (dD[m], p[1], q[1]), (dQ[m], p[2], q[2])):
                assert np.max(np.abs(ana - (a - b) / (2 * h))) < 1e-8


def _fd(fun, xyz, h):
    g = np.zeros_like(xyz)
    for a in range(len(xyz)):
        for k in range(3):
            x = xyz.copy()
            x[a, k] += h
            ep = fun(x)
            x[a, k] -= 2 * h
            g[a, k] = (ep - fun(x)) / (2 * h)
    return g


@pytest.fixture(scope="module")
def organic():
    rng = np.random.default_rng(71)
    num, xyz = workloads.organic_molecule(6, rng)
    assert 16 in num or 7 in num
    return num, xyz


@pytest.mark.parametrize("which", ["repulsion", "coulomb", "d4_twobody", "d4_threebody"])
def test_classical_terms(orc, organic, which):
    num, xyz = organic
    nat, nsh, _ = orc.dims(num)
    rng = np.random.default_rng(72)
    qsh = rng.standard_normal(nsh) * 0.2
    dpat = rng.standard_normal((nat, 3)) * 0.1
    qpat = rng.standard_normal((nat, 6)) * 0.1
    e0, g = orc.classical_gradient(num, xyz, which, qsh, dpat, qpat)
    fd = _fd(lambda x: orc.classical_gradient(num, x, which, qsh, dpat, qpat)[0], xyz, 1e-5)
    scale = max(np.max(np.abs(g)), 1e-6)
    assert np.max(np.abs(g - fd)) < 1e-8 * max(1.0, scale), which
    assert np.max(np.abs(g)) > 0


def test_integral_terms(orc, organic):
    num, xyz = organic
    nat, nsh, nao = orc.dims(num)
    rng = np.random.default_rng(73)
    P = rng.standard_normal((nao, nao)) * 0.1
    P = P + P.T
    W = rng.standard_normal((nao, nao)) * 0.1
    W = W + W.T
    vsh = rng.standard_normal(nsh) * 0.1
    vdp = rng.standard_normal((nat, 3)) * 0.1
    vqp = rng.standard_normal((nat, 6)) * 0.1
    e0, g = orc.integral_gradient(num, xyz, P, W, vsh, vdp, vqp)
    fd = _fd(lambda x: orc.integral_gradient(num, x, P, W, vsh, vdp, vqp)[0], xyz, 1e-5)
    assert np.max(np.abs(g - fd)) < 5e-9


def _cases():
    rng = np.random.default_rng(74)
    num, xyz = workloads.organic_molecule(5, rng)
    wn, wx = workloads.water_cluster(3, workloads.liquid_radius(3), rng)
    return [("organic", num, xyz, 0.0, 0), ("water3_cation", wn, wx, 1.0, 1)]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_full_gradient_vs_finite_difference(orc, case):
    _, num, xyz, charge, uhf = case
    r = orc.singlepoint(num, xyz, charge=charge, uhf=uhf, accuracy=1e-4, gradient=True)
    assert r["rc"] == 0
    g = r["gradient"]
    fd = _fd(lambda x: orc.singlepoint(num, x, charge=charge, uhf=uhf, accuracy=1e-4)["energy"],
             xyz, 1e-4)
    assert np.max(np.abs(g - fd)) < 5e-8, np.max(np.abs(g - fd))
    # invariances of a molecule's energy
    assert np.max(np.abs(g.sum(0))) < 1e-10
    assert np.max(np.abs(np.cross(xyz, g).sum(0))) < 1e-9
    sig = r["sigma"]
    assert np.allclose(sig, np.einsum("ai,aj->ij", g, xyz), atol=1e-12)
    assert np.max(np.abs(sig - sig.T)) < 1e-9


def test_gradient_does_not_change_the_energy(orc, organic):
    num, xyz = organic
    a = orc.singlepoint(num, xyz)
    b = orc.singlepoint(num, xyz, gradient=True)
    assert a["energy"] == b["energy"] and a["iterations"] == b["iterations"]
