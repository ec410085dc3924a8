"""Pins of the oracle functions that the reference does not contain (restated
from upstream tblite's published algorithms), by what the mathematics fixes
rather than by the oracle's own formulas:

* the modified Broyden mixer (Johnson's scheme, tblite's broyden `next`):
  its first step is linear mixing; on an affine map its iterates are those of
  Anderson acceleration with full history, which on a linear residual spans
  the Krylov space and terminates after n + 1 steps in exact arithmetic
  (GMRES equivalence) when the regularisation weight omega0 is zero; in one
  dimension its second step is the secant step, moved by omega0 only at
  order omega0^2 / (omega0^2 + omega_1^2);
* hscale (the H0 shell-pair scaling of GFN2): symmetric under exchange of the
  two shells, kshell(l, l) for a homonuclear pair of one shell, a Slater
  exponent factor 2 sqrt(zi zj) / (zi + zj) <= 1 (equality iff zi = zj,
  invariant under zj -> zi^2 / zj) and an electronegativity term quadratic
  in the difference;
* the Fock build, the Mulliken analysis and the energy expression together:
  the converged SCF must be a stationary point of the Mermin free energy
  F(T) = E - T S, so dF/dT = -S (envelope theorem).  A Fock term that is not
  the derivative of its energy term (a wrong sign or factor in any potential)
  moves the fixed point off the stationary point and breaks this relation
  (a planted sign error in the dipole potential gave relative deviations of
  0.19 and 5e-4 on the two molecules below, against ~1e-7 here);
* the Mulliken atomic dipoles: sum_A (q_A R_A + mu_A) equals
  sum_A n0_A R_A - tr(P r) with the origin-centred dipole operator, an
  identity of the partition that fixes the sign and the operator centre.
"""
import numpy as np
import pytest

import oracle
import tblite_amd
import workloads

KB = oracle.KB_EH


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


# ---------------------------------------------------------------- Broyden
def _affine(n, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, (n, n)), rng.standard_normal(n), rng.standard_normal(n)


def test_broyden_first_step_is_linear_mixing(orc):
    A, b, q0 = _affine(4, 1)
    for damp in (0.4, 0.1, 1.0):
        rc, qs, res = orc.broyden_linear(A, b, q0, damping=damp, nstep=1)
        assert rc == 0
        F0 = A @ q0 + b - q0
        assert np.array_equal(qs[1], q0 + damp * F0)


def test_broyden_secant_step_in_one_dimension(orc):
    """q2 is the root of the secant through (q0, F0) and (q1, F1); with the
    SCF's omega0 = 0.01 it moves by (1 - s) (alpha F1 + F1 dq / dF), where
    s = w1^2 / (omega0^2 + w1^2) >= 1 / (1 + 1e-4) since the weights are >= 1."""
    for a, b0, q0 in ((0.3, 1.0, 0.0), (-0.7, 2.0, 1.5), (0.9, -0.5, 3.0)):
        A = np.array([[a]])
        b = np.array([b0])
        qstar = b0 / (1.0 - a)
        rc, qs, res = orc.broyden_linear(A, b, np.array([q0]), nstep=2, omega0=0.0)
        assert rc == 0
        F = lambda q: a * q + b0 - q
        q1 = qs[1, 0]
        secant = q1 - F(q1) * (q1 - q0) / (F(q1) - F(q0))
        assert abs(secant - qstar) < 1e-12 * max(1.0, abs(qstar))
        assert abs(qs[2, 0] - qstar) < 1e-12 * max(1.0, abs(qstar))
        rc, qs, res = orc.broyden_linear(A, b, np.array([q0]), nstep=2, omega0=0.01)
        bound = (1e-4 / (1.0 + 1e-4)) * abs(0.4 * F(q1) + F(q1) * (q1 - q0) / (F(q1) - F(q0)))
        assert abs(qs[2, 0] - secant) <= bound * (1 + 1e-9) + 1e-15


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6])
def test_broyden_terminates_on_linear_problems(orc, n):
    """omega0 = 0: Anderson acceleration with full history on a linear
    residual reaches the solution after n + 1 mixer steps (one linear step
    plus n Krylov directions)."""
    for seed in range(3):
        A, b, q0 = _affine(n, 100 * n + seed)
        rc, qs, res = orc.broyden_linear(A, b, q0, nstep=n + 1, omega0=0.0)
        assert rc == 0
        assert res[n + 1] < 1e-8 * res[0], (n, seed, res)
        qstar = np.linalg.solve(np.eye(n) - A, b)
        assert np.max(np.abs(qs[n + 1] - qstar)) < 1e-8 * max(1.0, np.max(np.abs(qstar)))


def test_broyden_converges_with_scf_regularisation(orc):
    """omega0 = 0.01 (the SCF's value): the regularisation delays the
    termination by a few steps, the residual still falls below 1e-10 of its
    start within n + 4 steps.  (Iterating on past convergence amplifies
    rounding noise, as in any unsafeguarded Anderson scheme; the SCF stops
    at its convergence test long before that.)"""
    for n in (2, 5, 8, 12):
        A, b, q0 = _affine(n, 7 * n)
        rc, qs, res = orc.broyden_linear(A, b, q0, nstep=n + 4)
        assert rc == 0
        assert res.min() < 1e-10 * res[0], (n, res)


# ----------------------------------------------------------------- hscale
def test_hscale_symmetric_and_homonuclear(orc):
    p = tblite_amd.default_param()
    for ei in range(p.nelem):
        for ej in range(p.nelem):
            for ish in range(p.elem[ei].nsh):
                for jsh in range(p.elem[ej].nsh):
                    h = orc.hscale(ei, ish, ej, jsh)
                    assert h == pytest.approx(orc.hscale(ej, jsh, ei, ish), rel=1e-15)
        for ish in range(p.elem[ei].nsh):
            l = p.elem[ei].ang[ish]
            assert orc.hscale(ei, ish, ei, ish) == pytest.approx(p.kshell[l][l], rel=1e-15)


def _two_element_param(zeta_b, en_b):
    """Default parameters with C's 2s Slater exponent and electronegativity
    replaced, so that the H 1s - C 2s pair isolates the exponent and EN terms."""
    p = tblite_amd.default_param()
    p.elem[1].slater[0] = zeta_b
    p.elem[1].en = en_b
    return p


def test_hscale_exponent_and_electronegativity_terms():
    p0 = tblite_amd.default_param()
    zh, enh, k = p0.elem[0].slater[0], p0.elem[0].en, p0.kshell[0][0]

    def h(zb, enb):
        return oracle.Oracle(_two_element_param(zb, enb)).hscale(0, 0, 1, 0)

    # equal exponents and electronegativities: the bare shell-pair constant
    assert h(zh, enh) == pytest.approx(k, rel=1e-14)
    # the exponent factor: at most 1, smaller the further apart, and invariant
    # under zb -> zh^2 / zb (the factor depends on zb / zh + zh / zb only)
    f = [h(zh * t, enh) / k for t in (1.0, 1.3, 2.0, 3.5)]
    assert all(x <= 1.0 + 1e-15 for x in f)
    assert all(f[i + 1] < f[i] for i in range(3))
    for t in (1.3, 2.0, 3.5):
        assert h(zh * t, enh) == pytest.approx(h(zh / t, enh), rel=1e-13)
    # the EN term: h - k grows as the square of the difference, with no
    # linear part (same for either sign of the difference)
    d1 = h(zh, enh + 0.5) - k
    d2 = h(zh, enh + 1.0) - k
    d3 = h(zh, enh - 1.0) - k
    assert d1 > 0.0
    assert d2 == pytest.approx(4.0 * d1, rel=1e-12)
    assert d3 == pytest.approx(d2, rel=1e-12)


# --------------------------------------- Fock / Mulliken / energy together
def _entropy(focc):
    """Electronic entropy / kB of a closed-shell occupation (f/2 per spin)."""
    f = focc / 2.0
    m = (f > 1e-14) & (f < 1.0 - 1e-14)
    return -2.0 * np.sum(f[m] * np.log(f[m]) + (1.0 - f[m]) * np.log(1.0 - f[m]))


@pytest.mark.parametrize("which", ["water3", "organic6"])
def test_free_energy_is_stationary(orc, which):
    rng = np.random.default_rng(5)
    w3 = workloads.water_cluster(3, workloads.liquid_radius(3), rng)
    o6 = workloads.organic_molecule(6, rng)
    num, xyz = {"water3": w3, "organic6": o6}[which]
    T, dT = 20000.0, 20.0
    r0 = orc.singlepoint(num, xyz, accuracy=1e-5, temperature=T)
    rp = orc.singlepoint(num, xyz, accuracy=1e-5, temperature=T + dT)
    rm = orc.singlepoint(num, xyz, accuracy=1e-5, temperature=T - dT)
    assert r0["converged"] and rp["converged"] and rm["converged"]
    S = KB * _entropy(r0["focc"])
    dFdT = (rp["energy"] - rm["energy"]) / (2.0 * dT)
    assert S > 1e-7
    assert abs(dFdT + S) < 1e-5 * S, (which, dFdT, -S)


def test_mulliken_dipole_partition(orc):
    rng = np.random.default_rng(8)
    num, xyz = workloads.organic_molecule(9, rng)
    r = orc.singlepoint(num, xyz, wavefunction=True)
    S, D, Q, H = orc.hamiltonian(num, xyz)
    P = r["density"]
    p = tblite_amd.default_param()
    z2e = {p.elem[e].z: e for e in range(p.nelem)}
    n0 = np.array([sum(p.elem[z2e[z]].refocc[s] for s in range(p.elem[z2e[z]].nsh)) for z in num])
    # AO -> atom map
    ao_atom = np.concatenate([[a] * sum(2 * p.elem[z2e[z]].ang[s] + 1
                                        for s in range(p.elem[z2e[z]].nsh))
                              for a, z in enumerate(num)])
    # D[i, j, k] = <j| (r - R_atom(i))_k |i>, S[i, j] = <j|i>
    Rorig = D + xyz[ao_atom][:, None, :] * S[:, :, None]
    elec = np.einsum("ij,ijk->k", P, Rorig)
    lhs = (r["qat"][:, None] * xyz).sum(0) + r["dpat"].sum(0)
    rhs = (n0[:, None] * xyz).sum(0) - elec
    assert np.max(np.abs(lhs - rhs)) < 1e-9, (lhs, rhs)
    # the molecule is polar: the identity is not 0 = 0
    assert np.linalg.norm(lhs) > 1e-2
