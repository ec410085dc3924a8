"""Pins of the oracle's D4 dispersion.

Each check is fixed by something other than the oracle's own formula:

* Casimir-Polder reference C6: the trapezoid rule on the frequency grid is
  exact for a constant and a linear integrand (the integral is then the grid's
  span and half its squared span), and a single-pole alpha(i w) pair
  integrates to the London formula 3/2 a_A a_B w_A w_B / (w_A + w_B) up to the
  grid's quadrature error;
* reference weights: they sum to one at zero charge (the charge scale is 1
  there when the reference charges are zero), a CN far from all references
  falls back to the highest reference, and dg/dq equals the finite difference
  of g in q;
* the D4 coordination number of a dimer at the sum of covalent radii is half
  of the electronegativity factor (erf(0) = 0), and zero beyond the cutoff;
* two-body: potential = dE2/dq by finite differences of the whole energy, the
  dimer's undamped limit -(s6 C6 / r^6 + s8 C8 / r^8), pair additivity at large
  separation;
* three-body: equilateral (3 cos^3 60 + 1 = 11/8) and collinear (3 * 1 * 1 *
  (-1) + 1 = -2) angular factors in closed form, zero beyond the cutoff;
* singlepoint: energy = repulsion + three-body + electronic energy, and a
  parameter set with dispersion switched off gives back the energy without it
  (rotation invariance and size consistency of the D4 terms).
"""
import numpy as np
import pytest

import oracle
import tblite_amd
import workloads

FREQ = np.array([0.000001, 0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0,
                 1.2, 1.4, 1.6, 1.8, 2.0, 2.5, 3.0, 4.0, 5.0, 7.5, 10.0])


def _param_with_alpha(fa, fb):
    """Default parameters with H reference 0 set to alpha = fa(w) and C
    reference 0 set to alpha = fb(w)."""
    p = tblite_amd.default_param()
    for k, w in enumerate(FREQ):
        p.d4.elem[0].alpha[0][k] = fa(w)
        p.d4.elem[1].alpha[0][k] = fb(w)
    return p


def test_c6_quadrature_exact_for_constant_and_linear():
    o = oracle.Oracle(_param_with_alpha(lambda w: 1.0, lambda w: 1.0))
    span = FREQ[-1] - FREQ[0]
    assert o.d4_c6ref(1, 0, 6, 0) == pytest.approx(3.0 / np.pi * span, rel=1e-14)
    o = oracle.Oracle(_param_with_alpha(lambda w: 1.0, lambda w: w))
    assert o.d4_c6ref(1, 0, 6, 0) == pytest.approx(3.0 / np.pi * 0.5 * (FREQ[-1] ** 2 - FREQ[0] ** 2),
                                                   rel=1e-14)
    # symmetric in the pair
    assert o.d4_c6ref(6, 0, 1, 0) == pytest.approx(o.d4_c6ref(1, 0, 6, 0), rel=1e-15)


@pytest.mark.parametrize("aa,wa,ab,wb", [(4.5, 0.428, 4.5, 0.428), (12.0, 0.431, 5.3, 0.74),
                                          (2.75, 0.53, 19.6, 0.465)])
def test_c6_single_pole_london_formula(aa, wa, ab, wb):
    o = oracle.Oracle(_param_with_alpha(lambda w: aa / (1 + (w / wa) ** 2),
                                        lambda w: ab / (1 + (w / wb) ** 2)))
    london = 1.5 * aa * ab * wa * wb / (wa + wb)
    # trapezoid error of the 23-point grid for poles near 0.5 Eh: below 1 %
    assert o.d4_c6ref(1, 0, 6, 0) == pytest.approx(london, rel=1e-2)


def test_default_c6_magnitudes(orc):
    # free-atom H: exact London value 6.5 (alpha 4.5, 3/4 alpha^2 w = 6.5)
    assert orc.d4_c6ref(1, 0, 1, 0) == pytest.approx(6.5, rel=1e-2)
    for z in (1, 6, 7, 8, 16):
        gw, _ = orc.d4_weights(z, 0.0, 0.0)
        for r in range(len(gw)):
            assert orc.d4_c6ref(z, r, z, r) > 0.0
    assert orc.d4_c6ref(99, 0, 1, 0) == -1.0
    assert orc.d4_c6ref(1, 5, 1, 0) == -1.0


@pytest.mark.parametrize("z", [1, 6, 7, 8, 16])
@pytest.mark.parametrize("cn", [0.0, 0.6, 1.3, 2.2, 3.7])
def test_weights_sum_to_one_at_zero_charge(orc, z, cn):
    gw, dgw = orc.d4_weights(z, cn, 0.0)
    assert gw.sum() == pytest.approx(1.0, abs=1e-14)
    assert np.all(gw >= 0.0)


def test_weights_fall_back_to_highest_reference(orc):
    gw, _ = orc.d4_weights(6, 40.0, 0.0)  # every Gaussian underflows
    assert gw[-1] == 1.0 and np.all(gw[:-1] == 0.0)


@pytest.mark.parametrize("z,q", [(1, 0.3), (6, -0.4), (8, -0.7), (16, 0.9), (7, 0.05)])
def test_weight_charge_derivative_finite_difference(orc, z, q):
    h = 1e-6
    gw, dgw = orc.d4_weights(z, 1.1, q)
    gp, _ = orc.d4_weights(z, 1.1, q + h)
    gm, _ = orc.d4_weights(z, 1.1, q - h)
    assert np.allclose(dgw, (gp - gm) / (2 * h), rtol=1e-7, atol=1e-10)
    # a cation is less polarisable, an anion more
    g0, _ = orc.d4_weights(z, 1.1, 0.0)
    assert (gw.sum() < g0.sum()) == (q > 0)


def test_weights_below_minus_zeff_are_constant(orc):
    gw1, dgw1 = orc.d4_weights(1, 0.9, -1.5)
    gw2, _ = orc.d4_weights(1, 0.9, -3.0)
    assert np.allclose(gw1, gw2) and np.all(dgw1 == 0.0)
    assert gw1.sum() == pytest.approx(np.exp(3.0), rel=1e-14)


def test_d4_cn_dimer_closed_form(orc):
    p = tblite_amd.default_param()
    rc = p.elem[0].rcov + p.elem[3].rcov  # H + O
    den = p.d4.cn_k4 * np.exp(-(abs(p.elem[0].en - p.elem[3].en) + p.d4.cn_k5) ** 2 / p.d4.cn_k6)
    cn = orc.d4_cn([1, 8], [[0.0, 0.0, 0.0], [0.0, 0.0, rc]])
    assert np.allclose(cn, 0.5 * den, rtol=1e-14)
    cn = orc.d4_cn([1, 8], [[0.0, 0.0, 0.0], [0.0, 0.0, 30.5]])
    assert np.all(cn == 0.0)


def _mol(n=5, seed=3):
    return workloads.organic_molecule(n, np.random.default_rng(seed))


def test_twobody_potential_is_charge_derivative(orc):
    num, xyz = _mol()
    rng = np.random.default_rng(1)
    q = rng.normal(0.0, 0.3, len(num))
    e2, _, vat = orc.d4_eval(num, xyz, q)
    h = 1e-5
    for i in range(len(num)):
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (orc.d4_eval(num, xyz, qp)[0].sum() - orc.d4_eval(num, xyz, qm)[0].sum()) / (2 * h)
        assert vat[i] == pytest.approx(fd, rel=1e-6, abs=1e-12)


def test_twobody_dimer_asymptote(orc):
    p = tblite_amd.default_param()
    for r in (25.0, 40.0):
        e2, _, _ = orc.d4_eval([1, 1], [[0.0, 0.0, 0.0], [0.0, 0.0, r]], [0.0, 0.0])
        gw, _ = orc.d4_weights(1, 0.0, 0.0)  # CN = 0 beyond the CN range of H2 at this distance
        c6 = sum(gw[a] * gw[b] * orc.d4_c6ref(1, a, 1, b) for a in range(2) for b in range(2))
        rr = 3 * p.d4.elem[0].r4r2 ** 2
        r0 = p.d4.a1 * np.sqrt(rr) + p.d4.a2
        exact = -c6 * (p.d4.s6 / (r ** 6 + r0 ** 6) + p.d4.s8 * rr / (r ** 8 + r0 ** 8))
        assert e2.sum() == pytest.approx(exact, rel=1e-12)
        # undamped limit: the damping radius R0 ~ 8.5 bohr enters as (R0/r)^6
        undamped = -c6 * (p.d4.s6 / r ** 6 + p.d4.s8 * rr / r ** 8)
        assert e2.sum() == pytest.approx(undamped, rel=3 * (r0 / r) ** 6)
    e2, _, _ = orc.d4_eval([1, 1], [[0.0, 0.0, 0.0], [0.0, 0.0, 60.5]], [0.0, 0.0])
    assert np.all(e2 == 0.0)


def _triangle(z, r, shape):
    if shape == "equilateral":
        xyz = [[0, 0, 0], [r, 0, 0], [r / 2, r * np.sqrt(3) / 2, 0]]
    else:
        xyz = [[0, 0, 0], [r, 0, 0], [2 * r, 0, 0]]
    return [z] * 3, np.array(xyz, dtype=float)


@pytest.mark.parametrize("shape,ang", [("equilateral", 11.0 / 8.0), ("collinear", -2.0)])
def test_threebody_angular_factor(orc, shape, ang):
    p = tblite_amd.default_param()
    r = 9.0
    num, xyz = _triangle(1, r, shape)
    _, e3, _ = orc.d4_eval(num, xyz, np.zeros(3))
    cn = orc.d4_cn(num, xyz)
    c6 = {}
    for i in range(3):
        for j in range(3):
            gi, _ = orc.d4_weights(1, cn[i], 0.0)
            gj, _ = orc.d4_weights(1, cn[j], 0.0)
            c6[i, j] = sum(gi[a] * gj[b] * orc.d4_c6ref(1, a, 1, b) for a in range(2) for b in range(2))
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    r0 = p.d4.a1 * np.sqrt(3 * p.d4.elem[0].r4r2 ** 2) + p.d4.a2
    rprod = d[0, 1] * d[0, 2] * d[1, 2]
    fd = 1.0 / (1.0 + 6.0 * (r0 ** 3 / rprod) ** (p.d4.alp / 3))
    exact = p.d4.s9 * np.sqrt(c6[0, 1] * c6[0, 2] * c6[1, 2]) * ang * fd / rprod ** 3
    assert e3.sum() == pytest.approx(exact, rel=1e-12)
    assert (e3.sum() > 0) == (ang > 0)


def test_threebody_cutoff(orc):
    num, xyz = _triangle(1, 20.5, "collinear")  # outer pair at 41 bohr > 40
    _, e3, _ = orc.d4_eval(num, xyz, np.zeros(3))
    assert np.all(e3 == 0.0)


def test_dispersion_rotation_invariant_and_size_consistent(orc):
    num, xyz = _mol(6, 5)
    q = np.random.default_rng(2).normal(0, 0.2, len(num))
    e2, e3, vat = orc.d4_eval(num, xyz, q)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    e2r, e3r, vatr = orc.d4_eval(num, xyz @ R.T + 3.0, q)
    assert e2r.sum() == pytest.approx(e2.sum(), rel=1e-12)
    assert e3r.sum() == pytest.approx(e3.sum(), rel=1e-12)
    assert np.allclose(vatr, vat, rtol=1e-11, atol=1e-15)
    # two copies beyond every cutoff: energies add
    far = xyz + np.array([200.0, 0, 0])
    e2d, e3d, _ = orc.d4_eval(np.concatenate([num, num]), np.vstack([xyz, far]), np.concatenate([q, q]))
    assert e2d.sum() == pytest.approx(2 * e2.sum(), rel=1e-12)
    assert e3d.sum() == pytest.approx(2 * e3.sum(), rel=1e-12)


def test_singlepoint_energy_decomposition(orc):
    num, xyz = workloads.config0()[0]
    r = orc.singlepoint(num, xyz)
    assert r["energy"] == pytest.approx(r["erep"] + r["edisp3"] + r["eelec"], abs=1e-12)
    assert r["edisp2"] < 0.0 < r["edisp3"]
    # the two-body part at the converged charges equals the stage evaluation
    e2, e3, _ = orc.d4_eval(num, xyz, r["qat"])
    assert e2.sum() == pytest.approx(r["edisp2"], abs=1e-12)
    assert e3.sum() == pytest.approx(r["edisp3"], abs=1e-14)
    p = tblite_amd.default_param()
    p.d4.enabled = 0
    r0 = oracle.Oracle(p).singlepoint(num, xyz)
    assert r0["edisp2"] == 0.0 and r0["edisp3"] == 0.0
    # dispersion shifts the total by about its own size (plus small relaxation)
    assert abs((r["energy"] - r0["energy"]) - (r["edisp2"] + r["edisp3"])) < 1e-4
