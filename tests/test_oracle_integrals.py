"""Pins of the oracle's multipole integrals (multipole_cgto, multipole_3d,
transform0, shift_operator, get_hamiltonian).

* the known-answer vector of the reference's disabled device test
  (src/test.old:100-127, s-s case of the 1D moment recursion) and its
  transform0 invariant (src/test.old:152-176);
* brute-force Gauss-Hermite quadrature of overlap, dipole and traceless
  quadrupole integrals between contracted shells, with the spherical d
  functions and the quadrupole operator written out directly (an independent
  method: no recursion, no Cartesian->spherical matrix, no trace removal);
* invariants: unit self-overlap, symmetry, positive definiteness,
  tracelessness, translation invariance.
"""
import numpy as np
import pytest

import oracle
import workloads

S3 = np.sqrt(3.0)


def test_golden_vector_test_old(orc):
    # src/test.old:100-127: inputs rpi, rpj, ai, aj, s-s, s1d; expected s3d = 1,
    # d3d = the second displacement, q3d as listed (the moment is taken about
    # the second centre argument)
    rpi = [0.083114717636212157, -0.39033516689727543, -0.25671966537784952]
    rpj = [-0.804297356879628, 3.7772557251143746, 2.4842645706404505]
    ai, aj = 5.9405971566712825, 0.61389118221497996
    s1d = [1.0, 0.0, 0.076283605088381307, 0.0]
    s, d, q = oracle.multipole_3d(rpi, rpj, ai, aj, [0, 0, 0], [0, 0, 0], s1d)
    assert s == 1.0
    assert np.allclose(d, rpj, rtol=0, atol=1e-15)
    q_true = [0.72317784337193691, -3.0380367959679342, 14.343944417997701,
              -1.9980874279558183, 9.3837025721504457, 6.2478540620277627]
    assert np.allclose(q, q_true, rtol=1e-14, atol=1e-14)


def test_transform0_isotropic_d_vanishes():
    # src/test.old:152-176: an isotropic Cartesian d (xx = yy = zz) has no spherical d part
    cart = np.array([[0.6451174322], [0.6451174322], [0.6451174322], [0.0], [0.0], [0.0]])
    sph = oracle.transform0(2, 0, cart)
    assert np.all(sph == 0.0)


# ---------------------------------------------------------- quadrature
def _funcs(l, x, y, z):
    """Real spherical components in tblite order: s; p = y, z, x;
    d = z2, xz, yz, x2-y2, xy (normalised like the Cartesian xx component)."""
    if l == 0:
        return [np.ones_like(x)]
    if l == 1:
        return [y, z, x]
    return [z * z - 0.5 * (x * x + y * y), S3 * x * z, S3 * y * z, 0.5 * S3 * (x * x - y * y), S3 * x * y]


def quad_block(cg_j, Rj, cg_i, Ri, npt=10):
    """<phi_j | O | phi_i> for O = 1, (r - Ri)_k, and the traceless
    quadrupole 3/2 (r-Ri)_a (r-Ri)_b - 1/2 |r-Ri|^2 delta_ab, by product
    Gauss-Hermite quadrature on each Gaussian product."""
    lj, aj, cj = cg_j
    li, ai, ci = cg_i
    t, w = np.polynomial.hermite.hermgauss(npt)
    sj, si = 2 * lj + 1, 2 * li + 1
    S = np.zeros((si, sj))
    D = np.zeros((si, sj, 3))
    Q = np.zeros((si, sj, 6))
    Rj = np.asarray(Rj, float)
    Ri = np.asarray(Ri, float)
    for a_i, c_i in zip(ai, ci):
        for a_j, c_j in zip(aj, cj):
            e = a_i + a_j
            P = (a_i * Ri + a_j * Rj) / e
            K = np.exp(-a_i * a_j / e * np.sum((Ri - Rj) ** 2))
            g = t / np.sqrt(e)
            X, Y, Z = np.meshgrid(P[0] + g, P[1] + g, P[2] + g, indexing="ij")
            W = np.einsum("i,j,k->ijk", w, w, w) / e**1.5
            fi = _funcs(li, X - Ri[0], Y - Ri[1], Z - Ri[2])
            fj = _funcs(lj, X - Rj[0], Y - Rj[1], Z - Rj[2])
            xi, yi, zi = X - Ri[0], Y - Ri[1], Z - Ri[2]
            r2 = xi * xi + yi * yi + zi * zi
            ops_d = [xi, yi, zi]
            ops_q = [1.5 * xi * xi - 0.5 * r2, 1.5 * xi * yi, 1.5 * yi * yi - 0.5 * r2,
                     1.5 * xi * zi, 1.5 * yi * zi, 1.5 * zi * zi - 0.5 * r2]
            pref = c_i * c_j * K
            for a in range(si):
                for b in range(sj):
                    base = W * fi[a] * fj[b] * pref
                    S[a, b] += base.sum()
                    for k in range(3):
                        D[a, b, k] += (base * ops_d[k]).sum()
                    for k in range(6):
                        Q[a, b, k] += (base * ops_q[k]).sum()
    return S, D, Q


SHELL_CASES = [(1, 0, 1, 0), (8, 0, 1, 0), (8, 1, 1, 0), (8, 1, 8, 1), (8, 0, 8, 1),
               (16, 2, 1, 0), (16, 2, 8, 1), (16, 2, 16, 2), (16, 1, 6, 1), (7, 1, 16, 0)]


@pytest.mark.parametrize("zj,jsh,zi,ish", SHELL_CASES)
def test_multipole_cgto_vs_quadrature(orc, zj, jsh, zi, ish):
    rng = np.random.default_rng(zj * 100 + zi * 10 + jsh + ish)
    Ri = rng.uniform(-1, 1, 3)
    Rj = Ri - rng.uniform(-1.6, 1.6, 3)
    vec = Ri - Rj
    cg_j = orc.cgto(zj, jsh)
    cg_i = orc.cgto(zi, ish)
    S, D, Q = orc.multipole_cgto(zj, jsh, zi, ish, vec, intcut=1e9)
    lj, li = cg_j[0], cg_i[0]
    sj, si = 2 * lj + 1, 2 * li + 1
    # oracle layout: S(jao + sj*iao), D(k + 3*ij), Q(k + 6*ij)
    So = S[: sj * si].reshape(si, sj)
    Do = D[: 3 * sj * si].reshape(si, sj, 3)
    Qo = Q[: 6 * sj * si].reshape(si, sj, 6)
    Sq, Dq, Qq = quad_block(cg_j, Rj, cg_i, Ri)
    assert np.allclose(So, Sq, atol=1e-12), np.abs(So - Sq).max()
    assert np.allclose(Do, Dq, atol=1e-12), np.abs(Do - Dq).max()
    assert np.allclose(Qo, Qq, atol=1e-12), np.abs(Qo - Qq).max()


def test_hamiltonian_dipole_shift_vs_quadrature(orc):
    """The transposed block of get_hamiltonian carries the operator on the
    other centre (shift_operator): check D and Q of both blocks of an O-H pair."""
    num = np.array([8, 1])
    xyz = np.array([[0.1, -0.2, 0.3], [1.2, 0.9, -0.4]])
    S, D, Q, H = orc.hamiltonian(num, xyz)
    O, Hh = xyz
    # AO order: O s (0), O p y z x (1..3), H s (4); numpy D[i, j, k] = <phi_j|(r-R_at(i))_k|phi_i>
    for (zi, ish, Ri, i0), (zj, jsh, Rj, j0) in [((8, 1, O, 1), (1, 0, Hh, 4)), ((1, 0, Hh, 4), (8, 1, O, 1))]:
        cg_i, cg_j = orc.cgto(zi, ish), orc.cgto(zj, jsh)
        Sq, Dq, Qq = quad_block(cg_j, Rj, cg_i, Ri)
        si, sj = Sq.shape
        assert np.allclose(S[i0:i0 + si, j0:j0 + sj], Sq, atol=1e-12)
        assert np.allclose(D[i0:i0 + si, j0:j0 + sj], Dq, atol=1e-12)
        assert np.allclose(Q[i0:i0 + si, j0:j0 + sj], Qq, atol=1e-12)


@pytest.mark.parametrize("z", [1, 6, 7, 8, 16])
def test_self_overlap_is_one(orc, z):
    nsh = {1: 1, 6: 2, 7: 2, 8: 2, 16: 3}[z]
    for ish in range(nsh):
        cg = orc.cgto(z, ish)
        Sq, _, _ = quad_block(cg, [0, 0, 0], cg, [0, 0, 0])
        assert np.allclose(Sq, np.eye(2 * cg[0] + 1), atol=1e-12)


def _organic_small():
    rng = np.random.default_rng(7)
    return workloads.organic_molecule(6, rng)


@pytest.mark.parametrize("mol", ["water8", "organic"])
def test_matrix_invariants(orc, mol):
    num, xyz = workloads.config0()[0] if mol == "water8" else _organic_small()
    S, D, Q, H = orc.hamiltonian(num, xyz)
    assert np.allclose(S, S.T, atol=1e-14)
    assert np.allclose(H, H.T, atol=1e-14)
    assert np.all(np.abs(np.diag(S) - 1.0) < 1e-12)
    assert np.linalg.eigvalsh(S).min() > 0.0
    tr = Q[:, :, 0] + Q[:, :, 2] + Q[:, :, 5]
    assert np.max(np.abs(tr)) < 1e-12
    # translation invariance of everything (operators are relative to atoms)
    S2, D2, Q2, H2 = orc.hamiltonian(num, xyz + np.array([3.3, -1.7, 0.4]))
    for a, b in ((S, S2), (D, D2), (Q, Q2), (H, H2)):
        assert np.max(np.abs(a - b)) < 1e-11


def test_onsite_h0_is_selfenergy(orc):
    num, xyz = workloads.config0()[0]
    S, D, Q, H = orc.hamiltonian(num, xyz)
    se = orc.selfenergy(num, orc.cn(num, xyz))
    # AO -> shell map for water: O (s, p p p), H (s), H (s)
    aosh = []
    ish = 0
    for z in num:
        for l in ([0, 1] if z == 8 else [0]):
            aosh += [ish] * (2 * l + 1)
            ish += 1
    assert np.allclose(np.diag(H), se[aosh], atol=1e-14)
