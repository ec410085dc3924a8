"""STO-nG expansion table: pinned against Stewart's published fits
(J. Chem. Phys. 52, 431 (1970), 1s exponents and coefficients for zeta = 1)
and against its own normalisation."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "tblite-cuda_amd", "params", "sto_ng_table.h")


def _table(name):
    text = open(TABLE).read()
    body = text.split("double " + name)[1].split("};")[0]
    nums = [float(x) for x in re.findall(r"-?\d\.\d+e[+-]\d+", body)]
    return np.array(nums).reshape(6, 6, 6)


ALPHA = _table("sto_ng_alpha")
COEFF = _table("sto_ng_coeff")

# published values (Stewart 1970; the STO-3G / STO-6G 1s sets of the standard basis libraries)
STEWART = {
    2: ([0.851819, 0.151623], [0.430129, 0.678914], 4e-6),  # published to 6 digits
    3: ([2.227660584, 0.405771156, 0.109817510], [0.154328967, 0.535328142, 0.444634542], 5e-7),
    6: ([23.10303149, 4.235915534, 1.185056519, 0.4070988982, 0.1580884151, 0.06510953954],
        [0.009163596281, 0.04936149294, 0.1685383049, 0.3705627997, 0.4164915298,
         0.1303340841], 5e-5),
}


@pytest.mark.parametrize("ng", sorted(STEWART))
def test_1s_matches_stewart(ng):
    a_ref, c_ref, tol = STEWART[ng]
    a = ALPHA[0, ng - 1, :ng]
    c = COEFF[0, ng - 1, :ng]
    assert np.allclose(a, a_ref, rtol=tol, atol=0), (a, a_ref)
    assert np.allclose(c, c_ref, rtol=tol * 10, atol=tol * 10), (c, c_ref)


SHELLS = [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)]


@pytest.mark.parametrize("shell", range(6))
@pytest.mark.parametrize("ng", range(1, 7))
def test_contraction_normalised(shell, ng):
    a = ALPHA[shell, ng - 1, :ng]
    c = COEFF[shell, ng - 1, :ng]
    if np.all(a == 0):
        pytest.skip("no stable fit for this (shell, ng); the basis builder rejects it")
    l = SHELLS[shell][1]
    m = (2 * np.sqrt(np.outer(a, a)) / np.add.outer(a, a)) ** (l + 1.5)
    assert abs(c @ m @ c - 1.0) < 1e-12


def test_used_shells_are_accurate():
    """The shells used by the GFN2 data (1s/3, 2s/4, 2p/4, 3s/4, 3p/4, 3d/3) reproduce the
    Slater function to an overlap above 0.9997 (radial quadrature, independent of the fit)."""
    from math import factorial, gamma

    r = np.linspace(1e-6, 60, 200001)
    for shell, ng in [(0, 3), (1, 4), (2, 4), (3, 4), (4, 4), (5, 3)]:
        n, l = SHELLS[shell]
        sto = 2.0**n * np.sqrt(2.0 / factorial(2 * n)) * r ** (n - 1) * np.exp(-r)
        a = ALPHA[shell, ng - 1, :ng]
        c = COEFF[shell, ng - 1, :ng]
        g = sum(ci * np.sqrt(2 * (2 * ai) ** (l + 1.5) / gamma(l + 1.5)) * r**l * np.exp(-ai * r * r)
                for ai, ci in zip(a, c))
        ov = np.trapezoid(sto * g * r * r, r)
        assert 0.9997 < ov < 1.0, (shell, ng, ov)
