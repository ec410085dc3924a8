"""Generate the STO-nG expansion table used to build the GFN2 contracted basis.

tblite expands each Slater shell (n, l, zeta) into ng primitive Gaussians
r^l exp(-alpha r^2) following Stewart, J. Chem. Phys. 52, 431 (1970):
least-squares fits of a zeta = 1 Slater function, exponents scaled by zeta^2
(`src/basis/type.h:18-23` shows the resulting cgto_type; the expansion routine
itself, slater_to_gauss, is not part of the reference).  Stewart's tables are
not in this container, so this script repeats his fit: for normalised radial
functions, minimising the squared difference is the same as maximising the
overlap <STO|G>, and for fixed exponents the optimal coefficients follow from
one linear solve, c ~ M^-1 b, with overlap^2 = b^T M^-1 b.  The exponents are
optimised on that reduced objective and the contraction is normalised to unit
self-overlap.

The STO-<->GTO radial overlap uses the closed form
  int_0^inf r^k exp(-a r^2 - r) dr
      = Gamma(k+1) (2a)^(-(k+1)/2) exp(1/(8a)) D_{-(k+1)}(1/sqrt(2a))
with D the parabolic cylinder function.

Pinned by tests/test_sto_ng.py against Stewart's published STO-2G, STO-3G and
STO-6G 1s values.

Run:  python gen_sto_ng.py  > sto_ng_table.h
"""
from math import factorial, gamma, sqrt
from multiprocessing import Pool

import numpy as np
from scipy import integrate, optimize
from scipy.special import pbdv

SHELLS = [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)]
NG_RANGE = range(1, 7)


def _sto_norm(n):
    return 2.0**n * sqrt(2.0 / factorial(2 * n))


def _gto_norm(a, l):
    return sqrt(2.0 * (2.0 * a) ** (l + 1.5) / gamma(l + 1.5))


def _radint(k, a):
    z = 1.0 / sqrt(2.0 * a)
    if z > 6.0:
        # exp(z^2/4) D(z) loses precision for diffuse primitives: integrate directly
        v, _ = integrate.quad(lambda r: r**k * np.exp(-a * r * r - r), 0.0, np.inf,
                              epsabs=0.0, epsrel=2e-14, limit=500)
        return v
    return gamma(k + 1) * (2 * a) ** (-(k + 1) / 2) * np.exp(z * z / 4) * pbdv(-(k + 1), z)[0]


def ovl_sto_gto(n, l, a):
    return _sto_norm(n) * _gto_norm(a, l) * _radint(n + l + 1, a)


def ovl_gto_gto(a, b, l):
    return (2.0 * np.sqrt(a * b) / (a + b)) ** (l + 1.5)


def _objective(loga, n, l):
    if np.any(np.abs(loga) > 8.0):
        return 0.0
    a = np.exp(loga)
    b = np.array([ovl_sto_gto(n, l, x) for x in a])
    m = ovl_gto_gto(a[:, None], a[None, :], l)
    try:
        return -float(b @ np.linalg.solve(m, b))
    except np.linalg.LinAlgError:
        return 0.0


def fit(n, l, ng, prev=None):
    starts = [np.log(0.3 * s ** np.linspace(ng / 2.0, -ng / 2.0, ng)) for s in (3.0, 4.0, 5.0)]
    if prev is not None:
        lp = np.log(prev)
        starts += [np.concatenate([[lp[0] + 1.5], lp]), np.concatenate([lp, [lp[-1] - 1.2]]),
                   np.concatenate([lp[:1], [0.5 * (lp[0] + lp[1])], lp[1:]]) if len(lp) > 1 else lp]
    best = None
    for x0 in starts:
        if len(x0) != ng:
            continue
        r = optimize.minimize(_objective, x0, args=(n, l), method="Nelder-Mead",
                              options=dict(xatol=1e-13, fatol=1e-17, maxiter=60000, maxfev=60000))
        r = optimize.minimize(_objective, r.x, args=(n, l), method="BFGS",
                              options=dict(gtol=1e-14))
        if best is None or r.fun < best.fun:
            best = r
    a = np.exp(best.x)
    b = np.array([ovl_sto_gto(n, l, x) for x in a])
    m = ovl_gto_gto(a[:, None], a[None, :], l)
    c = np.linalg.solve(m, b)
    c /= np.sqrt(c @ m @ c)
    order = np.argsort(-a)
    if np.max(np.abs(c)) > 5.0 or -best.fun >= 1.0:
        # degenerate fit (coalescing exponents with cancelling coefficients): leave the
        # entry empty so that the basis builder rejects it
        return np.zeros(ng), np.zeros(ng), -1.0
    return a[order], c[order], -best.fun


def _fit_shell(nl):
    n, l = nl
    out, prev = [], None
    for ng in NG_RANGE:
        a, c, ov = fit(n, l, ng, prev)
        prev = a
        out.append((a, c, ov))
    return out


def main():
    print("/* STO-nG expansion of zeta=1 Slater shells: generated by gen_sto_ng.py, do not edit.")
    print(" * sto_ng_alpha[shell][ng-1][i]: exponents, sto_ng_coeff: coefficients of normalised")
    print(" * primitives, contraction normalised to unit self-overlap.")
    print(" * shell index: 0=1s 1=2s 2=2p 3=3s 4=3p 5=3d. */")
    print("#ifndef TBLITE_AMD_STO_NG_TABLE_H")
    print("#define TBLITE_AMD_STO_NG_TABLE_H")
    print("#define STO_NG_NSHELL 6")
    print("#define STO_NG_MAXG 6")
    with Pool(len(SHELLS)) as pool:
        results = pool.map(_fit_shell, SHELLS)
    al, co = [], []
    for (n, l), res in zip(SHELLS, results):
        ra, rc = [], []
        for ng, (a, c, ov) in zip(NG_RANGE, res):
            ra.append(list(a) + [0.0] * (6 - ng))
            rc.append(list(c) + [0.0] * (6 - ng))
            print(f"/* n={n} l={l} ng={ng} overlap={ov:.15f} */")
        al.append(ra)
        co.append(rc)

    def emit(name, tab):
        print(f"static const double {name}[STO_NG_NSHELL][STO_NG_MAXG][STO_NG_MAXG] = {{")
        for sh in tab:
            print("  {")
            for row in sh:
                print("    {" + ", ".join(f"{v:.17e}" for v in row) + "},")
            print("  },")
        print("};")

    emit("sto_ng_alpha", al)
    emit("sto_ng_coeff", co)
    print("#endif")


if __name__ == "__main__":
    main()
