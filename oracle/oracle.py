"""TEST INFRASTRUCTURE -- ctypes binding of the CPU oracle (gfn2_oracle.c).

Only tests/, the smoke check and bench.py's CPU-baseline leg use this, and only
as the checker.  Builds the oracle with its Makefile when the library is
missing (gcc is available wherever the tests run).
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgfn2_oracle.so")
sys.path.insert(0, os.path.join(ROOT, "tblite-cuda_amd"))
from tblite_amd import Param  # noqa: E402  (struct layout of include/tblite_amd.h)

KB_EH = 3.166808578545117e-06


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB


class Result(C.Structure):
    _fields_ = [
        ("energy", C.c_double), ("erep", C.c_double), ("eelec", C.c_double),
        ("edisp2", C.c_double), ("edisp3", C.c_double), ("iterations", C.c_int), ("converged", C.c_int), ("status", C.c_int),
        ("qsh", C.POINTER(C.c_double)), ("qat", C.POINTER(C.c_double)),
        ("dpat", C.POINTER(C.c_double)), ("qpat", C.POINTER(C.c_double)),
        ("emo", C.POINTER(C.c_double)), ("focc", C.POINTER(C.c_double)),
        ("density", C.POINTER(C.c_double)), ("coeff", C.POINTER(C.c_double)),
        ("gradient", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double)),
    ]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        L.or_new_calc.restype = C.c_void_p
        L.or_new_calc.argtypes = [C.POINTER(Param)]
        L.or_free_calc.argtypes = [C.c_void_p]
        L.or_dims.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_singlepoint.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                     C.POINTER(Result)]
        L.or_singlepoint_pbc.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.c_double, C.c_int, C.c_double,
                                         C.c_double, C.c_int, C.c_double, C.POINTER(Result)]
        L.or_pbc_tables.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                    C.POINTER(C.c_double), C.c_int]
        L.or_adjlist_pbc.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.c_int]
        L.or_pbc_terms.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)] + [C.POINTER(C.c_double)] * 10
        L.or_set_ewald_alpha_scale.argtypes = [C.c_double]
        L.or_singlepoint_many.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int,
                                          C.c_double, C.c_int, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_multipole_3d.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_double] * 2 + \
            [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_double)] * 4
        L.or_transform0.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_multipole_cgto.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(C.c_double), C.c_double] + [C.POINTER(C.c_double)] * 3
        L.or_get_cgto.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_broyden_linear.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int,
                                        C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_hscale.restype = C.c_double
        L.or_hscale.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_ncoord.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
        L.or_repulsion.argtypes = L.or_ncoord.argtypes
        L.or_get_cutoff.restype = C.c_double
        L.or_get_cutoff.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_double]
        L.or_adjlist.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
        L.or_selfenergy.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.or_hamiltonian.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.c_double] + [C.POINTER(C.c_double)] * 4
        L.or_hamiltonian_pbc.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.c_double] + [C.POINTER(C.c_double)] * 4
        L.or_coulomb_matrices.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                          C.POINTER(C.c_double)] + [C.POINTER(C.c_double)] * 5
        L.or_sygvd.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 4
        L.or_fermi_filling.argtypes = [C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_entropy.restype = C.c_double
        L.or_entropy.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double]
        L.or_coulomb_eval.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)] + \
            [C.POINTER(C.c_double)] * 8
        L.or_d4_ncoord.argtypes = L.or_ncoord.argtypes
        L.or_d4_weights.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_d4_c6ref.restype = C.c_double
        L.or_d4_c6ref.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_d4_eval.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)] + [C.POINTER(C.c_double)] * 5
        L.or_multipole_grad_cgto.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.POINTER(C.c_double), C.c_double] + [C.POINTER(C.c_double)] * 6
        L.or_classical_gradient.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                            C.c_int] + [C.POINTER(C.c_double)] * 5
        L.or_integral_gradient.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                           C.c_double] + [C.POINTER(C.c_double)] * 7
        _lib = L
    return _lib


class Oracle:
    def __init__(self, param=None):
        self.L = lib()
        self.calc = self.L.or_new_calc(C.byref(param) if param is not None else None)
        if not self.calc:
            raise RuntimeError("oracle calculator construction failed")

    def __del__(self):
        try:
            self.L.or_free_calc(self.calc)
        except Exception:
            pass

    @staticmethod
    def _mol(num, xyz):
        return (np.ascontiguousarray(num, dtype=np.int32),
                np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3))

    def dims(self, num):
        num = np.ascontiguousarray(num, dtype=np.int32)
        nsh, nao = C.c_int(), C.c_int()
        if self.L.or_dims(self.calc, len(num), _ip(num), C.byref(nsh), C.byref(nao)):
            raise ValueError("unknown element")
        return len(num), nsh.value, nao.value

    def singlepoint(self, num, xyz, charge=0.0, uhf=0, accuracy=1.0, temperature=300.0,
                    max_iter=250, damping=0.4, wavefunction=False, gradient=False, lattice=None):
        """wavefunction=True also returns density and coeff (nao x nao numpy
        arrays indexed [row, column]; coeff[:, k] is level k); gradient=True
        returns gradient (nat, 3) = dE/dR in Eh/bohr and sigma (3, 3).
        lattice (3, 3), rows = lattice vectors in bohr: a periodic cell at the
        Gamma point (no gradient)."""
        num, xyz = self._mol(num, xyz)
        nat, nsh, nao = self.dims(num)
        out = dict(qsh=np.zeros(nsh), qat=np.zeros(nat), dpat=np.zeros((nat, 3)),
                   qpat=np.zeros((nat, 6)), emo=np.zeros(nao), focc=np.zeros(nao))
        if wavefunction:
            # column-major in C: allocated as the transpose, read back as [row, col]
            out["density"] = np.zeros((nao, nao))
            out["coeff"] = np.zeros((nao, nao))
        if gradient:
            out["gradient"] = np.zeros((nat, 3))
            out["sigma"] = np.zeros((3, 3))
        r = Result()
        for k in out:
            setattr(r, k, _dp(out[k]))
        if lattice is None:
            rc = self.L.or_singlepoint(self.calc, nat, _ip(num), _dp(xyz), charge, uhf, accuracy,
                                       temperature * KB_EH, max_iter, damping, C.byref(r))
        else:
            lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
            rc = self.L.or_singlepoint_pbc(self.calc, nat, _ip(num), _dp(xyz), _dp(lat), charge,
                                           uhf, accuracy, temperature * KB_EH, max_iter, damping,
                                           C.byref(r))
        out.update(energy=r.energy, erep=r.erep, eelec=r.eelec, edisp2=r.edisp2,
                   edisp3=r.edisp3, iterations=r.iterations,
                   converged=bool(r.converged), status=r.status, rc=rc)
        if wavefunction:
            out["density"] = out["density"].T.copy()
            out["coeff"] = out["coeff"].T.copy()
        if gradient:
            out["sigma"] = out["sigma"].T.copy()
        return out

    def singlepoint_many(self, mols, accuracy=1.0, temperature=300.0, max_iter=250,
                         damping=0.4, nthreads=0):
        mol_nat = np.array([len(m[0]) for m in mols], dtype=np.int32)
        num = np.concatenate([m[0] for m in mols]).astype(np.int32)
        xyz = np.ascontiguousarray(np.concatenate([m[1] for m in mols]), dtype=np.float64)
        n = len(mols)
        e = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        self.L.or_singlepoint_many(self.calc, n, _ip(mol_nat), _ip(num), _dp(xyz), accuracy,
                                   temperature * KB_EH, max_iter, damping, nthreads, _dp(e),
                                   _ip(it), _ip(st))
        return e, it, st

    # ---- periodic cells (lattice: (3, 3), rows = lattice vectors, bohr)
    def set_ewald_alpha_scale(self, s):
        self.L.or_set_ewald_alpha_scale(float(s))

    def pbc_tables(self, lattice, xyz):
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        info = np.zeros(6)
        if self.L.or_pbc_tables(_dp(lat), len(xyz), _dp(xyz), _dp(info), None, 0, None, 0):
            raise ValueError("degenerate lattice")
        ntr, ng = int(info[4]), int(info[5])
        tr = np.zeros((ntr, 3))
        g = np.zeros((max(ng, 1), 3))
        self.L.or_pbc_tables(_dp(lat), len(xyz), _dp(xyz), _dp(info), _dp(tr), ntr, _dp(g), ng)
        return dict(volume=info[0], alpha=info[1], rreal=info[2], gmax=info[3], trans=tr, recip=g[:ng])

    def adjlist_pbc(self, xyz, lattice, cutoff):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
        nat = len(xyz)
        inl = np.zeros(nat, dtype=np.int32)
        nnl = np.zeros(nat, dtype=np.int32)
        n = -self.L.or_adjlist_pbc(nat, _dp(xyz), _dp(lat), cutoff, _ip(inl), _ip(nnl), None, None, 0)
        nlat = np.zeros(max(n, 1), dtype=np.int32)
        nltr = np.zeros(max(n, 1), dtype=np.int32)
        self.L.or_adjlist_pbc(nat, _dp(xyz), _dp(lat), cutoff, _ip(inl), _ip(nnl), _ip(nlat),
                              _ip(nltr), max(n, 1))
        return inl, nnl, nlat[:n], nltr[:n]

    def pbc_terms(self, num, xyz, lattice):
        """Classical terms of a cell: cn, erep (per atom), cn4, amat (nsh, nsh),
        sd (nat, nat, 3) [i, j] = amat_sd(:, j, i), dd (nat, 3, nat, 3),
        sq (nat, nat, 6), mrad, dmat (nat, nat), e3 (per atom)."""
        num, xyz = self._mol(num, xyz)
        nat, nsh, nao = self.dims(num)
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
        o = dict(cn=np.zeros(nat), erep=np.zeros(nat), cn4=np.zeros(nat), amat=np.zeros((nsh, nsh)),
                 sd=np.zeros((nat, nat, 3)), dd=np.zeros((nat, 3, nat, 3)), sq=np.zeros((nat, nat, 6)),
                 mrad=np.zeros(nat), dmat=np.zeros((nat, nat)), e3=np.zeros(nat))
        rc = self.L.or_pbc_terms(self.calc, nat, _ip(num), _dp(xyz), _dp(lat),
                                 *[_dp(o[k]) for k in ("cn", "erep", "cn4", "amat", "sd", "dd", "sq",
                                                       "mrad", "dmat", "e3")])
        if rc:
            raise ValueError("or_pbc_terms failed")
        return o

    def broyden_linear(self, A, b, q0, damping=0.4, memory=250, nstep=10, omega0=0.01):
        """Mixer iterates on F(q) = A q + b: visited q's (nstep+1, n) and
        their residual norms |F(q) - q|; omega0 as in the SCF by default."""
        A = np.asfortranarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        n = len(b)
        qs = np.zeros((nstep + 1, n))
        res = np.zeros(nstep + 1)
        rc = self.L.or_broyden_linear(n, A.ctypes.data_as(C.POINTER(C.c_double)), _dp(b), _dp(q0),
                                      damping, omega0, memory, nstep, _dp(qs), _dp(res))
        return rc, qs, res

    def hscale(self, ei, ish, ej, jsh):
        """H0 scaling of shell ish of element index ei with shell jsh of ej."""
        return self.L.or_hscale(self.calc, ei, ish, ej, jsh)

    def cn(self, num, xyz):
        num, xyz = self._mol(num, xyz)
        out = np.zeros(len(num))
        self.L.or_ncoord(self.calc, len(num), _ip(num), _dp(xyz), _dp(out))
        return out

    def repulsion(self, num, xyz):
        num, xyz = self._mol(num, xyz)
        out = np.zeros(len(num))
        self.L.or_repulsion(self.calc, len(num), _ip(num), _dp(xyz), _dp(out))
        return out

    def cutoff(self, num, accuracy=1.0):
        num = np.ascontiguousarray(num, dtype=np.int32)
        return self.L.or_get_cutoff(self.calc, len(num), _ip(num), accuracy)

    def adjlist(self, xyz, cutoff):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        nat = len(xyz)
        inl = np.zeros(nat, dtype=np.int32)
        nnl = np.zeros(nat, dtype=np.int32)
        cap = nat * (nat + 1) // 2 + 1
        nlat = np.zeros(cap, dtype=np.int32)
        nnz = self.L.or_adjlist(nat, _dp(xyz), cutoff, _ip(inl), _ip(nnl), _ip(nlat), cap)
        return inl, nnl, nlat[:nnz]

    def selfenergy(self, num, cn):
        num = np.ascontiguousarray(num, dtype=np.int32)
        _, nsh, _ = self.dims(num)
        out = np.zeros(nsh)
        self.L.or_selfenergy(self.calc, len(num), _ip(num), _dp(np.ascontiguousarray(cn, dtype=np.float64)), _dp(out))
        return out

    def hamiltonian(self, num, xyz, accuracy=1.0):
        num, xyz = self._mol(num, xyz)
        _, _, nao = self.dims(num)
        S = np.zeros((nao, nao))
        D = np.zeros((nao, nao, 3))
        Q = np.zeros((nao, nao, 6))
        H = np.zeros((nao, nao))
        self.L.or_hamiltonian(self.calc, len(num), _ip(num), _dp(xyz), accuracy, _dp(S), _dp(D), _dp(Q), _dp(H))
        return S, D, Q, H

    def hamiltonian_pbc(self, num, xyz, lattice, accuracy=1.0):
        num, xyz = self._mol(num, xyz)
        _, _, nao = self.dims(num)
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
        S = np.zeros((nao, nao))
        D = np.zeros((nao, nao, 3))
        Q = np.zeros((nao, nao, 6))
        H = np.zeros((nao, nao))
        self.L.or_hamiltonian_pbc(self.calc, len(num), _ip(num), _dp(xyz), _dp(lat), accuracy,
                                  _dp(S), _dp(D), _dp(Q), _dp(H))
        return S, D, Q, H

    def coulomb_matrices(self, num, xyz):
        num, xyz = self._mol(num, xyz)
        nat, nsh, _ = self.dims(num)
        amat = np.zeros((nsh, nsh))
        sd = np.zeros((nat, nat, 3))
        dd = np.zeros((nat, 3, nat, 3))
        sq = np.zeros((nat, nat, 6))
        mrad = np.zeros(nat)
        self.L.or_coulomb_matrices(self.calc, nat, _ip(num), _dp(xyz), _dp(amat), _dp(sd), _dp(dd), _dp(sq), _dp(mrad))
        return amat, sd, dd, sq, mrad

    def coulomb_eval(self, num, xyz, qsh, dpat, qpat):
        num, xyz = self._mol(num, xyz)
        nat, nsh, _ = self.dims(num)
        e = np.zeros(nat)
        vsh = np.zeros(nsh)
        vat = np.zeros(nat)
        vdp = np.zeros((nat, 3))
        vqp = np.zeros((nat, 6))
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.L.or_coulomb_eval(self.calc, nat, _ip(num), _dp(xyz), _dp(f(qsh)), _dp(f(dpat)), _dp(f(qpat)),
                               _dp(e), _dp(vsh), _dp(vat), _dp(vdp), _dp(vqp))
        return e, vsh, vat, vdp, vqp

    def d4_cn(self, num, xyz):
        num, xyz = self._mol(num, xyz)
        out = np.zeros(len(num))
        if self.L.or_d4_ncoord(self.calc, len(num), _ip(num), _dp(xyz), _dp(out)):
            raise ValueError("unknown element")
        return out

    def d4_weights(self, z, cn, q):
        gw, dgw = np.zeros(7), np.zeros(7)
        nref = self.L.or_d4_weights(self.calc, int(z), float(cn), float(q), _dp(gw), _dp(dgw))
        if nref < 0:
            raise ValueError("no D4 data")
        return gw[:nref].copy(), dgw[:nref].copy()

    def d4_c6ref(self, zi, ri, zj, rj):
        return self.L.or_d4_c6ref(self.calc, int(zi), int(ri), int(zj), int(rj))

    def d4_eval(self, num, xyz, qat):
        """Per-atom D4 two-body energy at charges qat, three-body energy (zero
        charges) and the two-body potential dE2/dq."""
        num, xyz = self._mol(num, xyz)
        nat = len(num)
        e2, e3, vat = np.zeros(nat), np.zeros(nat), np.zeros(nat)
        if self.L.or_d4_eval(self.calc, nat, _ip(num), _dp(xyz),
                             _dp(np.ascontiguousarray(qat, dtype=np.float64)), _dp(e2), _dp(e3), _dp(vat)):
            raise ValueError("D4 disabled or unknown element")
        return e2, e3, vat

    def cgto(self, z, ish):
        ang, npr = C.c_int(), C.c_int()
        a = np.zeros(6)
        c = np.zeros(6)
        if self.L.or_get_cgto(self.calc, z, ish, C.byref(ang), C.byref(npr), _dp(a), _dp(c)):
            raise ValueError("no such shell")
        return ang.value, a[:npr.value].copy(), c[:npr.value].copy()

    def multipole_cgto(self, zj, jsh, zi, ish, vec, intcut=25.0):
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        S = np.zeros(25)
        D = np.zeros(75)
        Q = np.zeros(150)
        self.L.or_multipole_cgto(self.calc, zj, jsh, zi, ish, float(vec @ vec), _dp(vec), intcut,
                                 _dp(S), _dp(D), _dp(Q))
        return S, D, Q


    def multipole_grad_cgto(self, zj, jsh, zi, ish, vec, intcut=25.0):
        """S(25), D(75), Q(150) as multipole_cgto and their derivatives by
        vec: dS (3, 25), dD (3, 75), dQ (3, 150) (component-major inside)."""
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        S, D, Q = np.zeros(25), np.zeros(75), np.zeros(150)
        dS, dD, dQ = np.zeros(75), np.zeros(225), np.zeros(450)
        if self.L.or_multipole_grad_cgto(self.calc, zj, jsh, zi, ish, float(vec @ vec), _dp(vec), intcut,
                                         _dp(S), _dp(D), _dp(Q), _dp(dS), _dp(dD), _dp(dQ)):
            raise ValueError("no such shell")
        return S, D, Q, dS.reshape(-1, 3).T.copy(), dD.reshape(-1, 3).T.copy(), dQ.reshape(-1, 3).T.copy()

    def classical_gradient(self, num, xyz, which, qsh, dpat, qpat):
        """Energy and gradient (nat, 3) of one classical term at fixed
        moments: which = 'repulsion', 'coulomb', 'd4_twobody', 'd4_threebody'."""
        w = {"repulsion": 0, "coulomb": 1, "d4_twobody": 2, "d4_threebody": 3}[which]
        num, xyz = self._mol(num, xyz)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        e = np.zeros(1)
        g = np.zeros((len(num), 3))
        if self.L.or_classical_gradient(self.calc, len(num), _ip(num), _dp(xyz), w, _dp(f(qsh)),
                                        _dp(f(dpat)), _dp(f(qpat)), _dp(e), _dp(g)):
            raise ValueError("term unavailable")
        return e[0], g

    def integral_gradient(self, num, xyz, P, W, vsh, vdp, vqp, accuracy=1.0):
        """Energy and gradient of the integral terms at fixed P, W (nao x nao,
        symmetric) and potentials vsh (nsh), vdp (nat, 3), vqp (nat, 6)."""
        num, xyz = self._mol(num, xyz)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        e = np.zeros(1)
        g = np.zeros((len(num), 3))
        self.L.or_integral_gradient(self.calc, len(num), _ip(num), _dp(xyz), accuracy,
                                    _dp(f(np.asarray(P).T)), _dp(f(np.asarray(W).T)), _dp(f(vsh)),
                                    _dp(f(vdp)), _dp(f(vqp)), _dp(e), _dp(g))
        return e[0], g


def sygvd(F, S):
    F = np.asfortranarray(F, dtype=np.float64)
    S = np.asfortranarray(S, dtype=np.float64)
    n = F.shape[0]
    w = np.zeros(n)
    Cm = np.zeros((n, n), order="F")
    rc = lib().or_sygvd(n, F.ctypes.data_as(C.POINTER(C.c_double)), S.ctypes.data_as(C.POINTER(C.c_double)),
                        _dp(w), Cm.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, w, Cm


def multipole_3d(rpj, rpi, aj, ai, lj, li, s1d):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    g = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    s = np.zeros(1)
    d = np.zeros(3)
    q = np.zeros(6)
    s1 = np.zeros(9)
    s1[:len(s1d)] = s1d
    lib().or_multipole_3d(_dp(f(rpj)), _dp(f(rpi)), aj, ai, _ip(g(lj)), _ip(g(li)), _dp(s1), _dp(s), _dp(d), _dp(q))
    return s[0], d, q


def transform0(lj, li, cart):
    cart = np.asfortranarray(cart, dtype=np.float64)
    sj, si = 2 * lj + 1, 2 * li + 1
    sph = np.zeros((sj, si), order="F")
    lib().or_transform0(lj, li, cart.ctypes.data_as(C.POINTER(C.c_double)),
                        sph.ctypes.data_as(C.POINTER(C.c_double)))
    return sph


def fermi_filling(nel, kt, emo):
    emo = np.ascontiguousarray(emo, dtype=np.float64)
    occ = np.zeros(len(emo))
    homo = C.c_int()
    ef = C.c_double()
    lib().or_fermi_filling(nel, kt, len(emo), _dp(emo), C.byref(homo), _dp(occ), C.byref(ef))
    return homo.value, occ, ef.value


def entropy(occ, kt):
    occ = np.ascontiguousarray(occ, dtype=np.float64)
    return lib().or_entropy(len(occ), _dp(occ), kt)
