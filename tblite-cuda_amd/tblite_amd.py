"""ctypes binding of the engine's C ABI (include/tblite_amd.h).

Loads libtblite_amd.so built in this directory (see build.py).  The binding
raises if the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBNAME = os.path.join(HERE, "libtblite_amd.so")

MAX_SHELL = 3
MAX_ELEM = 16

STATUS = {
    0: "converged",
    1: "not converged",
    2: "cholesky",
    4: "eigen",
    8: "nan",
    16: "mixer",
    32: "too large",
}


class Element(C.Structure):
    _fields_ = [
        ("z", C.c_int), ("nsh", C.c_int),
        ("ang", C.c_int * MAX_SHELL), ("pqn", C.c_int * MAX_SHELL), ("ngauss", C.c_int * MAX_SHELL),
        ("levels", C.c_double * MAX_SHELL), ("slater", C.c_double * MAX_SHELL),
        ("refocc", C.c_double * MAX_SHELL), ("shpoly", C.c_double * MAX_SHELL),
        ("kcn", C.c_double * MAX_SHELL), ("lgam", C.c_double * MAX_SHELL),
        ("gam", C.c_double), ("gam3", C.c_double), ("zeff", C.c_double), ("arep", C.c_double),
        ("en", C.c_double), ("dkernel", C.c_double), ("qkernel", C.c_double),
        ("mprad", C.c_double), ("mpvcn", C.c_double), ("rcov", C.c_double),
        ("atomic_rad", C.c_double),
    ]


D4_MAXREF = 7
D4_NFREQ = 23


class D4Element(C.Structure):
    _fields_ = [
        ("nref", C.c_int), ("cn", C.c_double * D4_MAXREF), ("q", C.c_double * D4_MAXREF),
        ("alpha", (C.c_double * D4_NFREQ) * D4_MAXREF), ("zeff", C.c_double), ("eta", C.c_double),
        ("r4r2", C.c_double),
    ]


class D4Param(C.Structure):
    _fields_ = [
        ("enabled", C.c_int), ("s6", C.c_double), ("s8", C.c_double), ("a1", C.c_double),
        ("a2", C.c_double), ("s9", C.c_double), ("alp", C.c_double), ("ga", C.c_double),
        ("gc", C.c_double), ("wf", C.c_double), ("cn_k4", C.c_double), ("cn_k5", C.c_double),
        ("cn_k6", C.c_double), ("cn_kn", C.c_double), ("cn_cutoff", C.c_double),
        ("disp2_cutoff", C.c_double), ("disp3_cutoff", C.c_double),
        ("elem", D4Element * MAX_ELEM),
    ]


class Param(C.Structure):
    _fields_ = [
        ("kshell", (C.c_double * 3) * 3), ("enscale", C.c_double), ("wexp", C.c_double),
        ("kpol", C.c_double), ("rep_kexp", C.c_double), ("rep_kexp_light", C.c_double),
        ("rep_rexp", C.c_double), ("rep_cutoff", C.c_double), ("gexp", C.c_double),
        ("thirdorder_shell", C.c_double * 3), ("dmp3", C.c_double), ("dmp5", C.c_double),
        ("mp_shift", C.c_double), ("mp_kexp", C.c_double), ("mp_rmax", C.c_double),
        ("cn_ka", C.c_double), ("cn_kb", C.c_double), ("cn_rshift", C.c_double),
        ("cn_cutoff", C.c_double), ("nelem", C.c_int), ("elem", Element * MAX_ELEM),
        ("d4", D4Param),
    ]


class Tuning(C.Structure):
    """tblite_amd_tuning: SCF path thresholds and the routing cost model."""
    _fields_ = [
        ("lds_max_ld", C.c_int), ("grid_min_ld", C.c_int), ("auto_grid", C.c_int),
        ("grid_graph", C.c_int), ("wg_a", C.c_double), ("wg_p", C.c_double),
        ("grid_b", C.c_double), ("grid_c", C.c_double),
        ("integral_chunk", C.c_int), ("pair_chunk_log2", C.c_int), ("atm_chunk_log2", C.c_int),
        ("fine_buckets", C.c_int), ("big_lds_min_kb", C.c_int),
    ]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

_PROTOS = {
    "tblite_amd_param_default": (_I, [C.POINTER(Param)]),
    "tblite_amd_new_context": (_P, [_I]),
    "tblite_amd_delete_context": (None, [C.POINTER(_P)]),
    "tblite_amd_get_error": (_I, [_P, C.c_char_p, _I]),
    "tblite_amd_new_structure": (_P, [_P, _I, _IP, _DP, _D, _I]),
    "tblite_amd_new_structure_periodic": (_P, [_P, _I, _IP, _DP, _D, _I, _DP, _IP]),
    "tblite_amd_singlepoint_flat_periodic": (_I, [_P, _P, _I, _IP, _IP, _DP, _DP, _DP, _IP, _D, _DP,
                                                  _IP, _IP]),
    "tblite_amd_batch_get_images": (_I, [_P, _I, _IP, _I, _DP, _I]),
    "tblite_amd_delete_structure": (None, [C.POINTER(_P)]),
    "tblite_amd_get_structure_dims": (_I, [_P, _P, _IP, _IP, _IP]),
    "tblite_amd_new_gfn2_calculator": (_P, [_P, C.POINTER(Param)]),
    "tblite_amd_delete_calculator": (None, [C.POINTER(_P)]),
    "tblite_amd_set_calculator_accuracy": (_I, [_P, _D]),
    "tblite_amd_set_calculator_max_iter": (_I, [_P, _I]),
    "tblite_amd_set_calculator_mixer_damping": (_I, [_P, _D]),
    "tblite_amd_set_calculator_temperature": (_I, [_P, _D]),
    "tblite_amd_new_wavefunction": (_P, [_P]),
    "tblite_amd_delete_wavefunction": (None, [C.POINTER(_P)]),
    "tblite_amd_get_wavefunction_array": (_I, [_P, C.c_char_p, _DP, _IP]),
    "tblite_amd_xtb_singlepoint": (_I, [_P, _P, _P, _P, _D, _DP, _IP, _I]),
    "tblite_amd_xtb_singlepoint_batch": (_I, [_P, _I, C.POINTER(_P), _P, C.POINTER(_P), _D, _DP, _IP, _IP]),
    "tblite_amd_singlepoint_flat": (_I, [_P, _P, _I, _IP, _IP, _DP, _DP, _IP, _D, _DP, _IP, _IP]),
    "tblite_amd_singlepoint_flat_noalloc": (_I, [_P, _P, _I, _IP, _IP, _DP, _DP, _IP, _D, _DP, _IP, _IP]),
    "tblite_amd_singlepoint_flat_gradient": (_I, [_P, _P, _I, _IP, _IP, _DP, _DP, _IP, _D, _DP, _DP, _DP,
                                                  _IP, _IP]),
    "tblite_amd_xtb_singlepoint_gradient": (_I, [_P, _P, _P, _P, _D, _DP, _DP, _DP, _IP, _I]),
    "tblite_amd_batch_get_gradient": (_I, [_P, _I, _DP, _DP]),
    "tblite_amd_batch_prepare": (_I, [_P, _P, _I, _IP, _IP, _DP, _IP, _D]),
    "tblite_amd_batch_set_geometry": (_I, [_P, _DP, _I]),
    "tblite_amd_batch_geometry_device": (_P, [_P]),
    "tblite_amd_batch_run": (_I, [_P]),
    "tblite_amd_batch_results": (_I, [_P, _DP, _IP, _IP]),
    "tblite_amd_batch_prepare_gradient": (_I, [_P, _P, _I, _IP, _IP, _DP, _IP, _D]),
    "tblite_amd_batch_run_gradient": (_I, [_P]),
    "tblite_amd_batch_gradients": (_I, [_P, _DP, _DP]),
    "tblite_amd_batch_gradient_device": (_P, [_P]),
    "tblite_amd_batch_dims": (_I, [_P, _I, _IP, _IP, _IP]),
    "tblite_amd_batch_get_cn": (_I, [_P, _I, _DP]),
    "tblite_amd_batch_get_adjlist": (_I, [_P, _I, _IP, _IP, _IP, _I]),
    "tblite_amd_batch_get_integrals": (_I, [_P, _I, _DP, _DP, _DP, _DP]),
    "tblite_amd_batch_get_selfenergy": (_I, [_P, _I, _DP]),
    "tblite_amd_batch_get_charges": (_I, [_P, _I, _DP, _DP, _DP, _DP, _DP, _DP]),
    "tblite_amd_batch_get_energy_parts": (_I, [_P, _I, _DP, _DP]),
    "tblite_amd_batch_get_dispersion": (_I, [_P, _I, _DP, _DP]),
    "tblite_amd_last_timings": (_I, [_P, _DP, _I]),
    "tblite_amd_last_bucket_timings": (_I, [_P, _DP, _IP, _IP, _IP, _I]),
    "tblite_amd_last_profile": (_I, [_P, C.POINTER(C.c_longlong)]),
    "tblite_amd_get_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "tblite_amd_set_tuning": (_I, [_P, C.POINTER(Tuning)]),
    "tblite_amd_molecule_path": (_I, [_P, _I]),
    "tblite_amd_read_structure_file": (_I, [C.c_char_p, C.c_char_p, _I, _IP, _IP, _DP, _DP, _IP]),
    "tblite_amd_write_json": (_I, [C.c_char_p, _I, _IP, _DP, _D, _I, _DP, _DP, _DP, _I, _DP, _DP]),
    "tblite_amd_new_server": (_P, [_P, _P, _I, _D]),
    "tblite_amd_delete_server": (None, [C.POINTER(_P)]),
    "tblite_amd_server_submit": (_I, [_P, _I, _IP, _DP, _D, _I, C.POINTER(C.c_longlong)]),
    "tblite_amd_server_poll": (_I, [_P, _I, C.POINTER(C.c_longlong), _DP, _IP, _IP, _DP]),
    "tblite_amd_server_stats": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _IP, _IP,
                                     C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "tblite_amd_workspace_estimate": (_I, [C.POINTER(Param), _I, _IP, _IP, _I, _DP, _DP]),
}

PROFILE_SLOTS = ["init", "mix", "potential", "fock", "transform", "jacobi", "occupation",
                 "density", "mulliken", "energy", "sweeps", "iterations",
                 "tridiagonal", "bisection", "inverse_iteration", "back_transform",
                 "tri_column", "tri_reflector", "tri_symv", "tri_correct", "tri_panel", "clusters",
                 "band_stage1", "band_stage2", "band_q2"]

EXPORTED = sorted(_PROTOS)

_lib = None


def load(path=None):
    """Load and type the engine library; raises OSError if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIBNAME
    if not os.path.exists(p):
        raise OSError(f"{p} not built: run `python __graft_entry__.py build` (no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ip(a):
    return a.ctypes.data_as(_IP) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def default_param():
    lib = load()
    p = Param()
    lib.tblite_amd_param_default(C.byref(p))
    return p


def read_structure(path, fmt=None):
    """(num, xyz_bohr (nat, 3), charge, uhf) from an xyz or Turbomole coord file."""
    lib = load()
    nat, uhf = C.c_int(), C.c_int()
    charge = C.c_double()
    p = str(path).encode()
    f = fmt.encode() if fmt else None
    if lib.tblite_amd_read_structure_file(p, f, 0, C.byref(nat), None, None, C.byref(charge),
                                          C.byref(uhf)) != 0:
        raise ValueError(f"cannot read structure from {path}")
    num = np.zeros(nat.value, dtype=np.int32)
    xyz = np.zeros((nat.value, 3))
    if lib.tblite_amd_read_structure_file(p, f, nat.value, C.byref(nat), _ip(num), _dp(xyz),
                                          C.byref(charge), C.byref(uhf)) != 0:
        raise ValueError(f"cannot read structure from {path}")
    return num, xyz, charge.value, uhf.value


def write_json(path, num, xyz, energy, iterations=0, qat=None, dpat=None, qpat=None, emo=None,
               focc=None):
    """Write one singlepoint result (as from Engine.singlepoint) as JSON."""
    f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    num = np.ascontiguousarray(num, dtype=np.int32)
    xyz, qat, dpat, qpat, emo, focc = (f(a) for a in (xyz, qat, dpat, qpat, emo, focc))
    nao = 0 if emo is None else len(emo)
    rc = load().tblite_amd_write_json(str(path).encode(), len(num), _ip(num), _dp(xyz),
                                      float(energy), int(iterations), _dp(qat), _dp(dpat),
                                      _dp(qpat), nao, _dp(emo), _dp(focc))
    if rc != 0:
        raise ValueError(f"cannot write {path}")


class EngineError(RuntimeError):
    pass


class Engine:
    """One device context with one GFN2 calculator."""

    def __init__(self, device=0, param=None, accuracy=1.0, max_iter=250, damping=0.4,
                 temperature=300.0, tuning=None):
        self.lib = load()
        self.ctx = self.lib.tblite_amd_new_context(device)
        if not self.ctx:
            raise EngineError("could not create a tblite_amd context (no HIP device?)")
        self.calc = self.lib.tblite_amd_new_gfn2_calculator(
            self.ctx, C.byref(param) if param is not None else None)
        if not self.calc:
            raise EngineError(self.error())
        self._check(self.lib.tblite_amd_set_calculator_accuracy(self.calc, accuracy))
        self._check(self.lib.tblite_amd_set_calculator_max_iter(self.calc, max_iter))
        self._check(self.lib.tblite_amd_set_calculator_mixer_damping(self.calc, damping))
        self._check(self.lib.tblite_amd_set_calculator_temperature(self.calc, temperature))
        self.accuracy = accuracy
        if tuning is not None:
            # a dict of tblite_amd_tuning fields or the path of a JSON file
            # written by autotune.py
            if isinstance(tuning, str):
                import json
                with open(tuning) as f:
                    tuning = json.load(f)
            self.set_tuning(**tuning)

    def error(self):
        buf = C.create_string_buffer(1024)
        self.lib.tblite_amd_get_error(self.ctx, buf, 1024)
        return buf.value.decode(errors="replace")

    def _check(self, rc, allow=()):
        if rc != 0 and rc not in allow:
            raise EngineError(f"tblite_amd error {rc}: {self.error()}")
        return rc

    def close(self):
        if getattr(self, "calc", None):
            h = C.c_void_p(self.calc)
            self.lib.tblite_amd_delete_calculator(C.byref(h))
            self.calc = None
        if getattr(self, "ctx", None):
            h = C.c_void_p(self.ctx)
            self.lib.tblite_amd_delete_context(C.byref(h))
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def singlepoint_flat(self, mol_nat, num, xyz, charge=None, uhf=None, accuracy=None):
        mol_nat = np.ascontiguousarray(mol_nat, dtype=np.int32)
        num = np.ascontiguousarray(num, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        nmol = len(mol_nat)
        self._nmol = nmol
        energies = np.zeros(nmol)
        iters = np.zeros(nmol, dtype=np.int32)
        status = np.zeros(nmol, dtype=np.int32)
        ch = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        uh = None if uhf is None else np.ascontiguousarray(uhf, dtype=np.int32)
        rc = self.lib.tblite_amd_singlepoint_flat(
            self.ctx, self.calc, nmol, _ip(mol_nat), _ip(num), _dp(xyz), _dp(ch), _ip(uh),
            self.accuracy if accuracy is None else accuracy, _dp(energies), _ip(iters), _ip(status))
        self._check(rc, allow=(-3,))
        return energies, iters, status

    def singlepoint_flat_periodic(self, mol_nat, num, xyz, lattice, charge=None, uhf=None,
                                  accuracy=None):
        """Batch of periodic cells; lattice (nmol, 3, 3), rows = lattice vectors (bohr)."""
        mol_nat = np.ascontiguousarray(mol_nat, dtype=np.int32)
        num = np.ascontiguousarray(num, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(-1)
        nmol = len(mol_nat)
        self._nmol = nmol
        energies = np.zeros(nmol)
        iters = np.zeros(nmol, dtype=np.int32)
        status = np.zeros(nmol, dtype=np.int32)
        ch = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        uh = None if uhf is None else np.ascontiguousarray(uhf, dtype=np.int32)
        rc = self.lib.tblite_amd_singlepoint_flat_periodic(
            self.ctx, self.calc, nmol, _ip(mol_nat), _ip(num), _dp(xyz), _dp(lat), _dp(ch), _ip(uh),
            self.accuracy if accuracy is None else accuracy, _dp(energies), _ip(iters), _ip(status))
        self._check(rc, allow=(-3,))
        return energies, iters, status

    def images(self, imol):
        """Translation index of every neighbour-list entry and the translation table of cell imol."""
        ntr = self._check_nonneg(self.lib.tblite_amd_batch_get_images(self.ctx, imol, None, 0, None, 0))
        nnz = self._check_nonneg(self.lib.tblite_amd_batch_get_adjlist(self.ctx, imol, None, None, None, 0))
        nltr = np.zeros(max(nnz, 1), dtype=np.int32)
        tr = np.zeros((max(ntr, 1), 3))
        self._check_nonneg(self.lib.tblite_amd_batch_get_images(self.ctx, imol, _ip(nltr), len(nltr),
                                                                _dp(tr), len(tr)))
        return nltr[:nnz], tr[:ntr]

    def singlepoint_flat_gradient(self, mol_nat, num, xyz, charge=None, uhf=None, accuracy=None):
        """Batched singlepoint with analytic gradients: energies, gradients
        (total atoms, 3) dE/dR in Eh/bohr, sigmas (nmol, 3, 3) strain
        derivatives, iterations, status."""
        mol_nat = np.ascontiguousarray(mol_nat, dtype=np.int32)
        num = np.ascontiguousarray(num, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        nmol = len(mol_nat)
        self._nmol = nmol
        energies = np.zeros(nmol)
        grads = np.zeros((len(num), 3))
        sigmas = np.zeros((nmol, 3, 3))
        iters = np.zeros(nmol, dtype=np.int32)
        status = np.zeros(nmol, dtype=np.int32)
        ch = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        uh = None if uhf is None else np.ascontiguousarray(uhf, dtype=np.int32)
        rc = self.lib.tblite_amd_singlepoint_flat_gradient(
            self.ctx, self.calc, nmol, _ip(mol_nat), _ip(num), _dp(xyz), _dp(ch), _ip(uh),
            self.accuracy if accuracy is None else accuracy, _dp(energies), _dp(grads),
            _dp(sigmas), _ip(iters), _ip(status))
        self._check(rc, allow=(-3,))
        # column-major 3 x 3 per molecule -> numpy [row, column]
        return energies, grads, sigmas.transpose(0, 2, 1).copy(), iters, status

    def gradient(self, imol):
        """Gradient (nat, 3) and sigma (3, 3) of molecule imol of the last
        gradient batch."""
        nat, _, _ = self.dims(imol)
        g = np.zeros((nat, 3))
        sg = np.zeros((3, 3))
        self._check(self.lib.tblite_amd_batch_get_gradient(self.ctx, imol, _dp(g), _dp(sg)))
        return g, sg.T.copy()

    def prepare(self, mol_nat, num, charge=None, uhf=None, accuracy=None, gradient=False):
        """Plan a device-resident batch from its composition (gradient=True:
        also the gradient buffers, for run_gradient())."""
        mol_nat = np.ascontiguousarray(mol_nat, dtype=np.int32)
        num = np.ascontiguousarray(num, dtype=np.int32)
        ch = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64)
        uh = None if uhf is None else np.ascontiguousarray(uhf, dtype=np.int32)
        self._nmol = len(mol_nat)
        self._natoms = len(num)
        fn = self.lib.tblite_amd_batch_prepare_gradient if gradient else self.lib.tblite_amd_batch_prepare
        self._check(fn(self.ctx, self.calc, len(mol_nat), _ip(mol_nat), _ip(num), _dp(ch), _ip(uh),
                       self.accuracy if accuracy is None else accuracy))

    def run_gradient(self):
        self._check(self.lib.tblite_amd_batch_run_gradient(self.ctx))

    def gradients(self):
        """Gradients (total atoms, 3) and sigmas (nmol, 3, 3) of the last run."""
        g = np.zeros((self._natoms, 3))
        sg = np.zeros((self._nmol, 3, 3))
        self._check(self.lib.tblite_amd_batch_gradients(self.ctx, _dp(g), _dp(sg)))
        return g, sg.transpose(0, 2, 1).copy()

    def set_geometry(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self._check(self.lib.tblite_amd_batch_set_geometry(self.ctx, _dp(xyz), 0))

    def run(self):
        self._check(self.lib.tblite_amd_batch_run(self.ctx))

    def results(self):
        e = np.zeros(self._nmol)
        it = np.zeros(self._nmol, dtype=np.int32)
        st = np.zeros(self._nmol, dtype=np.int32)
        self._check(self.lib.tblite_amd_batch_results(self.ctx, _dp(e), _ip(it), _ip(st)), allow=(-3,))
        return e, it, st

    def singlepoint(self, num, xyz, charge=0.0, uhf=0, accuracy=None, wavefunction=True,
                    gradient=False, lattice=None):
        """One molecule through tblite_amd_xtb_singlepoint (the per-molecule path).

        With wavefunction=True a wavefunction handle is passed and every array
        of it is returned (tblite's layout: density and coeff are nao x nao,
        column-major, so coeff[:, r] is level r after the transpose below);
        with False the call is energy-only (no handle, as a caller that wants
        the energy alone would make it).  gradient=True calls
        tblite_amd_xtb_singlepoint_gradient instead and also returns gradient
        (nat, 3) in Eh/bohr and sigma (3, 3).  lattice (3, 3), rows = lattice
        vectors in bohr, makes the structure a periodic cell."""
        num = np.ascontiguousarray(num, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        if lattice is None:
            mol = self.lib.tblite_amd_new_structure(self.ctx, len(num), _ip(num), _dp(xyz), charge, uhf)
        else:
            lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
            per = np.ones(3, dtype=np.int32)
            mol = self.lib.tblite_amd_new_structure_periodic(self.ctx, len(num), _ip(num), _dp(xyz),
                                                             charge, uhf, _dp(lat), _ip(per))
        if not mol:
            raise EngineError(self.error())
        wfn = self.lib.tblite_amd_new_wavefunction(self.ctx) if wavefunction else None
        e = C.c_double()
        it = C.c_int()
        try:
            acc = self.accuracy if accuracy is None else accuracy
            if gradient:
                g = np.zeros((len(num), 3))
                sg = np.zeros((3, 3))
                rc = self.lib.tblite_amd_xtb_singlepoint_gradient(
                    self.ctx, mol, self.calc, wfn, acc, C.byref(e), _dp(g), _dp(sg), C.byref(it), 0)
            else:
                rc = self.lib.tblite_amd_xtb_singlepoint(self.ctx, mol, self.calc, wfn, acc,
                                                         C.byref(e), C.byref(it), 0)
            out = {"energy": e.value, "iterations": it.value, "rc": rc}
            if gradient:
                out["gradient"] = g
                out["sigma"] = sg.T.copy()
            if wavefunction and rc in (0, -3):
                for name in ("qsh", "qat", "dpat", "qpat", "emo", "focc", "n0sh", "n0at",
                             "density", "coeff"):
                    cnt = C.c_int()
                    self._check(self.lib.tblite_amd_get_wavefunction_array(
                        wfn, name.encode(), None, C.byref(cnt)))
                    arr = np.zeros(cnt.value)
                    self._check(self.lib.tblite_amd_get_wavefunction_array(
                        wfn, name.encode(), _dp(arr), C.byref(cnt)))
                    out[name] = arr
                nao = len(out["emo"])
                # column-major (Fortran) nao x nao -> numpy [row, column]
                out["density"] = out["density"].reshape(nao, nao).T.copy()
                out["coeff"] = out["coeff"].reshape(nao, nao).T.copy()
        finally:
            h = C.c_void_p(mol)
            self.lib.tblite_amd_delete_structure(C.byref(h))
            if wfn:
                h = C.c_void_p(wfn)
                self.lib.tblite_amd_delete_wavefunction(C.byref(h))
        return out

    def dims(self, imol):
        nat, nsh, nao = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.tblite_amd_batch_dims(self.ctx, imol, C.byref(nat), C.byref(nsh), C.byref(nao)))
        return nat.value, nsh.value, nao.value

    def cn(self, imol):
        nat, _, _ = self.dims(imol)
        out = np.zeros(nat)
        self._check(self.lib.tblite_amd_batch_get_cn(self.ctx, imol, _dp(out)))
        return out

    def selfenergy(self, imol):
        _, nsh, _ = self.dims(imol)
        out = np.zeros(nsh)
        self._check(self.lib.tblite_amd_batch_get_selfenergy(self.ctx, imol, _dp(out)))
        return out

    def adjlist(self, imol):
        nat, _, _ = self.dims(imol)
        inl = np.zeros(nat, dtype=np.int32)
        nnl = np.zeros(nat, dtype=np.int32)
        nnz = self.lib.tblite_amd_batch_get_adjlist(self.ctx, imol, _ip(inl), _ip(nnl), None, 0)
        if nnz < 0:
            raise EngineError(self.error())
        nlat = np.zeros(max(nnz, 1), dtype=np.int32)
        nnz = self._check_nonneg(self.lib.tblite_amd_batch_get_adjlist(
            self.ctx, imol, _ip(inl), _ip(nnl), _ip(nlat), len(nlat)))
        return inl, nnl, nlat[:nnz]

    def _check_nonneg(self, rc):
        if rc < 0:
            raise EngineError(f"tblite_amd error {rc}: {self.error()}")
        return rc

    def integrals(self, imol):
        _, _, nao = self.dims(imol)
        S = np.zeros((nao, nao))
        D = np.zeros((nao, nao, 3))
        Q = np.zeros((nao, nao, 6))
        H = np.zeros((nao, nao))
        self._check(self.lib.tblite_amd_batch_get_integrals(self.ctx, imol, _dp(S), _dp(D), _dp(Q), _dp(H)))
        return S, D, Q, H

    def charges(self, imol):
        nat, nsh, nao = self.dims(imol)
        out = dict(qsh=np.zeros(nsh), qat=np.zeros(nat), dpat=np.zeros((nat, 3)),
                   qpat=np.zeros((nat, 6)), emo=np.zeros(nao), focc=np.zeros(nao))
        self._check(self.lib.tblite_amd_batch_get_charges(
            self.ctx, imol, _dp(out["qsh"]), _dp(out["qat"]), _dp(out["dpat"]), _dp(out["qpat"]),
            _dp(out["emo"]), _dp(out["focc"])))
        return out

    def energy_parts(self, imol):
        erep, eelec = C.c_double(), C.c_double()
        self._check(self.lib.tblite_amd_batch_get_energy_parts(self.ctx, imol, C.byref(erep), C.byref(eelec)))
        return erep.value, eelec.value

    def dispersion(self, imol):
        """D4 two-body (at the final charges, part of eelec) and three-body energy."""
        e2, e3 = C.c_double(), C.c_double()
        self._check(self.lib.tblite_amd_batch_get_dispersion(self.ctx, imol, C.byref(e2), C.byref(e3)))
        return e2.value, e3.value

    def profile(self):
        """Per-molecule SCF phase clocks (needs TBLITE_AMD_PROFILE=1 at context creation)."""
        out = np.zeros((self._nmol if hasattr(self, "_nmol") else 1) * len(PROFILE_SLOTS), dtype=np.int64)
        rc = self.lib.tblite_amd_last_profile(self.ctx, out.ctypes.data_as(C.POINTER(C.c_longlong)))
        if rc < 0:
            raise EngineError(self.error())
        return out.reshape(-1, len(PROFILE_SLOTS))

    def timings(self):
        t = np.zeros(5)
        self._check(self.lib.tblite_amd_last_timings(self.ctx, _dp(t), 5))
        return t

    def tuning(self):
        """Current tuning as a dict (see tblite_amd_tuning)."""
        t = Tuning()
        self._check(self.lib.tblite_amd_get_tuning(self.ctx, C.byref(t)))
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        """Change some tuning fields (keyword names as in tblite_amd_tuning)."""
        t = Tuning()
        self._check(self.lib.tblite_amd_get_tuning(self.ctx, C.byref(t)))
        for k, v in kw.items():
            if k not in dict(Tuning._fields_):
                raise KeyError(k)
            setattr(t, k, v)
        self._check(self.lib.tblite_amd_set_tuning(self.ctx, C.byref(t)))

    def molecule_path(self, imol):
        """SCF path of molecule imol of the last planned batch: 0 LDS, 1 one
        workgroup in global memory, 2 multi-workgroup."""
        r = self.lib.tblite_amd_molecule_path(self.ctx, imol)
        if r < 0:
            raise EngineError("no such molecule in the last batch")
        return r

    def bucket_timings(self):
        """SCF size buckets of the last run: list of dicts (ms, ldmax, nmol, nthreads)."""
        cap = 16
        ms = np.zeros(cap)
        ld, nm, nt = (np.zeros(cap, dtype=np.int32) for _ in range(3))
        nb = self.lib.tblite_amd_last_bucket_timings(self.ctx, _dp(ms), _ip(ld), _ip(nm), _ip(nt), cap)
        if nb < 0:
            raise EngineError(self.error())
        return [{"ms": float(ms[k]), "ldmax": int(ld[k]), "nmol": int(nm[k]), "nthreads": int(nt[k])}
                for k in range(min(nb, cap))]


class Server:
    """Continuous-batching request loop on the device of an Engine: submit()
    queues a molecule and returns its ticket, poll() collects finished results
    (ticket, energy, iterations, status, latency_ms) and keeps every slot busy."""

    def __init__(self, engine, slots=4, workspace_bytes=16e9):
        self.eng = engine
        self.lib = engine.lib
        self.srv = self.lib.tblite_amd_new_server(engine.ctx, engine.calc, slots, float(workspace_bytes))
        if not self.srv:
            raise EngineError(engine.error())
        self._cap = 4096
        self._t = np.zeros(self._cap, dtype=np.int64)
        self._e = np.zeros(self._cap)
        self._i = np.zeros(self._cap, dtype=np.int32)
        self._s = np.zeros(self._cap, dtype=np.int32)
        self._l = np.zeros(self._cap)

    def submit(self, num, xyz, charge=0.0, uhf=0):
        num = np.ascontiguousarray(num, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        t = C.c_longlong(0)
        self.eng._check(self.lib.tblite_amd_server_submit(self.srv, len(num), _ip(num), _dp(xyz),
                                                          float(charge), int(uhf), C.byref(t)))
        return t.value

    def poll(self):
        k = self.lib.tblite_amd_server_poll(
            self.srv, self._cap, self._t.ctypes.data_as(C.POINTER(C.c_longlong)), _dp(self._e),
            _ip(self._i), _ip(self._s), _dp(self._l))
        if k < 0:
            raise EngineError(self.eng.error())
        return [(int(self._t[j]), float(self._e[j]), int(self._i[j]), int(self._s[j]), float(self._l[j]))
                for j in range(k)]

    def stats(self):
        sub, comp, bat, grid = C.c_longlong(0), C.c_longlong(0), C.c_longlong(0), C.c_longlong(0)
        q, b = C.c_int(0), C.c_int(0)
        self.eng._check(self.lib.tblite_amd_server_stats(self.srv, C.byref(sub), C.byref(comp),
                                                         C.byref(q), C.byref(b), C.byref(bat),
                                                         C.byref(grid)))
        return {"submitted": sub.value, "completed": comp.value, "queued": q.value,
                "busy_slots": b.value, "batches": bat.value, "grid_requests": grid.value}

    def close(self):
        if getattr(self, "srv", None):
            h = C.c_void_p(self.srv)
            self.lib.tblite_amd_delete_server(C.byref(h))
            self.srv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
