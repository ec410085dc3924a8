"""The drop-in entry point tblite_amd_xtb_singlepoint (xtb_singlepoint,
src/xtb.cu:188-434) and the wavefunction it fills (wavefunction_type,
src/wavefunction/type.h:81-99), read back through
tblite_amd_get_wavefunction_array, on each of the three SCF paths.

Against the oracle on the same inputs: energy |dE| <= 1e-6 Eh, shell charges
and atomic multipoles 1e-5 e, iterations +-1, orbital energies 1e-6 Eh, the
density matrix (get_density_matrix, src/wavefunction/type.cu:17-35) 1e-5
elementwise.  The coefficients of degenerate levels are fixed only up to a
rotation within the level, so coeff is checked by what defines it: C^T S C = I
(1e-8), P = C diag(f) C^T (1e-10), and, for levels separated from their
neighbours by more than 1e-3 Eh, overlap |c_gpu^T S c_oracle| = 1 (1e-6)."""
import ctypes as C

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid_engine():
    # multi-workgroup path above 64 padded AOs, so molecules the oracle
    # finishes in seconds run through it
    eng = tblite_amd.Engine(tuning={"auto_grid": 0, "grid_min_ld": 64})
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def cases():
    rng = np.random.default_rng(71)
    return {
        "lds": workloads.water_cluster(9, workloads.liquid_radius(9), rng),      # 54 AOs, path 0
        "global": workloads.organic_molecule(26, rng),                          # > 96 AOs, path 1
        "charged_lds": workloads.water_cluster(5, workloads.liquid_radius(5), rng),
    }


def _check(r, ref, S, tag):
    assert r["rc"] == 0, (tag, r["rc"])
    assert abs(r["energy"] - ref["energy"]) < 1e-6, (tag, r["energy"], ref["energy"])
    assert abs(r["iterations"] - ref["iterations"]) <= 1, tag
    for k in ("qsh", "qat"):
        assert np.max(np.abs(r[k] - ref[k])) < 1e-5, (tag, k)
    assert np.max(np.abs(r["dpat"] - ref["dpat"].ravel())) < 1e-5, tag
    assert np.max(np.abs(r["qpat"] - ref["qpat"].ravel())) < 1e-5, tag
    assert np.max(np.abs(r["emo"] - ref["emo"])) < 1e-6, tag
    assert np.max(np.abs(r["focc"] - ref["focc"])) < 1e-6, tag
    assert abs(r["focc"].sum() - ref["focc"].sum()) < 1e-9, tag
    P, Cg, f = r["density"], r["coeff"], r["focc"]
    n = len(f)
    assert P.shape == (n, n) and Cg.shape == (n, n)
    assert np.max(np.abs(P - ref["density"])) < 1e-5, (tag, np.max(np.abs(P - ref["density"])))
    assert np.max(np.abs(Cg.T @ S @ Cg - np.eye(n))) < 1e-8, tag
    assert np.max(np.abs((Cg * f) @ Cg.T - P)) < 1e-10, tag
    e = ref["emo"]
    gap = np.minimum(np.r_[np.inf, np.diff(e)], np.r_[np.diff(e), np.inf])
    iso = np.nonzero(gap > 1e-3)[0]
    ov = np.abs(np.einsum("ik,ij,jk->k", Cg[:, iso], S, ref["coeff"][:, iso]))
    assert np.max(np.abs(ov - 1.0)) < 1e-6, (tag, np.max(np.abs(ov - 1.0)))


@pytest.mark.parametrize("name", ["lds", "global", "charged_lds"])
def test_xtb_singlepoint_wavefunction(engine, orc, cases, name):
    num, xyz = cases[name]
    charge, uhf = (1.0, 1) if name.startswith("charged") else (0.0, 0)
    nao = orc.dims(num)[2]
    print(f"[gpu] xtb_singlepoint + wavefunction: {name}, nao={nao}", flush=True)
    r = engine.singlepoint(num, xyz, charge=charge, uhf=uhf)
    path = engine.molecule_path(0)
    assert path == {"lds": 0, "global": 1, "charged_lds": 0}[name]
    ref = orc.singlepoint(num, xyz, charge=charge, uhf=uhf, wavefunction=True)
    S = orc.hamiltonian(num, xyz)[0]
    _check(r, ref, S, name)


def test_xtb_singlepoint_wavefunction_grid(grid_engine, orc, cases):
    num, xyz = cases["global"]
    print(f"[gpu] xtb_singlepoint + wavefunction on the multi-workgroup path, "
          f"nao={orc.dims(num)[2]}", flush=True)
    r = grid_engine.singlepoint(num, xyz)
    assert grid_engine.molecule_path(0) == 2
    ref = orc.singlepoint(num, xyz, wavefunction=True)
    S = orc.hamiltonian(num, xyz)[0]
    _check(r, ref, S, "grid")


def test_batch_entry_fills_wavefunctions(engine, cases):
    """tblite_amd_xtb_singlepoint_batch with one wavefunction handle per
    molecule (one NULL) gives the same arrays as the one-molecule calls."""
    lib = engine.lib
    names = ["lds", "global"]
    mols, wfns = [], []
    for k in names:
        num = np.ascontiguousarray(cases[k][0], dtype=np.int32)
        xyz = np.ascontiguousarray(cases[k][1], dtype=np.float64)
        mols.append(lib.tblite_amd_new_structure(engine.ctx, len(num), tblite_amd._ip(num),
                                                 tblite_amd._dp(xyz), 0.0, 0))
        wfns.append(lib.tblite_amd_new_wavefunction(engine.ctx))
    wfns.append(None)
    mols.append(mols[0])
    n = len(mols)
    marr = (C.c_void_p * n)(*mols)
    warr = (C.c_void_p * n)(*wfns)
    e = np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    print("[gpu] xtb_singlepoint_batch with wavefunctions", flush=True)
    try:
        rc = lib.tblite_amd_xtb_singlepoint_batch(engine.ctx, n, marr, engine.calc, warr, 1.0,
                                                  tblite_amd._dp(e), tblite_amd._ip(it),
                                                  tblite_amd._ip(st))
        assert rc == 0, engine.error()
        for m, k in enumerate(names):
            one = engine.singlepoint(*cases[k])
            assert abs(one["energy"] - e[m]) < 1e-9
            for arr in ("qsh", "density", "coeff", "emo"):
                cnt = C.c_int()
                assert lib.tblite_amd_get_wavefunction_array(wfns[m], arr.encode(), None,
                                                             C.byref(cnt)) == 0
                out = np.zeros(cnt.value)
                lib.tblite_amd_get_wavefunction_array(wfns[m], arr.encode(), tblite_amd._dp(out),
                                                      C.byref(cnt))
                ref = one[arr]
                if arr in ("density", "coeff"):
                    out = out.reshape(ref.shape).T
                if arr == "coeff":
                    # same vectors up to sign
                    out = out * np.sign(np.sum(out * ref, axis=0))
                assert np.max(np.abs(out - ref)) < 1e-8, (k, arr)
    finally:
        for h in mols[:2]:
            p = C.c_void_p(h)
            lib.tblite_amd_delete_structure(C.byref(p))
        for h in wfns[:2]:
            p = C.c_void_p(h)
            lib.tblite_amd_delete_wavefunction(C.byref(p))


def test_wavefunction_array_errors(engine):
    """An unfilled wavefunction or an unknown array name is an error, not an
    empty result."""
    lib = engine.lib
    w = lib.tblite_amd_new_wavefunction(engine.ctx)
    try:
        cnt = C.c_int(-1)
        assert lib.tblite_amd_get_wavefunction_array(w, b"density", None, C.byref(cnt)) < 0
        assert "empty" in engine.error()
        assert lib.tblite_amd_get_wavefunction_array(w, b"nonsense", None, C.byref(cnt)) < 0
        assert "unknown" in engine.error()
    finally:
        p = C.c_void_p(w)
        lib.tblite_amd_delete_wavefunction(C.byref(p))
