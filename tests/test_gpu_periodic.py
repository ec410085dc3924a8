"""Periodic cells on the GPU against the oracle (Gamma point).

Cells of 4, 12, 14, 16, 18 and 40 rigid waters at liquid density (cubic and
triclinic) run through tblite_amd_singlepoint_flat_periodic in one batch;
every stage the oracle
exposes is compared: the translation table and the neighbour list with its
images bit for bit, CN, the Gamma-point S/D/Q/H, the D4 two- and three-body
energies, charges and the total energy (tolerances of the molecular tests:
1e-12 for integer-free setup quantities, 1e-11 for integrals, 1e-6 Eh for
energies, 1e-5 e for charges, iterations within one).  The periodic
structure handle goes through tblite_amd_xtb_singlepoint.  Cells up to 96
AOs run batched on the LDS-resident SCF kernels and up to 1024 AOs on the
one-workgroup global-memory kernel (their cell instantiations read the
Ewald-summed AES2 planes), larger ones on the multi-workgroup path; a cell
gives the same energy on all three.
"""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cells():
    small = workloads.periodic_cells(count=1, seed=6, first=4)
    mid = workloads.periodic_cells(count=3, seed=5, first=12)
    big = workloads.periodic_cells(count=1, seed=7, first=18)
    large = [workloads.water_box(40, np.random.default_rng(12), shear=0.1)]
    return small + mid + big + large


@pytest.fixture(scope="module")
def run(engine, cells):
    mol_nat = [len(c[0]) for c in cells]
    num = np.concatenate([c[0] for c in cells])
    xyz = np.concatenate([c[1] for c in cells])
    lat = np.stack([c[2] for c in cells])
    print(f"[gpu] periodic cells nat={mol_nat}", flush=True)
    e, it, st = engine.singlepoint_flat_periodic(mol_nat, num, xyz, lat)
    return e, it, st


@pytest.fixture(scope="module")
def refs(orc, cells):
    return [orc.singlepoint(n, x, lattice=l) for n, x, l in cells]


def test_cells_energy_and_charges(engine, orc, cells, run, refs):
    e, it, st = run
    assert list(st) == [0] * len(cells)
    for k, (ref, (num, xyz, lat)) in enumerate(zip(refs, cells)):
        assert ref["rc"] == 0
        assert abs(e[k] - ref["energy"]) < 1e-6, (k, e[k], ref["energy"])
        assert abs(int(it[k]) - ref["iterations"]) <= 1
        nao = orc.dims(num)[2]
        assert engine.molecule_path(k) == (0 if nao <= 96 else 1), (k, nao)
        ch = engine.charges(k)
        assert np.abs(ch["qsh"] - ref["qsh"]).max() < 1e-5
        assert np.abs(ch["qat"] - ref["qat"]).max() < 1e-5
        assert np.abs(ch["dpat"] - ref["dpat"]).max() < 1e-5
        assert np.abs(ch["qpat"] - ref["qpat"]).max() < 1e-5
        assert abs(ch["focc"].sum() - (ch["qsh"].sum() * 0 + np.sum(ref["focc"]))) < 1e-8
        e2, e3 = engine.dispersion(k)
        assert abs(e2 - ref["edisp2"]) < 1e-8
        assert abs(e3 - ref["edisp3"]) < 1e-10


def test_cell_images_bitexact(engine, orc, cells, run):
    for k, (num, xyz, lat) in enumerate(cells):
        inl, nnl, nlat = engine.adjlist(k)
        nltr, tr = engine.images(k)
        ref = orc.pbc_tables(lat, xyz)
        assert tr.shape == ref["trans"].shape
        assert np.array_equal(tr, ref["trans"])
        rinl, rnnl, rnlat, rnltr = orc.adjlist_pbc(xyz, lat, orc.cutoff(num, engine.accuracy))
        assert np.array_equal(inl, rinl) and np.array_equal(nnl, rnnl)
        assert np.array_equal(nlat, rnlat) and np.array_equal(nltr, rnltr)


def test_cell_setup_and_integrals(engine, orc, cells, run):
    for k, (num, xyz, lat) in enumerate(cells):
        t = orc.pbc_terms(num, xyz, lat)
        assert np.abs(engine.cn(k) - t["cn"]).max() < 1e-12
        S, D, Q, H = engine.integrals(k)
        rS, rD, rQ, rH = orc.hamiltonian_pbc(num, xyz, lat, engine.accuracy)
        assert np.abs(S - rS).max() < 1e-11
        assert np.abs(H - rH).max() < 1e-11
        assert np.abs(D - rD).max() < 1e-11
        assert np.abs(Q - rQ).max() < 1e-11


def test_periodic_structure_singlepoint(engine, orc, cells, refs):
    num, xyz, lat = cells[0]
    out = engine.singlepoint(num, xyz, lattice=lat)
    assert out["rc"] == 0
    assert abs(out["energy"] - refs[0]["energy"]) < 1e-6
    assert np.abs(out["qat"] - refs[0]["qat"]).max() < 1e-5
    # the density reproduces the electron count through the Gamma-point overlap
    S, _, _, _ = orc.hamiltonian_pbc(num, xyz, lat)
    nel = float(np.sum(out["density"] * S))
    assert abs(nel - (out["n0at"].sum())) < 1e-8


def test_cell_gradient_and_bad_flags_rejected(engine, cells):
    from tblite_amd import _dp, _ip
    num, xyz, lat = cells[0]
    out = engine.singlepoint(num, xyz, lattice=lat, gradient=True, wavefunction=False)
    assert out["rc"] == -1
    assert "periodic" in engine.error()
    per = np.array([1, 1, 0], dtype=np.int32)
    latf = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    xyzf = np.ascontiguousarray(xyz, dtype=np.float64)
    numi = np.ascontiguousarray(num, dtype=np.int32)
    h = engine.lib.tblite_amd_new_structure_periodic(engine.ctx, len(num), _ip(numi), _dp(xyzf), 0.0,
                                                     0, _dp(latf), _ip(per))
    assert not h
    assert "three-dimensional" in engine.error()


@pytest.mark.parametrize("tuning,path", [({"lds_max_ld": 64}, 1),
                                         ({"lds_max_ld": 64, "grid_min_ld": 64}, 2)])
def test_cell_same_on_every_scf_path(cells, refs, tuning, path):
    import tblite_amd
    num, xyz, lat = cells[1]  # 72 AOs: LDS-resident by default
    eng = tblite_amd.Engine(tuning=dict(auto_grid=0, **tuning))
    try:
        e, it, st = eng.singlepoint_flat_periodic([len(num)], num, xyz, lat[None])
        assert eng.molecule_path(0) == path
    finally:
        eng.close()
    assert st[0] == 0
    assert abs(e[0] - refs[1]["energy"]) < 1e-6
    assert abs(int(it[0]) - refs[1]["iterations"]) <= 1


def test_mixed_batch_of_cells_and_molecules(engine, orc, cells, refs):
    """tblite_amd_xtb_singlepoint_batch over periodic and molecular structure
    handles in one call: each result equals its own reference (cells on the
    cell kernels, molecules on the molecular ones), wavefunctions included."""
    import ctypes as C
    import tblite_amd
    lib = engine.lib
    mol = workloads.config0()[0]
    items = [("cell", cells[0]), ("mol", mol), ("cell", cells[4]), ("mol", mol)]
    handles, wfns = [], []
    try:
        for kind, it in items:
            num = np.ascontiguousarray(it[0], dtype=np.int32)
            xyz = np.ascontiguousarray(it[1], dtype=np.float64)
            if kind == "cell":
                lat = np.ascontiguousarray(it[2], dtype=np.float64).reshape(9)
                per = np.ones(3, dtype=np.int32)
                h = lib.tblite_amd_new_structure_periodic(engine.ctx, len(num), tblite_amd._ip(num),
                                                          tblite_amd._dp(xyz), 0.0, 0,
                                                          tblite_amd._dp(lat), tblite_amd._ip(per))
            else:
                h = lib.tblite_amd_new_structure(engine.ctx, len(num), tblite_amd._ip(num),
                                                 tblite_amd._dp(xyz), 0.0, 0)
            assert h, engine.error()
            handles.append(h)
            wfns.append(lib.tblite_amd_new_wavefunction(engine.ctx))
        n = len(items)
        marr = (C.c_void_p * n)(*handles)
        warr = (C.c_void_p * n)(*wfns)
        e = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        print("[gpu] mixed batch of cells and molecules", flush=True)
        rc = lib.tblite_amd_xtb_singlepoint_batch(engine.ctx, n, marr, engine.calc, warr, 1.0,
                                                  tblite_amd._dp(e), tblite_amd._ip(it),
                                                  tblite_amd._ip(st))
        assert rc == 0, engine.error()
        ref_mol = orc.singlepoint(*mol)
        expect = [refs[0], ref_mol, refs[4], ref_mol]
        for k, r in enumerate(expect):
            assert st[k] == 0
            assert abs(e[k] - r["energy"]) < 1e-6, (k, e[k], r["energy"])
            cnt = C.c_int()
            assert lib.tblite_amd_get_wavefunction_array(wfns[k], b"qat", None, C.byref(cnt)) == 0
            q = np.zeros(cnt.value)
            lib.tblite_amd_get_wavefunction_array(wfns[k], b"qat", tblite_amd._dp(q), C.byref(cnt))
            assert np.abs(q - r["qat"]).max() < 1e-5
    finally:
        for h in handles:
            hh = C.c_void_p(h)
            lib.tblite_amd_delete_structure(C.byref(hh))
        for w in wfns:
            if w:
                ww = C.c_void_p(w)
                lib.tblite_amd_delete_wavefunction(C.byref(ww))


def test_diamond_cell(engine, orc):
    """A dense cell of another element: cubic diamond (8 C, a = 3.567 A,
    9261 translations within the 60 bohr table), against the oracle."""
    a = 3.567 * workloads.AATOAU
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.25, .25, .25],
                     [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    xyz = frac * a
    lat = np.eye(3) * a
    num = np.full(8, 6, dtype=np.int32)
    ref = orc.singlepoint(num, xyz, lattice=lat)
    e, it, st = engine.singlepoint_flat_periodic([8], num, xyz, lat[None])
    assert st[0] == 0 and ref["rc"] == 0
    assert abs(e[0] - ref["energy"]) < 1e-6
    assert abs(int(it[0]) - ref["iterations"]) <= 1
    ch = engine.charges(0)
    assert np.abs(ch["qsh"] - ref["qsh"]).max() < 1e-5
