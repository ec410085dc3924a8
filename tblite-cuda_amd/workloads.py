"""Seeded synthetic molecules for the singlepoint workloads.

All generators work in angstrom and return bohr coordinates (one conversion,
1 bohr = 0.52917721090380 angstrom) as float64 arrays of shape (nat, 3), with
integer atomic numbers.  Every workload is fully determined by its seed
(numpy PCG64), so CPU and GPU runs consume identical bits.

Workloads (sizes as used by the tests and the benchmark):
  water_cluster(n, radius, rng)  rigid waters, O uniform in a sphere, O-O >= 2.6 A
  config0()                      one (H2O)8 in a 4.5 A sphere, seed 0
  config1(count=4096)            (H2O)n, n ~ U{2..16}, liquid-density radius, seed 1
  config2()                      one (H2O)700 at liquid density, seed 2
  config3(count=2048)            tree-grown C/N/O/S organics, 4..60 heavy atoms, seed 3
  cells(count=512)               periodic water cells, n ~ U{4..16} molecules
                                 at liquid density, half of them triclinic, seed 8
  config4(count=20000)           request stream, 50 % water / 50 % organic,
                                 atoms log-uniform in 6..400, request k from
                                 child k of SeedSequence(4)
"""
import numpy as np

AUTOAA = 0.52917721090380
AATOAU = 1.0 / AUTOAA
ROH = 0.9572
HOH = np.deg2rad(104.52)
LIQUID_DENSITY = 0.0334  # molecules / A^3


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _water_frame():
    s, c = np.sin(HOH / 2), np.cos(HOH / 2)
    return np.array([[0.0, 0.0, 0.0], [ROH * s, 0.0, ROH * c], [-ROH * s, 0.0, ROH * c]])


def liquid_radius(n):
    return (3.0 * n / (4.0 * np.pi * LIQUID_DENSITY)) ** (1.0 / 3.0)


def _sphere_points(n, radius, dmin, rng, max_tries=200000):
    """Random sequential addition in a sphere with a cell list for the
    minimum-distance test; restarts from scratch if it jams."""
    cell = dmin
    for _ in range(50):
        pts = []
        grid = {}
        tries = 0
        while len(pts) < n and tries < max_tries:
            tries += 1
            p = rng.uniform(-radius, radius, 3)
            if p @ p > radius * radius:
                continue
            key = tuple(np.floor(p / cell).astype(int))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            d = p - pts[j]
                            if d @ d < dmin * dmin:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                grid.setdefault(key, []).append(len(pts))
                pts.append(p)
        if len(pts) == n:
            return np.array(pts)
    raise RuntimeError("sphere packing jammed")


def water_cluster(n, radius, rng):
    """(H2O)n: returns (num, xyz_bohr)."""
    opos = _sphere_points(n, radius, 2.6, rng)
    frame = _water_frame()
    xyz = []
    num = []
    for i in range(n):
        rot = _random_rotation(rng)
        xyz.append(frame @ rot.T + opos[i])
        num += [8, 1, 1]
    return np.array(num, dtype=np.int32), np.concatenate(xyz) * AATOAU


VALENCE = {6: 4, 7: 3, 8: 2, 16: 2}
VALENCE_ELECTRONS = {1: 1, 6: 4, 7: 5, 8: 6, 16: 6}


def _unit(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def organic_molecule(nheavy, rng, max_restart=1000, h_tries=200, allow_open=False):
    """Tree-grown molecule: attach C/N/O/S (0.70/0.15/0.10/0.05) at 1.50 A to a
    random heavy atom with free valence, reject pairs closer than 1.2 A, then
    saturate the free valences with H at 1.09 A (same rejection).  Positions
    live in a preallocated array; the random draws are the same as a
    list-based implementation (same calls, same order)."""
    elems = [6, 7, 8, 16]
    cdf = np.cumsum([0.70, 0.15, 0.10, 0.05])
    cdf /= cdf[-1]

    def pick():
        # same draw as rng.choice(elems, p=prob): one uniform, right-sided search
        return elems[int(np.searchsorted(cdf, rng.random(), side="right"))]

    cap = 5 * nheavy + 8
    for _ in range(max_restart):
        num = [pick()]
        pos = np.zeros((cap, 3))
        npos = 1
        free = [VALENCE[num[0]]]
        failed = False
        while len(num) < nheavy:
            cand = [i for i in range(len(num)) if free[i] > 0]
            if not cand:
                failed = True
                break
            z = pick()
            placed = False
            for _try in range(100):
                host = cand[rng.integers(len(cand))]
                p = pos[host] + 1.50 * _unit(rng)
                dv = pos[:npos] - p
                d2 = np.einsum("ij,ij->i", dv, dv)
                d2[host] = 100.0
                if d2.min() >= 1.44:
                    placed = True
                    break
            if not placed:
                failed = True
                break
            num.append(z)
            pos[npos] = p
            npos += 1
            free.append(VALENCE[z] - 1)
            free[host] -= 1
        if failed:
            continue
        for i in range(nheavy):
            while free[i] > 0 and not failed:
                for _try in range(h_tries):
                    p = pos[i] + 1.09 * _unit(rng)
                    dv = pos[:npos] - p
                    d2 = np.einsum("ij,ij->i", dv, dv)
                    d2[i] = 100.0
                    if d2.min() >= 1.44:
                        break
                else:
                    if allow_open:
                        # no room for this hydrogen: the valence stays open
                        free[i] -= 1
                        continue
                    failed = True
                    break
                num.append(1)
                pos[npos] = p
                npos += 1
                free[i] -= 1
        if failed:
            continue
        if allow_open and sum(VALENCE_ELECTRONS[z] for z in num) % 2 == 1:
            # closed shell: drop the last hydrogen to make the electron count even
            num.pop()
            npos -= 1
        return np.array(num, dtype=np.int32), pos[:npos].copy() * AATOAU
    raise RuntimeError("organic growth failed")


def config0():
    rng = np.random.default_rng(0)
    return [water_cluster(8, 4.5, rng)]


def config1(count=4096):
    rng = np.random.default_rng(1)
    out = []
    for _ in range(count):
        n = int(rng.integers(2, 17))
        out.append(water_cluster(n, liquid_radius(n), rng))
    return out


def config2():
    rng = np.random.default_rng(2)
    return [water_cluster(700, liquid_radius(700), rng)]


def config3(count=2048):
    rng = np.random.default_rng(3)
    out = []
    for _ in range(count):
        nh = int(rng.integers(4, 61))
        out.append(organic_molecule(nh, rng))
    return out


def _config4_request(seed_seq):
    """One request of the serving stream from its own seed sequence."""
    rng = np.random.default_rng(seed_seq)
    natoms = int(np.round(np.exp(rng.uniform(np.log(6), np.log(400)))))
    if rng.random() < 0.5:
        n = max(2, int(round(natoms / 3)))
        return water_cluster(n, liquid_radius(n), rng)
    # grow until heavy + H reaches natoms (heavy ~ natoms / 2.55).  Large
    # trees are crowded: a hydrogen gets 2000 placement attempts, a valence
    # with no room is left open, and an odd electron count is made even by
    # dropping the last hydrogen (closed shell)
    nh = max(2, int(round((natoms - 2) / 2.55)))
    return organic_molecule(nh, rng, h_tries=2000, allow_open=True)


def config4(count=20000, workers=1):
    """Request k is drawn from the k-th child of SeedSequence(4), so the
    stream is the same whatever the number of worker processes."""
    children = np.random.SeedSequence(4).spawn(count)
    if workers <= 1:
        return [_config4_request(c) for c in children]
    import multiprocessing

    with multiprocessing.get_context("fork").Pool(workers) as pool:
        return pool.map(_config4_request, children, chunksize=16)


def flatten(mols):
    """List of (num, xyz) -> (mol_nat, num, xyz) flat arrays."""
    mol_nat = np.array([len(m[0]) for m in mols], dtype=np.int32)
    num = np.concatenate([m[0] for m in mols]).astype(np.int32)
    xyz = np.ascontiguousarray(np.concatenate([m[1] for m in mols]), dtype=np.float64)
    return mol_nat, num, xyz


def water_box(n, rng, density=LIQUID_DENSITY, shear=0.0):
    """Periodic cell of n rigid waters at the given number density
    (molecules / A^3): O uniform in the cell with the O-O minimum image at
    least 2.6 A, uniform orientations.  The cell is cubic, or sheared by
    `shear` (a_2 gains shear * L along x, a_3 gains shear * L along y) for a
    triclinic case.  Returns (num, xyz_bohr, lattice_bohr) with
    lattice[k] = k-th lattice vector."""
    L = (n / density) ** (1.0 / 3.0)
    lat = np.array([[L, 0.0, 0.0], [shear * L, L, 0.0], [0.0, shear * L, L]])
    inv = np.linalg.inv(lat)
    pts = []
    for _ in range(200000):
        if len(pts) == n:
            break
        f = rng.uniform(0.0, 1.0, 3)
        p = f @ lat
        ok = True
        for q in pts:
            d = (p - q) @ inv
            d -= np.round(d)
            d = d @ lat
            if d @ d < 2.6 * 2.6:
                ok = False
                break
        if ok:
            pts.append(p)
    if len(pts) < n:
        raise RuntimeError("water_box: could not place the molecules")
    frame = _water_frame()
    num, xyz = [], []
    for p in pts:
        R = _random_rotation(rng)
        for k, z in enumerate((8, 1, 1)):
            num.append(z)
            xyz.append(p + frame[k] @ R.T)
    return (np.array(num, dtype=np.int32), np.array(xyz) * AATOAU, lat * AATOAU)


def periodic_cells(count=4, seed=5, first=4):
    """Seeded water cells: first, first + 2, ... molecules at liquid density,
    every other one triclinic (shear 0.15)."""
    rng = np.random.default_rng(seed)
    return [water_box(first + 2 * k, rng, shear=0.15 if k % 2 else 0.0) for k in range(count)]


def cells(count=512, seed=8):
    """Seeded periodic water cells: n ~ U{4..16} molecules at liquid density,
    every other cell triclinic (shear 0.15)."""
    rng = np.random.default_rng(seed)
    return [water_box(int(rng.integers(4, 17)), rng, shear=0.15 if k % 2 else 0.0)
            for k in range(count)]
