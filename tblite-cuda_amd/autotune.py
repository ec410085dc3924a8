"""Autotuner for the SCF path routing (tblite_amd_tuning).

Measures, on the device at hand:

* the cost of one SCF iteration of a single molecule on one workgroup
  (matrices in global memory) and across the device (multi-workgroup path),
  for water clusters of several sizes; fits wg_a * ld^wg_p and
  grid_b * ld + grid_c (seconds), which the planner's cost model uses to
  send molecules of small batches to the multi-workgroup path when that
  finishes the batch sooner;
* the largest padded AO dimension kept in LDS (lds_max_ld): a batch of
  small water clusters timed with each candidate;
* the work tiles (integral_chunk, pair_chunk_log2, atm_chunk_log2,
  fine_buckets, big_lds_min_kb): coordinate descent, each candidate of one
  tile size timed on a mixed batch (small water clusters and organics) with
  the others held, the fastest kept.

    python tblite-cuda_amd/autotune.py --out tuning.json [--quick]

writes the tuning as JSON; ``Engine(tuning="tuning.json")`` (or
``Engine.set_tuning(**json.load(...))``) applies it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import tblite_amd  # noqa: E402
import workloads  # noqa: E402


def _time_one(eng, num, xyz, reps):
    mol_nat = np.array([len(num)])
    eng.singlepoint_flat(mol_nat, num, xyz)  # warm-up (kernel load, workspace)
    best, iters = np.inf, 1
    for _ in range(reps):
        t = time.perf_counter()
        e, it, st = eng.singlepoint_flat(mol_nat, num, xyz)
        best = min(best, time.perf_counter() - t)
        iters = max(1, int(it[0]))
        if st[0] != 0:
            raise RuntimeError(f"calibration molecule failed with status {st[0]}")
    return best / iters, e[0]


def path_costs(eng, sizes=(40, 70, 100, 140), reps=2, seed=9, log=print):
    """Seconds per SCF iteration of one (H2O)n cluster alone on the device,
    on one workgroup and on the multi-workgroup path, for each n."""
    rng = np.random.default_rng(seed)
    base = eng.tuning()
    rows = []
    try:
        for nw in sizes:
            num, xyz = workloads.water_cluster(nw, workloads.liquid_radius(nw), rng)
            ld = (6 * nw + 3) // 4 * 4
            eng.set_tuning(auto_grid=0, grid_min_ld=1024)
            t_wg, e_wg = _time_one(eng, num, xyz, reps)
            eng.set_tuning(auto_grid=0, grid_min_ld=64)
            t_grid, e_grid = _time_one(eng, num, xyz, reps)
            if abs(e_wg - e_grid) > 1e-6:
                raise RuntimeError(f"paths disagree at ld {ld}: {e_wg} vs {e_grid}")
            log(f"[autotune] ld {ld}: one workgroup {t_wg * 1e3:.2f} ms/iteration, "
                f"multi-workgroup {t_grid * 1e3:.2f} ms/iteration")
            rows.append((ld, t_wg, t_grid))
    finally:
        eng.set_tuning(**{k: base[k] for k in ("auto_grid", "grid_min_ld")})
    return rows


def fit_costs(rows):
    """wg_a * ld^wg_p (log-log least squares) and grid_b * ld + grid_c."""
    ld = np.array([r[0] for r in rows], dtype=float)
    twg = np.array([r[1] for r in rows])
    tgr = np.array([r[2] for r in rows])
    p, loga = np.polyfit(np.log(ld), np.log(twg), 1)
    b, c = np.polyfit(ld, tgr, 1)
    return {"wg_a": float(np.exp(loga)), "wg_p": float(p), "grid_b": float(max(b, 0.0)),
            "grid_c": float(c)}


def lds_limit(eng, candidates=(64, 80, 96), count=1024, reps=2, log=print):
    """The lds_max_ld candidate with the fastest batch of small water clusters."""
    rng = np.random.default_rng(3)
    mols = []
    for _ in range(count):
        n = int(rng.integers(2, 17))
        mols.append(workloads.water_cluster(n, workloads.liquid_radius(n), rng))
    mol_nat, num, xyz = workloads.flatten(mols)
    base = eng.tuning()["lds_max_ld"]
    best = None
    try:
        for cand in candidates:
            eng.set_tuning(lds_max_ld=cand)
            eng.prepare(mol_nat, num)
            eng.set_geometry(xyz)
            eng.run()
            ts = []
            for _ in range(reps):
                t = time.perf_counter()
                eng.run()
                ts.append(time.perf_counter() - t)
            log(f"[autotune] lds_max_ld {cand}: {min(ts) * 1e3:.1f} ms for {count} molecules")
            if best is None or min(ts) < best[1]:
                best = (cand, min(ts))
    finally:
        eng.set_tuning(lds_max_ld=base)
    return best[0]


TILE_CANDIDATES = {
    "integral_chunk": (64, 128, 32),
    "pair_chunk_log2": (15, 13, 17),
    "atm_chunk_log2": (18, 16, 20),
    "fine_buckets": (0, 1),
    "big_lds_min_kb": (0, 96),
}


def tile_batch(count=768, seed=17):
    """Mixed batch for the tile search: water clusters of 2-16 molecules and
    organics of 4-40 heavy atoms (LDS, global-memory and three-body work)."""
    rng = np.random.default_rng(seed)
    mols = []
    for k in range(count):
        if k % 2 == 0:
            n = int(rng.integers(2, 17))
            mols.append(workloads.water_cluster(n, workloads.liquid_radius(n), rng))
        else:
            mols.append(workloads.organic_molecule(int(rng.integers(4, 41)), rng))
    return workloads.flatten(mols)


def _time_batch(eng, batch, reps):
    mol_nat, num, xyz = batch
    eng.prepare(mol_nat, num)
    eng.set_geometry(xyz)
    eng.run()
    e0, it0, st0 = eng.results()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        eng.run()
        ts.append(time.perf_counter() - t)
    return min(ts), e0


def tile_sizes(eng, batch=None, reps=2, min_gain=0.02, log=print):
    """Coordinate descent over the work tiles; every candidate must give the
    energies of the first (1e-9 Eh: only summation orders change), and
    replaces the current setting only when it is faster by more than
    min_gain (run-to-run noise is about 1 %; one molecule per CU on the
    global-memory path measured 0.3 % faster on 768 molecules and 48 %
    slower on a device-filling batch of 2048)."""
    batch = batch if batch is not None else tile_batch()
    base = eng.tuning()
    best = {k: base[k] for k in TILE_CANDIDATES}
    eng.set_tuning(**best)
    t_best, e_ref = _time_batch(eng, batch, reps)
    log(f"[autotune] tiles {best}: {t_best * 1e3:.1f} ms")
    try:
        for name, cands in TILE_CANDIDATES.items():
            for c in cands:
                if c == best[name]:
                    continue
                trial = dict(best, **{name: c})
                eng.set_tuning(**trial)
                t, e = _time_batch(eng, batch, reps)
                de = float(np.max(np.abs(e - e_ref)))
                if de > 1e-9:
                    raise RuntimeError(f"tile {name}={c} changed energies by {de:.2e} Eh")
                log(f"[autotune] {name}={c}: {t * 1e3:.1f} ms (best {t_best * 1e3:.1f})")
                if t < t_best * (1.0 - min_gain):
                    best, t_best = trial, t
    finally:
        eng.set_tuning(**best)
    return best


def autotune(eng, quick=False, log=print):
    sizes = (40, 100) if quick else (30, 60, 100, 140)
    tuning = fit_costs(path_costs(eng, sizes=sizes, reps=1 if quick else 2, log=log))
    tuning["lds_max_ld"] = 96 if quick else lds_limit(eng, log=log)
    eng.set_tuning(**tuning)
    tuning.update(tile_sizes(eng, batch=tile_batch(count=256 if quick else 2048),
                             reps=1 if quick else 2, log=log))
    return tuning


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--quick", action="store_true", help="two calibration sizes, no LDS sweep")
    args = ap.parse_args()
    eng = tblite_amd.Engine()
    t = autotune(eng, quick=args.quick)
    with open(args.out, "w") as f:
        json.dump(t, f, indent=1)
    print(json.dumps(t))


if __name__ == "__main__":
    main()
