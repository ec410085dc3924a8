"""Autotuner for the SCF path routing (tblite_amd_tuning).

Measures, on the device at hand:

* the cost of one SCF iteration of a single molecule on one workgroup
  (matrices in global memory) and across the device (multi-workgroup path),
  for water clusters of several sizes; fits wg_a * ld^wg_p and
  grid_b * ld + grid_c (seconds), which the planner's cost model uses to
  send molecules of small batches to the multi-workgroup path when that
  finishes the batch sooner;
* the largest padded AO dimension kept in LDS (lds_max_ld): a batch of
  small water clusters timed with each candidate.

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


def autotune(eng, quick=False, log=print):
    sizes = (40, 100) if quick else (30, 60, 100, 140)
    tuning = fit_costs(path_costs(eng, sizes=sizes, reps=1 if quick else 2, log=log))
    tuning["lds_max_ld"] = 96 if quick else lds_limit(eng, log=log)
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
