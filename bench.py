#!/usr/bin/env python
"""Batched GFN2-xTB singlepoint throughput on one MI355X.

A step is one full singlepoint pass (CN, repulsion, Coulomb matrices,
neighbour list, integrals/H0, SCF to convergence) over the whole batch, with
the coordinates already resident in HBM; the batch is the workload's seeded
molecules (default: config 3, 2048 seeded tree-grown C/N/O/S organics with
4..60 heavy atoms, the workload the headline metric is quoted on).  Prints one
JSON line.  Beside the GPU number it reports:
  roofline      FP64 work of the dominant SCF kernel (the kernel instantiation
                carrying the most work) / its launches' mean duration from HIP
                events recorded on their streams
  cpu_baseline  the CPU oracle (FP64 restatement) on bounded samples of the
                same molecules, timed on this host: OpenMP over molecules on
                all cores (value) and a per-molecule loop on one thread
  per_molecule  the plain one-molecule-per-call path (tblite_amd_xtb_singlepoint)
                on an evenly spaced sample, and the batched/per-molecule ratio
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tblite-cuda_amd"))

import tblite_amd  # noqa: E402
import workloads  # noqa: E402

PEAK_FP64 = 78.6e12  # MI355X dense FP64 (matrix and vector), FLOP/s

# HBM bytes per launch of the dominant SCF kernel (fetched + written), from
# two rocprofv3 counter runs on the same workload (FETCH_SIZE, WRITE_SIZE;
# FETCH_SIZE scaled x2 and WRITE_SIZE x1 by a 1 GiB device-copy calibration;
# mean over the kernel's launches in one batch).  Summaries with the full
# commands: profiles/k_scf_config1.md, profiles/k_scf_config3.md.
MEASURED_TRAFFIC = {
    ("config1", "k_scf<512, 2, 3, 6, true>"): 16.59e9 + 0.93e9,
    ("config3", "k_scf<512, 1, 0, 0, false>"): 1873.0e9 + 946.8e9,
}

WORKLOADS = {
    "config1": ("4096 seeded (H2O)n clusters, n ~ U{2..16}, liquid density", workloads.config1),
    "config0": ("one seeded (H2O)8 cluster", workloads.config0),
    "config2": ("one seeded (H2O)700 cluster at liquid density (4200 AOs, multi-workgroup path)",
                workloads.config2),
    "config3": ("2048 seeded tree-grown C/N/O/S organics, 4..60 heavy atoms", workloads.config3),
    "config4": ("serving: seeded requests, 50% water clusters / 50% organics, 6..400 atoms "
                "log-uniform, Poisson arrivals", workloads.config4),
}


def scf_flops(nao, nocc, iters):
    """Algorithmic FP64 flops of the SCF per molecule: per iteration
    (16/3) n^3 + 2 n^2 o + 40 n^2 (reduction, tridiagonalisation,
    back-transforms, density, Fock, Mulliken), plus n^3/3 once for the
    factorisation of S; o = MOs with occupation above 1e-12."""
    n = nao.astype(np.float64)
    o = nocc.astype(np.float64)
    per_iter = (16.0 / 3.0) * n**3 + 2.0 * n**2 * o + 40.0 * n**2
    return iters * per_iter + n**3 / 3.0


def scf_kernel_name(path, nthreads):
    """Kernel instantiation of an SCF bucket from its molecules' path (0: LDS
    matrices, 1: global-memory matrices) and its workgroup size."""
    if path == 1:
        return ("k_scf<512, 1, 0, 0, false> (global-memory matrices, two-stage band reduction "
                "to 512 AOs)")
    maxt = 2 if nthreads == 512 else 1
    maxb, maxc = (2, 4) if nthreads == 128 else (3, 6)
    return f"k_scf<{nthreads}, {maxt}, {maxb}, {maxc}, true> (LDS matrices, Jacobi)"


def bucket_members(buckets, ld, path):
    """Molecule indices of each SCF bucket.  The engine's buckets cover
    disjoint ranges of the padded AO dimension and are listed largest first,
    so bucket k holds the one-workgroup molecules (path 0 or 1) with
    ldmax(k+1) < ld <= ldmax(k)."""
    out = []
    for k, bk in enumerate(buckets):
        lo = buckets[k + 1]["ldmax"] if k + 1 < len(buckets) else 0
        sel = np.nonzero((ld > lo) & (ld <= bk["ldmax"]) & (path != 2))[0]
        if len(sel) != bk["nmol"]:
            raise RuntimeError(f"bucket {k}: {len(sel)} molecules in ({lo}, {bk['ldmax']}] "
                               f"but the engine launched {bk['nmol']}")
        out.append(sel)
    return out


def copy_probe(mib, reps=3):
    """Plain device copy of a known size (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)."""
    import torch

    n = mib * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "reps": reps}))


def replay(srv, mols, arrivals, tag=""):
    """Submit mols[k] at host time arrivals[k] (s from start), poll until all are
    done; returns per-request latency (ms, arrival to collection) and the span
    from the first arrival to the last completion (s).  Prints a progress line
    about every 20 s."""
    n = len(mols)
    tickets = {}
    lat = np.zeros(n)
    done = 0
    k = 0
    t0 = time.perf_counter()
    t_report = t0
    while done < n:
        now = time.perf_counter() - t0
        while k < n and arrivals[k] <= now:
            tickets[srv.submit(*mols[k])] = k
            k += 1
        res = srv.poll()
        tnow = time.perf_counter() - t0
        for t, e, it, st, _ in res:
            m = tickets.pop(t)
            lat[m] = (tnow - arrivals[m]) * 1e3
            done += 1
        if tnow + t0 - t_report > 20.0:
            t_report = tnow + t0
            print(f"[bench] {tag} {tnow:.0f} s: {k}/{n} submitted, {done} done", file=sys.stderr,
                  flush=True)
        if not res and k < n and arrivals[k] > tnow + 0.0005:
            time.sleep(min(arrivals[k] - tnow, 0.0005))
    return lat, time.perf_counter() - t0 - arrivals[0]


AO_PER_ELEMENT = {1: 1, 6: 4, 7: 4, 8: 4, 16: 9}


def latency_row(lat, mols, offered, span, arrivals):
    """p50 / p99 over all requests and per size class (LDS-resident sizes,
    padded AO dimension <= 96, and larger), and the drain time after the last
    arrival (close to the slowest latency when the server kept up)."""
    nao = np.array([sum(AO_PER_ELEMENT[int(z)] for z in m[0]) for m in mols])
    small = ((nao + 3) // 4 * 4) <= 96
    pct = lambda a, q: round(float(np.percentile(a, q)), 2) if len(a) else None
    return {"offered_per_s": round(float(offered), 2), "requests": len(mols),
            "throughput_per_s": round(len(mols) / span, 2),
            "arrival_window_s": round(float(arrivals[-1] - arrivals[0]), 2),
            "drain_s": round(float(span - (arrivals[-1] - arrivals[0])), 2),
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "small_p50_ms": pct(lat[small], 50), "small_p99_ms": pct(lat[small], 99),
            "large_p50_ms": pct(lat[~small], 50), "large_p99_ms": pct(lat[~small], 99)}


def serve_bench(args):
    """Config 4: closed-loop saturation throughput, then open-loop Poisson
    arrivals at fractions of it (or at the rates given) with p50 / p99
    latency, beside the plain one-molecule-per-call loop."""
    t0 = time.perf_counter()
    workers = min(16, os.cpu_count() or 1)
    mols = workloads.config4(args.requests, workers=workers)
    t_gen = time.perf_counter() - t0
    nat = np.array([len(m[0]) for m in mols])
    print(f"[bench] config4: {len(mols)} requests, atoms {nat.min()}..{nat.max()} "
          f"(generated in {t_gen:.1f} s on {workers} processes)", file=sys.stderr, flush=True)
    eng = tblite_amd.Engine(device=0)
    srv = tblite_amd.Server(eng, slots=args.slots, workspace_bytes=args.workspace_gb * 1e9)
    # warm-up on a few requests (kernel loading, workspace first touch)
    replay(srv, mols[:16], np.zeros(16), "warm-up")
    sat = None
    if args.saturation is not None:
        sat = args.saturation
    else:
        lat, span = replay(srv, mols, np.zeros(len(mols)), "closed loop")
        sat = len(mols) / span
        print(f"[bench] saturation: {sat:.1f} singlepoints/s (all {len(mols)} queued at t=0), "
              f"p50 {np.percentile(lat, 50):.0f} ms p99 {np.percentile(lat, 99):.0f} ms",
              file=sys.stderr, flush=True)
    rng = np.random.default_rng(40)
    sweep = []
    if args.rates:
        rates = [float(r) for r in args.rates.split(",") if r]
        fracs = [r / sat for r in rates]
    else:
        fracs = [float(f) for f in args.fractions.split(",") if f] if args.fractions else []
    for frac in fracs:
        rate = frac * sat
        m = min(len(mols), max(100, int(rate * args.seconds_per_rate)))
        arr = np.cumsum(rng.exponential(1.0 / rate, m))
        arr -= arr[0]
        lat, span = replay(srv, mols[:m], arr, f"rate {rate:.1f}/s")
        row = latency_row(lat, mols[:m], rate, span, arr)
        row["fraction_of_saturation"] = frac
        sweep.append(row)
        print(f"[bench] rate {rate:.1f}/s: {row}", file=sys.stderr, flush=True)
    stats = srv.stats()
    srv.close()
    k = min(args.per_molecule_sample, len(mols))
    per_mol = None
    if k > 0:
        pick = _spread(len(mols), k)
        t = time.perf_counter()
        for m in pick:
            eng.singlepoint(mols[m][0], mols[m][1], wavefunction=False)
        per_mol = len(pick) / (time.perf_counter() - t)
    best = sweep[-1] if sweep else {"p50_ms": None, "p99_ms": None, "offered_per_s": None}
    line = {
        "metric": "singlepoints/s + p99 latency, batched vs per-molecule, |dE|<=1e-6 Eh",
        "value": round(sat, 2),
        "unit": "singlepoints/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(1e3 * len(mols) / sat, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"config4: {WORKLOADS['config4'][0]}", "requests": len(mols),
                   "slots": args.slots, "workspace_GB": args.workspace_gb,
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)},
        "latency_ms": {"p50": best["p50_ms"], "p99": best["p99_ms"],
                       "at_offered_per_s": best["offered_per_s"]},
        "roofline": None,
        "cpu_baseline": None,
        "extra": {"sweep": sweep, "server": stats, "saturation_given": args.saturation is not None,
                  "per_molecule_path": {"value": round(per_mol, 2) if per_mol else None,
                                        "unit": "singlepoints/s", "sample": k,
                                        "speedup_served": round(sat / per_mol, 2) if per_mol else None}},
    }
    print(json.dumps(line))


def _spread(nmol, count):
    """count molecule indices evenly spaced over the batch."""
    return np.unique(np.linspace(0, nmol - 1, max(1, count)).round().astype(int))


def cpu_baseline(mols, seconds, threads):
    """The CPU oracle timed on this host on bounded, evenly spaced samples of
    the workload: (ii) OpenMP over molecules on `threads` cores (the reported
    value) and (i) a per-molecule loop on one thread.  Each sample is sized
    from a short calibration run to about `seconds` of CPU work.  Returns the
    indices and results of the larger sample (for the parity check) and the
    baseline record."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    orc = oracle.Oracle()

    def timed(idx, nthreads):
        t = time.perf_counter()
        e, it, st = orc.singlepoint_many([mols[m] for m in idx], nthreads=nthreads)
        return e, it, time.perf_counter() - t

    def sized(nthreads, secs):
        probe = _spread(len(mols), 2 * nthreads)
        _, _, dt = timed(probe, nthreads)
        count = int(min(len(mols), max(len(probe), len(probe) * secs / max(dt, 1e-3))))
        idx = _spread(len(mols), count)
        e, it, dt = timed(idx, nthreads)
        return idx, e, it, dt

    idx, e, it, dt = sized(threads, seconds)
    idx1, _, _, dt1 = sized(1, seconds / 2)
    rec = {
        "value": round(len(idx) / dt, 3),
        "unit": "singlepoints/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(idx)} evenly spaced molecules of the same workload, CPU oracle (FP64 C "
                  f"restatement of tblite's singlepoint, not tblite), OpenMP over molecules, "
                  f"{dt:.1f} s",
        "single_thread": {"value": round(len(idx1) / dt1, 3), "unit": "singlepoints/s", "cores": 1,
                          "sample": f"{len(idx1)} evenly spaced molecules, per-molecule loop on "
                                    f"one thread, {dt1:.1f} s"},
    }
    return idx, e, it, rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="config3", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--per-molecule-sample", type=int, default=64)
    ap.add_argument("--requests", type=int, default=1000, help="config4: number of requests")
    ap.add_argument("--slots", type=int, default=4, help="config4: concurrent batches")
    ap.add_argument("--workspace-gb", type=float, default=48.0, help="config4: device workspace")
    ap.add_argument("--seconds-per-rate", type=float, default=15.0,
                    help="config4: open-loop requests per rate = rate x this (at most --requests)")
    ap.add_argument("--fractions", default="0.25,0.5,0.8",
                    help="config4: open-loop rates as fractions of the saturation rate")
    ap.add_argument("--rates", default="",
                    help="config4: open-loop rates in singlepoints/s (instead of --fractions)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="config4: hardware queues of the process (GPU_MAX_HW_QUEUES, read when "
                         "the HIP runtime starts; 0 leaves the environment alone); the server "
                         "runs one stream per slot")
    ap.add_argument("--saturation", type=float, default=None,
                    help="config4: saturation rate (singlepoints/s) from an earlier run; skips "
                         "the closed-loop measurement")
    ap.add_argument("--gradient-steps", type=int, default=2,
                    help="timed steps of singlepoint + analytic gradient (0: skip)")
    ap.add_argument("--cell-steps", type=int, default=2,
                    help="timed steps of the periodic-cell leg (512 water cells; 0: skip)")
    ap.add_argument("--cells-only", action="store_true",
                    help="run only the periodic-cell leg and print its JSON")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); molecules are independent, "
              "so a multi-GPU run is a split of the batch with no exchange", file=sys.stderr)
        sys.exit(2)

    if args.cells_only:
        eng = tblite_amd.Engine(device=0)
        print(json.dumps(cell_leg(eng, max(1, args.cell_steps), not args.no_cpu)))
        eng.close()
        return
    if args.workload == "config4":
        if args.hw_queues > 0:
            os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, args.hw_queues))
        serve_bench(args)
        return
    desc, gen = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    mols = gen()
    t_gen = time.perf_counter() - t0
    mol_nat, num, xyz = workloads.flatten(mols)
    nmol = len(mols)
    print(f"[bench] {args.workload}: {nmol} molecules, {len(num)} atoms (generated in {t_gen:.1f} s)",
          file=sys.stderr, flush=True)

    eng = tblite_amd.Engine(device=0)
    t0 = time.perf_counter()
    eng.prepare(mol_nat, num)
    eng.set_geometry(xyz)
    t_prep = time.perf_counter() - t0

    for _ in range(args.warmup):
        eng.run()
    step_ms = []
    phase = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        eng.run()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        phase.append(eng.timings())
    t_total = time.perf_counter() - t_start
    e, it, st = eng.results()
    phase = np.array(phase)
    ms_per_step = t_total / args.steps * 1e3
    value = nmol / (t_total / args.steps)
    print(f"[bench] {ms_per_step:.2f} ms/step, {value:.1f} singlepoints/s; phases (ms): "
          f"setup {phase[:, 0].mean():.3f} adjlist {phase[:, 1].mean():.3f} "
          f"integrals {phase[:, 2].mean():.3f} scf {phase[:, 3].mean():.3f}",
          file=sys.stderr, flush=True)

    # algorithmic FP64 work of the SCF, per molecule; roofline of the dominant kernel
    nao = np.array([eng.dims(m)[2] for m in range(nmol)])
    nocc = np.array([int(np.count_nonzero(eng.charges(m)["focc"] > 1e-12)) for m in range(nmol)])
    fl = scf_flops(nao, nocc, it.astype(np.float64))
    ld = (nao + 3) & ~3
    path = np.array([eng.molecule_path(m) for m in range(nmol)])
    buckets = eng.bucket_timings()
    members = bucket_members(buckets, ld, path)
    per_bucket = []
    for bk, sel in zip(buckets, members):
        f = float(fl[sel].sum())
        per_bucket.append({"kernel": scf_kernel_name(int(path[sel[0]]), bk["nthreads"]),
                           "ld_range": [int(ld[sel].min()), int(ld[sel].max())],
                           "molecules": int(len(sel)), "ms": round(bk["ms"], 3), "flops": f,
                           "tflops": round(f / (bk["ms"] / 1e3) / 1e12, 4) if bk["ms"] > 0 else 0.0})
    # the dominant kernel is the instantiation that carries the most FP64 work
    # (buckets run concurrently on separate streams, so a small bucket's wall
    # time can stretch while it waits for CUs held by a large one); buckets
    # sharing an instantiation are launches of the same kernel, so its average
    # launch duration is the mean over them (what rocprofv3 reports)
    kern = {}
    for b in per_bucket:
        k = kern.setdefault(b["kernel"], {"ms": [], "flops": []})
        k["ms"].append(b["ms"])
        k["flops"].append(b["flops"])
    big = path == 2  # multi-workgroup molecules: a kernel sequence on the main stream
    if big.any():
        kern["multi-workgroup SCF (kernel sequence, main stream)"] = {
            "ms": [float(phase[:, 3].mean())], "flops": [float(fl[big].sum())]}
    dname, dk = max(kern.items(), key=lambda kv: sum(kv[1]["flops"]))
    avg_ms = float(np.mean(dk["ms"]))
    avg_flops = float(np.mean(dk["flops"]))
    achieved = avg_flops / (avg_ms / 1e3)
    scf_s = phase[:, 3].mean() / 1e3
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved / 1e12, 4),
        "peak": PEAK_FP64 / 1e12,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64, 5),
        "traffic": next((v for (w, k), v in MEASURED_TRAFFIC.items()
                         if w == args.workload and dname.startswith(k)), None),
        "kernel": dname,
        "launches_per_step": len(dk["ms"]),
        "avg_launch_ms": round(avg_ms, 3),
        "flops_per_launch": avg_flops,
        "all_scf_tflops": round(float(fl.sum()) / scf_s / 1e12, 4),
        "buckets": per_bucket,
    }

    # plain one-molecule-per-call path (tblite_amd_xtb_singlepoint) on an
    # evenly spaced sample of the batch
    k = min(args.per_molecule_sample, nmol)
    pick = _spread(nmol, k) if k > 0 else []
    k = len(pick)
    t = time.perf_counter()
    for m in pick:
        r = eng.singlepoint(mols[m][0], mols[m][1], wavefunction=False)
        if r["rc"] != 0:
            raise RuntimeError(f"per-molecule path: molecule {m} returned {r['rc']}")
    per_mol = k / (time.perf_counter() - t) if k > 0 else float("nan")
    extra = {
        "prepare_s": round(t_prep, 3),
        "step_ms_p50": round(float(np.percentile(step_ms, 50)), 3),
        "step_ms_p99": round(float(np.percentile(step_ms, 99)), 3),
        "phase_ms": {name: round(float(phase[:, i].mean()), 3)
                     for i, name in enumerate(["setup", "adjlist", "integrals", "scf", "total"])},
        "scf_iterations_mean": round(float(it.mean()), 3),
        "status_nonzero": int(np.count_nonzero(st)),
        "per_molecule_path": {"value": round(per_mol, 2) if k > 0 else None,
                              "unit": "singlepoints/s", "sample": k,
                              "speedup_batched": round(value / per_mol, 2) if k > 0 else None},
    }

    # singlepoint + analytic gradient of the same batch (device-resident)
    if args.gradient_steps > 0:
        eng.prepare(mol_nat, num, gradient=True)
        eng.set_geometry(xyz)
        eng.run_gradient()
        t = time.perf_counter()
        for _ in range(args.gradient_steps):
            eng.run_gradient()
        g_ms = (time.perf_counter() - t) / args.gradient_steps * 1e3
        g, sg = eng.gradients()
        e_g, it_g, st_g = eng.results()
        extra["gradient"] = {
            "value": round(nmol / (g_ms / 1e3), 2), "unit": "singlepoints with gradients/s",
            "ms_per_step": round(g_ms, 3), "gradient_ms": round(g_ms - ms_per_step, 3),
            "steps": args.gradient_steps,
            "max_abs_dE_vs_energy_only": float(np.abs(e_g - e).max()),
            "max_abs_sum_g": float(max(np.abs(g[o:o + n].sum(0)).max()
                                       for o, n in zip(np.cumsum(mol_nat) - mol_nat, mol_nat)))}
        print(f"[bench] singlepoint + gradient: {g_ms:.2f} ms/step "
              f"({nmol / (g_ms / 1e3):.1f}/s)", file=sys.stderr, flush=True)

    if args.cell_steps > 0:
        extra["periodic"] = cell_leg(eng, args.cell_steps, not args.no_cpu)

    cpu = None
    if args.workload == "config2":
        # one 4200-AO molecule: the CPU oracle needs minutes per SCF iteration,
        # beyond a bounded baseline sample
        args.no_cpu = True
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        idx, e_cpu, it_cpu, cpu = cpu_baseline(mols, args.cpu_seconds, threads)
        extra["parity_sample"] = {
            "molecules": int(len(idx)),
            "max_abs_dE": float(np.abs(e[idx] - e_cpu).max()),
            "max_abs_diter": int(np.abs(it[idx].astype(int) - np.asarray(it_cpu, dtype=int)).max())}
        cpu["cpu_model"] = _cpu_model()

    line = {
        "metric": "singlepoints/s + p99 latency, batched vs per-molecule, |dE|<=1e-6 Eh",
        "value": round(value, 2),
        "unit": "singlepoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "molecules": nmol,
                   "atoms": int(len(num)), "accuracy": 1.0, "etemp_K": 300.0,
                   "max_iter": 250, "mixer_damping": 0.4},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


def cell_leg(eng, steps, with_cpu):
    """Periodic water cells (Gamma point): one flat call per step, timed on
    the host around the call (plan, uploads and result copies included) and
    on the device (phase events); a parity and CPU sample against the oracle."""
    cl = workloads.cells()
    mol_nat = [len(c[0]) for c in cl]
    num = np.concatenate([c[0] for c in cl])
    xyz = np.concatenate([c[1] for c in cl])
    lat = np.stack([c[2] for c in cl])
    e, it, st = eng.singlepoint_flat_periodic(mol_nat, num, xyz, lat)
    ms, dev = [], []
    for _ in range(steps):
        t = time.perf_counter()
        e, it, st = eng.singlepoint_flat_periodic(mol_nat, num, xyz, lat)
        ms.append((time.perf_counter() - t) * 1e3)
        dev.append(float(eng.timings()[4]))
    paths = [eng.molecule_path(k) for k in range(len(cl))]
    out = {"value": round(len(cl) / (np.mean(ms) / 1e3), 2), "unit": "cells/s",
           "workload": f"{len(cl)} water cells, 4..16 molecules at liquid density, "
                       "half triclinic, seed 8",
           "ms_per_step": round(float(np.mean(ms)), 3),
           "device_ms": round(float(np.mean(dev)), 3),
           "lds_path_cells": int(paths.count(0)), "multi_workgroup_cells": int(paths.count(2)),
           "status_nonzero": int(np.count_nonzero(st))}
    print(f"[bench] periodic cells: {out['ms_per_step']:.2f} ms/step "
          f"({out['value']:.1f} cells/s, device {out['device_ms']:.2f} ms)",
          file=sys.stderr, flush=True)
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        o = oracle.Oracle()
        pick = np.linspace(0, len(cl) - 1, 6).astype(int)
        t = time.perf_counter()
        de = []
        for k in pick:
            r = o.singlepoint(cl[k][0], cl[k][1], lattice=cl[k][2])
            de.append(abs(r["energy"] - e[k]))
        dt = time.perf_counter() - t
        out["parity_sample"] = {"cells": len(pick), "max_abs_dE": float(max(de))}
        out["cpu_baseline"] = {"value": round(len(pick) / dt, 3), "unit": "cells/s",
                               "kind": "port", "sample": f"{len(pick)} evenly spaced cells, "
                               "CPU oracle one cell at a time (its three-body term on OpenMP)"}
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
