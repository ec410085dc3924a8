"""Serving loop (continuous batching): every request's result is the same
when it runs alone and when it runs in a mix of concurrently admitted
requests, and the mix is right against the oracle.

Tolerances: alone vs mix |dE| <= 1e-9 Eh (the same kernels on the same
inputs; a different batch only changes the size bucket, i.e. the workgroup
shape and with it the summation order); mix vs oracle |dE| <= 1e-6 Eh."""
import time

import numpy as np
import pytest

import tblite_amd
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def requests_mix():
    rng = np.random.default_rng(44)
    mols = []
    for k, nw in enumerate((2, 5, 9, 16)):
        mols.append(workloads.water_cluster(nw, workloads.liquid_radius(nw), rng))
    for nh in (4, 10, 22, 35):
        mols.append(workloads.organic_molecule(nh, rng))
    order = rng.permutation(len(mols))
    return [mols[i] for i in order]


def _serve(engine, mols, slots, arrive_in_two_waves):
    srv = tblite_amd.Server(engine, slots=slots, workspace_bytes=4e9)
    tickets = {}
    half = len(mols) // 2 if arrive_in_two_waves else len(mols)
    for m in range(half):
        tickets[srv.submit(*mols[m])] = m
    results = {}
    # the second wave arrives while the first batches are on the device
    polled = srv.poll()
    for t, e, it, st, lat in polled:
        results[tickets[t]] = (e, it, st, lat)
    busy_at_second_wave = srv.stats()["busy_slots"]
    for m in range(half, len(mols)):
        tickets[srv.submit(*mols[m])] = m
    t0 = time.time()
    while len(results) < len(mols):
        for t, e, it, st, lat in srv.poll():
            results[tickets[t]] = (e, it, st, lat)
        assert time.time() - t0 < 120, "server did not finish"
        time.sleep(0.001)
    stats = srv.stats()
    srv.close()
    return results, stats, busy_at_second_wave


def test_alone_equals_mix(engine, requests_mix):
    mols = requests_mix
    alone = []
    for num, xyz in mols:
        e, it, st = engine.singlepoint_flat(np.array([len(num)]), num, xyz)
        assert st[0] == 0
        alone.append((float(e[0]), int(it[0])))
    print(f"[gpu] serving mix of {len(mols)} requests", flush=True)
    res, stats, busy = _serve(engine, mols, slots=2, arrive_in_two_waves=True)
    print(f"[gpu] serving stats {stats}, busy slots when the second wave arrived: {busy}", flush=True)
    assert stats["completed"] == len(mols)
    assert busy >= 1  # the second wave was admitted while work was on the device
    for m in range(len(mols)):
        e, it, st, lat = res[m]
        assert st == 0, (m, st)
        assert abs(e - alone[m][0]) <= 1e-9, (m, e, alone[m][0])
        assert it == alone[m][1]
        assert lat > 0.0


def test_mix_vs_oracle(engine, orc, requests_mix):
    mols = requests_mix[:4]
    res, stats, _ = _serve(engine, mols, slots=3, arrive_in_two_waves=False)
    for m, (num, xyz) in enumerate(mols):
        r = orc.singlepoint(num, xyz)
        assert abs(res[m][0] - r["energy"]) <= 1e-6, (m, res[m][0], r["energy"])


def test_submit_rejects_unknown_element(engine):
    srv = tblite_amd.Server(engine, slots=1, workspace_bytes=1e9)
    with pytest.raises(tblite_amd.EngineError):
        srv.submit(np.array([92, 1]), np.array([[0.0, 0, 0], [0, 0, 2.0]]))
    srv.close()


def test_results_reported_as_molecules_finish(engine):
    """One batch holding small molecules and one large one: each small result is
    reported when its molecule finishes, not when the batch ends."""
    rng = np.random.default_rng(45)
    big = workloads.water_cluster(100, workloads.liquid_radius(100), rng)
    small = [workloads.water_cluster(nw, workloads.liquid_radius(nw), rng) for nw in (2, 3, 4, 5, 6, 8)]
    srv = tblite_amd.Server(engine, slots=1, workspace_bytes=4e9)
    tickets = {srv.submit(*big): "big"}
    for k, m in enumerate(small):
        tickets[srv.submit(*m)] = k
    lat = {}
    t0 = time.time()
    while len(lat) < len(tickets):
        for t, e, it, st, ms in srv.poll():
            assert st == 0
            lat[tickets[t]] = ms
        assert time.time() - t0 < 120, "server did not finish"
        time.sleep(0.0005)
    stats = srv.stats()
    srv.close()
    print(f"[gpu] early reporting: big {lat['big']:.1f} ms, small max "
          f"{max(lat[k] for k in range(len(small))):.1f} ms", flush=True)
    assert stats["batches"] == 1
    assert max(lat[k] for k in range(len(small))) < 0.5 * lat["big"]


def test_large_request_under_light_load_takes_multi_workgroup_path(engine):
    """A lone 600-AO request on a two-slot server runs on the multi-workgroup
    path (the cost model puts it about 2.7x ahead of one workgroup) on a
    worker thread while the other slot serves small requests; its energy
    equals the one-workgroup path's within 1e-8 Eh and the small results
    arrive first."""
    rng = np.random.default_rng(46)
    big = workloads.water_cluster(100, workloads.liquid_radius(100), rng)
    small = [workloads.water_cluster(nw, workloads.liquid_radius(nw), rng) for nw in (2, 4, 7)]
    e_ref, it_ref, st_ref = engine.singlepoint_flat(np.array([len(big[0])]), big[0], big[1])
    assert st_ref[0] == 0 and engine.molecule_path(0) == 1
    srv = tblite_amd.Server(engine, slots=2, workspace_bytes=8e9)
    # a first large-class request through the same slot leaves its completion
    # record behind (the multi-workgroup launch must not report it)
    warm = workloads.water_cluster(20, workloads.liquid_radius(20), rng)
    tw = srv.submit(*warm)
    t0 = time.time()
    got = []
    while not got:
        got = [r for r in srv.poll() if r[0] == tw]
        assert time.time() - t0 < 60, "server did not finish"
        time.sleep(0.0005)
    assert srv.stats()["grid_requests"] == 0
    tickets = {srv.submit(*big): "big"}
    for k, m in enumerate(small):
        tickets[srv.submit(*m)] = k
    res = {}
    t0 = time.time()
    print("[gpu] serving: lone large request on the multi-workgroup path", flush=True)
    while len(res) < len(tickets):
        for t, e, it, st, ms in srv.poll():
            res[tickets[t]] = (e, it, st, ms)
        assert time.time() - t0 < 120, "server did not finish"
        time.sleep(0.0005)
    stats = srv.stats()
    srv.close()
    print(f"[gpu] stats {stats}, big {res['big'][3]:.0f} ms", flush=True)
    assert stats["grid_requests"] == 1
    e, it, st, ms = res["big"]
    assert st == 0
    assert abs(e - e_ref[0]) < 1e-8, (e, e_ref[0])
    assert abs(it - int(it_ref[0])) <= 1
    assert max(res[k][3] for k in range(len(small))) < ms
