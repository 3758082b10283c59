"""Serving loop: results against the complex128 oracle (small and 21-22
qubit lanes), the same result alone and in the mix, admission under a small
HBM budget and while lanes run, and Poisson arrivals end to end against both
one-at-a-time baselines."""
import numpy as np
import pytest
import torch

from unitair_amd.serving import (Request, Server, poisson_workload, run_one_at_a_time,
                                 run_one_at_a_time_fused)
from oracle import circuits as oc

pytestmark = pytest.mark.gpu


def oracle_answer(r: Request):
    psi = np.zeros(2 ** r.n, dtype=np.complex128)
    psi[0] = 1
    psi = oc.run_closed_form(r.circuit().gates, psi)
    p = np.abs(psi) ** 2
    bit = (np.arange(2 ** r.n) >> (r.n - 1)) & 1
    return float(p.sum()), float(np.sum(p * (1 - 2 * bit)))


def test_serving_matches_oracle(device):
    reqs = [Request(i, n, d, seed=100 + i) for i, (n, d) in
            enumerate([(10, 12), (12, 30), (10, 25), (13, 11), (16, 14), (12, 10)])]
    Server(device=device).run(reqs, realtime=False)
    for r in reqs:
        norm, z0 = oracle_answer(r)
        assert abs(r.norm2 - 1.0) < 1e-5
        assert abs(r.z0 - z0) < 1e-5, (r.n, r.depth)


def test_alone_equals_mix(device):
    target = Request(0, 12, 37, seed=7)
    alone = Server(device=device).run([Request(0, 12, 37, seed=7)], realtime=False)[0]
    others = [Request(1 + i, n, d, seed=50 + i) for i, (n, d) in
              enumerate([(12, 15), (12, 80), (11, 20), (12, 37), (14, 9), (12, 5)])]
    mix = Server(device=device, max_layers_per_step=5).run([target] + others, realtime=False)
    assert mix[0].norm2 == alone.norm2 and mix[0].z0 == alone.z0


def test_admission_under_small_budget(device):
    fp = 2 ** 14 * 8
    server = Server(hbm_budget=3 * fp, max_slots=4, device=device)
    reqs = [Request(i, 14, 10 + i, seed=i) for i in range(7)]
    peak = []
    orig = server.try_admit

    def tracking(r, now):
        ok = orig(r, now)
        peak.append(server.used_bytes)
        return ok
    server.try_admit = tracking
    server.run(reqs, realtime=False)
    assert max(peak) <= 3 * fp
    assert all(abs(r.norm2 - 1) < 1e-5 for r in reqs)
    with pytest.raises(ValueError):
        Server(hbm_budget=fp // 2, device=device).run([Request(0, 14, 3, seed=1)],
                                                     realtime=False)


def test_poisson_arrivals_and_one_at_a_time(device):
    reqs = poisson_workload(10, rate=200.0, seed=3, n_range=(10, 16), depth_range=(10, 40))
    Server(device=device).run(reqs, realtime=True)
    base = poisson_workload(10, rate=200.0, seed=3, n_range=(10, 16), depth_range=(10, 40))
    run_one_at_a_time(base, device=device, realtime=True)
    for a, b in zip(reqs, base):
        assert a.done >= a.arrival and b.done >= b.arrival
        assert abs(a.z0 - b.z0) < 1e-5 and abs(a.norm2 - b.norm2) < 1e-5


def test_large_lanes_match_oracle(device):
    """21- and 22-qubit lanes (two slots each, several steps, graph replay)
    against the complex128 oracle."""
    reqs = [Request(i, n, d, seed=300 + i) for i, (n, d) in
            enumerate([(21, 6), (22, 5), (21, 9), (12, 20)])]
    Server(device=device, max_layers_per_step=4).run(reqs, realtime=False)
    for r in reqs:
        norm, z0 = oracle_answer(r)
        assert abs(r.norm2 - 1.0) < 1e-5
        assert abs(r.z0 - z0) < 1e-5, (r.n, r.depth)


def test_admission_while_lanes_run_and_fused_baseline(device):
    """Requests arriving while a lane's steps are in flight join freed or
    free slots (at most two steps in flight, no device-wide waits); each
    result equals the one-at-a-time run on the same fused programs within
    the complex64 tolerance (the two runs cut the layers into different
    steps, so stage grouping, hence rounding, may differ)."""
    reqs = poisson_workload(16, rate=400.0, seed=8, n_range=(10, 15), depth_range=(10, 60))
    server = Server(device=device, max_slots=4, max_layers_per_step=3)
    server.run(reqs, realtime=True)
    alone = poisson_workload(16, rate=400.0, seed=8, n_range=(10, 15), depth_range=(10, 60))
    run_one_at_a_time_fused(alone, device=device, realtime=True)
    for a, b in zip(reqs, alone):
        assert a.done >= a.arrival and b.done >= b.arrival
        assert abs(a.norm2 - b.norm2) < 1e-6 and abs(a.z0 - b.z0) < 1e-6
    assert not server.lanes and server.used_bytes == 0 and server.active == 0


def test_compiled_lanes_match_oracle(device):
    """Server(compiled=True): 13- and 14-qubit lanes on compiled passes
    (captured into graphs) against the complex128 oracle."""
    reqs = [Request(i, n, d, seed=500 + i) for i, (n, d) in
            enumerate([(13, 8), (14, 8), (13, 8)])]
    Server(device=device, max_layers_per_step=4, compiled=True).run(reqs, realtime=False)
    for r in reqs:
        norm, z0 = oracle_answer(r)
        assert abs(r.norm2 - 1.0) < 1e-5
        assert abs(r.z0 - z0) < 1e-5, (r.n, r.depth)


def test_background_compiled_lanes_match_oracle(device):
    """Server(compiled="background"): lanes start on interpreted passes while
    their structure compiles on a host thread, then switch; every result
    against the complex128 oracle."""
    from unitair_amd.serving import _BACKGROUND_JOBS
    reqs = [Request(i, n, d, seed=700 + i) for i, (n, d) in
            enumerate([(13, 30), (15, 30), (13, 24), (12, 40)])]
    Server(device=device, max_layers_per_step=3, compiled="background").run(reqs,
                                                                          realtime=False)
    for r in reqs:
        norm, z0 = oracle_answer(r)
        assert abs(r.norm2 - 1.0) < 1e-5
        assert abs(r.z0 - z0) < 1e-5, (r.n, r.depth)
    jobs = [j for (n, k), j in _BACKGROUND_JOBS.items() if n in (12, 13, 15)]
    assert jobs
    for j in jobs:
        j.result(timeout=120)            # the compiles themselves succeed
    # a second run of the same shapes starts on compiled passes
    again = [Request(10 + i, n, d, seed=800 + i) for i, (n, d) in enumerate([(13, 9), (15, 6)])]
    Server(device=device, max_layers_per_step=3, compiled="background").run(again,
                                                                          realtime=False)
    for r in again:
        norm, z0 = oracle_answer(r)
        assert abs(r.z0 - z0) < 1e-5, (r.n, r.depth)
