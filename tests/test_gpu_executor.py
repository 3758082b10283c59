"""The native circuit executor with per-gate HIP events (separate start/end
events, or one shared event at each gate boundary): same results as without
events, and per-gate times that add up to no more than the run."""
import time

import pytest
import torch

from unitair_amd._hip import EventPool
from unitair_amd.circuits import CircuitRunner
from unitair_amd.initializations import seeded_state
from unitair_amd.workloads import random_circuit_24q

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shared", [False, True])
def test_events_do_not_change_results_and_time_every_gate(device, shared):
    circ = random_circuit_24q(seed=5, n=16)
    psi = seeded_state(16, (), seed=3, dtype=torch.complex64, device=device)
    plain = CircuitRunner(circ, dtype=torch.complex64, device=device).run(psi.clone())
    runner = CircuitRunner(circ, dtype=torch.complex64, device=device)
    pool = EventPool(len(runner.kinds), shared=shared)
    x = psi.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runner.run(x, events=pool)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    assert torch.equal(x, plain)
    ms = pool.elapsed_ms()
    assert len(ms) == len(circ.gates) and all(t >= 0 for t in ms)
    assert sum(ms) <= wall_ms * 1.01 + 0.05
    if shared:
        h = pool.handles
        assert all(h[2 * i + 1] == h[2 * i + 2] for i in range(len(ms) - 1))
