"""Dense-block passes (ua_run_dense_blocks): runs of 4/5-qubit dense gates
applied several per pass on the matrix cores, against the complex128 oracle
and against gate-by-gate application."""
import numpy as np
import pytest
import torch

from unitair_amd.circuits import CircuitRunner
from unitair_amd.workloads import Circuit, Gate, fused_dense_circuit, haar_unitary
from unitair_amd.initializations import seeded_state_np
from oracle import unitair_np as U
from oracle import circuits as oc

pytestmark = pytest.mark.gpu

# fidelity error / max-abs error against the complex128 oracle
BAR = {torch.complex64: (1e-9, 1e-5), torch.complex128: (1e-13, 1e-12)}


def mixed_circuit(n, seed, depth=24):
    """4/5-qubit Haar blocks on random subsets, with 1-qubit, CZ, CNOT and
    phase gates in between (they cut the passes)."""
    rng = np.random.default_rng(seed)
    gates = []
    for i in range(depth):
        r = rng.random()
        if r < 0.7:
            k = int(rng.integers(4, 6))
            qs = tuple(int(q) for q in rng.choice(n, k, replace=False))
            gates.append(Gate("dense", qs, matrix=haar_unitary(2 ** k, rng)))
        elif r < 0.8:
            gates.append(Gate("dense", (int(rng.integers(n)),), matrix=haar_unitary(2, rng)))
        elif r < 0.9:
            a, b = (int(q) for q in rng.choice(n, 2, replace=False))
            gates.append(Gate("cz" if rng.random() < 0.5 else "cnot", (a, b)))
        else:
            gates.append(Gate("phase", (), angles=rng.uniform(0, 2 * np.pi, 2 ** n)
                              .astype(np.float32)))
    return Circuit(n, gates)


def run_blocks(circ, psi, dt, device):
    r = CircuitRunner(circ, dtype=dt, device=device, in_place=True, dense_blocks=True)
    x = torch.from_numpy(psi).to(device, dt)
    return r.run(x).cpu().numpy(), r


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,seed", [(14, 1), (15, 2), (17, 3), (18, 4)])
def test_blocks_match_oracle(device, dt, n, seed):
    circ = mixed_circuit(n, seed)
    psi = seeded_state_np(n, (2,), seed=seed)
    got, r = run_blocks(circ, psi, dt, device)
    assert any(k.startswith("blocks") and len(k) > 7 for k in r.kinds)  # some pass holds 2+
    for b in range(2):
        ref = oc.run_closed_form(circ.gates, psi[b])
        fe, ma = BAR[dt]
        assert U.fidelity_error(ref, got[b]) < fe
        assert U.max_abs_error(ref, got[b]) < ma


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_blocks_equal_gate_by_gate(device, dt):
    n = 20
    circ = fused_dense_circuit(seed=7, n=n, n5=12, n4=6)
    psi = seeded_state_np(n, (), seed=9)
    got, r = run_blocks(circ, psi, dt, device)
    per_gate = CircuitRunner(circ, dtype=dt, device=device, in_place=True)
    want = per_gate.run(torch.from_numpy(psi).to(device, dt)).cpu().numpy()
    tol = 1e-5 if dt == torch.complex64 else 1e-12
    assert np.max(np.abs(got - want)) < tol
    assert len(r.kinds) < len(circ.gates)


def test_full_dense_block_slice_c64_and_c128(device):
    # the 64 + 16 block workload at 20 qubits against the complex128 oracle
    n = 20
    circ = fused_dense_circuit(seed=4, n=n)
    psi = seeded_state_np(n, (), seed=24)
    ref = oc.run_closed_form(circ.gates, psi)
    for dt in (torch.complex64, torch.complex128):
        got, _ = run_blocks(circ, psi, dt, device)
        fe, ma = BAR[dt]
        assert U.fidelity_error(ref, got) < fe and U.max_abs_error(ref, got) < ma


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_blocks_28q_roundtrip(device, dt):
    # U then U^dagger (reversed) returns the input; norm kept in between
    n = 28
    rng = np.random.default_rng(11)
    gates = []
    for qs in [(0, 1, 2, 3, 4), (23, 24, 25, 26, 27), (5, 17, 9, 27), (2, 20, 14, 8, 11)]:
        gates.append(Gate("dense", qs, matrix=haar_unitary(2 ** len(qs), rng)))
    fwd = Circuit(n, gates)
    inv = Circuit(n, [Gate("dense", g.qubits, matrix=np.conj(np.asarray(g.matrix)).T)
                      for g in reversed(gates)])
    x0 = torch.randn(2 ** n, dtype=dt, device=device)
    x0 /= torch.linalg.vector_norm(x0)
    x = x0.clone()
    a = CircuitRunner(fwd, dtype=dt, device=device, dense_blocks=True)
    b = CircuitRunner(inv, dtype=dt, device=device, dense_blocks=True)
    a.run(x)
    norm = float(torch.linalg.vector_norm(x))
    diff_mid = float(torch.max(torch.abs(x - x0)) / torch.max(torch.abs(x0)))
    b.run(x)
    err = float(torch.max(torch.abs(x - x0)) / torch.max(torch.abs(x0)))
    assert diff_mid > 0.1                        # the gates did something
    assert abs(norm - 1) < (1e-5 if dt == torch.complex64 else 1e-12)
    assert err < (1e-4 if dt == torch.complex64 else 1e-11)


def test_blocks_misaligned_state_runs_gate_by_gate(device):
    """A state that starts one complex64 element past a 16-byte boundary cannot
    take the pair loads of a block pass: the same gate list then runs gate by
    gate, with the same result."""
    from unitair_amd import _lib
    n = 16
    circ = fused_dense_circuit(seed=12, n=n, n5=4, n4=2)
    psi = seeded_state_np(n, (), seed=13)
    runner = CircuitRunner(circ, dtype=torch.complex64, device=device, dense_blocks=True)
    buf = torch.zeros(2 ** n + 1, dtype=torch.complex64, device=device)
    view = buf[1:]
    view.copy_(torch.from_numpy(psi).to(device, torch.complex64))
    assert view.data_ptr() % 16 == 8
    lib = _lib.load()
    rc = lib.ua_run_dense_blocks(_lib.ptr(view), _lib.UA_COMPLEX64, n, 1, runner._gates,
                                 len(circ.gates), None, _lib.stream_handle(device))
    assert rc == 0
    ref = oc.run_closed_form(circ.gates, psi)
    fe, ma = BAR[torch.complex64]
    got = view.cpu().numpy()
    assert U.fidelity_error(ref, got) < fe and U.max_abs_error(ref, got) < ma
