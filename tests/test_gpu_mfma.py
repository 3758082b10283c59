"""Matrix-core path for complex 4/5-qubit gates: against the oracle, against
the vector path, in place, for every target layout class."""
import numpy as np
import pytest
import torch

from unitair_amd import _lib
from unitair_amd import simulation as sim
from unitair_amd.circuits import CircuitRunner
from unitair_amd.workloads import fused_dense_circuit
from unitair_amd.initializations import seeded_state_np
from oracle import unitair_np as U
from oracle import circuits as oc

pytestmark = pytest.mark.gpu
TOL = {torch.complex64: 1e-5, torch.complex128: 1e-12}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,qubits", [(12, (0, 1, 2, 3, 4)), (12, (7, 8, 9, 10, 11)),
                                      (13, (11, 0, 6, 3, 9)), (14, (2, 13, 5, 8)),
                                      (11, (10, 9, 8, 7)), (15, (4, 14, 0, 10, 7)),
                                      # every target on index bit >= 6: the direct kernel
                                      (13, (0, 3, 5, 6)), (16, (9, 1, 4, 0, 7))])
def test_mfma_matches_oracle(device, dt, n, qubits):
    rng = np.random.default_rng(n + len(qubits))
    k = len(qubits)
    op = rng.standard_normal((2 ** k, 2 ** k)) + 1j * rng.standard_normal((2 ** k, 2 ** k))
    st = seeded_state_np(n, (3,), seed=n)
    out = sim.apply_operator(torch.from_numpy(op).to(device, dt), qubits,
                             torch.from_numpy(st).to(device, dt)).cpu().numpy()
    ref = U.apply_operator(op.astype(np.complex64).astype(np.complex128) if dt == torch.complex64
                           else op, qubits, st.astype(np.complex64).astype(np.complex128)
                           if dt == torch.complex64 else st)
    assert rel(out, ref) < TOL[dt]


@pytest.mark.parametrize("qubits", [(15, 3, 9, 0, 12), (0, 2, 4, 6, 8), (3, 0, 8, 4)])
def test_mfma_equals_vector_path_closely(device, qubits):
    lib = _lib.load()
    rng = np.random.default_rng(1)
    n = 16
    D = 2 ** len(qubits)
    op = torch.from_numpy(rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D)))
    st = torch.from_numpy(seeded_state_np(n, (2,), seed=2))
    for dt in (torch.complex64, torch.complex128):
        a = sim.apply_operator(op.to(device, dt), qubits, st.to(device, dt))
        lib.ua_set_tuning(_lib.UA_TUNE_MFMA, 0)
        try:
            b = sim.apply_operator(op.to(device, dt), qubits, st.to(device, dt))
        finally:
            lib.ua_set_tuning(_lib.UA_TUNE_MFMA, 1)
        assert rel(a.cpu().numpy(), b.cpu().numpy()) < TOL[dt]


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_fused_dense_circuit_20q_slice(device, dt):
    circ = fused_dense_circuit(seed=4, n=20, n5=12, n4=4)
    psi = seeded_state_np(20, (), seed=8)
    ref = oc.run_closed_form(circ.gates, psi)
    out = CircuitRunner(circ, dtype=dt, device=device).run(
        torch.from_numpy(psi).to(device, dt).reshape(1, -1)).cpu().numpy()[0]
    fid = float(U.fidelity_error(ref, out))
    assert fid < TOL[dt] and U.max_abs_error(ref, out) < TOL[dt]
