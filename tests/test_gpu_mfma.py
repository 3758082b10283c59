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


# achieved on MI355X: complex64 fidelity error ~1e-12 (max-abs ~5e-9),
# complex128 ~2e-16; the bars leave two to three orders of margin
SLICE_BAR = {torch.complex64: (1e-9, 1e-6), torch.complex128: (1e-13, 1e-12)}


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_fused_dense_circuit_20q_slice(device, dt):
    """The dense-block workload's full gate list (64 Haar 32x32 + 16 Haar
    16x16 blocks) at 20 qubits, against the complex128 closed form."""
    circ = fused_dense_circuit(seed=4, n=20)
    assert circ.counts() == {"dense5": 64, "dense4": 16}
    psi = seeded_state_np(20, (), seed=8)
    ref = oc.run_closed_form(circ.gates, psi)
    out = CircuitRunner(circ, dtype=dt, device=device).run(
        torch.from_numpy(psi).to(device, dt).reshape(1, -1)).cpu().numpy()[0]
    fid, mae = float(U.fidelity_error(ref, out)), U.max_abs_error(ref, out)
    assert fid < SLICE_BAR[dt][0] and mae < SLICE_BAR[dt][1], (fid, mae)


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("qubits", [(0, 1, 2, 3, 4), (23, 24, 25, 26, 27), (27, 3, 14, 0, 22),
                                    (2, 27, 9, 16)])
def test_mfma_28q_unitary_roundtrip(device, dt, qubits):
    """4- and 5-qubit Haar blocks on a 28-qubit state (the dense-block
    configuration's size; the highest and lowest strides): the norm is kept
    and U then U^H returns the input, in place through the matrix cores."""
    from unitair_amd.gates import haar_unitary
    from unitair_amd.initializations import rand_state
    import unitair_amd as ua
    rng = np.random.default_rng(sum(qubits))
    u = torch.from_numpy(haar_unitary(2 ** len(qubits), rng)).to(device, dt)
    uh = u.conj().T.contiguous()
    psi = rand_state(28, device=device, dtype=dt, seed=5)
    lib = _lib.load()
    x = psi.clone()
    q = _lib.int_array(list(qubits))
    for m in (u, uh):
        rc = lib.ua_apply_operator(_lib.ptr(x), _lib.ptr(x), _lib.ptr(m), _lib.DTYPE_CODES[dt],
                                   28, len(qubits), q, 1, 2 ** 28, 0, _lib.stream_handle(device))
        _lib.check(rc, "apply_operator")
        if m is u:
            assert abs(float(ua.states.norm_squared_f64(x)) - 1.0) < (1e-5 if dt == torch.complex64
                                                                     else 1e-12)
    err = float((x - psi).abs().max()) / float(psi.abs().max())
    assert err < (1e-5 if dt == torch.complex64 else 1e-12), err
    del x, psi
    torch.cuda.empty_cache()
