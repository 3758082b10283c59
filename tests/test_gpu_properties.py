"""Property tests on the GPU in the style of the reference's hypothesis suite
(reference tests/test_operations.py:23-200): drawn qubit counts, gate sizes,
batch structures and values, checked against the oracle and against the
reference's own invariants (act_first_qubits == apply_operator on the
leading qubits, a batch entry equals the unbatched call, phase inverse and
complement, apply_all_qubits batching, swap involutory / symmetric / equal to
permute_qubits).  Derandomized so every run draws the same examples."""
import math

import numpy as np
import pytest
import torch
from hypothesis import HealthCheck, given, settings
import hypothesis.strategies as st

from unitair_amd import simulation as sim
from oracle import unitair_np as U

pytestmark = pytest.mark.gpu

SETTINGS = settings(max_examples=60, deadline=None, derandomize=True,
                    suppress_health_check=list(HealthCheck))
TOL = {torch.complex64: 2e-5, torch.float32: 2e-5}


@st.composite
def op_and_state(draw, max_qubits=8, max_op_qubits=5, real_ok=True):
    """A state (*batch, 2^n) and an operator (*op_batch, 2^k, 2^k); op_batch
    is empty or equal to the state's batch, or the state is unbatched."""
    n = draw(st.integers(1, max_qubits))
    k = draw(st.integers(1, min(n, max_op_qubits)))
    batch = tuple(draw(st.lists(st.integers(1, 3), max_size=3)))
    mode = draw(st.sampled_from(["shared", "same", "op_only"]))
    op_batch = () if mode == "shared" else batch
    state_batch = () if mode == "op_only" else batch
    dt = draw(st.sampled_from([torch.complex64, torch.float32] if real_ok else [torch.complex64]))
    seed = draw(st.integers(0, 2 ** 31))
    rng = np.random.default_rng(seed)

    def arr(shape):
        x = rng.uniform(-10, 10, shape)
        if dt.is_complex:
            x = x + 1j * rng.uniform(-10, 10, shape)
        return x.astype(np.complex64 if dt.is_complex else np.float32)

    return n, k, arr(op_batch + (2 ** k, 2 ** k)), arr(state_batch + (2 ** n,)), dt


def gpu(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.max(np.abs(a - b)) / max(1e-30, float(np.max(np.abs(b))))) if b.size else 0.0


@SETTINGS
@given(data=op_and_state(), qubit_seed=st.integers(0, 2 ** 31))
def test_apply_operator_matches_oracle(device, data, qubit_seed):
    n, k, op, st_, dt = data
    qubits = [int(q) for q in np.random.default_rng(qubit_seed).permutation(n)[:k]]
    out = sim.apply_operator(gpu(op, device), qubits, gpu(st_, device)).cpu().numpy()
    ref = U.apply_operator(op.astype(np.complex128), qubits, st_.astype(np.complex128))
    assert out.shape == ref.shape
    assert rel(out, ref) < TOL[dt] * 2 ** k


@SETTINGS
@given(data=op_and_state())
def test_act_first_qubits_equals_apply_operator(device, data):
    # reference tests/test_operations.py:43-67
    n, k, op, st_, dt = data
    a = sim.act_first_qubits(gpu(op, device), gpu(st_, device))
    b = sim.apply_operator(gpu(op, device), range(k), gpu(st_, device))
    assert torch.equal(a, b)


@SETTINGS
@given(data=op_and_state(), pick=st.integers(0, 2 ** 31))
def test_batch_entry_equals_unbatched_call(device, data, pick):
    # reference tests/test_operations.py:70-122
    n, k, op, st_, dt = data
    batch = op.shape[:-2] if op.ndim > 2 else st_.shape[:-1]
    if not batch:
        return
    out = sim.act_first_qubits(gpu(op, device), gpu(st_, device)).cpu().numpy()
    rng = np.random.default_rng(pick)
    idx = tuple(int(rng.integers(0, d)) for d in batch)
    op_e = op[idx] if op.ndim > 2 else op
    st_e = st_[idx] if st_.ndim > 1 else st_
    single = sim.act_first_qubits(gpu(op_e, device), gpu(st_e, device)).cpu().numpy()
    assert rel(out[idx], single) < TOL[dt]


@SETTINGS
@given(data=op_and_state(max_op_qubits=1, real_ok=False), pick=st.integers(0, 2 ** 31))
def test_apply_all_qubits_batching(device, data, pick):
    # reference tests/test_operations.py:125-164
    n, k, op, st_, dt = data
    out = sim.apply_all_qubits(gpu(op, device), gpu(st_, device)).cpu().numpy()
    ref = U.apply_all_qubits(op.astype(np.complex128), st_.astype(np.complex128))
    assert rel(out, ref) < TOL[dt] * n * 4
    if st_.ndim > 1:
        rng = np.random.default_rng(pick)
        idx = tuple(int(rng.integers(0, d)) for d in st_.shape[:-1])
        op_e = op[idx] if op.ndim > 2 else op
        single = sim.apply_all_qubits(gpu(op_e, device), gpu(st_[idx], device)).cpu().numpy()
        assert rel(out[idx], single) < TOL[dt]


@SETTINGS
@given(n=st.integers(1, 8), batch=st.lists(st.integers(1, 3), max_size=3),
       seed=st.integers(0, 2 ** 31), scale=st.sampled_from([1.0, 10.0, 1000.0]))
def test_apply_phase_inverse_and_complement(device, n, batch, seed, scale):
    # reference tests/test_operations.py:23-40
    rng = np.random.default_rng(seed)
    shape = tuple(batch) + (2 ** n,)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ang = rng.uniform(-scale, scale, shape).astype(np.float32)
    x, a = gpu(psi, device), gpu(ang, device)
    back = sim.apply_phase(-a, sim.apply_phase(a, x)).cpu().numpy()
    assert np.allclose(back, psi, atol=1e-4)
    comp = sim.apply_phase(2 * math.pi - a, sim.apply_phase(a, x)).cpu().numpy()
    assert np.allclose(comp, psi, atol=1e-4 + float(np.abs(ang).sum()) * 1e-3)
    out = sim.apply_phase(a, x).cpu().numpy()
    assert rel(out, U.apply_phase(ang, psi.astype(np.complex128))) < 1e-5


@SETTINGS
@given(n=st.integers(2, 8), batch=st.lists(st.integers(1, 3), max_size=2),
       ij=st.tuples(st.integers(0, 7), st.integers(0, 7)), seed=st.integers(0, 2 ** 31))
def test_swap_properties(device, n, batch, ij, seed):
    # reference tests/test_operations.py:165-200
    i, j = ij
    if i == j or i >= n or j >= n:
        return
    rng = np.random.default_rng(seed)
    shape = tuple(batch) + (2 ** n,)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    x = gpu(psi, device)
    sw = sim.swap(x, (i, j))
    assert torch.equal(sim.swap(sw, (i, j)), x)
    assert torch.equal(sw, sim.swap(x, (j, i)))
    perm = list(range(n))
    perm[i], perm[j] = j, i
    assert torch.equal(sw, sim.permute_qubits(perm, x))
    assert np.array_equal(sw.cpu().numpy(), U.swap(psi, (i, j)))
