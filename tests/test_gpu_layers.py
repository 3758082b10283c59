"""1-qubit layers on the GPU: apply_to_qubits (fused per qubit, tiled passes)
and act_last_qubit, against the oracle's restatement of the reference."""
import warnings

import numpy as np
import pytest
import torch

from unitair_amd import _lib
from unitair_amd import simulation as sim
from unitair_amd.gates.matrix_algebra import fuse_single_qubit_operators
from oracle import unitair_np as U

pytestmark = pytest.mark.gpu

TOL = {torch.complex64: 2e-5, torch.complex128: 1e-12, torch.float64: 1e-12}


def rand(shape, dt, rng):
    x = rng.standard_normal(shape)
    if dt.is_complex:
        x = x + 1j * rng.standard_normal(shape)
    return x


def cast(x, dt):
    return x.astype(np.complex64 if dt == torch.complex64 else
                    np.complex128 if dt == torch.complex128 else np.float64)


def rel(out, ref):
    ref = np.asarray(ref, dtype=np.complex128)
    return float(np.max(np.abs(np.asarray(out, dtype=np.complex128) - ref)) /
                 max(1e-30, float(np.max(np.abs(ref)))))


CASES = [
    # n, qubits (repeats are fused)
    (5, [0, 3, 0, 4]),
    (12, [11, 0, 5, 7, 11]),
    (12, []),
    (13, [12, 0, 5, 7, 3, 3]),
    (16, list(range(16))),
    (17, [16, 15, 14, 13]),        # only high passes
    (20, [0, 1, 2, 19, 9, 12, 0]),
    (21, list(range(21)) + [4, 17]),
]


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,qubits", CASES)
def test_apply_to_qubits_matches_oracle(device, dt, n, qubits):
    rng = np.random.default_rng(n * 31 + len(qubits))
    ops = [cast(rand((2, 2), dt, rng), dt) for _ in qubits]
    st = cast(rand((2, 2 ** n), dt, rng), dt)
    out = sim.apply_to_qubits([torch.from_numpy(o).to(device) for o in ops], qubits,
                              torch.from_numpy(st).to(device))
    ref = U.apply_to_qubits([o.astype(np.complex128) for o in ops], qubits,
                            st.astype(np.complex128))
    assert out.shape == ref.shape and out.dtype == dt
    assert rel(out.cpu().numpy(), ref) < TOL[dt] * max(1, len(qubits))


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_apply_to_qubits_batched_operators(device, dt):
    rng = np.random.default_rng(3)
    n, qubits = 14, [13, 2, 6, 2, 9]
    ops = [cast(rand((3, 2, 2), dt, rng), dt) for _ in qubits]
    for st_shape in [(3, 2 ** n), (2 ** n,)]:
        st = cast(rand(st_shape, dt, rng), dt)
        out = sim.apply_to_qubits([torch.from_numpy(o).to(device) for o in ops], qubits,
                                  torch.from_numpy(st).to(device)).cpu().numpy()
        assert out.shape == (3, 2 ** n)
        for r in range(3):
            ref = U.apply_to_qubits([o[r].astype(np.complex128) for o in ops], qubits,
                                    (st[r] if st.ndim == 2 else st).astype(np.complex128))
            assert rel(out[r], ref) < TOL[dt] * len(qubits)


def test_apply_to_qubits_broadcast_and_real_fall_back_to_gate_rules(device):
    rng = np.random.default_rng(4)
    n = 6
    # operator batch (2, 1) against state batch (2, 3): applied gate by gate
    ops = [rand((2, 1, 2, 2), torch.complex128, rng) for _ in range(3)]
    st = rand((2, 3, 2 ** n), torch.complex128, rng)
    out = sim.apply_to_qubits([torch.from_numpy(o).to(device) for o in ops], [1, 4, 1],
                              torch.from_numpy(st).to(device)).cpu().numpy()
    assert rel(out, U.apply_to_qubits(ops, [1, 4, 1], st)) < 1e-12
    # real states
    ops = [rng.standard_normal((2, 2)) for _ in range(4)]
    st = rng.standard_normal((2, 2 ** 13))
    out = sim.apply_to_qubits([torch.from_numpy(o).to(device) for o in ops], [0, 12, 5, 0],
                              torch.from_numpy(st).to(device)).cpu().numpy()
    assert rel(out, U.apply_to_qubits(ops, [0, 12, 5, 0], st)) < 1e-12


def test_apply_to_qubits_equals_apply_all_qubits(device):
    # reference tests/test_unitary.py:38-52
    rng = np.random.default_rng(5)
    for n in (1, 7, 12, 18):
        u = torch.from_numpy(rand((2, 2), torch.complex128, rng)).to(device)
        psi = torch.from_numpy(rand((2 ** n,), torch.complex128, rng)).to(device)
        a = sim.apply_all_qubits(u, psi)
        b = sim.apply_to_qubits([u] * n, range(n), psi)
        assert rel(b.cpu().numpy(), a.cpu().numpy()) < 1e-12


def test_apply_to_qubits_gradients(device):
    rng = np.random.default_rng(6)
    n = 5
    ops = [torch.from_numpy(rand((2, 2), torch.complex128, rng)).to(device).requires_grad_()
           for _ in range(3)]
    psi = torch.from_numpy(rand((2 ** n,), torch.complex128, rng)).to(device).requires_grad_()
    w = torch.from_numpy(rand((2 ** n,), torch.complex128, rng)).to(device)
    (sim.apply_to_qubits(ops, [2, 0, 2], psi) * w).real.sum().backward()
    g_ops = [o.grad.cpu().numpy() for o in ops]
    g_psi = psi.grad.cpu().numpy()
    # same gradients through one apply_operator per gate
    ops2 = [o.detach().clone().requires_grad_() for o in ops]
    psi2 = psi.detach().clone().requires_grad_()
    x = psi2
    for q, o in zip([2, 0, 2], ops2):
        x = sim.apply_operator(o, [q], x)
    (x * w).real.sum().backward()
    for a, b in zip(g_ops, ops2):
        assert rel(a, b.grad.cpu().numpy()) < 1e-12
    assert rel(g_psi, psi2.grad.cpu().numpy()) < 1e-12


def test_apply_to_qubits_errors(device):
    psi = torch.zeros(2 ** 4, dtype=torch.complex64, device=device)
    u = torch.eye(2, dtype=torch.complex64, device=device)
    with pytest.raises(ValueError):
        sim.apply_to_qubits([u], [4], psi)
    with pytest.raises(ValueError):
        sim.apply_to_qubits([torch.eye(4, dtype=torch.complex64, device=device)], [0], psi)
    with pytest.raises(RuntimeError):
        sim.apply_to_qubits([u.to(torch.complex128)], [0], psi)


def test_apply_to_qubits_no_gates_copies(device):
    psi = torch.randn(3, 2 ** 12, dtype=torch.complex64, device=device)
    out = sim.apply_to_qubits([], [], psi)
    assert torch.equal(out, psi) and out.data_ptr() != psi.data_ptr()


def test_layer_entry_point_rejects_bad_arguments(device):
    lib = _lib.load()
    psi = torch.zeros(2 ** 12, dtype=torch.complex64, device=device)
    out = torch.empty_like(psi)
    mats = torch.zeros(2, 2, 2, dtype=torch.complex64, device=device)
    args = lambda qs, ob: (_lib.ptr(out), _lib.ptr(psi), _lib.ptr(mats), _lib.DTYPE_CODES[
        torch.complex64], 12, len(qs), _lib.int_array(qs), 1, 2 ** 12, ob,
        _lib.stream_handle(device))
    assert lib.ua_apply_1q_layer(*args([3, 3], 0)) == 1        # repeated qubit
    assert lib.ua_apply_1q_layer(*args([3, 12], 0)) == 1       # out of range
    assert lib.ua_apply_1q_layer(*args([3, 4], 5)) == 1        # bad operator stride


def test_fuse_single_qubit_operators_on_device(device):
    rng = np.random.default_rng(7)
    A, B, C = (torch.from_numpy(rand((2, 2), torch.complex128, rng)).to(device)
               for _ in range(3))
    f = fuse_single_qubit_operators([1, 0, 1], [A, B, C])
    assert list(f) == [1, 0]
    assert torch.allclose(f[1], C @ A) and f[0] is B


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128, torch.float32, torch.float64])
@pytest.mark.parametrize("batch,count", [((), 7), ((3,), 5), ((2, 2), 150)])
def test_fuse_single_qubit_operators_native_products(device, dt, batch, count):
    """ua_fuse_1q_ops (chained past 64 operators) against products on the host."""
    rng = np.random.default_rng(count + len(batch))
    qubits = [int(q) for q in rng.integers(0, 4, count)]
    npdt = {torch.complex64: np.complex64, torch.complex128: np.complex128,
            torch.float32: np.float32, torch.float64: np.float64}[dt]
    ops = [rand(batch + (2, 2), dt, rng).astype(npdt) for _ in range(count)]
    f = fuse_single_qubit_operators(qubits, [torch.from_numpy(o).to(device) for o in ops])
    ref = {}
    for q, o in zip(qubits, ops):
        o = o.astype(np.complex128)
        ref[q] = o @ ref[q] if q in ref else o
    assert list(f) == list(ref)
    tol = 1e-4 if dt in (torch.complex64, torch.float32) else 1e-11
    for q in ref:
        got = f[q].cpu().numpy().astype(np.complex128)
        assert got.shape == ref[q].shape
        assert np.max(np.abs(got - ref[q])) <= tol * max(1.0, np.max(np.abs(ref[q])))


def test_apply_to_qubits_native_fusion_equals_gate_by_gate(device):
    """The fused launch (ua_fuse_1q_ops + ua_apply_1q_layer) against the
    reference's gate-by-gate rule, batched operators, many repeats."""
    rng = np.random.default_rng(31)
    n = 14
    qubits = [int(q) for q in rng.integers(0, n, 90)]
    ops = [cast(rand((2, 2, 2), torch.complex128, rng), torch.complex128) for _ in qubits]
    st = cast(rand((2, 2 ** n), torch.complex128, rng), torch.complex128)
    out = sim.apply_to_qubits([torch.from_numpy(o).to(device) for o in ops], qubits,
                              torch.from_numpy(st).to(device)).cpu().numpy()
    for r in range(2):
        ref = U.apply_to_qubits([o[r] for o in ops], qubits, st[r])
        assert rel(out[r], ref) < 1e-11


def test_act_last_qubit(device):
    # reference tests/test_unitary.py:54-92
    rng = np.random.default_rng(8)
    op = rand((2, 2), torch.complex64, rng).astype(np.complex64)
    for n in range(1, 6):
        st = rand((2 ** n,), torch.complex64, rng).astype(np.complex64)
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            out = sim.act_last_qubit(torch.from_numpy(op).to(device),
                                     torch.from_numpy(st).to(device))
        assert any("outdated" in str(x.message) for x in w)
        assert rel(out.cpu().numpy(), U.act_last_qubit(op.astype(np.complex128),
                                                       st.astype(np.complex128))) < 2e-6
    batch = rand((4, 3, 2, 2, 2, 2, 2), torch.complex64, rng).astype(np.complex64)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = sim.act_last_qubit(torch.from_numpy(op).to(device),
                                 torch.from_numpy(batch).to(device)).cpu().numpy()
    assert rel(out, U.act_last_qubit(op.astype(np.complex128),
                                     batch.astype(np.complex128))) < 2e-6
