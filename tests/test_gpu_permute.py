"""Qubit permutations and sign/flip gates on the GPU against the oracle's
restatement of the reference (operations.py:506-806).  Pure data movement
and +-1 factors are compared bit for bit."""
import itertools

import numpy as np
import pytest
import torch

from unitair_amd import _lib
from unitair_amd import simulation as sim
from oracle import unitair_np as U
from oracle import closed_form

pytestmark = pytest.mark.gpu

DTYPES = [torch.complex64, torch.complex128, torch.float32, torch.float64]
NP = {torch.complex64: np.complex64, torch.complex128: np.complex128,
      torch.float32: np.float32, torch.float64: np.float64}


def rand(shape, dt, rng):
    x = rng.standard_normal(shape)
    if dt.is_complex:
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(NP[dt])


def dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 10, 11, 13, 16])
def test_permute_qubits_bit_exact(device, dt, n):
    rng = np.random.default_rng(n)
    st = rand((3, 2 ** n), dt, rng)
    for _ in range(3):
        perm = [int(p) for p in rng.permutation(n)]
        out = sim.permute_qubits(perm, dev(st, device)).cpu().numpy()
        assert np.array_equal(out, U.permute_qubits(perm, st))


def test_permute_qubits_24q_round_trip_and_reference(device):
    rng = np.random.default_rng(24)
    n = 24
    st = rand((2 ** n,), torch.complex64, rng)
    perm = [int(p) for p in rng.permutation(n)]
    x = dev(st, device)
    y = sim.permute_qubits(perm, x)
    inv = [0] * n
    for i, p in enumerate(perm):
        inv[p] = i
    assert torch.equal(sim.permute_qubits(inv, y), x)
    assert np.array_equal(y.cpu().numpy(), U.permute_qubits(perm, st))


def test_swap_matches_reference_invariants(device):
    rng = np.random.default_rng(6)
    n = 6
    st = rand((2, 3, 2 ** n), torch.complex64, rng)
    x = dev(st, device)
    for i, j in itertools.permutations(range(n), 2):
        sw = sim.swap(x, (i, j))
        assert torch.equal(sim.swap(sw, (i, j)), x)
        assert torch.equal(sw, sim.swap(x, (j, i)))
        assert np.array_equal(sw.cpu().numpy(), U.swap(st, (i, j)))
    assert torch.equal(sim.swap(x, (-1, 0)), sim.swap(x, (n - 1, 0)))
    assert sim.swap(x, (2, 2)) is x
    with pytest.raises(ValueError):
        sim.swap(x, (0, n))
    # real 2-qubit example of the reference's tests/test_unitary.py:94-101
    r = torch.tensor([.3, 0., 0., -0.13], device=device)
    assert torch.equal(sim.swap(r, (0, 1)), r)


@pytest.mark.parametrize("n", [4, 13])
def test_roll_qubits(device, n):
    rng = np.random.default_rng(n)
    st = rand((2, 2 ** n), torch.complex128, rng)
    x = dev(st, device)
    for k in range(-3, n + 3):
        assert np.array_equal(sim.roll_qubits(x, k).cpu().numpy(), U.roll_qubits(st, k))


@pytest.mark.parametrize("dt", DTYPES)
def test_multi_cz(device, dt):
    rng = np.random.default_rng(7)
    n = 12
    st = rand((2, 2 ** n), dt, rng)
    x = dev(st, device)
    for pairs in ([0, 5], [[0, 5], [11, 3], [4, 5]], [[1, 2], [2, 1], [7, 9]],
                  [[q, (q + 1) % n] for q in range(n)]):
        p = torch.tensor(pairs, device=device)
        out = sim.multi_cz(p, x).cpu().numpy()
        assert out.dtype == NP[dt]
        assert np.array_equal(out, U.multi_cz(np.array(pairs), st).astype(NP[dt]))
    # a negative index puts the pair beyond the state: no effect, as in the reference
    assert torch.equal(sim.multi_cz(torch.tensor([[-1, 3]], device=device), x), x)


def test_multi_cz_errors(device):
    x = torch.zeros(2 ** 4, dtype=torch.complex64, device=device)
    with pytest.raises(ValueError):
        sim.multi_cz(torch.tensor([[0, 4]], device=device), x)
    with pytest.raises(ValueError):
        sim.multi_cz(torch.tensor([[2, 2]], device=device), x)
    with pytest.raises(ValueError):
        sim.multi_cz(torch.tensor([[0, 1]]), x)            # pairs on the CPU
    with pytest.raises(ValueError):
        sim.multi_cz(torch.zeros(2, 2, 2, dtype=torch.long, device=device), x)


@pytest.mark.parametrize("dt", DTYPES)
def test_multi_controlled_z(device, dt):
    rng = np.random.default_rng(8)
    n = 13
    st = rand((2 ** n,), dt, rng)
    x = dev(st, device)
    for qs in ([3], [0, 12], [5, 1, 9, 11], []):
        out = sim.multi_controlled_z(qs, x).cpu().numpy()
        assert np.array_equal(out, U.multi_controlled_z(qs, st).astype(NP[dt]))
    with pytest.raises(RuntimeError):
        sim.multi_controlled_z([1, 1], x)


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_multi_controlled_x(device, dt):
    rng = np.random.default_rng(9)
    n = 12
    st = rand((2, 2 ** n), dt, rng)
    x = dev(st, device)
    for ctrl, t in (([0], 3), ([11, 1], 2), ([0, 1, 2, 5], 11), ([], 7)):
        out = sim.multi_controlled_x(x, ctrl, t).cpu().numpy()
        m = np.eye(2 ** (len(ctrl) + 1))
        m[-2:, -2:] = [[0, 1], [1, 0]]
        exact = closed_form.apply_dense(m, list(ctrl) + [t], st.astype(np.complex128))
        assert np.array_equal(out, exact.astype(NP[dt]))
        ref = U.multi_controlled_x(st.astype(np.complex128), ctrl, t)
        assert np.max(np.abs(out - ref)) <= 4e-7 * np.max(np.abs(ref))
    with pytest.raises(RuntimeError):
        sim.multi_controlled_x(x, [2, 3], 3)
    with pytest.raises(ValueError):
        sim.multi_controlled_x(x, [2], n)


def test_gradients_are_the_adjoint_maps(device):
    rng = np.random.default_rng(10)
    n = 6
    st = rand((2 ** n,), torch.complex128, rng)
    w = rand((2 ** n,), torch.complex128, rng)
    perm = [3, 0, 5, 1, 4, 2]
    cases = [
        (lambda v: sim.permute_qubits(perm, v), lambda g: U.permute_qubits(
            [perm.index(i) for i in range(n)], g)),
        (lambda v: sim.roll_qubits(v, 2), lambda g: U.roll_qubits(g, -2)),
        (lambda v: sim.multi_cz(torch.tensor([[0, 3], [1, 2]], device=device), v),
         lambda g: U.multi_cz(np.array([[0, 3], [1, 2]]), g)),
        (lambda v: sim.multi_controlled_x(v, [0, 4], 1),
         lambda g: closed_form.apply_dense(np.array(
             [[1, 0, 0, 0, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0, 0, 0], [0, 0, 1, 0, 0, 0, 0, 0],
              [0, 0, 0, 1, 0, 0, 0, 0], [0, 0, 0, 0, 1, 0, 0, 0], [0, 0, 0, 0, 0, 1, 0, 0],
              [0, 0, 0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 0, 0, 1, 0]]), [0, 4, 1], g)),
    ]
    for fwd, adj in cases:
        x = dev(st, device).requires_grad_()
        (fwd(x) * dev(w, device)).real.sum().backward()
        # d/dx Re sum(w * M x) = conj(M^T w) for torch's conjugate convention
        expect = np.conj(adj(w))
        assert np.allclose(x.grad.cpu().numpy(), expect, atol=1e-14)


def test_entry_points_reject_bad_arguments(device):
    lib = _lib.load()
    x = torch.zeros(2 ** 4, dtype=torch.complex64, device=device)
    out = torch.empty_like(x)
    s = _lib.stream_handle(device)
    c64 = _lib.DTYPE_CODES[torch.complex64]
    assert lib.ua_permute_qubits(_lib.ptr(out), _lib.ptr(x), c64, 4, _lib.int_array([0, 1, 1, 3]),
                                 1, 16, s) == 1
    assert lib.ua_permute_qubits(_lib.ptr(x), _lib.ptr(x), c64, 4, _lib.int_array([0, 1, 2, 3]),
                                 1, 16, s) == 1
    assert lib.ua_apply_sign_gates(_lib.ptr(out), _lib.ptr(x), c64, 4, 1, _lib.int_array([2, 2]),
                                   0, 0, 1, 16, s) == 1
    assert lib.ua_apply_controlled_flip(_lib.ptr(out), _lib.ptr(x), c64, 4, 0b0101, 2, 1, 16,
                                        s) == 1
