"""The general adjoint (unitair_amd.adjoint.expectation): random circuits of
exp_x / exp_y / exp_z rotations (unitair/gates/gates.py:7-91) with shared
parameters and coefficients, CZ, CNOT and fixed gates, and a random real
diagonal observable (innerprod.py:49-59).  Gradients are checked against the
exact parameter-shift rule on the complex128 oracle, central differences,
and torch autograd through the CPU restatement of the reference."""
import numpy as np
import pytest
import torch

from unitair_amd import adjoint
from oracle import unitair_torch as T
import adjoint_oracle as AO

pytestmark = pytest.mark.gpu


def _run(device, params, gates, obs, n, dtype, psi0=None, compiled=False):
    p = torch.from_numpy(params).to(device, torch.float64 if dtype == torch.complex128
                                    else torch.float32).requires_grad_(True)
    o = torch.from_numpy(obs).to(device).requires_grad_(True)
    init = None if psi0 is None else torch.from_numpy(psi0).to(device)
    loss = adjoint.expectation(p, gates, o, n, dtype=dtype, initial_state=init,
                               compiled=compiled)
    loss.sum().backward()
    return loss.detach().cpu().numpy(), p.grad.double().cpu().numpy(), o.grad.cpu().numpy()


@pytest.mark.parametrize("n,ngates,P,B,seed", [(4, 30, 6, 2, 0), (5, 60, 12, 3, 1),
                                               (7, 80, 20, 2, 2), (13, 50, 16, 2, 3)])
def test_general_adjoint_matches_parameter_shift_c128(device, n, ngates, P, B, seed):
    rng = np.random.default_rng(seed)
    gates = AO.random_rotation_circuit(rng, n, ngates, P)
    params = rng.uniform(-np.pi, np.pi, (B, P))
    obs = rng.standard_normal(2 ** n)
    loss, g, go = _run(device, params, gates, obs, n, torch.complex128)
    for b in range(B):
        assert abs(loss[b] - AO.general_loss(params[b], gates, obs, n)) < 1e-11
        for i in range(P):
            ps = AO.general_param_shift(params[b], gates, obs, n, i)
            assert abs(g[b, i] - ps) < 1e-10, (b, i)
    # d loss / d obs_x = sum_b |psi_b(x)|^2
    probs = sum(np.abs(AO.run_gates(params[b], gates, n)) ** 2 for b in range(B))
    assert np.allclose(go, probs, atol=1e-12)


def test_general_adjoint_central_differences_and_batched_observable(device):
    n, P, B = 6, 10, 3
    rng = np.random.default_rng(9)
    gates = AO.random_rotation_circuit(rng, n, 50, P)
    params = rng.uniform(-np.pi, np.pi, (B, P))
    obs = rng.standard_normal((B, 2 ** n))
    psi0 = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    psi0 /= np.linalg.norm(psi0)
    loss, g, go = _run(device, params, gates, obs, n, torch.complex128, psi0=psi0)
    h = 1e-4
    for b in range(B):
        assert abs(loss[b] - AO.general_loss(params[b], gates, obs[b], n, psi0)) < 1e-11
        assert np.allclose(go[b], np.abs(AO.run_gates(params[b], gates, n, psi0)) ** 2,
                           atol=1e-12)
        for i in range(P):
            pp, pm = params[b].copy(), params[b].copy()
            pp[i] += h
            pm[i] -= h
            fd = (AO.general_loss(pp, gates, obs[b], n, psi0) -
                  AO.general_loss(pm, gates, obs[b], n, psi0)) / (2 * h)
            assert abs(g[b, i] - fd) < 1e-6


def test_general_adjoint_matches_torch_autograd_through_restatement(device):
    """The same circuit built from unitair's exp gates and applied with the
    torch restatement of apply_operator, differentiated by torch autograd."""
    n, P, B = 5, 8, 2
    rng = np.random.default_rng(4)
    gates = AO.random_rotation_circuit(rng, n, 40, P)
    params = rng.uniform(-np.pi, np.pi, (B, P))
    obs = rng.standard_normal(2 ** n)

    def exp_t(kind, t):
        c, s = torch.cos(t).to(torch.complex128), torch.sin(t).to(torch.complex128)
        z = torch.zeros_like(c)
        if kind == "x":
            rows = [[c, -1j * s], [-1j * s, c]]
        elif kind == "y":
            rows = [[c, -s], [s, c]]
        else:
            rows = [[c - 1j * s, z], [z, c + 1j * s]]
        return torch.stack([torch.stack(r, -1) for r in rows], -2)

    p = torch.from_numpy(params).requires_grad_(True)
    psi = torch.zeros(B, 2 ** n, dtype=torch.complex128)
    psi[:, 0] = 1
    for g in gates:
        if g[0] in ("x", "y", "z"):
            psi = T.apply_operator(exp_t(g[0], g[3] * p[:, g[2]]), [g[1]], psi)
        elif g[0] in ("cz", "cnot"):
            m = AO.CZ if g[0] == "cz" else AO.CNOT
            psi = T.apply_operator(torch.from_numpy(m), [g[1], g[2]], psi)
        else:
            psi = T.apply_operator(torch.from_numpy(g[2]), list(g[1]), psi)
    T.diag_expectation_value(torch.from_numpy(obs), psi).sum().backward()
    _, g_gpu, _ = _run(device, params, gates, obs, n, torch.complex128)
    assert np.allclose(g_gpu, p.grad.numpy(), atol=1e-11)


@pytest.mark.parametrize("n", [14, 17, 20])
def test_general_adjoint_high_qubits_c128(device, n):
    """X and Y rotations on qubits whose index bits are >= 12 go through the
    windowed high-bit gradient kernels (all windows, with the other high
    bits split across workgroups): checked against parameter shift."""
    rng = np.random.default_rng(n)
    P = 6
    gates = [("y", 0, 0, 0.7), ("x", 1, 1, 1.1), ("cnot", 0, n - 1), ("x", 2, 2, -0.4),
             ("y", n - 1, 3, 0.9), ("cz", 1, n - 2), ("x", 0, 4, 1.3), ("y", 3, 5, 0.5),
             ("z", 2, 0, 0.8), ("x", n - 13, 1, 0.6)]
    params = rng.uniform(-np.pi, np.pi, (1, P))
    obs = rng.standard_normal(2 ** n)
    loss, g, _ = _run(device, params, gates, obs, n, torch.complex128)
    assert abs(loss[0] - AO.general_loss(params[0], gates, obs, n)) < 1e-10
    for i in range(P):
        assert abs(g[0, i] - AO.general_param_shift(params[0], gates, obs, n, i)) < 1e-9, i


def test_general_adjoint_complex64_tolerance(device):
    """complex64 gradients against the complex128 ones on a 12-qubit, 300-gate
    circuit (float64 reductions, checkpointed psi)."""
    n, P, B = 12, 40, 4
    rng = np.random.default_rng(12)
    gates = AO.random_rotation_circuit(rng, n, 300, P)
    params = rng.uniform(-np.pi, np.pi, (B, P))
    obs = rng.standard_normal(2 ** n)
    l64, g64, _ = _run(device, params, gates, obs, n, torch.complex64)
    l128, g128, _ = _run(device, params, gates, obs, n, torch.complex128)
    scale = np.abs(obs).max()
    assert np.max(np.abs(l64 - l128)) < 2e-5 * scale
    assert np.max(np.abs(g64 - g128)) < 5e-5 * scale


@pytest.mark.parametrize("n,dtype", [(13, torch.complex128), (14, torch.complex64)])
def test_general_adjoint_compiled_passes(device, n, dtype):
    """compiled=True (every pass a compiled kernel) against parameter shift
    on the oracle (complex128) and against the interpreted passes; a second
    call reuses the compiled kernels."""
    rng = np.random.default_rng(n + 40)
    P, B = 10, 2
    gates = AO.random_rotation_circuit(rng, n, 60, P)
    params = rng.uniform(-np.pi, np.pi, (B, P))
    obs = rng.standard_normal(2 ** n)
    loss, g, go = _run(device, params, gates, obs, n, dtype, compiled=True)
    loss_i, g_i, go_i = _run(device, params, gates, obs, n, dtype)
    scale = np.abs(obs).max()
    tol = {torch.complex64: 2e-5, torch.complex128: 1e-11}[dtype] * scale
    assert np.max(np.abs(loss - loss_i)) < tol and np.max(np.abs(g - g_i)) < tol
    assert np.allclose(go, go_i, atol=tol)
    if dtype == torch.complex128:
        for i in range(P):
            ps = AO.general_param_shift(params[0], gates, obs, n, i)
            assert abs(g[0, i] - ps) < 1e-9, i
    loss2, g2, _ = _run(device, params, gates, obs, n, dtype, compiled=True)
    assert np.array_equal(loss2, loss) and np.array_equal(g2, g)
