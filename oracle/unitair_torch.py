"""Torch CPU restatement of unitair's gate-application op sequence.

TEST INFRASTRUCTURE ONLY: the timed CPU baseline in bench.py ("port" of the
reference's torch path, run on the host cores) and, being differentiable,
the source of autograd gradients that the adjoint tests compare with.
It replays the reference's torch calls step for step:
  apply_operator   operations.py:45-186  view -> permute().contiguous() ->
                   einsum('ab...,b...->a...') (operations.py:322) ->
                   inverse permute().contiguous() -> view
  apply_phase      operations.py:41-42   exp(-1j*angles) * state
  apply_all_qubits operations.py:332-413 act on qubit 0, then (swap 0<->i,
                   act on qubit 0) for i = 1..n-1, roll by -1, contiguous
  apply_to_qubits  operations.py:416-503 (fuse per qubit, swap, act, permute)
  permute / swap   operations.py:506-537, 600-654
  multi_cz         operations.py:657-748
  sampling         simulation/measurement.py:45-48
  reductions       states/innerprod.py:4-65
"""
from __future__ import annotations

from typing import Sequence

import torch


def _to_front(n: int, qubits: Sequence[int]):
    perm = list(qubits) + [q for q in range(n) if q not in qubits]   # utils.py:5-27
    inv = [0] * n
    for i, p in enumerate(perm):                                      # utils.py:30-49
        inv[p] = i
    return perm, inv


def act_first_qubits_tensor(op: torch.Tensor, st: torch.Tensor, n: int, k: int):
    """operations.py:258-329."""
    D = 2 ** k
    sbd, obd = st.dim() - n, op.dim() - 2
    sb = st.shape[:sbd]
    if obd > 0 and sbd == 0:
        st = st.expand(*op.shape[:obd], *st.shape)
        sbd, sb = obd, op.shape[:obd]
    st = st.permute(list(range(sbd, st.dim())) + list(range(sbd)))
    o = op.permute(list(range(obd, op.dim())) + list(range(obd)))
    old = st.shape
    res = torch.einsum("ab...,b...->a...", o, st.reshape((D,) + (2,) * (n - k) + tuple(sb)))
    res = res.reshape(old)
    d = res.dim()
    return res.permute(list(range(d - sbd, d)) + list(range(d - sbd)))


def apply_operator(op: torch.Tensor, qubits: Sequence[int], state: torch.Tensor):
    """operations.py:45-186."""
    N = state.shape[-1]
    n = N.bit_length() - 1
    k = len(qubits)
    perm, inv = _to_front(n, list(qubits))
    sbd = state.dim() - 1
    st = state.reshape(state.shape[:-1] + (2,) * n)
    st = st.permute(list(range(sbd)) + [sbd + p for p in perm]).contiguous()
    st = act_first_qubits_tensor(op, st, n, k)
    sbd = st.dim() - n
    st = st.permute(list(range(sbd)) + [sbd + p for p in inv]).contiguous()
    return st.reshape(st.shape[:sbd] + (N,))


def apply_phase(angles: torch.Tensor, state: torch.Tensor):
    """operations.py:15-42."""
    return torch.exp(-1.0j * angles) * state


def norm_squared(state: torch.Tensor):
    """innerprod.py:29-46."""
    return torch.sum((state.conj() * state).real, dim=-1)


def diag_expectation_value(diag: torch.Tensor, state: torch.Tensor):
    """innerprod.py:49-59."""
    return torch.sum((state.conj() * state).real * diag, dim=-1)


def apply_all_qubits(op: torch.Tensor, state: torch.Tensor):
    """operations.py:332-413 (swap_tensor :520-537, roll_qubits_tensor
    :565-597, to_vector_layout states/conversions.py:48-92)."""
    N = state.shape[-1]
    n = N.bit_length() - 1
    st = state.reshape(state.shape[:-1] + (2,) * n)
    st = act_first_qubits_tensor(op, st, n, 1)
    for i in range(1, n):
        sbd = st.dim() - n
        st = st.transpose(sbd, sbd + i)
        st = act_first_qubits_tensor(op, st, n, 1)
    sbd = st.dim() - n
    steps = (-1) % n
    if steps:
        ident = list(range(sbd, sbd + n))
        st = st.permute(list(range(sbd)) + ident[-steps:] + ident[:-steps])
    return st.contiguous().reshape(st.shape[:sbd] + (N,))


def apply_to_qubits(operators, qubits, state: torch.Tensor):
    """operations.py:416-503 with fuse_single_qubit_operators
    (gates/matrix_algebra.py:6-51)."""
    N = state.shape[-1]
    n = N.bit_length() - 1
    fused = {}
    for q, op in zip(qubits, operators):
        fused[q] = torch.matmul(op, fused[q]) if q in fused else op
    st = state.reshape(state.shape[:-1] + (2,) * n)
    if 0 in fused:
        st = act_first_qubits_tensor(fused.pop(0), st, n, 1)
    perm = list(range(n))
    for q, gate in fused.items():
        perm[q], perm[0] = perm[0], q
        sbd = st.dim() - n
        st = st.transpose(sbd, sbd + q)
        st = act_first_qubits_tensor(gate, st, n, 1)
    _, inv = _to_front(n, perm)
    sbd = st.dim() - n
    st = st.permute(list(range(sbd)) + [sbd + p for p in inv])
    return st.contiguous().reshape(st.shape[:sbd] + (N,))


def permute_qubits(permutation, state: torch.Tensor):
    """operations.py:600-654 (contiguous, as to_vector_layout leaves it)."""
    N = state.shape[-1]
    n = N.bit_length() - 1
    st = state.reshape(state.shape[:-1] + (2,) * n)
    sbd = st.dim() - n
    st = st.permute(list(range(sbd)) + [sbd + p for p in permutation])
    return st.contiguous().reshape(state.shape)


def swap(state: torch.Tensor, qubit_pair):
    """operations.py:506-537."""
    n = state.shape[-1].bit_length() - 1
    perm = list(range(n))
    i, j = qubit_pair
    perm[i], perm[j] = perm[j], perm[i]
    return permute_qubits(perm, state)


def multi_cz(qubit_pairs: torch.Tensor, state: torch.Tensor):
    """operations.py:657-748 (batched-pairs branch)."""
    n = state.shape[-1].bit_length() - 1
    pairs = qubit_pairs.view(-1, 2)
    crit = 2 ** (n - pairs[:, 0] - 1) + 2 ** (n - pairs[:, 1] - 1)
    ints = torch.arange(2 ** n)
    match = torch.bitwise_and(ints.view(-1, 1), crit.view(1, -1)) == crit
    return torch.prod(1 - 2 * match, dim=-1) * state


def sample_indices(state: torch.Tensor, num_samples: int):
    """measurement.py:45-48: Categorical(probs=abs_squared(state)).sample."""
    probs = abs_squared(state)
    return torch.distributions.Categorical(probs=probs).sample([num_samples])


def abs_squared(state: torch.Tensor):
    """innerprod.py:4-26."""
    return (state.conj() * state).real


def inner_product(a: torch.Tensor, b: torch.Tensor):
    """innerprod.py:62-65."""
    return torch.sum(a.conj() * b, dim=-1)
