"""CPU restatement of unitair's gate-application algorithm in numpy.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it, and
never as the thing measured or shipped.  The product (unitair_amd) never
imports anything under oracle/.

Each function restates the reference step it cites (paths relative to the
reference's src/unitair/).  Pins (tests/test_oracle.py): the documented
known answers in README.rst / docs/tutorial (tests/golden/), a brute-force
2^n x 2^n matrix built from the index definition, the closed forms in
tests/test_unitary.py:12-27, the invariants in tests/test_operations.py, and
agreement with the independent C closed form (oracle/closed_form.c).
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np


def count_qubits(state: np.ndarray) -> int:
    """states/shapes.py:13-27: last dim must be 2**n."""
    length = state.shape[-1]
    n = round(math.log2(length))
    if 2 ** n != length:
        raise ValueError(f"state size {state.shape} is not (*batch, 2**n)")
    return n


def count_qubits_gate_matrix(op: np.ndarray) -> int:
    """states/shapes.py:31-47."""
    length = op.shape[-1]
    k = round(math.log2(length))
    if 2 ** k != length:
        raise RuntimeError(f"gate size {op.shape} is not (*batch, 2**k, 2**k)")
    return k


def permutation_to_front(n: int, entries: Sequence[int]) -> List[int]:
    """utils.py:5-27: `entries` first, in order, then the rest ascending."""
    rest = [q for q in range(n) if q not in entries]
    if len(set(entries)) != len(entries):
        raise ValueError("duplicate qubit")
    return list(entries) + rest


def inverse_list_permutation(perm: Sequence[int]) -> List[int]:
    """utils.py:30-49."""
    inv = [0] * len(perm)
    for i, p in enumerate(perm):
        inv[p] = i
    return inv


def act_first_qubits_tensor(op: np.ndarray, state_t: np.ndarray, n: int, k: int):
    """simulation/operations.py:258-329 on a tensor-layout state
    (*batch, 2, ..., 2): contract the op with the first k qubit dims; batch
    dims of op and state are moved last and broadcast as einsum's ellipsis."""
    D = 2 ** k
    sbd = state_t.ndim - n
    obd = op.ndim - 2
    sb = state_t.shape[:sbd]
    if obd > 0 and sbd == 0:                       # operations.py:304-309
        state_t = np.broadcast_to(state_t, op.shape[:obd] + state_t.shape)
        sbd = obd
        sb = op.shape[:obd]
    st = np.moveaxis(state_t, list(range(sbd)), list(range(n, n + sbd)))
    o = np.moveaxis(op, list(range(obd)), list(range(2, 2 + obd)))
    old = st.shape
    view = st.reshape((D,) + (2,) * (n - k) + tuple(sb))
    res = np.einsum("ab...,b...->a...", o, view)
    if res.size != int(np.prod(old)):
        raise RuntimeError("operator batch does not broadcast against the state")
    res = res.reshape(old)
    return np.moveaxis(res, list(range(n, n + sbd)), list(range(sbd)))


def apply_operator(op: np.ndarray, qubits: Sequence[int], state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:45-186: permute targets to the front, act,
    permute back."""
    n = count_qubits(state)
    qubits = list(qubits)
    if not set(qubits).issubset(range(n)):
        raise ValueError("qubits out of range")
    k = count_qubits_gate_matrix(op)
    if len(qubits) != k:
        raise ValueError("qubit count does not match the operator")
    sbd = state.ndim - 1
    st = state.reshape(state.shape[:-1] + (2,) * n)
    perm = permutation_to_front(n, qubits)
    inv = inverse_list_permutation(perm)
    st = np.transpose(st, list(range(sbd)) + [sbd + p for p in perm])
    st = act_first_qubits_tensor(op, np.ascontiguousarray(st), n, k)
    sbd = st.ndim - n
    st = np.transpose(st, list(range(sbd)) + [sbd + p for p in inv])
    return np.ascontiguousarray(st).reshape(st.shape[:sbd] + (2 ** n,))


def act_first_qubits(op: np.ndarray, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:236-255."""
    n = count_qubits(state)
    k = count_qubits_gate_matrix(op)
    st = state.reshape(state.shape[:-1] + (2,) * n)
    st = act_first_qubits_tensor(op, st, n, k)
    return np.ascontiguousarray(st).reshape(st.shape[:st.ndim - n] + (2 ** n,))


def apply_phase(angles: np.ndarray, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:15-42: exp(-1j*angles) * state.  The factor is
    formed in the angles' precision (complex64 for float32 angles)."""
    angles = np.asarray(angles)
    if angles.dtype == np.float32 or angles.dtype == np.complex64:
        factor = np.exp(np.complex64(-1j) * angles.astype(np.complex64))
    else:
        factor = np.exp(-1j * angles.astype(np.complex128))
    return factor * state


def apply_all_qubits(op: np.ndarray, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:332-413: act on qubit 0, then for i = 1..n-1
    swap qubit dims (0, i) and act on dim 0 again, finally roll qubits by -1."""
    if count_qubits_gate_matrix(op) != 1:
        raise ValueError("apply_all_qubits needs a 1-qubit operator")
    n = count_qubits(state)
    st = state.reshape(state.shape[:-1] + (2,) * n)
    st = act_first_qubits_tensor(op, st, n, 1)
    for i in range(1, n):
        b = st.ndim - n
        st = np.swapaxes(st, b, b + i)                      # swap_tensor :520-537
        st = act_first_qubits_tensor(op, np.ascontiguousarray(st), n, 1)
    b = st.ndim - n
    order = list(range(b)) + [b + (q + 1) % n for q in range(n)]   # roll :565-597
    st = np.transpose(st, order)
    return np.ascontiguousarray(st).reshape(st.shape[:b] + (2 ** n,))


def fuse_single_qubit_operators(qubits, operators):
    """gates/matrix_algebra.py:6-51: gates on the same qubit are multiplied,
    later gate on the left."""
    fused = {}
    for q, op in zip(qubits, operators):
        fused[q] = np.matmul(op, fused[q]) if q in fused else op
    return fused


def apply_to_qubits(operators, qubits, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:416-503: fuse per qubit; act on qubit 0 first;
    for every other qubit swap its dim with dim 0 and act on dim 0, tracking
    the permutation; finally permute the qubit dims back."""
    n = count_qubits(state)
    st = state.reshape(state.shape[:-1] + (2,) * n)
    fused = fuse_single_qubit_operators(qubits, operators)
    if 0 in fused:
        st = act_first_qubits_tensor(fused.pop(0), st, n, 1)
    perm = list(range(n))
    for q, gate in fused.items():
        perm[q], perm[0] = perm[0], q
        b = st.ndim - n
        st = np.swapaxes(st, b, b + q)                      # swap_tensor :520-537
        st = act_first_qubits_tensor(gate, np.ascontiguousarray(st), n, 1)
    inv = inverse_list_permutation(perm)
    b = st.ndim - n
    st = np.transpose(st, list(range(b)) + [b + p for p in inv])   # :626-654
    return np.ascontiguousarray(st).reshape(st.shape[:b] + (2 ** n,))


def act_last_qubit(op: np.ndarray, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:189-233: einsum('ab, ...b -> ...a') over the
    last qubit's tensor dim."""
    n = count_qubits(state)
    st = state.reshape(state.shape[:-1] + (2 ** (n - 1), 2))
    out = np.einsum("ab,...b->...a", op, st)
    return out.reshape(state.shape)


def permute_qubits(permutation, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:600-654: permute the qubit dims of the tensor
    layout (output dim q = input dim permutation[q])."""
    n = count_qubits(state)
    st = state.reshape(state.shape[:-1] + (2,) * n)
    b = st.ndim - n
    st = np.transpose(st, list(range(b)) + [b + p for p in permutation])
    return np.ascontiguousarray(st).reshape(state.shape)


def swap(state: np.ndarray, qubit_pair) -> np.ndarray:
    """simulation/operations.py:506-537: transpose of two qubit dims."""
    n = count_qubits(state)
    st = state.reshape(state.shape[:-1] + (2,) * n)
    b = st.ndim - n
    q1, q2 = (q + b if q >= 0 else q for q in qubit_pair)
    return np.ascontiguousarray(np.swapaxes(st, q1, q2)).reshape(state.shape)


def roll_qubits(state: np.ndarray, num_steps: int = 1) -> np.ndarray:
    """simulation/operations.py:540-597."""
    n = count_qubits(state)
    k = num_steps % n
    ident = list(range(n))
    return permute_qubits(ident[-k:] + ident[:-k] if k else ident, state)


def multi_cz(qubit_pairs: np.ndarray, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:657-748: phase = prod over pairs of
    (1 - 2 [x has both bits 2^(n-1-c), 2^(n-1-t)])."""
    n = count_qubits(state)
    pairs = np.asarray(qubit_pairs).reshape(-1, 2)
    crit = 2 ** (n - pairs[:, 0] - 1) + 2 ** (n - pairs[:, 1] - 1)
    ints = np.arange(2 ** n)
    match = (ints[:, None] & crit[None, :]) == crit[None, :]
    return np.prod(1 - 2 * match, axis=-1) * state


def multi_controlled_z(qubits, state: np.ndarray) -> np.ndarray:
    """simulation/operations.py:751-783: move the qubits last, multiply by
    diag(1, ..., 1, -1) repeated, move them back."""
    n = count_qubits(state)
    qubits = list(qubits)
    inert = [q for q in range(n) if q not in set(qubits)]
    sub = np.ones(2 ** len(qubits))
    sub[-1] = -1.0
    factors = np.tile(sub, 2 ** len(inert))
    perm = inert + qubits
    rev = [0] * n
    for q, i in enumerate(perm):
        rev[i] = q
    return permute_qubits(rev, factors * permute_qubits(perm, state))


def multi_controlled_x(state: np.ndarray, controls, target: int) -> np.ndarray:
    """simulation/operations.py:786-806: H on the target, C...CZ, H again."""
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    state = apply_operator(h, [target], state)
    state = multi_controlled_z(list(controls) + [target], state)
    return apply_operator(h, [target], state)


# ---------------------------------------------------------------- reductions

def abs_squared(state):
    """states/innerprod.py:4-26."""
    return (np.conj(state) * state).real


def norm_squared(state):
    """states/innerprod.py:29-46, accumulated in float64."""
    return np.sum(abs_squared(np.asarray(state, dtype=np.complex128)), axis=-1)


def diag_expectation_value(diag, state):
    """states/innerprod.py:49-59, accumulated in float64/complex128."""
    return np.sum(abs_squared(np.asarray(state, dtype=np.complex128)) * diag, axis=-1)


def inner_product(a, b):
    """states/innerprod.py:62-65, accumulated in complex128."""
    return np.sum(np.conj(np.asarray(a, np.complex128)) * np.asarray(b, np.complex128), axis=-1)


# ------------------------------------------------- brute force (pins above)

def full_matrix(op: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """The 2^n x 2^n matrix of `op` on `qubits`, entry by entry from the index
    definition: qubit q is bit n-1-q of the state index and qubits[0] is the
    most significant bit of the gate index.  Pure loops; n <= 8."""
    k = len(qubits)
    N = 2 ** n
    M = np.zeros((N, N), dtype=np.complex128)
    bits = [n - 1 - q for q in qubits]
    mask = sum(1 << b for b in bits)
    for i in range(N):
        r = 0
        for b in bits:
            r = (r << 1) | ((i >> b) & 1)
        for c in range(2 ** k):
            j = i & ~mask
            for m, b in enumerate(bits):
                if (c >> (k - 1 - m)) & 1:
                    j |= 1 << b
            M[i, j] += op[r, c]
    return M


def fidelity_error(ref: np.ndarray, x: np.ndarray) -> np.ndarray:
    """1 - |<ref, x>|^2 / (|ref|^2 |x|^2) per state, in float64."""
    ref = np.asarray(ref, np.complex128)
    x = np.asarray(x, np.complex128)
    ip = np.sum(np.conj(ref) * x, axis=-1)
    return 1.0 - np.abs(ip) ** 2 / (np.sum(np.abs(ref) ** 2, -1) * np.sum(np.abs(x) ** 2, -1))


def max_abs_error(ref: np.ndarray, x: np.ndarray) -> float:
    return float(np.max(np.abs(np.asarray(ref, np.complex128) - np.asarray(x, np.complex128))))
