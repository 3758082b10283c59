"""Gate application on ROCm GPUs (mirror of ``unitair.simulation``'s hot path).

``apply_operator``, ``act_first_qubits``, ``apply_phase`` and
``apply_all_qubits`` keep the reference's names, argument meaning, qubit
ordering, batch broadcasting, validation order and exception types
(unitair/simulation/operations.py:15-413).  Validation and broadcast
analysis run here on the host; the arithmetic runs in libunitair_amd.so.
Inputs are never modified and every result is a new contiguous tensor.

Tensors must be on a ROCm GPU.  There is no CPU path and no fallback: a
missing library raises ImportError on first use.
"""
from __future__ import annotations

import math
from typing import Iterable, List, Sequence, Tuple

import torch

from .. import _lib
from .._device import on_device, prepare, require_gpu, same_device
from ..states import count_qubits, count_qubits_gate_matrix

__all__ = ["apply_operator", "act_first_qubits", "apply_phase", "apply_all_qubits",
           "apply_to_qubits", "act_last_qubit", "swap", "permute_qubits", "roll_qubits",
           "multi_cz", "multi_controlled_z", "multi_controlled_x", "measure"]

_SUPPORTED = tuple(_lib.DTYPE_CODES)


# ------------------------------------------------------------------ helpers --

def _flat_batch(t: torch.Tensor, inner: int):
    """View `t` as (rows, inner) with contiguous rows at one element stride;
    returns (tensor, row stride).  Copies when no single stride fits."""
    rows = t.numel() // inner if inner else 0
    v = t.reshape(rows, inner)
    if inner > 1 and v.stride(1) != 1:
        v = v.contiguous()
    return v, (v.stride(0) if rows > 1 else inner)


def _check_operator(operator: torch.Tensor, state: torch.Tensor, what: str):
    if operator.dim() < 2:
        raise RuntimeError(
            f"Gate with size {tuple(operator.size())} is incorrectly shaped for an "
            f"operator batch. Expected size is\n  (*optional_batch_dims, 2^k, 2^k)\n"
            f"with k the number of qubits on which the gate acts.")
    if operator.shape[-2] != operator.shape[-1]:
        raise RuntimeError(f"{what}: operator matrices must be square, got size "
                           f"{tuple(operator.size())}")
    require_gpu(state, what)
    same_device(operator, state, what)
    if operator.dtype != state.dtype:
        raise RuntimeError(f"{what}: expected scalar type {state.dtype} for the operator "
                           f"but found {operator.dtype}")
    if state.dtype not in _SUPPORTED:
        raise TypeError(f"{what}: dtype {state.dtype} is not supported "
                        f"(float32, float64, complex64, complex128)")


def _distinct(qubits: List[int]) -> None:
    seen = set()
    for q in qubits:
        if q in seen:
            # the reference fails inside list.index (unitair/utils.py:26)
            raise ValueError(f"{q} is not in list")
        seen.add(q)


def _plan_batch(op_batch: Tuple[int, ...], state_batch: Tuple[int, ...], n: int,
                k: int, qubits: List[int]):
    """Batch structure of one application, following act_first_qubits_tensor
    (operations.py:295-329): the op batch dims are right-aligned, as torch.einsum
    aligns ellipsis dims, against (non-target qubit dims..., state batch dims).

    Returns (out_batch, expand_state, selectors, n_sel_dims, op_batch_part,
    op_batch) where `op_batch` is the op batch without leading size-1 dims
    that reach past every ellipsis dim, its first `n_sel_dims` dims line up
    with non-target qubits, `selectors` lists the (op_dim, qubit) pairs among
    them of size 2 (the operator is chosen by that qubit's value), and
    `op_batch_part` is the rest, aligned with the state batch."""
    expand_state = len(op_batch) > 0 and len(state_batch) == 0
    sb = tuple(op_batch) if expand_state else tuple(state_batch)
    remaining = [q for q in range(n) if q not in set(qubits)]
    ell = [2] * (n - k) + list(sb)            # state's ellipsis dims
    ob = list(op_batch)
    if len(ob) > len(ell):
        extra, ob = ob[:len(ob) - len(ell)], ob[len(ob) - len(ell):]
        if any(d != 1 for d in extra):
            raise RuntimeError(
                f"operator batch {tuple(op_batch)} does not broadcast against state "
                f"batch {tuple(state_batch)}")
    lead = len(ell) - len(ob)                 # first ellipsis dim the op covers
    selectors = []
    for i, d in enumerate(ob):
        e = ell[lead + i]
        if d != 1 and d != e:
            raise RuntimeError(
                f"einsum(): operator batch dim of size {d} does not match size {e} "
                f"(operator batch {tuple(op_batch)}, state batch {tuple(state_batch)})")
        pos = lead + i
        if pos < n - k and d == 2:
            selectors.append((i, remaining[pos]))
    n_sel_dims = max(0, (n - k) - lead)
    return sb, expand_state, selectors, n_sel_dims, tuple(ob[n_sel_dims:]), tuple(ob)


def _multiplexed(op: torch.Tensor, n_sel_dims: int, sel_dims: List[int], k: int):
    """Block-diagonal operator on (selector qubits + targets) equivalent to
    choosing op[s] by the selector qubits' values."""
    D = 2 ** k
    # drop size-1 selector dims, keep the size-2 ones in order
    keep = [slice(None) if i in sel_dims else 0 for i in range(n_sel_dims)]
    sub = op[tuple(keep)] if n_sel_dims else op
    m = len(sel_dims)
    rest = sub.shape[m:-2]
    sub = sub.reshape((2 ** m,) + tuple(rest) + (D, D))
    big = torch.zeros(tuple(rest) + (2 ** m, D, 2 ** m, D), dtype=op.dtype, device=op.device)
    for s in range(2 ** m):
        big[..., s, :, s, :] = sub[s]
    return big.reshape(tuple(rest) + (2 ** m * D, 2 ** m * D))


def _op_for_batch(op: torch.Tensor, op_batch: Tuple[int, ...], sb: Tuple[int, ...],
                  D: int):
    """Returns (op tensor, op batch stride in elements) for a flat batch of
    prod(sb) entries."""
    if len(op_batch) == 0 or math.prod(op_batch) == 1 and all(d == 1 for d in op_batch):
        o = op.reshape(D, D)
        if not o.is_contiguous():
            o = o.contiguous()
        return o, 0
    if tuple(op_batch) == tuple(sb):
        o = op.reshape(-1, D, D)
        if not o.is_contiguous():
            o = o.contiguous()
        return o, D * D
    o = op.expand(tuple(sb) + (D, D)).contiguous().reshape(-1, D, D)
    return o, D * D


def _needs_grad(*ts) -> bool:
    return torch.is_grad_enabled() and any(t.requires_grad for t in ts)


class _ApplyDense(torch.autograd.Function):
    """out[b] = op[b or 0] on `qubits` of state[b or 0].

    Backward (derivation: out = U x on each group of amplitudes):
      grad_state = U^H grad_out          (ua_apply_operator, same kernel)
      grad_op    = sum grad_out x^H      (ua_operator_gradient)
    """

    @staticmethod
    def forward(ctx, op, state, qubits, n, k, out_batch, expand_state, op_bstride):
        ctx.save_for_backward(op, state)
        ctx.meta = (tuple(qubits), n, k, tuple(out_batch), expand_state, op_bstride)
        return _launch_operator(op, qubits, state, n, k, out_batch, expand_state, op_bstride)

    @staticmethod
    def backward(ctx, grad_out):
        op, state = ctx.saved_tensors
        qubits, n, k, out_batch, expand_state, op_bstride = ctx.meta
        g = grad_out.contiguous()
        grad_op = grad_state = None
        if ctx.needs_input_grad[1]:
            op_h = op.conj().transpose(-1, -2).contiguous()
            full = _launch_operator(op_h, qubits, g, n, k, out_batch, False, op_bstride)
            grad_state = full.reshape(-1, 2 ** n).sum(0).reshape(state.shape) \
                if expand_state else full.reshape(state.shape)
        if ctx.needs_input_grad[0]:
            grad_op = _operator_gradient(g, state, op, qubits, n, k, out_batch, expand_state,
                                         op_bstride)
        return grad_op, grad_state, None, None, None, None, None, None


def _operator_gradient(g, state, op, qubits, n, k, out_batch, expand_state, op_bstride):
    lib = _lib.load()
    batch = math.prod(out_batch)
    N = 2 ** n
    grad = torch.empty_like(op)
    if batch == 0:
        return grad.zero_()
    if k > _lib.UA_MAX_GRAD_QUBITS:
        raise NotImplementedError(f"operator gradients support up to "
                                  f"{_lib.UA_MAX_GRAD_QUBITS} qubits")
    s_flat, s_stride = (state.reshape(N), 0) if expand_state else _flat_batch(state, N)
    op_count = 1 if op_bstride == 0 else batch
    # a negative size flags bad arguments, which the entry point then reports
    nbytes = lib.ua_operator_gradient_workspace_bytes(n, k, batch, op_count)
    work = torch.empty(nbytes, dtype=torch.uint8, device=state.device) if nbytes > 0 else None
    with on_device(state.device):
        rc = lib.ua_operator_gradient(
            _lib.ptr(grad), _lib.ptr(g), _lib.ptr(s_flat.contiguous() if expand_state else s_flat),
            _lib.DTYPE_CODES[state.dtype], n, k, _lib.int_array(qubits), batch, op_count,
            s_stride, _lib.ptr(work) if work is not None else None,
            _lib.stream_handle(state.device))
    _lib.check(rc, "operator gradient")
    return grad


def _launch_operator(op, qubits, state, n, k, out_batch, expand_state, op_bstride):
    lib = _lib.load()
    op, state = prepare(op), prepare(state)
    if not op.is_contiguous():
        op = op.contiguous()
    N = 2 ** n
    if expand_state:
        s_flat, s_stride = state.reshape(N), 0
        if not s_flat.is_contiguous():
            s_flat = s_flat.contiguous()
    else:
        s_flat, s_stride = _flat_batch(state, N)
    batch = math.prod(out_batch)
    out = torch.empty(tuple(out_batch) + (N,), dtype=state.dtype, device=state.device)
    if batch == 0:
        return out
    q = _lib.int_array(qubits)
    with on_device(state.device):
        rc = lib.ua_apply_operator(
            _lib.ptr(out), _lib.ptr(s_flat), _lib.ptr(op), _lib.DTYPE_CODES[state.dtype],
            n, k, q, batch, s_stride, op_bstride, _lib.stream_handle(state.device))
    _lib.check(rc, "apply_operator")
    return out


def _apply_dense(operator: torch.Tensor, qubits: List[int], state: torch.Tensor,
                 n: int, k: int, what: str) -> torch.Tensor:
    _check_operator(operator, state, what)
    operator, state = prepare(operator), prepare(state)
    sb, expand_state, selectors, n_sel, op_batch, ob = _plan_batch(
        tuple(operator.shape[:-2]), tuple(state.shape[:-1]), n, k, qubits)
    operator = operator.reshape(ob + tuple(operator.shape[-2:]))
    if n_sel:
        sel_dims = [i for i, _ in selectors]
        operator = _multiplexed(operator, n_sel, sel_dims, k)
        qubits = [q for _, q in selectors] + list(qubits)
        k = len(qubits)
    D = 2 ** k
    op, op_bstride = _op_for_batch(operator, op_batch, sb, D)
    if _needs_grad(op, state):
        if not expand_state:
            state = state.contiguous()
        return _ApplyDense.apply(op, state, list(qubits), n, k, tuple(sb), expand_state,
                                 op_bstride)
    return _launch_operator(op, qubits, state, n, k, sb, expand_state, op_bstride)


# --------------------------------------------------------------- the surface --

def apply_operator(operator: torch.Tensor, qubits: Iterable[int], state: torch.Tensor):
    """Apply `operator` to `qubits` of a vector-layout state (operations.py:45-148).

    operator: (*op_batch, 2^k, 2^k); qubits: k distinct qubits, the first one
    is the most significant bit of the gate index; state: (*state_batch, 2^n).
    Batching: identical batch dims pair up entry by entry; an unbatched
    operator acts on every state; a batched operator on an unbatched state
    yields one output per operator.  More generally the op batch is
    right-aligned against (non-target qubit dims, state batch dims) exactly
    as the reference's einsum aligns it.
    """
    n = count_qubits(state)
    qubits = list(qubits)
    if not set(qubits).issubset(range(n)):
        raise ValueError(f"qubits={qubits} is not consistent with state vector with "
                         f"{n} qubits.")
    k = count_qubits_gate_matrix(operator)
    if len(qubits) != k:
        raise ValueError(f"Cannot apply operator with {k} to the {len(qubits)} "
                         f"qubit sequence {qubits}.")
    _distinct(qubits)
    return _apply_dense(operator, [int(q) for q in qubits], state, n, k, "apply_operator")


def act_first_qubits(operator: torch.Tensor, state: torch.Tensor):
    """Apply a k-qubit operator to qubits 0..k-1 (operations.py:236-255)."""
    n = count_qubits(state)
    k = count_qubits_gate_matrix(operator)
    if n < k:
        raise ValueError(f"Attempted to apply a {k}-qubit gate to {n} qubit(s).")
    return _apply_dense(operator, list(range(k)), state, n, k, "act_first_qubits")


def apply_all_qubits(operator: torch.Tensor, state: torch.Tensor) -> torch.Tensor:
    """Apply one 2x2 operator to every qubit (operations.py:332-413)."""
    k = count_qubits_gate_matrix(operator)
    if k != 1:
        raise ValueError(f"Expected operator on 1 qubit, found a {k} qubit operator.")
    n = count_qubits(state)
    _check_operator(operator, state, "apply_all_qubits")
    operator, state = prepare(operator), prepare(state)
    obatch, sbatch = tuple(operator.shape[:-2]), tuple(state.shape[:-1])
    sb, expand_state, selectors, n_sel, op_batch, _ = _plan_batch(obatch, sbatch, n, 1, [0])
    if selectors:
        # The reference lines such dims up with qubit dims of a state whose
        # qubits it keeps swapping; they are rejected here.
        raise ValueError(
            f"apply_all_qubits: operator batch {obatch} has dims beyond the state "
            f"batch {sbatch} that would be matched against qubit dims")
    operator = operator.reshape(op_batch + (2, 2))
    op, op_bstride = _op_for_batch(operator, op_batch, sb, 2)
    if _needs_grad(op, state):
        # differentiable form: one dense application per qubit
        x = state
        for q in range(n):
            x = _ApplyDense.apply(op, x if (q > 0 or not expand_state) else state, [q], n, 1,
                                  tuple(sb), expand_state and q == 0, op_bstride)
        return x
    lib = _lib.load()
    N = 2 ** n
    if expand_state:
        s_flat, s_stride = state.reshape(N).contiguous(), 0
    else:
        s_flat, s_stride = _flat_batch(state, N)
    batch = math.prod(sb)
    out = torch.empty(tuple(sb) + (N,), dtype=state.dtype, device=state.device)
    if batch == 0:
        return out
    with on_device(state.device):
        rc = lib.ua_apply_all_qubits(
            _lib.ptr(out), _lib.ptr(s_flat), _lib.ptr(op), _lib.DTYPE_CODES[state.dtype],
            n, batch, s_stride, op_bstride, _lib.stream_handle(state.device))
    _lib.check(rc, "apply_all_qubits")
    return out


def apply_to_qubits(operators: Iterable[torch.Tensor], qubits: Iterable[int],
                    state: torch.Tensor) -> torch.Tensor:
    """Apply 1-qubit gates to the listed qubits (operations.py:416-503).

    Gates on the same qubit are multiplied first, in order
    (fuse_single_qubit_operators, unitair/gates/matrix_algebra.py:6-51); the
    fused gates, which act on distinct qubits and commute, then run as one
    1-qubit layer in tiled passes (a few passes over the state instead of one
    per qubit).  operators: (*batch_dims, 2, 2) each, all with the same batch
    dims; zip(operators, qubits) pairs them as in the reference.  Batched
    operators with the state's batch dims (or with an unbatched state) pair up
    entry by entry; other broadcasts apply the fused gates one by one with the
    reference's apply_operator rules.
    """
    from ..gates.matrix_algebra import _native_ok, fuse_1q_native, fuse_single_qubit_operators
    n = count_qubits(state)
    qubits, operators = list(qubits), list(operators)
    raw = operators[:len(qubits)]
    if (raw and not _needs_grad(state) and _native_ok(raw) and
            raw[0].dtype == state.dtype and raw[0].device == state.device and
            all(0 <= int(q) < n for q in qubits[:len(raw)])):
        # no autograd: the gates on each qubit are multiplied by one native
        # launch straight into the 1-qubit layer's operand layout
        return _apply_1q_layer_fused(raw, [int(q) for q in qubits[:len(raw)]], state, n)
    fused = fuse_single_qubit_operators(qubits, operators)
    for q in fused:
        if not (0 <= int(q) < n):
            raise ValueError(f"Expected index in {{-num_qubits, ..., num_qubits - 1}}.\n"
                             f"Num_qubits: {n}, index: {q}.")
    qs = [int(q) for q in fused]
    ops = list(fused.values())
    for op in ops:
        if count_qubits_gate_matrix(op) != 1:
            raise ValueError(f"Expected operator on 1 qubit, found a "
                             f"{count_qubits_gate_matrix(op)} qubit operator.")
        _check_operator(op, state, "apply_to_qubits")
    obatch = tuple(ops[0].shape[:-2]) if ops else ()
    if any(tuple(op.shape[:-2]) != obatch for op in ops):
        raise RuntimeError("apply_to_qubits: all operators must have the same batch dims")
    sbatch = tuple(state.shape[:-1])
    pairable = obatch == () or obatch == sbatch or sbatch == ()
    if ops and (not pairable or _needs_grad(state, *ops)):
        x = state
        for q, op in zip(qs, ops):
            x = apply_operator(op, [q], x)
        return x
    require_gpu(state, "apply_to_qubits")
    if state.dtype not in _SUPPORTED:
        raise TypeError(f"apply_to_qubits: dtype {state.dtype} is not supported")
    state = prepare(state)
    N = 2 ** n
    k = len(qs)
    if obatch and not sbatch:
        out_batch, s_flat, s_stride = obatch, state.reshape(N).contiguous(), 0
    else:
        out_batch = sbatch
        s_flat, s_stride = _flat_batch(state, N)
    # (k, *obatch, 2, 2) -> (*obatch, k, 2, 2): target j of entry b at 4 (b k + j)
    mats = (torch.stack([prepare(op) for op in ops], dim=-3).contiguous() if ops
            else torch.empty(0, dtype=state.dtype, device=state.device))
    batch = math.prod(out_batch)
    out = torch.empty(tuple(out_batch) + (N,), dtype=state.dtype, device=state.device)
    if batch == 0:
        return out
    lib = _lib.load()
    with on_device(state.device):
        rc = lib.ua_apply_1q_layer(
            _lib.ptr(out), _lib.ptr(s_flat), _lib.ptr(mats), _lib.DTYPE_CODES[state.dtype], n,
            k, _lib.int_array(qs), batch, s_stride, 4 * k if obatch else 0,
            _lib.stream_handle(state.device))
    _lib.check(rc, "apply_to_qubits")
    return out


def _apply_1q_layer_fused(ops: List[torch.Tensor], qubits: List[int], state: torch.Tensor,
                          n: int) -> torch.Tensor:
    """apply_to_qubits without autograd: ua_fuse_1q_ops multiplies the gates of
    each qubit into a (*batch, k, 2, 2) operand, then ua_apply_1q_layer runs the
    layer.  ops: equal-shape (*obatch, 2, 2) device tensors of the state's dtype."""
    from ..gates.matrix_algebra import fuse_1q_native
    for op in ops:
        _check_operator(op, state, "apply_to_qubits")
    slot_of = {}
    for q in qubits:
        slot_of.setdefault(q, len(slot_of))
    qs = list(slot_of)
    obatch = tuple(ops[0].shape[:-2])
    sbatch = tuple(state.shape[:-1])
    if not (obatch == () or obatch == sbatch or sbatch == ()):
        # other broadcasts: the reference's per-gate rules on the fused gates
        fused = torch.empty(obatch + (len(qs), 2, 2), dtype=state.dtype, device=state.device)
        fuse_1q_native([prepare(o).contiguous() for o in ops], [slot_of[q] for q in qubits],
                       len(qs), fused)
        x = state
        for s, q in enumerate(qs):
            x = apply_operator(fused[..., s, :, :], [q], x)
        return x
    if state.dtype not in _SUPPORTED:
        raise TypeError(f"apply_to_qubits: dtype {state.dtype} is not supported")
    state = prepare(state)
    N = 2 ** n
    if obatch and not sbatch:
        out_batch, s_flat, s_stride = obatch, state.reshape(N).contiguous(), 0
    else:
        out_batch = sbatch
        s_flat, s_stride = _flat_batch(state, N)
    batch = math.prod(out_batch)
    out = torch.empty(tuple(out_batch) + (N,), dtype=state.dtype, device=state.device)
    if batch == 0:
        return out
    mats = torch.empty(obatch + (len(qs), 2, 2), dtype=state.dtype, device=state.device)
    fuse_1q_native([prepare(o).contiguous() for o in ops], [slot_of[q] for q in qubits],
                   len(qs), mats)
    lib = _lib.load()
    with on_device(state.device):
        rc = lib.ua_apply_1q_layer(
            _lib.ptr(out), _lib.ptr(s_flat), _lib.ptr(mats), _lib.DTYPE_CODES[state.dtype], n,
            len(qs), _lib.int_array(qs), batch, s_stride, 4 * len(qs) if obatch else 0,
            _lib.stream_handle(state.device))
    _lib.check(rc, "apply_to_qubits")
    return out


def act_last_qubit(single_qubit_operator: torch.Tensor, state: torch.Tensor) -> torch.Tensor:
    """Apply a 2x2 operator to the last qubit (operations.py:189-213); like the
    reference, warns that the function is outdated.  The operator has no batch
    dims (the reference contracts it with einsum('ab, ...b -> ...a'))."""
    import warnings
    warnings.warn(
        'act_last_qubit and act_last_qubit_tensor are outdated. These\n'
        'functions will be removed from Unitair in a later release or may\n'
        'be revised to meet our current standards.\n'
        'Please consider using apply_operator instead.')
    n = count_qubits(state)
    if single_qubit_operator.dim() != 2:
        raise RuntimeError("einsum(): the number of subscripts in the equation (2) does not "
                           "match the number of dimensions "
                           f"({single_qubit_operator.dim()}) for operand 0 and no ellipsis "
                           "was given")
    return apply_operator(single_qubit_operator, [n - 1], state)


# ------------------------------------------- permutations and sign gates --

def _linear_launch(kind: str, state: torch.Tensor, n: int, args) -> torch.Tensor:
    """Out-of-place launch of one of the index-map kernels on (*batch, 2^n)."""
    N = 2 ** n
    s_flat, s_stride = _flat_batch(state, N)
    batch = s_flat.shape[0]
    out = torch.empty(tuple(state.shape[:-1]) + (N,), dtype=state.dtype, device=state.device)
    if batch == 0:
        return out
    lib = _lib.load()
    code = _lib.DTYPE_CODES[state.dtype]
    stream = _lib.stream_handle(state.device)
    with on_device(state.device):
        if kind == "permute":
            rc = lib.ua_permute_qubits(_lib.ptr(out), _lib.ptr(s_flat), code, n,
                                       _lib.int_array(args), batch, s_stride, stream)
        elif kind == "sign":
            pairs, use_all, mask = args
            flat = [b for pr in pairs for b in pr]
            rc = lib.ua_apply_sign_gates(_lib.ptr(out), _lib.ptr(s_flat), code, n, len(pairs),
                                         _lib.int_array(flat), use_all, mask, batch, s_stride,
                                         stream)
        else:
            ctrl, tbit = args
            rc = lib.ua_apply_controlled_flip(_lib.ptr(out), _lib.ptr(s_flat), code, n, ctrl,
                                              tbit, batch, s_stride, stream)
    _lib.check(rc, kind)
    return out


class _IndexMap(torch.autograd.Function):
    """A permutation or +-1 diagonal of the amplitudes; the backward pass
    applies the adjoint map (the inverse permutation; the same sign or
    involutive flip)."""

    @staticmethod
    def forward(ctx, state, kind, n, args, adj_args):
        ctx.kind, ctx.n, ctx.adj_args = kind, n, adj_args
        return _linear_launch(kind, state, n, args)

    @staticmethod
    def backward(ctx, g):
        return _linear_launch(ctx.kind, prepare(g.contiguous()), ctx.n, ctx.adj_args), \
            None, None, None, None


def _index_map(state: torch.Tensor, what: str, kind: str, n: int, args, adj_args):
    require_gpu(state, what)
    if state.dtype not in _SUPPORTED:
        raise TypeError(f"{what}: dtype {state.dtype} is not supported")
    state = prepare(state)
    if _needs_grad(state):
        return _IndexMap.apply(state, kind, n, args, adj_args)
    return _linear_launch(kind, state, n, args)


def _inverse(perm: List[int]) -> List[int]:
    inv = [0] * len(perm)
    for i, p in enumerate(perm):
        inv[p] = i
    return inv


def permute_qubits(permutation: Iterable[int], state_vector: torch.Tensor) -> torch.Tensor:
    """Permute qubits (operations.py:600-654): output qubit q is input qubit
    permutation[q], as torch.permute on the qubit dims of the tensor layout;
    the same permutation for every batch entry."""
    n = count_qubits(state_vector)
    perm = [int(p) for p in permutation]
    if sorted(perm) != list(range(n)):
        raise RuntimeError(f"permute_qubits: {perm} is not a permutation of the {n} qubits")
    return _index_map(state_vector, "permute_qubits", "permute", n, perm, _inverse(perm))


def swap(state: torch.Tensor, qubit_pair: Tuple[int, int]) -> torch.Tensor:
    """Swap two qubits (operations.py:506-537).  Indices follow
    get_qubit_indices (states/shapes.py:72-116): -n <= q < n."""
    n = count_qubits(state)
    q1, q2 = (int(q) for q in qubit_pair)
    for q in (q1, q2):
        if q >= n or q < -n:
            raise ValueError(f"Expected index in {{-num_qubits, ..., num_qubits - 1}}.\n"
                             f"Num_qubits: {n}, index: {list(qubit_pair)}.")
    if q1 == q2:
        return state
    q1, q2 = q1 % n, q2 % n
    perm = list(range(n))
    perm[q1], perm[q2] = perm[q2], perm[q1]
    return _index_map(state, "swap", "permute", n, perm, perm)


def roll_qubits(state: torch.Tensor, num_steps: int = 1) -> torch.Tensor:
    """Cyclic permutation of the qubits (operations.py:540-597): with
    k = num_steps mod n, output qubit i is input qubit (i - k) mod n (the
    reference's permute(identity[-k:] + identity[:-k])), so rolling by 1
    moves <0100| to <0010|."""
    n = count_qubits(state)
    k = num_steps % n
    if k == 0:
        return state
    ident = list(range(n))
    perm = ident[-k:] + ident[:-k]
    return _index_map(state, "roll_qubits", "permute", n, perm, _inverse(perm))


def multi_cz(qubit_pairs: torch.Tensor, state_vector: torch.Tensor,
             num_bits_memory_cutoff=None) -> torch.Tensor:
    """CZ on every listed pair (operations.py:657-748); pairs commute, a pair
    listed twice cancels.  num_bits_memory_cutoff only steered the reference's
    memory use and is accepted for compatibility."""
    n = count_qubits(state_vector)
    if qubit_pairs.dim() == 1:
        qubit_pairs = qubit_pairs.view(1, 2)
    if state_vector.device.type != qubit_pairs.device.type:
        raise ValueError("state_vector and qubit_qubit_pairs must have same device.")
    elif qubit_pairs.dim() != 2:
        raise ValueError("qubit_pairs must have size (2,) or (num_pairs, 2).")
    rows = qubit_pairs.tolist()          # one device->host copy for every check
    if any(v >= n for r in rows for v in r):
        raise ValueError("Control/target indices for CZ gate must be less than num_bits.")
    if any(r[0] == r[1] for r in rows):
        raise ValueError("Control and target qubits are not distinct.")
    # a negative index puts its bit beyond the state, so the reference never
    # matches that pair: it has no effect
    pairs = [(n - 1 - int(r[0]), n - 1 - int(r[1])) for r in rows if r[0] >= 0 and r[1] >= 0]
    args = (pairs, 0, 0)
    return _index_map(state_vector, "multi_cz", "sign", n, args, args)


def multi_controlled_z(qubits: Iterable[int], state_vector: torch.Tensor) -> torch.Tensor:
    """C...CZ on `qubits` (operations.py:751-783): -1 where all of them are 1.
    With no qubits every amplitude changes sign, as in the reference."""
    n = count_qubits(state_vector)
    qubits = [int(q) for q in qubits]
    if len(set(qubits)) != len(qubits):
        raise RuntimeError(f"multi_controlled_z: repeated qubit in {qubits}")
    if any(q < 0 or q >= n for q in qubits):
        raise IndexError(f"multi_controlled_z: qubits {qubits} out of range for {n} qubits")
    mask = 0
    for q in qubits:
        mask |= 1 << (n - 1 - q)
    args = ([], 1, mask)
    return _index_map(state_vector, "multi_controlled_z", "sign", n, args, args)


def multi_controlled_x(state_vector: torch.Tensor, controls: Iterable[int],
                       target: int) -> torch.Tensor:
    """C...CX (operations.py:786-806): flips `target` where every control is
    1.  The reference computes H . C...CZ . H; here the flip is exact."""
    n = count_qubits(state_vector)
    controls = [int(q) for q in controls]
    target = int(target)
    if not 0 <= target < n:
        raise ValueError(f"qubits={[target]} is not consistent with state vector with "
                         f"{n} qubits.")
    if len(set(controls + [target])) != len(controls) + 1:
        raise RuntimeError(f"multi_controlled_x: repeated qubit in {controls + [target]}")
    if any(q < 0 or q >= n for q in controls):
        raise IndexError(f"multi_controlled_x: controls {controls} out of range")
    ctrl = 0
    for q in controls:
        ctrl |= 1 << (n - 1 - q)
    args = (ctrl, n - 1 - target)
    return _index_map(state_vector, "multi_controlled_x", "flip", n, args, args)


# ------------------------------------------------------------------- phase --

_PHASE_OF = {torch.float32: torch.complex64, torch.float64: torch.complex128,
             torch.complex64: torch.complex64, torch.complex128: torch.complex128}


def _coalesce(shape: Sequence[int], *strides: Sequence[int]):
    """Merge neighbouring dims that every operand walks contiguously."""
    dims = [(s, [st[i] for st in strides]) for i, s in enumerate(shape) if s != 1]
    if not dims:
        return [1], [[0] for _ in strides]
    out = [dims[0]]
    for size, st in dims[1:]:
        psize, pst = out[-1]
        if all(p == s * size for p, s in zip(pst, st)):
            out[-1] = (psize * size, st)
        else:
            out.append((size, st))
    shape2 = [s for s, _ in out]
    strides2 = [[st[j] for _, st in out] for j in range(len(strides))]
    return shape2, strides2


class _ApplyPhase(torch.autograd.Function):
    """exp(-1j*angles) * state.  Backward (torch's conjugate-Wirtinger
    convention): grad_state = exp(+1j*conj(angles)) * grad_out, reduced to the
    state's shape; grad_angles = i*conj(out)*grad_out (its real part,
    Im(conj(grad_out)*out), for real angles), reduced to the angles' shape."""

    @staticmethod
    def forward(ctx, angles, state):
        out = _phase_launch(angles, state)
        ctx.save_for_backward(angles, state, out)
        return out

    @staticmethod
    def backward(ctx, g):
        angles, state, out = ctx.saved_tensors
        ga = gs = None
        if ctx.needs_input_grad[1]:
            back = _phase_launch(-angles.conj() if angles.is_complex() else -angles, g)
            gs = back.sum_to_size(state.shape)
            gs = gs if state.is_complex() else gs.real
            gs = gs.to(state.dtype)
        if ctx.needs_input_grad[0]:
            ga = _phase_angle_grad(angles, out, g)
            if ga is None:   # complex angles or an unusual broadcast: torch's rule as is
                full = 1j * out.conj() * g
                ga = full.sum_to_size(angles.shape)
                ga = (ga if angles.is_complex() else ga.real).to(angles.dtype)
        return ga, gs


def _phase_angle_grad(angles: torch.Tensor, out: torch.Tensor, g: torch.Tensor):
    """grad of real angles through exp(-1j*angles)*state by ua_phase_angle_grad:
    Im(out) Re(g) - Re(out) Im(g), summed over the leading dims that the angles
    broadcast over.  None for the cases it does not cover."""
    if angles.is_complex() or angles.dtype not in (torch.float32, torch.float64) or \
            out.dtype not in (torch.complex64, torch.complex128) or g.dtype != out.dtype or \
            g.shape != out.shape or angles.numel() == 0:
        return None
    # the angles' shape, leading 1s aside, must be a trailing block of the output's
    a_shape = tuple(angles.shape)
    while a_shape and a_shape[0] == 1:
        a_shape = a_shape[1:]
    o_shape = tuple(out.shape)
    if len(a_shape) > len(o_shape) or o_shape[len(o_shape) - len(a_shape):] != a_shape:
        return None
    cols = math.prod(a_shape)
    rows = out.numel() // cols if cols else 0
    grad = torch.empty(angles.shape, dtype=angles.dtype, device=angles.device)
    lib = _lib.load()
    with on_device(out.device):
        rc = lib.ua_phase_angle_grad(_lib.ptr(grad), _lib.DTYPE_CODES[angles.dtype],
                                     _lib.ptr(prepare(g).contiguous()),
                                     _lib.ptr(prepare(out).contiguous()),
                                     _lib.DTYPE_CODES[out.dtype], rows, cols,
                                     _lib.stream_handle(out.device))
    _lib.check(rc, "phase_angle_grad")
    return grad


def apply_phase(angles: torch.Tensor, state: torch.Tensor) -> torch.Tensor:
    """exp(-1j * angles) * state with torch broadcasting (operations.py:15-42).

    The phase factor has the precision of `angles` (complex64 for float32
    angles, even against a complex128 state, as in the reference); the product
    is formed in the promoted result dtype.
    """
    if not isinstance(angles, torch.Tensor):
        raise TypeError(f"exp(): argument 'input' must be Tensor, not {type(angles).__name__}")
    require_gpu(state, "apply_phase")
    if angles.device != state.device:
        if angles.dim() == 0 and angles.device.type == "cpu":
            angles = angles.to(state.device)
        else:
            same_device(angles, state, "apply_phase")
    angles, state = prepare(angles), prepare(state)
    if angles.dtype in (torch.float16, torch.bfloat16, torch.complex32) or \
            state.dtype in (torch.float16, torch.bfloat16, torch.complex32):
        raise TypeError("apply_phase: half-precision inputs are not supported")
    if angles.dtype not in _PHASE_OF:
        angles = angles.to(torch.float32)      # ints/bool: exp(-1j*x) is complex64
    if state.dtype not in _PHASE_OF:
        state = state.to(torch.float32)
    if _needs_grad(angles, state):
        return _ApplyPhase.apply(angles, state)
    return _phase_launch(angles, state)


def _phase_launch(angles: torch.Tensor, state: torch.Tensor) -> torch.Tensor:
    angles, state = prepare(angles), prepare(state)
    out_dtype = torch.promote_types(_PHASE_OF[angles.dtype], state.dtype)
    shape = torch.broadcast_shapes(angles.shape, state.shape)
    out = torch.empty(shape, dtype=out_dtype, device=state.device)
    if out.numel() == 0:
        return out
    ea, es = angles.expand(shape), state.expand(shape)
    dims, (sst, ast) = _coalesce(shape, es.stride(), ea.stride())
    if len(dims) > 8:
        es, ea = es.contiguous(), ea.contiguous()
        dims, (sst, ast) = _coalesce(shape, es.stride(), ea.stride())
    lib = _lib.load()
    code = _lib.DTYPE_CODES
    with on_device(state.device):
        rc = lib.ua_apply_phase(
            _lib.ptr(out), code[out_dtype], _lib.ptr(es), code[es.dtype], _lib.ptr(ea),
            code[ea.dtype], len(dims), _lib.i64_array(dims), _lib.i64_array(sst),
            _lib.i64_array(ast), _lib.stream_handle(state.device))
    _lib.check(rc, "apply_phase")
    return out


from ..measurement import measure  # noqa: E402  (unitair.simulation exports measure)

# submodules with the reference's layout (unitair/simulation/{operations,measurement,utils}.py)
from . import operations, measurement, utils  # noqa: E402,F401
