"""State layout, qubit counting and reductions (mirror of ``unitair.states``).

Shape/layout helpers follow unitair/states/shapes.py and conversions.py
(same names, argument meaning and exception types).  The reductions follow
unitair/states/innerprod.py but run as HIP kernels with float64 accumulation
and only accept tensors on a ROCm device.
"""
from __future__ import annotations

import enum
import math
from typing import List, Tuple, Union

import torch

from .. import _lib
from .._device import require_gpu, same_device, prepare

__all__ = [
    "StateLayout", "StateShapeError", "count_qubits", "count_qubits_gate_matrix",
    "hilbert_space_dim", "count_qubits_tensor", "count_batch_dims_tensor",
    "get_qubit_indices", "subset_roll_to_back", "subset_roll_to_front",
    "to_tensor_layout", "to_vector_layout", "abs_squared", "norm_squared",
    "diag_expectation_value", "inner_product", "norm_squared_f64",
    "inner_product_f64",
]


class StateLayout(str, enum.Enum):
    """The two layouts of a state (unitair/states/shapes.py:8-10)."""
    VECTOR = "vector"
    TENSOR = "tensor"


class StateShapeError(ValueError):
    """A tensor does not have the shape of a state (shapes.py:149-182)."""

    def __init__(self, data: torch.Tensor = None, expected_layout: StateLayout = None):
        super().__init__()
        self.data_size = tuple(data.size()) if data is not None else None
        self.layout = (StateLayout(expected_layout.lower())
                       if expected_layout is not None else None)

    def __str__(self):
        lines = ["The tensor does not have the shape of a quantum state."]
        if self.layout is StateLayout.VECTOR:
            lines.append("Vector layout is (*optional_batch_dims, 2^num_qubits).")
        elif self.layout is StateLayout.TENSOR:
            lines.append("Tensor layout is (*optional_batch_dims, 2, 2, ..., 2) "
                         "with one 2 per qubit.")
        if self.data_size is not None:
            lines.append(f"Given state has size {self.data_size}")
        return "\n".join(lines)


def count_qubits(state: torch.Tensor) -> int:
    """Number of qubits of a vector-layout state (shapes.py:13-27)."""
    length = state.size()[-1]
    n = round(math.log2(length))
    if 2 ** n != length:
        raise StateShapeError(data=state, expected_layout=StateLayout.VECTOR)
    return n


def count_qubits_gate_matrix(gate: torch.Tensor) -> int:
    """k for a gate of size (*batch, 2^k, 2^k) (shapes.py:31-47)."""
    length = gate.size()[-1]
    k = round(math.log2(length))
    if 2 ** k != length:
        raise RuntimeError(f"Gate matrix of size {tuple(gate.size())} does not act on "
                           f"a whole number of qubits.")
    return k


def hilbert_space_dim(state: torch.Tensor) -> int:
    """2**n for an n-qubit vector-layout state (shapes.py:50-52)."""
    return 2 ** count_qubits(state)


def count_qubits_tensor(state_tensor: torch.Tensor, num_batch_dims: int) -> int:
    """Qubits of a tensor-layout state (shapes.py:55-60)."""
    return state_tensor.dim() - num_batch_dims


def count_batch_dims_tensor(state_tensor: torch.Tensor, num_qubits: int) -> int:
    """Batch dims of a tensor-layout state (shapes.py:63-69)."""
    return state_tensor.dim() - num_qubits


def get_qubit_indices(index: Union[int, List[int], Tuple[int, ...], torch.Tensor],
                      state_tensor: torch.Tensor, num_qubits: int):
    """Torch dim of each qubit index of a tensor-layout state (shapes.py:72-116).

    Non-negative indices are shifted past the batch dims; negative ones are
    returned unchanged.  Lists/ints come back as lists/ints, tensors as tensors.
    """
    nbatch = state_tensor.dim() - num_qubits
    as_tensor = isinstance(index, torch.Tensor)
    idx = index if as_tensor else torch.tensor(index)
    if (idx >= num_qubits).any() or (idx < -num_qubits).any():
        raise ValueError(f"Expected index in {{-num_qubits, ..., num_qubits - 1}}.\n"
                         f"Num_qubits: {num_qubits}, index: {index}.")
    idx = idx + (idx >= 0) * nbatch
    return idx if as_tensor else idx.tolist()


def subset_roll_to_back(tensor: torch.Tensor, subset_num_dims: int) -> torch.Tensor:
    """Move the first `subset_num_dims` dims to the end (shapes.py:121-132)."""
    d = tensor.dim()
    return tensor.permute(list(range(subset_num_dims, d)) + list(range(subset_num_dims)))


def subset_roll_to_front(tensor: torch.Tensor, subset_num_dims: int) -> torch.Tensor:
    """Move the last `subset_num_dims` dims to the front (shapes.py:135-146)."""
    d = tensor.dim()
    return tensor.permute(list(range(d - subset_num_dims, d)) + list(range(d - subset_num_dims)))


def to_tensor_layout(state: torch.Tensor) -> torch.Tensor:
    """(*batch, 2^n) -> (*batch, 2, ..., 2) view (conversions.py:5-45)."""
    n = count_qubits(state)
    if not state.is_contiguous():
        state = state.contiguous()
    return state.view(state.size()[:-1] + torch.Size([2] * n))


def to_vector_layout(state_tensor: torch.Tensor, num_qubits: int) -> torch.Tensor:
    """(*batch, 2, ..., 2) -> (*batch, 2^n) view (conversions.py:48-92)."""
    if not state_tensor.is_contiguous():
        state_tensor = state_tensor.contiguous()
    batch = state_tensor.size()[:state_tensor.dim() - num_qubits]
    return state_tensor.view(batch + torch.Size([2 ** num_qubits]))


# ------------------------------------------------------------- reductions --

_REAL_OF = {torch.complex64: torch.float32, torch.complex128: torch.float64,
            torch.float32: torch.float32, torch.float64: torch.float64}


def _rows(t: torch.Tensor, shape: torch.Size):
    """Expand `t` to `shape` and view it as (rows, length); returns
    (tensor, row_stride, rows, length).  Row stride 0 means broadcast; a
    layout that no single stride describes is copied to contiguous."""
    if len(shape) == 0:
        return t.reshape(1), 0, 1, 1
    rows, length = math.prod(shape[:-1]), shape[-1]
    e = t.expand(shape).reshape(rows, length)
    if length > 1 and e.stride(1) != 1:
        e = e.contiguous()
    return e, (e.stride(0) if rows > 1 else 0), rows, length


def _reduce(kind: str, out_dtype: torch.dtype, a: torch.Tensor, b: torch.Tensor = None):
    lib = _lib.load()
    shape = torch.broadcast_shapes(a.shape, b.shape) if b is not None else a.shape
    ea, sa, rows, length = _rows(a, shape)
    if b is not None:
        eb, sb, _, _ = _rows(b, shape)
    out = torch.empty(shape[:-1], dtype=out_dtype, device=a.device)
    if rows == 0:
        return out
    ws = torch.empty(lib.ua_reduce_workspace_bytes(rows, length), dtype=torch.uint8,
                     device=a.device)
    stream = _lib.stream_handle(a.device)
    code = _lib.DTYPE_CODES
    if kind == "norm":
        rc = lib.ua_norm_squared(_lib.ptr(out), code[out_dtype], _lib.ptr(ea), code[ea.dtype],
                                 rows, length, sa, _lib.ptr(ws), stream)
    elif kind == "inner":
        rc = lib.ua_inner_product(_lib.ptr(out), code[out_dtype], _lib.ptr(ea), sa,
                                  _lib.ptr(eb), sb, code[ea.dtype], rows, length,
                                  _lib.ptr(ws), stream)
    else:  # diag: a = diag values, b = state
        rc = lib.ua_diag_expectation(_lib.ptr(out), code[out_dtype], _lib.ptr(ea),
                                     code[ea.dtype], sa, _lib.ptr(eb), code[eb.dtype], sb,
                                     rows, length, _lib.ptr(ws), stream)
    _lib.check(rc, f"{kind} reduction")
    return out


def _as_supported(t: torch.Tensor, what: str) -> torch.Tensor:
    t = prepare(t)
    if t.dtype not in _lib.DTYPE_CODES:
        if t.dtype.is_complex:
            raise TypeError(f"{what}: unsupported dtype {t.dtype}")
        t = t.to(torch.float32)
    return t


def _needs_grad(*ts) -> bool:
    return torch.is_grad_enabled() and any(isinstance(t, torch.Tensor) and t.requires_grad
                                           for t in ts)


def _real_scale(g: torch.Tensor, like: torch.Tensor) -> torch.Tensor:
    """A real gradient factor in the real dtype of `like`."""
    return g.to(_REAL_OF.get(like.dtype, like.dtype) if not g.is_complex() else g.dtype)


def _grad_to(x: torch.Tensor, full: torch.Tensor) -> torch.Tensor:
    """Reduce a full-shape gradient to x's shape and dtype (real inputs keep
    the real part, as torch does for a real leaf)."""
    full = full.sum_to_size(x.shape)
    if not x.is_complex() and full.is_complex():
        full = full.real
    return full.to(x.dtype)


class _AbsSquared(torch.autograd.Function):
    """|x|^2 elementwise (innerprod.py:4-26).  Backward: grad_x = 2 g x."""

    @staticmethod
    def forward(ctx, state):
        ctx.save_for_backward(state)
        return _abs_squared_launch(state)

    @staticmethod
    def backward(ctx, g):
        (state,) = ctx.saved_tensors
        return _grad_to(state, 2 * _real_scale(g, state) * state)


class _Reduce(torch.autograd.Function):
    """A row reduction on the HIP kernels with a torch-convention backward
    (conjugate Wirtinger for complex inputs), the same gradients as torch's
    autograd through the reference's expressions in innerprod.py:29-65:
      norm  out = sum |s|^2                 grad_s = 2 g s
      inner out = sum conj(a) b             grad_a = conj(g) b, grad_b = g a
      diag  out = sum |s|^2 d  (a = d)      grad_d = g |s|^2,
                                            grad_s = 2 Re(conj(d) g) s
    Broadcast inputs get their gradients summed back to their own shapes."""

    @staticmethod
    def forward(ctx, kind, out_dtype, a, b):
        ctx.kind = kind
        ctx.save_for_backward(a, b)
        return _reduce(kind, out_dtype, a, b)

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors
        g = g.unsqueeze(-1)
        ga = gb = None
        if ctx.kind == "norm":
            if ctx.needs_input_grad[2]:
                ga = _grad_to(a, 2 * _real_scale(g, a) * a)
        elif ctx.kind == "inner":
            if ctx.needs_input_grad[2]:
                ga = _grad_to(a, (g.conj() if g.is_complex() else g) * b)
            if ctx.needs_input_grad[3]:
                gb = _grad_to(b, g * a)
        else:
            d, s = a, b
            if ctx.needs_input_grad[2]:
                p = (s.conj() * s).real if s.is_complex() else s * s
                ga = _grad_to(d, g * p.to(g.real.dtype if g.is_complex() else g.dtype))
            if ctx.needs_input_grad[3]:
                gp = (d.conj() * g).real if (d.is_complex() or g.is_complex()) else d * g
                gb = _grad_to(s, 2 * gp.to(_REAL_OF.get(s.dtype, s.dtype)) * s)
        return None, None, ga, gb


def _reduce_op(kind: str, out_dtype: torch.dtype, a: torch.Tensor, b: torch.Tensor = None):
    if _needs_grad(a, b):
        return _Reduce.apply(kind, out_dtype, a, b)
    return _reduce(kind, out_dtype, a, b)


def _abs_squared_launch(state: torch.Tensor) -> torch.Tensor:
    state = state.contiguous()
    out = torch.empty(state.shape, dtype=_REAL_OF[state.dtype], device=state.device)
    if state.numel():
        lib = _lib.load()
        _lib.check(lib.ua_abs_squared(_lib.ptr(out), _lib.ptr(state),
                                      _lib.DTYPE_CODES[state.dtype], state.numel(),
                                      _lib.stream_handle(state.device)), "abs_squared")
    return out


def abs_squared(state: torch.Tensor) -> torch.Tensor:
    """|x|^2 elementwise, real dtype of the input (innerprod.py:4-26)."""
    require_gpu(state, "abs_squared")
    state = _as_supported(state, "abs_squared")
    if _needs_grad(state):
        return _AbsSquared.apply(state)
    return _abs_squared_launch(state)


def norm_squared(state: torch.Tensor) -> torch.Tensor:
    """<state|state> per batch entry (innerprod.py:29-46), float64 accumulation,
    result in the real dtype of the state."""
    require_gpu(state, "norm_squared")
    state = _as_supported(state, "norm_squared")
    return _reduce_op("norm", _REAL_OF[state.dtype], state)


def norm_squared_f64(state: torch.Tensor) -> torch.Tensor:
    """norm_squared with a float64 result."""
    require_gpu(state, "norm_squared_f64")
    return _reduce_op("norm", torch.float64, _as_supported(state, "norm_squared_f64"))


def inner_product(state_1: torch.Tensor, state_2: torch.Tensor) -> torch.Tensor:
    """<state_1|state_2>, conjugate-linear in the left entry (innerprod.py:62-65)."""
    require_gpu(state_1, "inner_product")
    same_device(state_1, state_2, "inner_product")
    a = _as_supported(state_1, "inner_product")
    b = _as_supported(state_2, "inner_product")
    dt = torch.promote_types(a.dtype, b.dtype)
    a, b = a.to(dt), b.to(dt)
    return _reduce_op("inner", dt, a, b)


def inner_product_f64(state_1: torch.Tensor, state_2: torch.Tensor) -> torch.Tensor:
    """inner_product with a complex128 result (float64 accumulation)."""
    require_gpu(state_1, "inner_product_f64")
    same_device(state_1, state_2, "inner_product_f64")
    a = _as_supported(state_1, "inner_product_f64")
    b = _as_supported(state_2, "inner_product_f64")
    dt = torch.promote_types(a.dtype, b.dtype)
    a, b = a.to(dt), b.to(dt)
    if not dt.is_complex:
        return _reduce_op("inner", torch.float64, a, b)
    return _reduce_op("inner", torch.complex128, a, b)


def diag_expectation_value(diag_values: torch.Tensor, state: torch.Tensor) -> torch.Tensor:
    """sum_x |state_x|^2 * diag_x (innerprod.py:49-59), float64 accumulation;
    the result dtype follows torch's promotion of abs_squared(state) * diag."""
    require_gpu(state, "diag_expectation_value")
    if not isinstance(diag_values, torch.Tensor):
        diag_values = torch.tensor(diag_values, device=state.device)
    same_device(state, diag_values, "diag_expectation_value")
    s = _as_supported(state, "diag_expectation_value")
    out_dtype = torch.promote_types(_REAL_OF[s.dtype], diag_values.dtype)
    d = _as_supported(diag_values, "diag_expectation_value")
    if out_dtype not in _lib.DTYPE_CODES:
        raise TypeError(f"diag_expectation_value: unsupported result dtype {out_dtype}")
    return _reduce_op("diag", out_dtype, d, s)


# submodules with the reference's layout (unitair/states/{shapes,conversions,innerprod}.py)
from . import shapes, conversions, innerprod  # noqa: E402,F401
