"""Adjoint gradients of batched rotation circuits with a diagonal observable.

A circuit is a list of gates applied to every state b of a batch:
    ("x" | "y" | "z", qubit, index[, coef])
        unitair's exp_x / exp_y / exp_z(coef * params[b, index]) on `qubit`
        (= exp(-i t P), unitair/gates/gates.py:7-91), applied as
        apply_operator(gate, [qubit], state);
    ("cz", a, b), ("cnot", control, target)   (gates.py:170-213);
    ("dense", qubits, matrix)   a fixed 1- or 2-qubit gate (2x2 / 4x4).
The loss per state is diag_expectation_value(observable, psi)
(unitair/states/innerprod.py:49-59) for a real diagonal observable.

unitair differentiates such a loop of apply_operator calls with torch
autograd, keeping one saved state per gate.  Here the backward pass keeps two
states: psi, and the co-state lam = H psi carried back through the inverse
gates.  For U(t) = exp(-i t P) the derivative of <psi|H|psi> is
2 Im <lam|P|psi> at the point right after the gate.  Rotations on distinct
qubits commute, so a run of them ("group") shares one (lam, psi) point, and
one ua_pauli_gradients call per Pauli kind gives every qubit's value.  psi
and lam are stored as one (2B, 2^n) batch, so every inverse step is one fused
pass over both.  Un-applying gates in complex64 drifts, so psi is
re-synchronised from checkpoints taken every ceil(sqrt(groups)) groups.
Memory: two states plus ceil(sqrt(groups)) checkpoints.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._device import on_device
from .fusion import FusedProgram, PlanTemplate, _Item

_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64}
_PAULI = {"x": _lib.UA_PAULI_X, "y": _lib.UA_PAULI_Y, "z": _lib.UA_PAULI_Z}


def _rotation_data(kind: str, t: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    """Per-state gate data of exp(-i t P): (B, 2, 2) for X / Y, the (B, 2)
    diagonal for Z (gates.py:7-91)."""
    c, s = torch.cos(t), torch.sin(t)
    if kind == "z":
        return torch.stack([torch.complex(c, -s), torch.complex(c, s)], -1).to(dtype)
    z = torch.zeros_like(c)
    if kind == "x":
        a, b = torch.complex(c, z), torch.complex(z, -s)
        return torch.stack([torch.stack([a, b], -1), torch.stack([b, a], -1)], -2).to(dtype)
    return torch.stack([torch.stack([c, -s], -1), torch.stack([s, c], -1)], -2).to(dtype)


class _Spec:
    """A gate list split into segments: (fixed gates, rotation group)."""

    def __init__(self, gates: Sequence[tuple], n: int, num_params: int):
        self.n = n
        self.compiled = False
        self.key = None
        segments, fixed, rot = [], [], []
        for g in gates:
            kind = g[0]
            if kind in _PAULI:
                q, idx = int(g[1]), int(g[2])
                coef = float(g[3]) if len(g) > 3 else 1.0
                if not 0 <= q < n:
                    raise ValueError(f"rotation qubit {q} out of range for {n} qubits")
                if not 0 <= idx < num_params:
                    raise ValueError(f"parameter index {idx} out of range ({num_params})")
                if any(r[1] == q for r in rot):
                    segments.append((fixed, rot))
                    fixed, rot = [], []
                rot.append((kind, q, idx, coef))
            elif kind in ("cz", "cnot", "dense"):
                if rot:
                    segments.append((fixed, rot))
                    fixed, rot = [], []
                fixed.append(self._fixed_item(g))
            else:
                raise ValueError(f"unknown gate kind {kind!r}")
        if rot:
            segments.append((fixed, rot))
            fixed = []
        self.segments = segments
        self.trailing = fixed

    def _fixed_item(self, g):
        n = self.n
        if g[0] in ("cz", "cnot"):
            a, b = int(g[1]), int(g[2])
            if not (0 <= a < n and 0 <= b < n) or a == b:
                raise ValueError(f"{g[0]} qubits ({a}, {b}) invalid for {n} qubits")
            return _Item(g[0], (n - 1 - a, n - 1 - b))
        qubits = tuple(int(q) for q in g[1])
        m = np.asarray(g[2], dtype=np.complex128)
        k = len(qubits)
        if k not in (1, 2) or m.shape != (2 ** k, 2 ** k):
            raise ValueError("fixed dense gates act on 1 or 2 qubits with a 2^k x 2^k matrix")
        if len(set(qubits)) != k or not all(0 <= q < n for q in qubits):
            raise ValueError(f"dense gate qubits {qubits} invalid for {n} qubits")
        return _Item(f"dense{k}", tuple(n - 1 - q for q in qubits), m)


def _inverse_fixed(it: _Item) -> _Item:
    if it.kind in ("cz", "cnot"):
        return it
    return _Item(it.kind, it.bits, np.conj(it.data).T.copy())


def _rotation_table(spec, params, dtype, sign, repeat=1):
    """Per-state data of every rotation of the circuit, angles
    sign * coef * params[:, idx], made with a few batched ops per Pauli kind:
    {(segment, position in its group): (B * repeat, 2, 2) or (B * repeat, 2)
    contiguous tensor}."""
    table = {}
    for kind in _PAULI:
        keys, idxs, coefs = [], [], []
        for s, (_, rot) in enumerate(spec.segments):
            for r, (k, q, idx, coef) in enumerate(rot):
                if k == kind:
                    keys.append((s, r))
                    idxs.append(idx)
                    coefs.append(sign * coef)
        if not keys:
            continue
        dev = params.device
        ckey = (spec.key, kind, sign, params.dtype, str(dev))
        consts = _GROUP_INDEX.get(ckey) if spec.key is not None else None
        if consts is None:       # device constants, made once per gate list (see _group_index)
            consts = (torch.tensor(idxs, device=dev),
                      torch.tensor(coefs, dtype=params.dtype, device=dev))
            if spec.key is not None:
                _GROUP_INDEX[ckey] = consts
        t = (params[:, consts[0]] * consts[1]).t()                           # (R, B)
        data = _rotation_data(kind, t, dtype)                               # (R, B, ...)
        if repeat > 1:
            data = data.repeat((1, repeat) + (1,) * (data.dim() - 2))
        data = data.contiguous()
        for i, key in enumerate(keys):
            table[key] = data[i]
    return table


def _rot_items(s, rot, table, n):
    """Fused-pass items of rotation group s (its per-state data from table)."""
    return [_Item("diag1" if kind == "z" else "dense1", (n - 1 - q,), table[(s, r)], True)
            for r, (kind, q, idx, coef) in enumerate(rot)]


# Plans of the adjoint's programs, by gate-list structure and program: each
# call builds fresh programs (its own op records and data) from these
# without planning again.  Oldest entries are dropped past _TEMPLATE_LIMIT.
_TEMPLATES = {}
_TEMPLATE_LIMIT = 256


_KEY_IDS = {}


def _gates_key(gates, n: int, num_params: int) -> int:
    """A small id for everything of a gate list that the plans depend on
    (array operands by shape only)."""
    key = []
    for g in gates:
        parts = [g[0]]
        for x in g[1:]:
            if isinstance(x, (bool, int, float, str, np.integer, np.floating)):
                parts.append(x.item() if isinstance(x, (np.integer, np.floating)) else x)
            else:
                parts.append(("array", tuple(np.shape(x))))
        key.append(tuple(parts))
    return _KEY_IDS.setdefault((n, num_params, tuple(key)), len(_KEY_IDS))


_GROUP_INDEX = {}


def _group_index(spec, s, rot, dev):
    """Device index tensors of rotation group s (Pauli kind, qubit,
    parameter index, 2 * coefficient), made once per gate list and device:
    building them per call would copy from the host, which waits for the
    stream, once per group."""
    key = (spec.key, s, str(dev))
    hit = _GROUP_INDEX.get(key) if spec.key is not None else None
    if hit is None:
        hit = (torch.tensor([_PAULI[r[0]] for r in rot], device=dev),
               torch.tensor([r[1] for r in rot], device=dev),
               torch.tensor([r[2] for r in rot], device=dev),
               torch.tensor([2.0 * r[3] for r in rot], dtype=torch.float64, device=dev))
        if spec.key is not None:
            if len(_GROUP_INDEX) >= 4 * _TEMPLATE_LIMIT:
                _GROUP_INDEX.pop(next(iter(_GROUP_INDEX)))
            _GROUP_INDEX[key] = hit
    return hit


def _program(items, n, dtype, device, compiled=False, key=None):
    if not items:
        return None
    template = None
    if key is not None:
        template = _TEMPLATES.get(key)
        if template is None:
            template = PlanTemplate(items, n)
            if len(_TEMPLATES) >= _TEMPLATE_LIMIT:
                _TEMPLATES.pop(next(iter(_TEMPLATES)))
            _TEMPLATES[key] = template
    # 12-bit tiles also for compiled passes: measured faster here than the
    # 13-bit default of compiled programs (34.5 vs 38.4 ms per step at 16
    # qubits x 1024 states)
    return FusedProgram.from_items(items, n, dtype, device, tile_bits=12, compiled=compiled,
                                   template=template)


def _check_inputs(params, observable, n, dtype, device):
    if dtype not in _REAL:
        raise ValueError("adjoint circuits run complex64 or complex128 states")
    if params.dim() != 2:
        raise ValueError("params must have shape (batch, num_params)")
    if device.type != "cuda":
        raise RuntimeError("unitair_amd.adjoint runs on ROCm GPU tensors only")
    if n < 4:
        raise ValueError("the adjoint path needs at least 4 qubits")
    if observable.shape[-1] != 2 ** n or observable.dim() not in (1, 2) or \
            (observable.dim() == 2 and observable.shape[0] != params.shape[0]):
        raise ValueError("observable must have shape (2^n,) or (batch, 2^n)")
    if observable.is_complex():
        raise ValueError("the observable is a real diagonal")


class _Programs:
    """Every fused program of one parameter set, built before any of them
    runs (building uploads op records, which waits for the stream)."""

    def __init__(self, spec, p, dtype, dev, backward: bool):
        n = spec.n
        S = len(spec.segments)
        self.every = max(1, math.ceil(math.sqrt(max(S, 1))))
        fwd_table = _rotation_table(spec, p, dtype, +1.0)
        self.fwd = []
        for c0 in range(0, S, self.every):
            items = []
            for s in range(c0, min(S, c0 + self.every)):
                fixed, rot = spec.segments[s]
                items += list(fixed) + _rot_items(s, rot, fwd_table, n)
            self.fwd.append((c0, _program(items, n, dtype, dev, spec.compiled,
                                          (spec.key, "fwd", c0))))
        self.trail = _program(list(spec.trailing), n, dtype, dev, spec.compiled,
                              (spec.key, "trail"))
        self.inv_trail = self.inv = None
        if backward:
            inv_table = _rotation_table(spec, p, dtype, -1.0, repeat=2)
            self.inv_trail = _program([_inverse_fixed(it) for it in reversed(spec.trailing)],
                                      n, dtype, dev, spec.compiled, (spec.key, "inv_trail"))
            self.inv = []
            for s, (fixed, rot) in enumerate(spec.segments):
                items = _rot_items(s, rot, inv_table, n)
                items += [_inverse_fixed(it) for it in reversed(fixed)]
                self.inv.append(_program(items, n, dtype, dev, spec.compiled,
                                         (spec.key, "inv", s)))


# Algorithmic HBM bytes of the last forward and backward (state passes, copies,
# diagonal scaling and gradient reductions), for throughput reports.
last_traffic = {"forward_bytes": 0, "backward_bytes": 0}


def _bytes(x: torch.Tensor) -> int:
    return x.numel() * x.element_size()


def _run_counted(prog, x, key):
    prog.run(x)
    last_traffic[key] += len(prog.steps) * 2 * _bytes(x)


def _initial(B, n, dtype, dev, initial):
    psi = torch.zeros(B, 2 ** n, dtype=dtype, device=dev)
    if initial is None:
        psi[:, 0] = 1
    else:
        psi.copy_(initial.detach().to(dev, dtype).expand(B, 2 ** n))
    return psi


class _AdjointExpectation(torch.autograd.Function):
    """loss[b] = sum_x observable[(b,) x] |psi_b(x)|^2 (float64)."""

    @staticmethod
    def forward(ctx, params, observable, spec, dtype, initial):
        n = spec.n
        B = params.shape[0]
        dev = params.device
        p = params.detach().to(_REAL[dtype])
        obs = observable.detach().to(dev, torch.float64).contiguous()
        with on_device(dev):
            progs = _Programs(spec, p, dtype, dev, backward=True)
            psi = _initial(B, n, dtype, dev, initial)
            last_traffic["forward_bytes"] = _bytes(psi)
            ckpt = {}
            for c0, prog in progs.fwd:
                if c0 > 0:
                    ckpt[c0] = psi.clone()          # state before segment c0
                    last_traffic["forward_bytes"] += 2 * _bytes(psi)
                if prog is not None:
                    _run_counted(prog, psi, "forward_bytes")
            if progs.trail is not None:
                _run_counted(progs.trail, psi, "forward_bytes")
            loss = _diag_loss(obs, psi)
            last_traffic["forward_bytes"] += _bytes(psi) + _bytes(obs)
        ctx.save_for_backward(params, observable)
        ctx.state = (spec, dtype, obs, psi, ckpt, progs)
        return loss

    @staticmethod
    def backward(ctx, gloss):
        params, observable = ctx.saved_tensors
        spec, dtype, obs, psi, ckpt, progs = ctx.state
        ctx.state = None
        n, B, dev = spec.n, params.shape[0], params.device
        lib = _lib.load()
        grad_p = grad_obs = None
        with on_device(dev):
            if ctx.needs_input_grad[1]:
                from .states import abs_squared
                w = gloss.to(torch.float64).view(B, 1) * abs_squared(psi).to(torch.float64)
                grad_obs = (w if observable.dim() == 2 else w.sum(0)).to(observable.dtype)
            if not ctx.needs_input_grad[0]:
                return grad_p, grad_obs, None, None, None
            X = torch.empty(2 * B, 2 ** n, dtype=dtype, device=dev)
            X[:B].copy_(psi)
            # copy + diagonal scaling (read psi and obs, write lambda)
            last_traffic["backward_bytes"] = 4 * _bytes(psi) + _bytes(obs)
            code = _lib.DTYPE_CODES
            rc = lib.ua_apply_diagonal(_lib.ptr(X[B:]), _lib.ptr(psi), code[dtype],
                                       _lib.ptr(obs), code[obs.dtype],
                                       0 if obs.dim() == 1 else 2 ** n, n, B,
                                       _lib.stream_handle(dev))
            _lib.check(rc, "apply_diagonal")
            del psi
            if progs.inv_trail is not None:
                _run_counted(progs.inv_trail, X, "backward_bytes")
            ws = torch.empty(lib.ua_pauli_gradients_workspace(n, B), dtype=torch.uint8,
                             device=dev)
            vals = torch.empty(len(_PAULI), B, n, dtype=torch.float64, device=dev)
            grad = torch.zeros(B, params.shape[1], dtype=torch.float64, device=dev)
            for s in range(len(spec.segments) - 1, -1, -1):
                fixed, rot = spec.segments[s]
                for kind in sorted({r[0] for r in rot}):
                    out = vals[_PAULI[kind]]
                    rc = lib.ua_pauli_gradients(_lib.ptr(out), _lib.ptr(X[B:]), _lib.ptr(X[:B]),
                                                code[dtype], n, B, _PAULI[kind], _lib.ptr(ws),
                                                _lib.stream_handle(dev))
                    _lib.check(rc, "pauli_gradients")
                    last_traffic["backward_bytes"] += _bytes(X)
                # one gather-add per group: grad[:, idx] += 2 coef vals[kind, :, q]
                kinds, qs, idx, coef = _group_index(spec, s, rot, dev)
                grad.index_add_(1, idx, vals[kinds, :, qs].transpose(0, 1) * coef)
                if progs.inv[s] is not None:
                    _run_counted(progs.inv[s], X, "backward_bytes")
                if s in ckpt:
                    X[:B].copy_(ckpt[s])         # psi before segment s, as the forward had it
                    last_traffic["backward_bytes"] += 2 * _bytes(ckpt[s])
            grad_p = (grad * gloss.to(torch.float64).view(B, 1)).to(params.dtype)
        return grad_p, grad_obs, None, None, None


def _diag_loss(obs: torch.Tensor, psi: torch.Tensor) -> torch.Tensor:
    from .states import _reduce
    return _reduce("diag", torch.float64, obs, psi)


def expectation(params: torch.Tensor, gates: Sequence[tuple], observable: torch.Tensor,
                num_qubits: int, dtype: torch.dtype = torch.complex64,
                initial_state: Optional[torch.Tensor] = None,
                compiled: bool = False) -> torch.Tensor:
    """loss[b] = diag_expectation_value(observable, U(params[b]) |psi0>) in
    float64, differentiable in params (adjoint backward) and in the
    observable.  params: (batch, num_params) real; gates: see the module
    docstring; observable: (2^n,) or (batch, 2^n) real; initial_state:
    |0...0> by default, else (2^n,) or (batch, 2^n) (not differentiated).
    compiled: run the passes with 12 tile bits (n >= 12) as compiled kernels
    (FusedProgram(compiled=True)); the first call with a gate list compiles
    them, later calls with the same gate list reuse them."""
    n = int(num_qubits)
    if not isinstance(observable, torch.Tensor):
        observable = torch.as_tensor(observable, device=params.device)
    _check_inputs(params, observable, n, dtype, params.device)
    spec = _Spec(gates, n, params.shape[1])
    spec.compiled = compiled
    spec.key = _gates_key(gates, n, params.shape[1])
    return _AdjointExpectation.apply(params, observable, spec, dtype, initial_state)


def run_gates(params: torch.Tensor, gates: Sequence[tuple], num_qubits: int,
              dtype: torch.dtype = torch.complex64,
              initial_state: Optional[torch.Tensor] = None,
              compiled: bool = False) -> torch.Tensor:
    """Final states (batch, 2^n) of the circuit for every parameter row."""
    n = int(num_qubits)
    obs = torch.zeros(2 ** n, dtype=torch.float64, device=params.device)
    _check_inputs(params, obs, n, dtype, params.device)
    spec = _Spec(gates, n, params.shape[1])
    spec.key = _gates_key(gates, n, params.shape[1])
    B, dev = params.shape[0], params.device
    p = params.detach().to(_REAL[dtype])
    with on_device(dev):
        table = _rotation_table(spec, p, dtype, +1.0)
        items = []
        for s, (fixed, rot) in enumerate(spec.segments):
            items += list(fixed) + _rot_items(s, rot, table, n)
        items += list(spec.trailing)
        prog = _program(items, n, dtype, dev, compiled, (spec.key, "run"))
        psi = _initial(B, n, dtype, dev, initial_state)
        if prog is not None:
            prog.run(psi)
    return psi


# ------------------------------------------------- hardware-efficient ansatz --

def hardware_efficient_gates(num_qubits: int, layers: int) -> List[tuple]:
    """Per layer l: RZ(phi[l, q]) then RY(theta[l, q]) on every qubit, then CZ
    on (q, q+1 mod n); RY(t) = exp_y(t/2), RZ(t) = exp_z(t/2).  Parameter
    index of theta[l, q] is l*2n + q, of phi[l, q] is l*2n + n + q."""
    n = num_qubits
    gates = []
    pairs = [(q, (q + 1) % n) for q in range(n)] if n > 2 else [(0, 1)]
    for l in range(layers):
        gates += [("z", q, l * 2 * n + n + q, 0.5) for q in range(n)]
        gates += [("y", q, l * 2 * n + q, 0.5) for q in range(n)]
        gates += [("cz", a, b) for a, b in pairs]
    return gates


def sum_z_observable(num_qubits: int, device) -> torch.Tensor:
    """sum_i Z_i as a diagonal: n - 2 popcount(x)."""
    x = torch.arange(2 ** num_qubits, device=device, dtype=torch.int64)
    pop = torch.zeros_like(x)
    for b in range(num_qubits):
        pop += (x >> b) & 1
    return (num_qubits - 2 * pop).to(torch.float64)


def _check(theta, phi, n):
    if theta.shape != phi.shape or theta.dim() != 3 or theta.shape[2] != n:
        raise ValueError("theta and phi must both have shape (batch, layers, num_qubits)")
    if theta.device.type != "cuda":
        raise RuntimeError("unitair_amd.adjoint runs on ROCm GPU tensors only")
    if n < 4:
        raise ValueError("the adjoint path needs at least 4 qubits")


def _ansatz_params(theta, phi):
    B, L, n = theta.shape
    return torch.cat([theta, phi], dim=2).reshape(B, L * 2 * n)


def run_circuit(theta: torch.Tensor, phi: torch.Tensor, num_qubits: int,
                dtype: torch.dtype = torch.complex64, compiled: bool = False) -> torch.Tensor:
    """Final states (batch, 2^n) of the hardware-efficient circuit."""
    _check(theta, phi, num_qubits)
    return run_gates(_ansatz_params(theta, phi), hardware_efficient_gates(num_qubits,
                                                                          theta.shape[1]),
                     num_qubits, dtype, compiled=compiled)


def sum_z_loss(theta: torch.Tensor, phi: torch.Tensor, num_qubits: int,
               dtype: torch.dtype = torch.complex64, compiled: bool = False) -> torch.Tensor:
    """Differentiable loss[b] = sum_i <Z_i> (float64) of the hardware-efficient
    circuit on (batch, layers, n) angle tensors, with adjoint gradients
    (compiled: see expectation)."""
    _check(theta, phi, num_qubits)
    n, L = num_qubits, theta.shape[1]
    return expectation(_ansatz_params(theta, phi), hardware_efficient_gates(n, L),
                       sum_z_observable(n, theta.device), n, dtype, compiled=compiled)
