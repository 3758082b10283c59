"""Adjoint gradients of batched hardware-efficient circuits.

Circuit (per state b, n qubits, L layers), starting from |0...0>:
    layer l:  RZ(phi[b,l,q]) then RY(theta[b,l,q]) on every qubit q,
              then CZ on (q, q+1 mod n) for every q
with RY(t) = exp(-i t/2 Y) = unitair exp_y(t/2) and RZ(t) = exp(-i t/2 Z) =
exp_z(t/2) (unitair/gates/gates.py:35-91).  Loss per state: sum_i <Z_i> =
diag_expectation_value(n - 2*popcount(x), psi) (unitair/states/innerprod.py:
49-59).

unitair differentiates such a circuit with torch autograd through every
apply_operator call, keeping one saved state per gate.  Here the backward pass
keeps two states: psi, and the co-state lam = H psi carried back through the
inverse gates.  For a rotation exp(-i t/2 P) the derivative is
Im <lam|P|psi> at the point right after the gate; rotations of one kind on
different qubits commute, so one (lam, psi) point per layer and kind gives all
n derivatives (ua_pauli_gradients).  Un-applying gates in complex64 drifts, so
psi is re-synchronised from checkpoints saved every ceil(sqrt(L)) layers.

Forward and inverse layers run as fused passes (unitair_amd.fusion) with
per-state gate data; the reductions run in float64.
"""
from __future__ import annotations

import math

import torch

from . import _lib
from ._device import on_device
from .fusion import FusedProgram, _Item

_REAL = {torch.complex64: torch.float32, torch.complex128: torch.float64}


def _rot_tensors(theta: torch.Tensor, phi: torch.Tensor, dtype: torch.dtype, sign: float):
    """Per-state gate data, indexed [l][q] -> contiguous (B, 2, 2) RY and
    (B, 2) RZ diagonal, for angles sign*theta / sign*phi."""
    real = _REAL[dtype]
    th = (sign * theta).to(real) / 2
    ph = (sign * phi).to(real) / 2
    c, s = torch.cos(th), torch.sin(th)
    ry = torch.stack([torch.stack([c, -s], -1), torch.stack([s, c], -1)], -2).to(dtype)
    e = torch.polar(torch.ones_like(ph), -ph).to(dtype)       # e^{-i phi/2}
    rz = torch.stack([e, e.conj()], -1)
    # (B, L, n, ...) -> (L, n, B, ...) contiguous
    ry = ry.permute(1, 2, 0, 3, 4).contiguous()
    rz = rz.permute(1, 2, 0, 3).contiguous()
    return ry, rz


def _cz_ring(n):
    if n < 2:
        return []
    pairs = [(q, (q + 1) % n) for q in range(n)] if n > 2 else [(0, 1)]
    return [_Item("cz", (n - 1 - a, n - 1 - b)) for a, b in pairs]


class _Programs:
    """Fused programs for every layer of one parameter set."""

    def __init__(self, theta, phi, n, dtype, device):
        L = theta.shape[1]
        self.ry, self.rz = _rot_tensors(theta, phi, dtype, +1.0)
        self.ry_inv, self.rz_inv = _rot_tensors(theta, phi, dtype, -1.0)
        bit = lambda q: n - 1 - q
        self.fwd, self.ry_back, self.rz_cz_back = [], [], []
        for l in range(L):
            items = [_Item("diag1", (bit(q),), self.rz[l, q], True) for q in range(n)]
            items += [_Item("dense1", (bit(q),), self.ry[l, q], True) for q in range(n)]
            items += _cz_ring(n)
            self.fwd.append(FusedProgram.from_items(items, n, dtype, device))
            self.ry_back.append(FusedProgram.from_items(
                [_Item("dense1", (bit(q),), self.ry_inv[l, q], True) for q in range(n)],
                n, dtype, device))
            back = [_Item("diag1", (bit(q),), self.rz_inv[l, q], True) for q in range(n)]
            if l > 0:
                back += _cz_ring(n)          # the previous layer's CZ ring
            self.rz_cz_back.append(FusedProgram.from_items(back, n, dtype, device))
        self.cz_only = FusedProgram.from_items(_cz_ring(n), n, dtype, device) if n > 1 else None


def _workspace(n, batch, device):
    lib = _lib.load()
    return torch.empty(lib.ua_pauli_gradients_workspace(n, batch), dtype=torch.uint8,
                       device=device)


def _expect_sum_z(psi, n, ws):
    lib = _lib.load()
    out = torch.empty(psi.shape[0], dtype=torch.float64, device=psi.device)
    rc = lib.ua_expect_sum_z(_lib.ptr(out), _lib.ptr(psi), _lib.DTYPE_CODES[psi.dtype], n,
                             psi.shape[0], _lib.ptr(ws), _lib.stream_handle(psi.device))
    _lib.check(rc, "expect_sum_z")
    return out


def _pauli_grads(lam, psi, n, pauli, ws, out):
    lib = _lib.load()
    rc = lib.ua_pauli_gradients(_lib.ptr(out), _lib.ptr(lam), _lib.ptr(psi),
                                _lib.DTYPE_CODES[psi.dtype], n, psi.shape[0], pauli,
                                _lib.ptr(ws), _lib.stream_handle(psi.device))
    _lib.check(rc, "pauli_gradients")


def _check(theta, phi, n):
    if theta.shape != phi.shape or theta.dim() != 3 or theta.shape[2] != n:
        raise ValueError("theta and phi must both have shape (batch, layers, num_qubits)")
    if theta.device.type != "cuda":
        raise RuntimeError("unitair_amd.adjoint runs on ROCm GPU tensors only")
    if n < 4:
        raise ValueError("the adjoint path needs at least 4 qubits")


def run_circuit(theta: torch.Tensor, phi: torch.Tensor, num_qubits: int,
                dtype: torch.dtype = torch.complex64) -> torch.Tensor:
    """Final states (batch, 2^n) of the circuit, no gradients."""
    _check(theta, phi, num_qubits)
    B = theta.shape[0]
    with on_device(theta.device):
        progs = _Programs(theta.detach(), phi.detach(), num_qubits, dtype, theta.device)
        psi = torch.zeros(B, 2 ** num_qubits, dtype=dtype, device=theta.device)
        psi[:, 0] = 1
        for prog in progs.fwd:
            prog.run(psi)
    return psi


class AdjointSumZ(torch.autograd.Function):
    """loss[b] = sum_i <Z_i> of the circuit's final state; float64."""

    @staticmethod
    def forward(ctx, theta, phi, num_qubits, dtype):
        n = num_qubits
        _check(theta, phi, n)
        B, L = theta.shape[0], theta.shape[1]
        dev = theta.device
        with on_device(dev):
            progs = _Programs(theta.detach(), phi.detach(), n, dtype, dev)
            psi = torch.zeros(B, 2 ** n, dtype=dtype, device=dev)
            psi[:, 0] = 1
            every = max(1, math.ceil(math.sqrt(L)))
            ckpt = {}
            for l in range(L):
                if l > 0 and l % every == 0:
                    ckpt[l] = psi.clone()        # state before layer l
                progs.fwd[l].run(psi)
            ws = _workspace(n, B, dev)
            loss = _expect_sum_z(psi, n, ws)
        ctx.save_for_backward(theta, phi)
        ctx.state = (progs, psi, ckpt, ws, n, dtype)
        return loss

    @staticmethod
    def backward(ctx, gloss):
        theta, phi = ctx.saved_tensors
        progs, psi, ckpt, ws, n, dtype = ctx.state
        ctx.state = None
        B, L = theta.shape[0], theta.shape[1]
        dev = theta.device
        lib = _lib.load()
        with on_device(dev):
            lam = torch.empty_like(psi)
            rc = lib.ua_apply_sum_z(_lib.ptr(lam), _lib.ptr(psi), _lib.DTYPE_CODES[dtype], n, B,
                                    _lib.stream_handle(dev))
            _lib.check(rc, "apply_sum_z")
            gy = torch.empty(L, B, n, dtype=torch.float64, device=dev)
            gz = torch.empty(L, B, n, dtype=torch.float64, device=dev)
            if progs.cz_only is not None:
                progs.cz_only.run(psi)
                progs.cz_only.run(lam)
            for l in range(L - 1, -1, -1):
                _pauli_grads(lam, psi, n, _lib.UA_PAULI_Y, ws, gy[l])
                progs.ry_back[l].run(psi)
                progs.ry_back[l].run(lam)
                _pauli_grads(lam, psi, n, _lib.UA_PAULI_Z, ws, gz[l])
                progs.rz_cz_back[l].run(psi)
                progs.rz_cz_back[l].run(lam)
                if l in ckpt:
                    # re-synchronise psi: the checkpoint is the state before
                    # layer l; the previous layer's CZ ring is undone again
                    psi.copy_(ckpt[l])
                    progs.cz_only.run(psi)
        scale = gloss.to(torch.float64).view(1, B, 1)
        gtheta = (gy * scale).permute(1, 0, 2).to(theta.dtype)
        gphi = (gz * scale).permute(1, 0, 2).to(phi.dtype)
        return gtheta, gphi, None, None


def sum_z_loss(theta: torch.Tensor, phi: torch.Tensor, num_qubits: int,
               dtype: torch.dtype = torch.complex64) -> torch.Tensor:
    """Differentiable loss[b] = sum_i <Z_i> (float64) with adjoint gradients."""
    return AdjointSumZ.apply(theta, phi, num_qubits, dtype)
