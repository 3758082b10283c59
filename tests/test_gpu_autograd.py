"""Gradients: the drop-in's autograd (ua_apply_operator / ua_operator_gradient /
phase backward) against torch autograd through the torch restatement of the
reference (CPU, complex128); the adjoint path against exact parameter-shift
derivatives of the complex128 oracle, central finite differences, and torch
autograd through the restatement."""
import numpy as np
import pytest
import torch

from unitair_amd import simulation as sim
from unitair_amd import adjoint
from oracle import unitair_torch as T
import adjoint_oracle as AO

pytestmark = pytest.mark.gpu


def _grads_cpu(fn, *xs):
    xs = [x.detach().clone().requires_grad_(True) for x in xs]
    out = fn(*xs)
    w = torch.linspace(-1, 1, out.numel(), dtype=torch.float64).reshape(out.shape)
    loss = (out.real * w).sum() + (out.abs() ** 2).sum()
    return torch.autograd.grad(loss, xs)


def _grads_gpu(fn, device, *xs):
    xs = [x.detach().clone().to(device).requires_grad_(True) for x in xs]
    out = fn(*xs)
    w = torch.linspace(-1, 1, out.numel(), dtype=torch.float64, device=device).reshape(out.shape)
    loss = (out.real * w).sum() + (out.abs() ** 2).sum()
    return [g.cpu() for g in torch.autograd.grad(loss, xs)]


@pytest.mark.parametrize("ob,sb,qubits", [((), (3,), (2, 0)), ((3,), (3,), (1,)),
                                         ((4,), (), (0, 3)), ((), (), (3, 1, 2))])
def test_apply_operator_autograd(device, ob, sb, qubits):
    torch.manual_seed(1)
    n, k = 4, len(qubits)
    op = torch.randn(*ob, 2 ** k, 2 ** k, dtype=torch.complex128)
    st = torch.randn(*sb, 2 ** n, dtype=torch.complex128)
    ref = _grads_cpu(lambda o, s: T.apply_operator(o, qubits, s), op, st)
    got = _grads_gpu(lambda o, s: sim.apply_operator(o, qubits, s), device, op, st)
    for r, g in zip(ref, got):
        assert torch.allclose(r, g, atol=1e-10)


# Larger states: the operator-gradient reduction is split across slices of
# the groups (n=14..20), across operators (op batch 2 with 2^19 groups each),
# and tiled 4 x 4 for k >= 3 (k = 3 with 16 slices, k = 5 and k = 6 blocked).
@pytest.mark.parametrize("n,qubits,ob,sb", [(14, (5,), (), ()), (14, (13, 0), (), (3,)),
                                            (18, (2, 9, 17), (), ()),
                                            (20, (0,), (2,), (2,)),
                                            (16, (0, 4, 8, 12, 15), (2,), (2,)),
                                            (13, (1, 3, 5, 7, 9, 11), (), ()),
                                            (12, (3,), (4,), ())])
def test_operator_gradient_large_states(device, n, qubits, ob, sb):
    torch.manual_seed(n + len(qubits))
    k = len(qubits)
    op = torch.randn(*ob, 2 ** k, 2 ** k, dtype=torch.complex128)
    st = torch.randn(*sb, 2 ** n, dtype=torch.complex128)
    ref = _grads_cpu(lambda o, s: T.apply_operator(o, qubits, s), op, st)
    got = _grads_gpu(lambda o, s: sim.apply_operator(o, qubits, s), device, op, st)
    for r, g in zip(ref, got):
        scale = float(r.abs().max())
        assert float((r - g).abs().max()) <= 1e-12 * scale


def test_operator_gradient_complex64_large(device):
    torch.manual_seed(7)
    n = 20
    op = torch.randn(2, 2, dtype=torch.complex128)
    st = torch.randn(2 ** n, dtype=torch.complex128)
    ref = _grads_cpu(lambda o, s: T.apply_operator(o, [6], s),
                     op.to(torch.complex64).to(torch.complex128),
                     st.to(torch.complex64).to(torch.complex128))
    got = _grads_gpu(lambda o, s: sim.apply_operator(o, [6], s), device,
                     op.to(torch.complex64), st.to(torch.complex64))
    for r, g in zip(ref, got):
        scale = float(r.abs().max())
        assert float((r - g.to(torch.complex128)).abs().max()) <= 1e-5 * scale


@pytest.mark.parametrize("ashape,sshape", [((16,), (3, 16)), ((3, 16), (3, 16)), ((3, 1), (3, 16)),
                                           ((), (3, 16)), ((1, 16), (2, 3, 16))])
def test_apply_phase_autograd(device, ashape, sshape):
    torch.manual_seed(2)
    ang = torch.rand(ashape, dtype=torch.float64) * 6
    st = torch.randn(sshape, dtype=torch.complex128)
    ref = _grads_cpu(lambda a, s: T.apply_phase(a, s), ang, st)
    got = _grads_gpu(lambda a, s: sim.apply_phase(a, s), device, ang, st)
    for r, g in zip(ref, got):
        assert torch.allclose(r, g, atol=1e-10)


def test_apply_phase_angle_gradient_complex64(device):
    """The native angle gradient (float32 angles, complex64 state) against the
    float64 autograd of the restatement."""
    torch.manual_seed(4)
    ang = torch.rand(2 ** 12, dtype=torch.float32) * 6
    st = torch.randn(5, 2 ** 12, dtype=torch.complex64)
    ref = _grads_cpu(lambda a, s: T.apply_phase(a, s), ang.double(), st.to(torch.complex128))
    got = _grads_gpu(lambda a, s: sim.apply_phase(a, s), device, ang, st)
    assert got[0].dtype == torch.float32
    assert float((ref[0] - got[0].double()).abs().max()) < 1e-4 * float(ref[0].abs().max())


def test_apply_all_qubits_autograd(device):
    torch.manual_seed(3)
    op = torch.randn(2, 2, dtype=torch.complex128)
    st = torch.randn(2, 2 ** 5, dtype=torch.complex128)

    def ref_fn(o, s):
        for q in range(5):
            s = T.apply_operator(o, [q], s)
        return s
    ref = _grads_cpu(ref_fn, op, st)
    got = _grads_gpu(lambda o, s: sim.apply_all_qubits(o, s), device, op, st)
    for r, g in zip(ref, got):
        assert torch.allclose(r, g, atol=1e-9)


@pytest.mark.parametrize("n,L,B", [(4, 1, 2), (5, 2, 3), (6, 3, 1), (8, 4, 2), (13, 2, 2)])
def test_adjoint_matches_parameter_shift_c128(device, n, L, B):
    rng = np.random.default_rng(n * 10 + L)
    th = rng.uniform(0, 2 * np.pi, (B, L, n))
    ph = rng.uniform(0, 2 * np.pi, (B, L, n))
    t = torch.from_numpy(th).to(device).requires_grad_(True)
    p = torch.from_numpy(ph).to(device).requires_grad_(True)
    loss = adjoint.sum_z_loss(t, p, n, dtype=torch.complex128)
    for b in range(B):
        assert abs(float(loss[b]) - AO.loss(th[b], ph[b], n)) < 1e-11
    loss.sum().backward()
    gt, gp = t.grad.cpu().numpy(), p.grad.cpu().numpy()
    for b in range(B):
        for l in range(L):
            for q in range(n):
                assert abs(gt[b, l, q] - AO.param_shift(th[b], ph[b], n, "theta", l, q)) < 1e-10
                assert abs(gp[b, l, q] - AO.param_shift(th[b], ph[b], n, "phi", l, q)) < 1e-10


def test_adjoint_matches_torch_autograd_through_restatement(device):
    """Autograd through the torch restatement of the reference op sequence
    (apply_operator with unitair's exp_y/exp_z conventions), complex128."""
    n, L, B = 5, 2, 2
    rng = np.random.default_rng(7)
    th = torch.from_numpy(rng.uniform(0, 2 * np.pi, (B, L, n)))
    ph = torch.from_numpy(rng.uniform(0, 2 * np.pi, (B, L, n)))

    def ref_loss(th, ph):
        psi = torch.zeros(B, 2 ** n, dtype=torch.complex128)
        psi[:, 0] = 1
        cz = torch.diag(torch.tensor([1, 1, 1, -1], dtype=torch.complex128))
        for l in range(L):
            for q in range(n):
                e = torch.exp(-0.5j * ph[:, l, q])
                z = torch.zeros_like(e)
                rzm = torch.stack([torch.stack([e, z], -1), torch.stack([z, e.conj()], -1)], -2)
                psi = T.apply_operator(rzm, [q], psi)
            for q in range(n):
                c, s = torch.cos(th[:, l, q] / 2), torch.sin(th[:, l, q] / 2)
                rym = torch.stack([torch.stack([c, -s], -1), torch.stack([s, c], -1)], -2)
                psi = T.apply_operator(rym.to(torch.complex128), [q], psi)
            for q in range(n):
                psi = T.apply_operator(cz, [q, (q + 1) % n], psi)
        x = torch.arange(2 ** n)
        pop = torch.tensor([bin(int(v)).count("1") for v in x], dtype=torch.float64)
        return T.diag_expectation_value(n - 2 * pop, psi)

    thr, phr = th.clone().requires_grad_(True), ph.clone().requires_grad_(True)
    ref_loss(thr, phr).sum().backward()
    t, p = th.to(device).requires_grad_(True), ph.to(device).requires_grad_(True)
    adjoint.sum_z_loss(t, p, n, dtype=torch.complex128).sum().backward()
    assert torch.allclose(t.grad.cpu(), thr.grad, atol=1e-11)
    assert torch.allclose(p.grad.cpu(), phr.grad, atol=1e-11)


def test_adjoint_16q_batch1024_c64(device):
    """The 16-qubit, 1024-state, 10-layer configuration in complex64: loss
    and gradients of sampled states against the complex128 oracle (exact
    parameter shift) and central differences (h = 1e-3, complex128 oracle);
    the complex64 gradients of all states against the complex128 adjoint."""
    n, L, B = 16, 10, 1024
    rng = np.random.default_rng(16)
    th = rng.uniform(0, 2 * np.pi, (B, L, n))
    ph = rng.uniform(0, 2 * np.pi, (B, L, n))
    grads = {}
    for dt in (torch.complex64, torch.complex128):
        t = torch.from_numpy(th).to(device, torch.float32 if dt == torch.complex64
                                     else torch.float64).requires_grad_(True)
        p = torch.from_numpy(ph).to(device, t.dtype).requires_grad_(True)
        loss = adjoint.sum_z_loss(t, p, n, dtype=dt)
        loss.sum().backward()
        grads[dt] = (loss.detach().cpu().numpy(), t.grad.double().cpu().numpy(),
                     p.grad.double().cpu().numpy())
    l64, gt64, gp64 = grads[torch.complex64]
    l128, gt128, gp128 = grads[torch.complex128]
    assert np.max(np.abs(l64 - l128)) < 1e-4 * n
    # complex64: gradients of sum_i <Z_i> (|grad| <= 16) within 2e-4 of complex128
    assert np.max(np.abs(gt64 - gt128)) < 2e-4 and np.max(np.abs(gp64 - gp128)) < 2e-4
    for b in (0, 777):
        assert abs(l128[b] - AO.loss(th[b], ph[b], n)) < 1e-10
        for (l, q) in ((0, 0), (4, 7), (9, 15)):
            ps = AO.param_shift(th[b], ph[b], n, "theta", l, q)
            assert abs(gt128[b, l, q] - ps) < 1e-9
            assert abs(gt64[b, l, q] - ps) < 2e-4
            h = 1e-3
            tp, tm = th[b].copy(), th[b].copy()
            tp[l, q] += h
            tm[l, q] -= h
            fd = (AO.loss(tp, ph[b], n) - AO.loss(tm, ph[b], n)) / (2 * h)
            assert abs(fd - ps) < 1e-5
            pz = AO.param_shift(th[b], ph[b], n, "phi", l, q)
            assert abs(gp128[b, l, q] - pz) < 1e-9


# Reductions are differentiable like the reference's torch expressions
# (innerprod.py:4-65): a VQE-style loss built from the drop-in calls only.
@pytest.mark.parametrize("dt", [torch.complex128, torch.float64])
def test_reduction_gradients(device, dt):
    import unitair_amd as ua
    torch.manual_seed(11)
    a = torch.randn(3, 2 ** 5, dtype=dt)
    b = torch.randn(3, 2 ** 5, dtype=dt)
    d = torch.randn(2 ** 5, dtype=torch.float64)
    dc = torch.randn(2, 1, 2 ** 5, dtype=torch.complex128)
    cases = [
        (lambda x: T.norm_squared(x), lambda x: ua.norm_squared(x), (a,)),
        (lambda x: T.abs_squared(x), lambda x: ua.abs_squared(x), (a,)),
        (lambda x, y: T.inner_product(x, y), lambda x, y: ua.inner_product(x, y), (a, b)),
        (lambda x, y: T.inner_product(x, y), lambda x, y: ua.states.inner_product_f64(x, y),
         (a, b[:1])),
        (lambda h, x: T.diag_expectation_value(h, x),
         lambda h, x: ua.diag_expectation_value(h, x), (d, a)),
        (lambda h, x: T.diag_expectation_value(h, x),
         lambda h, x: ua.diag_expectation_value(h, x), (dc, a)),
    ]
    for ref_fn, fn, xs in cases:
        ref = _grads_cpu(ref_fn, *xs)
        got = _grads_gpu(fn, device, *xs)
        for r, g in zip(ref, got):
            assert r.shape == g.shape and r.dtype == g.dtype
            assert torch.allclose(r, g, atol=1e-10), fn


def test_vqe_style_loss_through_drop_in_calls(device):
    """diag_expectation_value(H, apply_operator(U(theta), q, psi)) with the
    gradient reaching theta, against the same expression on the CPU
    restatement (complex128)."""
    import unitair_amd as ua
    n = 6
    torch.manual_seed(5)
    psi = torch.randn(2, 2 ** n, dtype=torch.complex128)
    h = torch.randn(2 ** n, dtype=torch.float64)
    theta = torch.tensor([0.3, -1.1], dtype=torch.float64)

    def loss(apply, expect, t, st, hh):
        c, s = torch.cos(t / 2), torch.sin(t / 2)
        u = torch.stack([torch.stack([c, -s], -1), torch.stack([s, c], -1)], -2)
        out = apply(u.to(st.dtype), [2], st)
        return expect(hh, out).sum()

    tr = theta.clone().requires_grad_(True)
    loss(T.apply_operator, T.diag_expectation_value, tr, psi, h).backward()
    tg = theta.to(device).requires_grad_(True)
    loss(sim.apply_operator, ua.diag_expectation_value, tg, psi.to(device),
         h.to(device)).backward()
    assert torch.allclose(tg.grad.cpu(), tr.grad, atol=1e-11)
