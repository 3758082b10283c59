"""Host logic of the reductions' backward rules (unitair_amd.states._Reduce /
_AbsSquared): with the forward launch replaced by the reference's torch
expression on the CPU (innerprod.py:4-65), the backward formulas must give
torch autograd's gradients for every broadcast form and dtype."""
import pytest
import torch

from oracle import unitair_torch as T
import unitair_amd.states as S


@pytest.fixture
def cpu_forward(monkeypatch):
    def fake(kind, out_dtype, a, b=None):
        if kind == "norm":
            return T.norm_squared(a).to(out_dtype)
        if kind == "inner":
            return T.inner_product(a, b).to(out_dtype)
        return T.diag_expectation_value(a, b).to(out_dtype)
    monkeypatch.setattr(S, "_reduce", fake)
    monkeypatch.setattr(S, "_abs_squared_launch", T.abs_squared)


def _check(ref_fn, fn, *xs):
    xr = [x.clone().requires_grad_(True) for x in xs]
    xg = [x.clone().requires_grad_(True) for x in xs]
    w = None
    outs = []
    for f, ins in ((ref_fn, xr), (fn, xg)):
        out = f(*ins)
        if w is None:
            w = torch.linspace(-1, 1, out.numel(), dtype=torch.float64).reshape(out.shape)
        loss = (out.real * w).sum() + (out.abs() ** 2).sum()
        outs.append(torch.autograd.grad(loss, ins))
    for r, g in zip(*outs):
        assert r.shape == g.shape and r.dtype == g.dtype
        assert torch.allclose(r, g, atol=1e-12)


@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64, torch.float64])
def test_backward_rules(cpu_forward, dt):
    torch.manual_seed(0)
    a = torch.randn(2, 3, 8, dtype=dt)
    b = torch.randn(3, 8, dtype=dt)
    dreal = torch.randn(8, dtype=torch.float64)
    dcplx = torch.randn(2, 1, 8, dtype=torch.complex128)
    _check(T.norm_squared, lambda x: S._Reduce.apply("norm", S._REAL_OF[dt], x, None), a)
    _check(T.abs_squared, lambda x: S._AbsSquared.apply(x), a)
    _check(T.inner_product, lambda x, y: S._Reduce.apply("inner", dt, x, y), a, b)
    _check(T.inner_product, lambda x, y: S._Reduce.apply("inner", dt, x, y), b, a)
    for d in (dreal, dcplx):
        od = torch.promote_types(S._REAL_OF[dt], d.dtype)
        _check(T.diag_expectation_value,
               lambda h, x: S._Reduce.apply("diag", od, h, x), d, a)


def test_planted_mistake_fails(cpu_forward):
    """A wrong rule (grad_a = g b instead of conj(g) b) is caught."""
    torch.manual_seed(1)
    a = torch.randn(4, 8, dtype=torch.complex128)
    b = torch.randn(4, 8, dtype=torch.complex128)
    xr = [a.clone().requires_grad_(True), b.clone().requires_grad_(True)]
    out = T.inner_product(*xr)
    g = torch.randn(4, dtype=torch.complex128)
    ga, _ = torch.autograd.grad(out, xr, g)
    wrong = g.unsqueeze(-1) * b
    assert not torch.allclose(ga, wrong)
    xg = [a.clone().requires_grad_(True), b.clone().requires_grad_(True)]
    out2 = S._Reduce.apply("inner", torch.complex128, *xg)
    ga2, _ = torch.autograd.grad(out2, xg, g)
    assert torch.allclose(ga, ga2)
