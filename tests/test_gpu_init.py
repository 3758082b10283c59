"""State constructors on the device (unitair/initializations.py:46-128):
rand_state against the oracle's Philox/Box-Muller restatement and the
reference's distribution; uniform_superposition and unit_vector against the
reference's torch calls."""
import numpy as np
import pytest
import torch

import unitair_amd as ua
from unitair_amd import initializations as init
from oracle.sampling import rand_state_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt,tol", [(torch.complex64, 2e-6), (torch.complex128, 1e-13),
                                    (torch.float32, 2e-6), (torch.float64, 1e-13)])
@pytest.mark.parametrize("n", [2, 5, 10])
def test_rand_state_matches_generator_restatement(device, dt, tol, n):
    got = init.rand_state(n, (3,), device=device, dtype=dt, seed=987654321).cpu().numpy()
    ref = rand_state_np(n, 3, 987654321, dt.is_complex)
    assert got.shape == (3, 2 ** n)
    assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))


def test_rand_state_norms_repeatability_and_distribution(device):
    from scipy.stats import kstest
    n = 16
    s = init.rand_state(n, (4, 2), device=device)
    norms = (s.abs().double() ** 2).sum(-1)
    assert torch.allclose(norms, torch.ones_like(norms), atol=1e-6)
    assert not torch.equal(s[0, 0], s[0, 1])
    torch.manual_seed(5)
    a = init.rand_state(n, device=device)
    torch.manual_seed(5)
    b = init.rand_state(n, device=device)
    assert torch.equal(a, b)
    # sqrt(N) * Re(psi) ~ N(0, 1/2) as for normalised torch.randn complex entries
    x = (a.real.double() * np.sqrt(2 ** n) * np.sqrt(2)).cpu().numpy()
    assert kstest(x, "norm").pvalue > 1e-4
    r = init.rand_state(12, device=device, dtype=torch.float64)
    assert kstest((r * 2 ** 6).cpu().numpy(), "norm").pvalue > 1e-4


def test_rand_state_28_qubits_is_normalised(device):
    s = init.rand_state(28, device=device, dtype=torch.complex64, seed=3)
    assert abs(float(ua.states.norm_squared_f64(s)) - 1.0) < 1e-5


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128, torch.float32, torch.float64])
def test_uniform_superposition_and_unit_vector_match_reference_calls(device, dt):
    for n, bd in ((0, None), (3, None), (5, (2, 3)), (13, (2,))):
        got = init.uniform_superposition(n, bd, device=device, dtype=dt).cpu()
        ref = 2 ** (-n / 2.) * torch.ones(2 ** n, dtype=dt)
        if bd is not None:
            ref = ref.repeat(torch.Size(bd) + (1,))
        assert torch.equal(got, ref)
    for idx, n in ((0, 1), (5, 3), (-1, 4), (4095, 12)):
        got = init.unit_vector(idx, num_qubits=n, device=device, dtype=dt).cpu()
        ref = torch.zeros(2 ** n, dtype=dt)
        ref[idx] = 1
        assert torch.equal(got, ref)
    assert torch.equal(init.unit_vector(2, dim=6, device=device, dtype=dt).cpu(),
                       torch.tensor([0, 0, 1, 0, 0, 0], dtype=dt))
    v = init.unit_vector_from_bitstring("0110", device=device, dtype=dt)
    assert int(torch.nonzero(v).item()) == 6
    with pytest.raises(IndexError):
        init.unit_vector(16, num_qubits=4, device=device, dtype=dt)


def test_reference_shape_and_norm_properties(device):
    # reference tests/test_initializations.py:10-44 on the device
    for n in range(1, 6):
        for bd in ((), (3,), (2, 0), (1, 4, 2)):
            for dt in (torch.complex64, torch.float32):
                s = init.rand_state(num_qubits=n, batch_dims=bd, device=device, dtype=dt)
                assert s.shape == bd + (2 ** n,)
                assert ua.norm_squared(s).isclose(torch.tensor(1.0, device=device)).all()
            u = init.uniform_superposition(num_qubits=n, batch_dims=bd, device=device)
            assert u.shape == bd + (2 ** n,)
            assert u.real.isclose(torch.tensor(2 ** (-n / 2.), device=device)).all()
            assert (u.imag == 0).all()
