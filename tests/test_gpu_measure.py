"""Sampling on the GPU (ua_sample) against the oracle's Philox4x32-10 and
inverse-CDF draw rule, plus the distribution and histogram behaviour of
unitair's measure (unitair/simulation/measurement.py:9-140).  The draws are
random: parity with the reference is in distribution (chi-square), and with
the oracle's draw rule sample for sample."""
import numpy as np
import pytest
import torch

from unitair_amd import simulation as sim
from unitair_amd.measurement import sample_indices
from oracle.sampling import draw, uniform53

pytestmark = pytest.mark.gpu


def _seed(gen_seed):
    return int(torch.randint(0, 2 ** 62, (1,),
                             generator=torch.Generator().manual_seed(gen_seed)).item())


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128, torch.float32])
@pytest.mark.parametrize("n", [3, 10, 14])
def test_uniform_weights_draw_exactly_floor_u_n(device, dt, n):
    # all |psi_x|^2 = 1: every prefix is an exact integer, so sample s is
    # floor(U(s) * 2^n) exactly
    state = torch.ones(2 ** n, dtype=dt, device=device)
    k = 5000
    got = sample_indices(state, k, generator=torch.Generator().manual_seed(11)).cpu().numpy()
    seed = _seed(11)
    expect = np.array([int(uniform53(s, seed) * 2 ** n) for s in range(k)])
    assert np.array_equal(got, expect)


def test_draws_follow_the_inverse_cdf_of_the_state(device):
    rng = np.random.default_rng(3)
    n = 13
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    psi[rng.random(2 ** n) < 0.3] = 0           # zero-probability outcomes are never drawn
    k = 4000
    got = sample_indices(torch.from_numpy(psi).to(device), k,
                         generator=torch.Generator().manual_seed(5)).cpu().numpy()
    seed = _seed(5)
    probs = np.abs(psi) ** 2
    expect = np.array([draw(probs, uniform53(s, seed)) for s in range(k)])
    # float64 prefix sums in another order may move a draw that lands within
    # ~1e-16 of a boundary by one index; none is expected in 4000 draws
    mism = np.nonzero(got != expect)[0]
    assert len(mism) <= 1 and all(abs(got[i] - expect[i]) == 1 for i in mism)
    assert np.all(probs[got] > 0)


def test_chi_square_against_probabilities(device):
    from scipy.stats import chisquare
    rng = np.random.default_rng(4)
    n = 7
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    k = 200000
    got = sample_indices(torch.from_numpy(psi).to(device, torch.complex64), k).cpu().numpy()
    p = np.abs(psi.astype(np.complex64).astype(np.complex128)) ** 2
    counts = np.bincount(got, minlength=2 ** n)
    assert chisquare(counts, p / p.sum() * k).pvalue > 1e-4


def test_measure_histogram_format(device):
    n = 4
    basis = torch.zeros(2 ** n, dtype=torch.complex64, device=device)
    basis[0b0101] = 1
    h = sim.measure(basis, 300)
    assert h.num_qubits == 4 and dict(h.histogram) == {"0101": 300}
    assert h["0101"] == 300 and h["1111"] == 0
    assert sim.measure(basis, 300, raw_output=True) == {5: 300}
    # sorted by count, descending; keys are big-endian bit strings
    amp = torch.tensor([0.1, 0.7, 0.0, 0.2, 0.0, 0.0, 0.9, 0.05], dtype=torch.float64,
                       device=device)
    h = sim.measure(amp, 20000)
    counts = list(h.histogram.values())
    assert counts == sorted(counts, reverse=True)
    assert set(h.histogram) <= {"000", "001", "011", "110", "111"}
    assert next(iter(h.histogram)) == "110"
    assert sum(counts) == 20000


def test_measure_is_repeatable_under_manual_seed(device):
    psi = torch.randn(2 ** 9, dtype=torch.complex64, device=device)
    torch.manual_seed(123)
    a = sim.measure(psi, 1000, raw_output=True)
    torch.manual_seed(123)
    b = sim.measure(psi, 1000, raw_output=True)
    assert a == b


def test_large_state_support(device):
    # 24 qubits, support on even indices only: every draw is even and every
    # outcome's frequency stays close to its probability
    n = 24
    psi = torch.zeros(2 ** n, dtype=torch.complex64, device=device)
    psi[0::2] = 1.0
    psi[2 ** (n - 1)] = 2 ** 12                 # one outcome with ~1/3 of the mass
    k = 300000
    got = sample_indices(psi, k)
    assert int((got % 2).sum()) == 0
    p_big = 2 ** 24 / (2 ** 24 + (2 ** 23 - 1))
    frac = float((got == 2 ** (n - 1)).double().mean())
    assert abs(frac - p_big) < 5 * np.sqrt(p_big * (1 - p_big) / k)


def test_sample_count_limit_is_checked_before_any_allocation(device):
    state = torch.ones(8, dtype=torch.complex64, device=device)
    for bad in (2 ** 31, -1):
        with pytest.raises(ValueError):
            sample_indices(state, bad)
    assert sample_indices(state, 0).shape == (0,)


def test_measure_rejects_batches_and_cpu(device):
    with pytest.raises(TypeError):
        sim.measure(torch.ones(2, 8, dtype=torch.complex64, device=device), 10)
    with pytest.raises(RuntimeError):
        sim.measure(torch.ones(8, dtype=torch.complex64), 10)


def test_uniform_weights_25_qubits_global_prefix_search(device):
    # 8192 chunk prefixes: the search over them stays in global memory
    n = 25
    state = torch.ones(2 ** n, dtype=torch.complex64, device=device)
    k = 3000
    got = sample_indices(state, k, generator=torch.Generator().manual_seed(21)).cpu().numpy()
    seed = _seed(21)
    assert np.array_equal(got, [int(uniform53(s, seed) * 2 ** n) for s in range(k)])


@pytest.mark.parametrize("bad", ["zero", "nan", "inf"])
@pytest.mark.parametrize("n", [4, 13])
def test_invalid_distribution_raises_value_error(device, bad, n):
    """Categorical(probs) raises ValueError when the probabilities sum to
    zero or are not finite (measurement.py:45-48); so do measure and
    sample_indices, for a zero, a NaN and an infinite amplitude."""
    state = torch.zeros(2 ** n, dtype=torch.complex64, device=device)
    if bad == "nan":
        state[3] = float("nan")
    elif bad == "inf":
        state[:] = 1
        state[5] = float("inf")
    with pytest.raises(ValueError):
        sample_indices(state, 100)
    with pytest.raises(ValueError):
        sim.measure(state, 10)
