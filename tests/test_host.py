"""Host-side behaviour of unitair_amd that needs no GPU: validation order,
exception types (unitair/simulation/operations.py:45-148, states/shapes.py),
batch-broadcast analysis, layout helpers, and that CPU tensors are refused
rather than computed on a CPU path."""
import numpy as np
import pytest
import torch

import unitair_amd as ua
from unitair_amd import simulation as sim
from unitair_amd import states
from unitair_amd.simulation import _plan_batch


def test_count_qubits_and_state_shape_error():
    assert states.count_qubits(torch.zeros(7, 4, 1, 32)) == 5
    assert states.count_qubits(torch.zeros(2, 2, 2, 2, 2)) == 1
    with pytest.raises(states.StateShapeError) as e:
        states.count_qubits(torch.zeros(3, 6))
    assert isinstance(e.value, ValueError)
    assert "(3, 6)" in str(e.value)


def test_count_qubits_gate_matrix():
    assert states.count_qubits_gate_matrix(torch.zeros(5, 8, 8)) == 3
    with pytest.raises(RuntimeError):
        states.count_qubits_gate_matrix(torch.zeros(3, 3))


def test_get_qubit_indices_examples():
    # reference tests/test_state_shapes.py:88-105 and shapes.py docstring
    st = states.to_tensor_layout(torch.rand(5, 2 ** 4))
    assert states.get_qubit_indices(3, st, 4) == 4
    st = states.to_tensor_layout(torch.rand(2, 5, 7, 2 ** 5))
    got = states.get_qubit_indices(torch.tensor([2, -3, 1, 0]), st, 5)
    assert torch.equal(got, torch.tensor([5, -3, 4, 3]))
    st = states.to_tensor_layout(torch.rand(2, 2, 2 ** 4))
    assert states.get_qubit_indices([-2, -4, 0, 1], st, 4) == [-2, -4, 2, 3]
    with pytest.raises(ValueError):
        states.get_qubit_indices(4, st, 4)


def test_layout_round_trip():
    v = torch.rand(7, 1, 4, 2 ** 6)
    t = states.to_tensor_layout(v)
    assert t.shape == (7, 1, 4) + (2,) * 6
    assert torch.equal(states.to_vector_layout(t, 6), v)
    x = torch.rand(4, 3, 2, 2, 2, 2, 2)
    assert states.count_qubits(states.to_vector_layout(x, 3)) == 3


def test_subset_rolls():
    x = torch.rand(4, 5, 6, 7, 8, 9)
    assert states.subset_roll_to_back(x, 2).shape == (6, 7, 8, 9, 4, 5)
    assert states.subset_roll_to_front(x, 2).shape == (8, 9, 4, 5, 6, 7)


@pytest.mark.parametrize("qubits,exc", [((0, 4), ValueError), ((1,), ValueError),
                                        ((1, 1), ValueError)])
def test_apply_operator_validation(qubits, exc):
    state = torch.zeros(2 ** 4, dtype=torch.complex64)
    op = torch.zeros(4, 4, dtype=torch.complex64)
    with pytest.raises(exc):
        sim.apply_operator(op, qubits, state)


def test_apply_operator_validation_order():
    # state shape is checked first (operations.py:126), then qubit range, then
    # the gate size, then the qubit count
    with pytest.raises(states.StateShapeError):
        sim.apply_operator(torch.zeros(3, 3), (0,), torch.zeros(6))
    with pytest.raises(ValueError, match="not consistent"):
        sim.apply_operator(torch.zeros(3, 3), (9,), torch.zeros(8))
    with pytest.raises(RuntimeError):
        sim.apply_operator(torch.zeros(3, 3), (0,), torch.zeros(8))
    with pytest.raises(ValueError, match="Cannot apply"):
        sim.apply_operator(torch.zeros(4, 4), (0,), torch.zeros(8))


def test_operator_rank_and_dtype_errors():
    with pytest.raises(RuntimeError):          # 1-D "gate" (simulation/utils.py:13-19)
        sim.act_first_qubits(torch.zeros(2, dtype=torch.complex64),
                             torch.zeros(4, dtype=torch.complex64))
    with pytest.raises(ValueError):
        sim.act_first_qubits(torch.zeros(8, 8), torch.zeros(4))
    with pytest.raises(ValueError):
        sim.apply_all_qubits(torch.zeros(4, 4), torch.zeros(8))


def test_cpu_tensors_are_refused():
    op = torch.eye(2, dtype=torch.complex64)
    st = torch.zeros(8, dtype=torch.complex64)
    for fn in (lambda: sim.apply_operator(op, (0,), st), lambda: sim.act_first_qubits(op, st),
               lambda: sim.apply_all_qubits(op, st),
               lambda: sim.apply_phase(torch.zeros(8), st), lambda: ua.norm_squared(st)):
        with pytest.raises(RuntimeError, match="ROCm GPU"):
            fn()


@pytest.mark.parametrize("ob,sb,ok", [
    ((), (3,), True), ((3,), (3,), True), ((5,), (), True), ((5,), (3, 5), True),
    ((1, 5), (3, 5), True), ((1,), (3, 5), True), ((2,), (3,), False),
    ((2, 4), (4,), True), ((1, 1, 1, 1, 4), (4,), True), ((3, 1, 1, 1, 4), (4,), False),
    ((5,), (5, 1), False), ((4, 2, 2), (), True),
])
def test_plan_batch_matches_oracle(ob, sb, ok):
    """The host's broadcast analysis accepts exactly what the numpy
    restatement of the reference's einsum accepts."""
    from oracle import unitair_np
    n, k = 3, 1
    rng = np.random.default_rng(0)
    op = rng.standard_normal(ob + (2, 2)) + 0j
    st = rng.standard_normal(sb + (2 ** n,)) + 0j
    try:
        ref = unitair_np.apply_operator(op, [0], st)
        ref_ok = True
    except (RuntimeError, ValueError):
        ref_ok = False
    assert ref_ok == ok
    if ok:
        out_batch, *_ = _plan_batch(ob, sb, n, k, [0])
        assert tuple(out_batch) + (2 ** n,) == ref.shape
    else:
        with pytest.raises((RuntimeError, ValueError)):
            _plan_batch(ob, sb, n, k, [0])


def test_gate_conventions():
    from unitair_amd import gates
    assert torch.allclose(gates.exp_x(0.5), torch.tensor(
        [[0.8776, -0.4794j], [-0.4794j, 0.8776]], dtype=torch.complex64), atol=1e-4)
    assert gates.exp_y(0.1, dtype=torch.complex128).dtype == torch.complex128
    assert gates.exp_z(torch.zeros(3)).shape == (3, 2, 2)
    assert torch.equal(gates.cnot()[2:, 2:], torch.tensor([[0, 1], [1, 0]], dtype=torch.complex64))
    u = gates.haar_unitary(8, np.random.default_rng(1))
    assert np.allclose(u.conj().T @ u, np.eye(8))


def test_int_to_bin_and_strings():
    # reference tests/test_utils.py:46-53
    from unitair_amd.utils import bool_to_strings, int_to_bin
    for num_bits in range(1, 9):
        for x in range(2 ** num_bits):
            digits = int_to_bin(x, num_bits=num_bits)
            powers = torch.tensor([2 ** k for k in reversed(range(num_bits))])
            assert x == int((digits * powers).sum())
            assert int_to_bin(x, num_bits, str_output=True) == bool_to_strings(digits)
    with pytest.raises(ValueError):
        int_to_bin(8, 3)
    with pytest.raises(ValueError):
        int_to_bin(-1, 3)


def test_permutation_helpers():
    # reference tests/test_utils.py:28-44
    from unitair_amd.utils import inverse_list_permutation, permutation_to_front
    perm = permutation_to_front(6, [4, 1])
    assert perm == [4, 1, 0, 2, 3, 5]
    inv = inverse_list_permutation(perm)
    assert [perm[i] for i in inv] == list(range(6))


def test_measurement_histogram_host_behaviour():
    from collections import OrderedDict
    from unitair_amd.measurement import MeasurementHistogram
    h = MeasurementHistogram(3, OrderedDict([("101", 7), ("000", 2)]))
    assert h["101"] == 7 and h["111"] == 0
    with pytest.raises(KeyError):
        h["12"]
    assert h.num_distinct_samples == 2 and h.observed_samples == {"101", "000"}
    assert list(h.int_key_histogram().items()) == [(5, 7), (0, 2)]
    assert repr(h) == "101: 7\n000: 2"
    big = MeasurementHistogram(6, OrderedDict((format(i, "06b"), 60 - i) for i in range(55)))
    r = repr(big)
    assert r.count("\n") == 15 and r.endswith(" ... (40 lines omitted)")


def test_reference_module_paths():
    # the reference's package layout: unitair.states.{shapes,conversions,innerprod},
    # unitair.simulation.{operations,measurement,utils}, unitair.gates.{gates,matrix_algebra}
    import importlib
    for mod, names in {
        "unitair_amd.states.shapes": ["StateLayout", "StateShapeError", "count_qubits",
                                      "count_qubits_gate_matrix", "get_qubit_indices",
                                      "subset_roll_to_back", "subset_roll_to_front"],
        "unitair_amd.states.conversions": ["to_tensor_layout", "to_vector_layout"],
        "unitair_amd.states.innerprod": ["abs_squared", "norm_squared", "inner_product",
                                         "diag_expectation_value"],
        "unitair_amd.simulation.operations": ["apply_operator", "apply_phase", "apply_all_qubits",
                                              "apply_to_qubits", "act_first_qubits",
                                              "act_last_qubit", "swap", "permute_qubits",
                                              "roll_qubits", "multi_cz", "multi_controlled_z",
                                              "multi_controlled_x"],
        "unitair_amd.simulation.measurement": ["measure", "MeasurementHistogram"],
        "unitair_amd.simulation.utils": ["count_gate_batch_dims", "gate_batch_size"],
        "unitair_amd.gates.gates": ["exp_x", "exp_y", "exp_z", "hadamard", "pauli_x", "pauli_y",
                                    "pauli_z", "cnot", "cz"],
        "unitair_amd.gates.matrix_algebra": ["fuse_single_qubit_operators"],
        "unitair_amd.utils": ["int_to_bin", "bool_to_strings", "permutation_to_front",
                              "inverse_list_permutation"],
        "unitair_amd.simulation": ["apply_operator", "apply_phase", "apply_to_qubits",
                                   "apply_all_qubits", "act_last_qubit", "swap", "multi_cz",
                                   "multi_controlled_x", "multi_controlled_z",
                                   "act_first_qubits", "measure"],
        "unitair_amd": ["states", "simulation", "StateLayout", "unit_vector",
                        "uniform_superposition", "rand_state", "count_qubits",
                        "hilbert_space_dim", "diag_expectation_value", "inner_product",
                        "norm_squared", "abs_squared", "VECTOR_LAYOUT", "TENSOR_LAYOUT"],
    }.items():
        m = importlib.import_module(mod)
        for name in names:
            assert hasattr(m, name), (mod, name)
    from unitair_amd.simulation.utils import count_gate_batch_dims, gate_batch_size
    assert count_gate_batch_dims(torch.zeros(3, 5, 4, 4)) == 2
    assert gate_batch_size(torch.zeros(3, 5, 4, 4)) == torch.Size([3, 5])
    with pytest.raises(RuntimeError):
        count_gate_batch_dims(torch.zeros(4))
