"""The dense-block schedule (ua_plan_dense_blocks) on the CPU: which gates
share a matrix-core pass.  Host-only; no GPU call."""
import ctypes

import numpy as np
import pytest

from unitair_amd import _lib

C64, C128, F32 = _lib.UA_COMPLEX64, _lib.UA_COMPLEX128, _lib.UA_FLOAT32


def gate_array(spec):
    """spec: list of (kind, qubits)."""
    arr = (_lib.Gate * max(1, len(spec)))()
    for i, (kind, qs) in enumerate(spec):
        arr[i].kind = {"dense": _lib.UA_GATE_DENSE, "cz": _lib.UA_GATE_CZ,
                       "cnot": _lib.UA_GATE_CNOT, "phase": _lib.UA_GATE_PHASE}[kind]
        arr[i].num_targets = len(qs)
        for j, q in enumerate(qs):
            arr[i].qubits[j] = q
        arr[i].data = 1 if kind == "dense" else None  # never dereferenced by the planner
    return arr


def plan(spec, n, dtype):
    lib = _lib.load()
    arr = gate_array(spec)
    unit = (ctypes.c_int * max(1, len(spec)))()
    nu = ctypes.c_int()
    rc = lib.ua_plan_dense_blocks(dtype, n, arr, len(spec), unit, ctypes.byref(nu))
    return rc, [unit[i] for i in range(len(spec))], nu.value


def test_two_disjoint_5q_gates_share_a_pass():
    n = 28
    rc, unit, nu = plan([("dense", (0, 1, 2, 3, 4)), ("dense", (23, 24, 25, 26, 27))], n, C64)
    assert rc == 0 and unit == [0, 0] and nu == 1


def test_pass_limits_and_order():
    n = 28
    rng = np.random.default_rng(0)
    spec = []
    for _ in range(60):
        k = int(rng.integers(4, 6))
        spec.append(("dense", tuple(int(q) for q in rng.choice(n, k, replace=False))))
        if rng.random() < 0.2:
            spec.append(("cz", tuple(int(q) for q in rng.choice(n, 2, replace=False))))
    for dtype, maxg in ((C64, 3), (C128, 2)):
        rc, unit, nu = plan(spec, n, dtype)
        assert rc == 0
        assert unit == sorted(unit) and unit[-1] == nu - 1          # in order, no gaps
        for u in range(nu):
            members = [i for i in range(len(spec)) if unit[i] == u]
            assert members == list(range(members[0], members[-1] + 1))
            if len(members) > 1:
                assert len(members) <= maxg
                assert all(spec[i][0] == "dense" and len(spec[i][1]) in (4, 5) for i in members)
                bits = {0, 1}
                for i in members:
                    bits |= {n - 1 - q for q in spec[i][1]}
                assert len(bits) <= 12
        # greedy: a pass is cut only when the next gate does not fit
        for u in range(nu - 1):
            last = max(i for i in range(len(spec)) if unit[i] == u)
            nxt = last + 1
            members = [i for i in range(len(spec)) if unit[i] == u]
            if spec[nxt][0] != "dense" or len(spec[nxt][1]) < 4 or spec[members[0]][0] != "dense" \
                    or len(spec[members[0]][1]) < 4:
                continue
            bits = {0, 1}
            for i in members + [nxt]:
                bits |= {n - 1 - q for q in spec[i][1]}
            assert len(members) == maxg or len(bits) > 12


def test_small_states_and_other_gates_run_alone():
    rc, unit, nu = plan([("dense", (0, 1, 2, 3)), ("dense", (4, 5, 6, 7))], 13, C64)
    assert rc == 0 and nu == 2                       # below 14 qubits: one gate per launch
    rc, unit, nu = plan([("dense", (0, 1, 2, 3)), ("dense", (4, 5, 6, 7, 8))], 20, F32)
    assert rc == 0 and nu == 2                       # float32 state: no passes
    rc, unit, nu = plan([("dense", (0, 1, 2)), ("phase", ()), ("dense", (4, 5, 6, 7))], 20, C64)
    assert rc == 0 and unit == [0, 1, 2]


def test_invalid_gates_rejected():
    bad = _lib.UA_ERR_INVALID_ARGUMENT
    assert plan([("dense", (0, 0, 1, 2))], 20, C64)[0] == bad       # repeated qubit
    assert plan([("dense", (0, 1, 2, 20))], 20, C64)[0] == bad      # out of range
    assert plan([("cz", (3, 3))], 20, C64)[0] == bad
