"""CPU checks of the fused-pass compiler (unitair_amd.fusion): the plan keeps
every gate, respects the local-bit and cube limits, and the 1-qubit merging
step leaves the circuit's result unchanged (checked with the oracle)."""
import numpy as np
import pytest

from unitair_amd import fusion
from unitair_amd.circuits import Circuit, Gate
from unitair_amd.gates import haar_unitary
from unitair_amd.workloads import SeededAngles, layered_circuit_12q, random_circuit_24q
from oracle import circuits as oc


def random_mixed_circuit(n, ngates, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(ngates):
        r = rng.random()
        if r < 0.45:
            gates.append(Gate("dense", (int(rng.integers(n)),), matrix=haar_unitary(2, rng)))
        elif r < 0.6:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(Gate("dense", (int(a), int(b)), matrix=haar_unitary(4, rng)))
        elif r < 0.7:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(Gate("cnot", (int(a), int(b))))
        elif r < 0.8:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(Gate("cz", (int(a), int(b))))
        elif r < 0.9:
            gates.append(Gate("phase", (), angles=SeededAngles(n, rng.integers(2 ** 62))))
        else:
            q = rng.choice(n, 3, replace=False)
            gates.append(Gate("dense", tuple(int(x) for x in q), matrix=haar_unitary(8, rng)))
    return Circuit(n, gates)


def items_to_gates(items, n):
    gates = []
    for it in items:
        qubits = tuple(n - 1 - b for b in it.bits)
        if it.kind in ("dense1", "dense2", "wide"):
            gates.append(Gate("dense", qubits, matrix=it.data))
        elif it.kind == "diag1":
            gates.append(Gate("dense", qubits, matrix=np.diag(it.data)))
        elif it.kind == "phase":
            gates.append(Gate("phase", (), angles=it.data))
        else:
            gates.append(Gate(it.kind, qubits))
    return gates


@pytest.mark.parametrize("n,seed", [(5, 0), (8, 1), (12, 2), (14, 3)])
def test_merge_preserves_result(n, seed):
    circ = random_mixed_circuit(n, 60, seed)
    items = fusion._items_from_circuit(circ)
    merged = fusion._merge_1q(items)
    assert len(merged) <= len(items)
    rng = np.random.default_rng(seed + 100)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    a = oc.run_closed_form(circ.gates, psi)
    b = oc.run_closed_form(items_to_gates(merged, n), psi)
    assert np.allclose(a, b, atol=1e-12)


@pytest.mark.parametrize("n,seed", [(4, 0), (7, 1), (12, 2), (16, 3)])
def test_plan_respects_limits(n, seed):
    circ = random_mixed_circuit(n, 120, seed)
    items = fusion._merge_1q(fusion._items_from_circuit(circ))
    plan = fusion.plan_passes(items, n)
    m = min(n, fusion.TILE_BITS)
    flat = []
    for p in plan:
        if isinstance(p, fusion._Item):
            assert p.kind == "wide"
            flat.append(p)
            continue
        assert len(p.local) == m and p.local == sorted(p.local)
        assert p.local[:4] == [0, 1, 2, 3]
        for it in p.items:
            assert fusion._needs(it) <= set(p.local)
        stages = fusion._stages(p)
        seen = []
        for cube, ops in stages:
            if cube is None:            # diagonal stage: diagonal gates only, with a phase
                assert ops and all(it.kind in ("phase", "cz", "diag1", "cz_layer") for it in ops)
                assert ops[0].kind == "phase"
                seen += ops
                continue
            assert len(cube) == 4 and cube == sorted(cube) and all(0 <= t < m for t in cube)
            assert all(it.kind != "phase" for it in ops)
            tile_of = {s: i for i, s in enumerate(p.local)}
            for it in ops:
                assert {tile_of[b] for b in fusion._needs(it)} <= set(cube)
            seen += ops
        assert sorted(map(id, seen)) == sorted(map(id, p.items))
        flat += seen
    # every gate once, and any two gates that do not commute keep their order
    assert sorted(map(id, flat)) == sorted(map(id, items))
    pos = {id(x): i for i, x in enumerate(flat)}
    for i in range(len(items)):
        for j in range(i + 1, len(items)):
            if not fusion._commute(items[i], items[j]):
                assert pos[id(items[i])] < pos[id(items[j])]
    # and the reordered circuit gives the same state (oracle, complex128)
    rng = np.random.default_rng(seed + 7)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    a = oc.run_closed_form(items_to_gates(items, n), psi)
    b = oc.run_closed_form(items_to_gates(flat, n), psi)
    assert np.allclose(a, b, atol=1e-11)


def test_commutation_rule_against_matrices():
    """fusion._commute claims exact commutation; check each claim (and each
    refusal that matters) with explicit operators on 4 qubits."""
    from unitair_amd.fusion import _Item, _commute
    rng = np.random.default_rng(3)
    n = 4
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)

    def full(it):
        psi_cols = []
        for x in range(2 ** n):
            e = np.zeros(2 ** n, dtype=np.complex128)
            e[x] = 1
            psi_cols.append(oc.run_closed_form(items_to_gates([it], n), e))
        return np.stack(psi_cols, axis=1)

    kinds = []
    for b in range(n):
        kinds.append(_Item("dense1", (b,), H))
        kinds.append(_Item("diag1", (b,), np.exp(1j * rng.random(2))))
    for a in range(n):
        for b in range(n):
            if a != b:
                kinds += [_Item("cz", (a, b)), _Item("cnot", (a, b)),
                          _Item("dense2", (a, b), np.linalg.qr(
                              rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0])]
    kinds.append(_Item("phase", (), rng.uniform(0, 6, 2 ** n)))
    mats = [full(k) for k in kinds]
    for i, a in enumerate(kinds):
        for j, b in enumerate(kinds):
            if _commute(a, b):
                assert np.allclose(mats[i] @ mats[j], mats[j] @ mats[i], atol=1e-12), \
                    (a.kind, a.bits, b.kind, b.bits)
    # CNOTs sharing a control commute; a CNOT and a gate on its target do not
    assert _commute(_Item("cnot", (0, 1)), _Item("cnot", (0, 2)))
    assert not _commute(_Item("cnot", (0, 1)), _Item("dense1", (1,), H))


def test_24q_circuit_needs_few_passes():
    circ = random_circuit_24q(seed=2)
    items = fusion._merge_1q(fusion._items_from_circuit(circ))
    plan = fusion.plan_passes(items, 24)
    assert len(plan) < 40


def test_stage_segments_keep_the_stage_result():
    # _segment_stage groups runs of 1-qubit gates on cube bits by bit (and
    # multiplies same-bit neighbours); applying its output in kernel order
    # must give the original stage's result
    from unitair_amd.fusion import _Item, _segment_stage
    from oracle import closed_form
    rng = np.random.default_rng(12)
    n = 6

    def haar2():
        q, r = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
        return q * (np.diag(r) / abs(np.diag(r)))

    cube = [0, 2, 3, 5]                              # state bits of the cube
    items = [_Item("dense1", (2,), haar2()), _Item("diag1", (0,), np.exp(1j * rng.random(2))),
             _Item("dense1", (5,), haar2()), _Item("dense1", (2,), haar2()),
             _Item("cz", (2, 4)),
             _Item("dense1", (3,), haar2()), _Item("diag1", (1,), np.exp(1j * rng.random(2))),
             _Item("dense1", (0,), haar2()), _Item("dense2", (3, 0), np.linalg.qr(
                 rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]),
             _Item("dense1", (5,), haar2())]

    def apply(its, psi):
        for it in its:
            qs = [n - 1 - b for b in it.bits]
            if it.kind == "dense1":
                psi = closed_form.apply_dense(it.data, qs, psi)
            elif it.kind == "diag1":
                psi = closed_form.apply_dense(np.diag(it.data), qs, psi)
            elif it.kind == "dense2":
                psi = closed_form.apply_dense(it.data, qs, psi)
            elif it.kind == "cz":
                psi = closed_form.apply_dense(np.diag([1, 1, 1, -1]), qs, psi)
        return psi

    seg = _segment_stage(items, cube)
    kinds = [k for k, _ in seg]
    # runs: {2,0,5,2} -> layer; cz; {3} (the diag on bit 1, not a cube bit, ends
    # the run) -> layer; diag1 bit 1 stays an op; {0} -> layer; dense2; {5} -> layer
    assert kinds == ["layer", "op", "layer", "op", "layer", "op", "layer"]
    layer = seg[0][1]
    assert [len(l) for l in layer] == [1, 1, 0, 1]   # bit 2 (cube pos 1) merged to one gate
    flat = []
    for k, x in seg:
        flat.extend([g for lane in x for g in lane] if k == "layer" else [x])
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    assert np.allclose(apply(flat, psi), apply(items, psi), atol=1e-13)


def test_dense_gates_wider_than_the_in_place_limit_are_rejected_at_build():
    """Wide dense steps run in place through ua_apply_operator (1..5 target
    qubits); a 6-qubit gate is refused while the program is built, before
    any device buffer exists."""
    import torch
    from unitair_amd.fusion import FusedProgram
    rng = np.random.default_rng(0)
    circ = Circuit(8, [Gate("dense", (0,), matrix=haar_unitary(2, rng)),
                       Gate("dense", tuple(range(6)), matrix=haar_unitary(64, rng))])
    with pytest.raises(ValueError, match="1..5"):
        FusedProgram(circ, dtype=torch.complex64, device="cpu")
    items = [fusion._Item("wide", tuple(range(6)), haar_unitary(64, rng))]
    with pytest.raises(ValueError, match="1..5"):
        FusedProgram.from_items(items, 8, dtype=torch.complex64, device="cpu")


@pytest.mark.parametrize("n,seed", [(7, 1), (16, 3)])
def test_in_order_plan_keeps_the_gate_order(n, seed):
    """reorder=False (used by the serving loop, whose results must not
    depend on how a request's layers are cut into programs) keeps every gate
    in its place."""
    circ = random_mixed_circuit(n, 120, seed)
    items = fusion._merge_1q(fusion._items_from_circuit(circ))
    flat = []
    for p in fusion.plan_passes(items, n, reorder=False):
        if isinstance(p, fusion._Item):
            flat.append(p)
            continue
        seen = []
        for cube, ops in fusion._stages(p, reorder=False):
            seen += ops
        assert seen == p.items
        flat += p.items
    assert [id(x) for x in flat] == [id(x) for x in items]


def _template_matches_plan(items, n):
    from unitair_amd.fusion import PlanTemplate, _pass_structs
    t = PlanTemplate(items, n)
    steps, ops, pending = _pass_structs(items, n)
    assert [bytes(o) for o in ops] == [bytes(o) for o in t.op_structs]
    assert [i for i, _ in pending] == [i for i, _ in t.pending]
    for (_, it), (_, r) in zip(pending, t.pending):
        v = PlanTemplate.evaluate(r, items)
        assert v.kind == it.kind and v.bits == it.bits
        if it.data is None:
            assert v.data is None
        else:
            assert np.array_equal(np.asarray(v.data), np.asarray(it.data))
    return t


@pytest.mark.parametrize("n,seed", [(8, 1), (12, 2), (14, 3), (20, 4)])
def test_plan_template_reproduces_the_plan(n, seed):
    """A PlanTemplate's recipes (input items, their diagonals as 2x2
    matrices, products of runs on one bit) give back exactly the data the
    planner attaches to each op, for mixed circuits and for runs of
    diagonal and dense 1-qubit gates on shared qubits."""
    from unitair_amd.fusion import _Item, _items_from_circuit, _merge_1q
    items = _merge_1q(_items_from_circuit(random_mixed_circuit(n, 80, seed)))
    _template_matches_plan(items, n)
    rng = np.random.default_rng(seed)
    items = []
    for layer in range(5):
        for q in range(n):
            items.append(_Item("diag1", (q,), np.exp(1j * rng.uniform(0, 6, 2))))
            items.append(_Item("dense1", (q,), haar_unitary(2, rng)))
        items.append(_Item("cz", (layer % n, (layer + 1) % n)))
    t = _template_matches_plan(items, n)
    # the same kinds and bits with other data: accepted; other bits: refused
    other = [_Item(it.kind, it.bits, None if it.data is None else np.asarray(it.data) * 1j)
             for it in items]
    t.check(other)
    other[0] = _Item("diag1", ((other[0].bits[0] + 1) % n,), other[0].data)
    with pytest.raises(ValueError, match="differ from the template"):
        t.check(other)
