"""Fused passes (ua_run_fused_pass via unitair_amd.fusion) against the oracle."""
import numpy as np
import pytest
import torch

from unitair_amd.fusion import FusedProgram
from unitair_amd.gates import haar_unitary
from unitair_amd.initializations import seeded_state_np
from oracle import circuits as oc
from oracle import unitair_np as U
from test_fusion_plan import random_mixed_circuit

pytestmark = pytest.mark.gpu
TOL = {torch.complex64: 1e-5, torch.complex128: 1e-12}


def rel_err(out, ref):
    return float(np.max(np.abs(out - ref)) / max(1e-300, np.max(np.abs(ref))))


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,batch", [(4, 3), (5, 1), (8, 2), (11, 1), (12, 3), (13, 1), (16, 2), (20, 1)])
def test_fused_random_circuits(device, dt, n, batch):
    circ = random_mixed_circuit(n, 80, seed=n * 7 + batch)
    prog = FusedProgram(circ, dtype=dt, device=device)
    psi = seeded_state_np(n, (batch,), seed=n)
    ref = np.stack([oc.run_closed_form(circ.gates, p) for p in psi])
    x = torch.from_numpy(psi).to(device, dt)
    out = prog.run(x).cpu().numpy()
    assert rel_err(out, ref) < TOL[dt] * 10, (prog.num_passes, n)
    assert np.max(U.fidelity_error(ref, out)) < TOL[dt]


def test_fused_layered_12q_batch8_c128(device):
    from unitair_amd.workloads import layered_circuit_12q
    circ = layered_circuit_12q(seed=1)
    psi = seeded_state_np(12, (8,), seed=11)
    ref = np.stack([oc.run_closed_form(circ.gates, p) for p in psi])
    prog = FusedProgram(circ, dtype=torch.complex128, device=device)
    out = prog.run(torch.from_numpy(psi).to(device)).cpu().numpy()
    assert rel_err(out, ref) < 1e-12


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_fused_random_circuit_24q(device, dt):
    from test_gpu_parity import _circuit_24q_and_oracle
    circ, psi, ref = _circuit_24q_and_oracle()
    prog = FusedProgram(circ, dtype=dt, device=device)
    out = prog.run(torch.from_numpy(psi).to(device, dt)).cpu().numpy()
    fid = float(U.fidelity_error(ref, out))
    mae = U.max_abs_error(ref, out)
    print(f"fused 24q {dt}: {prog.num_passes} passes, fidelity error {fid:.3e}, max-abs {mae:.3e}")
    # achieved: complex64 ~3e-13 / ~6e-10, complex128 ~1e-16 / ~1e-17
    bar = {torch.complex64: (1e-10, 1e-7), torch.complex128: (1e-14, 1e-14)}[dt]
    assert fid < bar[0] and mae < bar[1], (fid, mae)


def test_fused_graph_replay_matches_eager(device):
    n = 14
    circ = random_mixed_circuit(n, 60, seed=5)
    prog = FusedProgram(circ, dtype=torch.complex64, device=device)
    psi = torch.from_numpy(seeded_state_np(n, (2,), seed=3)).to(device, torch.complex64)
    eager = prog.run(psi.clone())
    state = psi.clone()
    prog.capture(state)
    for _ in range(3):
        state.copy_(psi)                 # eager work between replays
        torch.cuda.synchronize()
        prog.replay()
        torch.cuda.synchronize()
        assert torch.equal(state, eager)


# ---- compiled passes (ua_compile_fused_pass / ua_run_compiled_pass): the same
# circuits, every pass with 12 tile bits run by its own compiled kernel

@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,batch", [(8, 2), (12, 3), (13, 1), (16, 2), (20, 1)])
def test_compiled_random_circuits(device, dt, n, batch):
    """Under 12 qubits the tiles are smaller and the passes stay interpreted."""
    circ = random_mixed_circuit(n, 80, seed=n * 7 + batch)
    prog = FusedProgram(circ, dtype=dt, device=device, compiled=True)
    assert (prog.num_compiled > 0) == (n >= 12)
    psi = seeded_state_np(n, (batch,), seed=n)
    ref = np.stack([oc.run_closed_form(circ.gates, p) for p in psi])
    out = prog.run(torch.from_numpy(psi).to(device, dt)).cpu().numpy()
    assert rel_err(out, ref) < TOL[dt] * 10, (prog.num_passes, n)
    assert np.max(U.fidelity_error(ref, out)) < TOL[dt]


@pytest.mark.parametrize("dt,phase_angles", [(torch.complex64, False),
                                             (torch.complex128, False),
                                             (torch.complex64, True)])
def test_compiled_random_circuit_24q(device, dt, phase_angles):
    from test_gpu_parity import _circuit_24q_and_oracle
    circ, psi, ref = _circuit_24q_and_oracle()
    prog = FusedProgram(circ, dtype=dt, device=device, compiled=True,
                        phase_angles=phase_angles)
    assert prog.num_compiled == prog.num_passes
    out = prog.run(torch.from_numpy(psi).to(device, dt)).cpu().numpy()
    fid = float(U.fidelity_error(ref, out))
    mae = U.max_abs_error(ref, out)
    print(f"compiled 24q {dt} (phase angles: {phase_angles}): {prog.num_passes} passes, "
          f"fidelity error {fid:.3e}, max-abs {mae:.3e}")
    bar = {torch.complex64: (1e-10, 1e-7), torch.complex128: (1e-14, 1e-14)}[dt]
    assert fid < bar[0] and mae < bar[1], (fid, mae)


def _per_state_items(n, batch, seed, device, dt):
    """Gate items of a random mixed circuit plus per-state 1-qubit matrices,
    diagonals and CZ-layer masks (the forms the adjoint and serving loops
    build), on state bits inside and outside the tile."""
    from unitair_amd.fusion import _Item, _items_from_circuit, _merge_1q
    rng = np.random.default_rng(seed)
    items = _merge_1q(_items_from_circuit(random_mixed_circuit(n, 50, seed=seed)))
    items = [it for it in items if it.kind != "wide"]
    for _ in range(40):
        r = rng.random()
        b = int(rng.integers(n))
        if r < 0.4:
            m = np.stack([haar_unitary(2, rng) for _ in range(batch)])
            items.append(_Item("dense1", (b,), torch.from_numpy(m).to(device, dt), True))
        elif r < 0.7:
            d = np.exp(1j * rng.uniform(0, 2 * np.pi, size=(batch, 2)))
            items.append(_Item("diag1", (b,), torch.from_numpy(d).to(device, dt), True))
        elif r < 0.85:
            masks = rng.integers(0, 2 ** (n - 1), size=batch, dtype=np.int64)
            items.append(_Item("cz_layer", (), torch.from_numpy(masks).to(device), True))
        else:
            a, c = rng.choice(n, 2, replace=False)
            items.append(_Item("cnot" if r < 0.93 else "cz", (int(a), int(c))))
    return items


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,batch,reorder", [(12, 4, True), (14, 3, True), (14, 2, False),
                                             (17, 2, True)])
def test_compiled_matches_interpreter_per_state_data(device, dt, n, batch, reorder):
    """Compiled passes against the interpreter (pinned against the oracle by
    the tests above and by the serving and adjoint tests) on the op kinds and
    per-state data those loops use."""
    items = _per_state_items(n, batch, seed=n * 3 + batch, device=device, dt=dt)
    interp = FusedProgram.from_items(items, n, dt, device, reorder=reorder)
    comp = FusedProgram.from_items(items, n, dt, device, reorder=reorder, compiled=True)
    assert comp.num_compiled == comp.num_passes
    psi = torch.from_numpy(seeded_state_np(n, (batch,), seed=5)).to(device, dt)
    a = interp.run(psi.clone()).cpu().numpy()
    c = comp.run(psi.clone()).cpu().numpy()
    assert rel_err(c, a) < {torch.complex64: 1e-5, torch.complex128: 1e-13}[dt]


def test_compiled_graph_replay_matches_eager(device):
    """Captured before any eager run: the program loaded its kernels when it
    was built, so the capture records launches only."""
    n = 14
    circ = random_mixed_circuit(n, 61, seed=5)    # a structure no other test compiles
    prog = FusedProgram(circ, dtype=torch.complex64, device=device, compiled=True)
    assert prog.num_compiled > 0
    psi = torch.from_numpy(seeded_state_np(n, (2,), seed=3)).to(device, torch.complex64)
    state = psi.clone()
    prog.capture(state)
    eager = prog.run(psi.clone())
    for _ in range(3):
        state.copy_(psi)
        torch.cuda.synchronize()
        prog.replay()
        torch.cuda.synchronize()
        assert torch.equal(state, eager)


def test_compiled_phase_angles_accuracy(device):
    """A compiled complex64 pass that reads phase angles forms exp(-1j angle)
    itself (reduction by pi/2 and polynomials up to |angle| 256, sincosf
    beyond): against float64 exp on angles in [-300, 300] and special values."""
    from unitair_amd.circuits import Circuit, Gate
    n = 12
    rng = np.random.default_rng(0)
    ang = rng.uniform(-300, 300, 2 ** n).astype(np.float32)
    ang[:10] = [0.0, np.pi / 2, -np.pi / 2, np.pi, 2 * np.pi, 256.0, -256.0, 256.5, 1e-30,
                -3 * np.pi / 4]
    ang[10:2000] = rng.uniform(0, 2 * np.pi, 1990).astype(np.float32)
    circ = Circuit(n, [Gate("phase", (), angles=ang)])
    prog = FusedProgram(circ, torch.complex64, device, compiled=True, phase_angles=True)
    assert prog.num_compiled == prog.num_passes == 1
    out = prog.run(torch.ones(1, 2 ** n, dtype=torch.complex64, device=device))
    err = np.abs(out.cpu().numpy()[0] - np.exp(-1j * ang.astype(np.float64)))
    assert np.max(err) < 3e-7, (int(np.argmax(err)), float(np.max(err)))


@pytest.mark.parametrize("n,tile_bits", [(16, 12), (16, 14), (17, 13)])
def test_compiled_tile_sizes(device, n, tile_bits):
    """Compiled passes on 12-, 13- and 14-bit complex64 tiles against the oracle."""
    circ = random_mixed_circuit(n, 90, seed=n + tile_bits)
    prog = FusedProgram(circ, dtype=torch.complex64, device=device, compiled=True,
                        tile_bits=tile_bits)
    fused = [st for st in prog.steps if st[0] == "fused"]
    assert fused and all(st[1].tile_bits == tile_bits for st in fused)
    assert prog.num_compiled == len(fused)
    psi = seeded_state_np(n, (2,), seed=9)
    ref = np.stack([oc.run_closed_form(circ.gates, p) for p in psi])
    out = prog.run(torch.from_numpy(psi).to(device, torch.complex64)).cpu().numpy()
    assert rel_err(out, ref) < 1e-4
    assert np.max(U.fidelity_error(ref, out)) < 1e-5
