"""CPU reference for the adjoint tests: the hardware-efficient circuit run
gate by gate with the oracle (complex128), its loss, and exact
parameter-shift derivatives (exact for exp(-i t/2 P) rotations)."""
import numpy as np

from oracle import closed_form


def ry(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)


def run(theta, phi, n):
    """theta, phi: (L, n) for one state -> final state (2^n,)."""
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1
    L = theta.shape[0]
    pairs = [(q, (q + 1) % n) for q in range(n)] if n > 2 else [(0, 1)]
    for l in range(L):
        for q in range(n):
            closed_form.apply_dense(rz(phi[l, q]), [q], psi, out=psi)
        for q in range(n):
            closed_form.apply_dense(ry(theta[l, q]), [q], psi, out=psi)
        for a, b in pairs:
            closed_form.apply_dense(CZ, [a, b], psi, out=psi)
    return psi


def loss(theta, phi, n):
    psi = run(theta, phi, n)
    x = np.arange(2 ** n)
    pop = np.array([bin(v).count("1") for v in x])
    return float(np.sum(np.abs(psi) ** 2 * (n - 2 * pop)))


def param_shift(theta, phi, n, which, l, q):
    th, ph = theta.copy(), phi.copy()
    arr = th if which == "theta" else ph
    arr[l, q] += np.pi / 2
    plus = loss(th, ph, n)
    arr[l, q] -= np.pi
    minus = loss(th, ph, n)
    return 0.5 * (plus - minus)


# ---- general rotation circuits (unitair_amd.adjoint.expectation's gate list)

def exp_pauli(kind, t):
    """unitair's exp_x / exp_y / exp_z(t) = exp(-i t P) (gates.py:7-91)."""
    c, s = np.cos(t), np.sin(t)
    if kind == "x":
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)
    if kind == "y":
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    return np.diag([c - 1j * s, c + 1j * s])


CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)


def run_gates(params, gates, n, psi0=None):
    """One parameter row -> final state, gate by gate with the oracle."""
    if psi0 is None:
        psi = np.zeros(2 ** n, dtype=np.complex128)
        psi[0] = 1
    else:
        psi = np.array(psi0, dtype=np.complex128)
    for g in gates:
        if g[0] in ("x", "y", "z"):
            coef = g[3] if len(g) > 3 else 1.0
            closed_form.apply_dense(exp_pauli(g[0], coef * params[g[2]]), [g[1]], psi, out=psi)
        elif g[0] == "cz":
            closed_form.apply_dense(CZ, [g[1], g[2]], psi, out=psi)
        elif g[0] == "cnot":
            closed_form.apply_dense(CNOT, [g[1], g[2]], psi, out=psi)
        else:
            closed_form.apply_dense(np.asarray(g[2]), list(g[1]), psi, out=psi)
    return psi


def general_loss(params, gates, obs, n, psi0=None):
    psi = run_gates(params, gates, n, psi0)
    return float(np.sum(np.abs(psi) ** 2 * obs))


def general_param_shift(params, gates, obs, n, idx, psi0=None):
    """d loss / d params[idx], exact: every gate using params[idx] is
    exp(-i coef t P), so each contributes coef * (L(t_g + pi/4) - L(t_g - pi/4))
    with only that gate's angle shifted (the shift rule for frequency 2)."""
    total = 0.0
    for gi, g in enumerate(gates):
        if g[0] in ("x", "y", "z") and g[2] == idx:
            coef = g[3] if len(g) > 3 else 1.0
            shifted = []
            for delta in (np.pi / 4, -np.pi / 4):
                gs = list(gates)
                # the shifted gate reads a private extra parameter
                gs[gi] = (g[0], g[1], len(params), 1.0)
                p2 = np.append(params, coef * params[idx] + delta)
                shifted.append(general_loss(p2, gs, obs, n, psi0))
            total += coef * (shifted[0] - shifted[1])
    return total


def random_rotation_circuit(rng, n, num_gates, num_params):
    """Rotations on random qubits with random coefficients (some parameters
    shared by several gates), CZ, CNOT and fixed Hadamards."""
    H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
    gates = []
    for _ in range(num_gates):
        r = rng.random()
        if r < 0.7:
            gates.append((str(rng.choice(["x", "y", "z"])), int(rng.integers(n)),
                          int(rng.integers(num_params)), float(rng.uniform(-1.5, 1.5))))
        elif r < 0.8:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(("cz", int(a), int(b)))
        elif r < 0.9:
            a, b = rng.choice(n, 2, replace=False)
            gates.append(("cnot", int(a), int(b)))
        else:
            gates.append(("dense", (int(rng.integers(n)),), H))
    return gates
