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
