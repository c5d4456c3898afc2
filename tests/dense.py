"""Dense state-vector simulator (numpy) used as an independent check.

Little-endian like the reference (qubit j is bit j of the basis index), so
its vectors compare directly with ``to_statevector`` of the reference
(crates/necstar/src/state/stabilizer_decomposed_state/statevector.rs:12-24).
It knows nothing about CH-forms: it fixes what the mathematics fixes.
"""
from __future__ import annotations

import numpy as np

H, X, Y, Z, S, SDG, SQRTX, SQRTXDG, CX, CZ, SWAP, T, TDG, CCX = range(14)

_s2 = 1 / np.sqrt(2)
ONE_QUBIT = {
    H: np.array([[_s2, _s2], [_s2, -_s2]], complex),
    X: np.array([[0, 1], [1, 0]], complex),
    Y: np.array([[0, -1j], [1j, 0]], complex),
    Z: np.array([[1, 0], [0, -1]], complex),
    S: np.array([[1, 0], [0, 1j]], complex),
    SDG: np.array([[1, 0], [0, -1j]], complex),
    SQRTX: 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]], complex),
    SQRTXDG: 0.5 * np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]], complex),
    T: np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], complex),
    TDG: np.array([[1, 0], [0, np.exp(-1j * np.pi / 4)]], complex),
}


def zero_state(n):
    sv = np.zeros(1 << n, complex)
    sv[0] = 1.0
    return sv


def _apply_1q(sv, n, U, q):
    t = sv.reshape((2,) * n)
    ax = n - 1 - q
    t = np.moveaxis(np.tensordot(U, t, axes=([1], [ax])), 0, ax)
    return t.reshape(-1)


def apply_gate(sv, n, g):
    op = g[0]
    if op in ONE_QUBIT:
        return _apply_1q(sv, n, ONE_QUBIT[op], g[1])
    idx = np.arange(1 << n)
    if op == CX:
        c, t = g[1], g[2]
        src = np.where((idx >> c) & 1, idx ^ (1 << t), idx)
        return sv[src]
    if op == CZ:
        a, b = g[1], g[2]
        sign = np.where(((idx >> a) & 1) & ((idx >> b) & 1), -1.0, 1.0)
        return sv * sign
    if op == SWAP:
        a, b = g[1], g[2]
        ba, bb = (idx >> a) & 1, (idx >> b) & 1
        src = idx ^ ((ba ^ bb) << a) ^ ((ba ^ bb) << b)
        return sv[src]
    if op == CCX:
        c1, c2, t = g[1], g[2], g[3]
        src = np.where(((idx >> c1) & 1) & ((idx >> c2) & 1), idx ^ (1 << t), idx)
        return sv[src]
    raise ValueError(op)


def run(n, gates):
    sv = zero_state(n)
    for g in gates:
        sv = apply_gate(sv, n, g)
    return sv


PAULI_1Q = {1: ONE_QUBIT[X], 2: ONE_QUBIT[Y], 3: ONE_QUBIT[Z]}


def apply_pauli(sv, n, pauli):
    kind, data = pauli
    items = list(enumerate(data)) if kind == "dense" else [(q, o) for o, q in data]
    for q, o in items:
        if o:
            sv = _apply_1q(sv, n, PAULI_1Q[o], q)
    return sv


def random_circuit(rng, n, n_clifford, n_t, gate_set=(H, S, CX, CZ), tdg=False):
    """Seeded random Clifford+T circuit: uniform gates from gate_set on
    uniform qubits (distinct pairs), T gates at uniformly random slots."""
    gates = []
    for _ in range(n_clifford):
        op = gate_set[rng.integers(len(gate_set))]
        if op in (CX, CZ, SWAP):
            a, b = rng.choice(n, 2, replace=False)
            gates.append((int(op), int(a), int(b)))
        else:
            gates.append((int(op), int(rng.integers(n))))
    if n_t:
        slots = sorted(rng.choice(n_clifford + n_t, n_t, replace=False))
        for s in slots:
            op = TDG if (tdg and rng.integers(2)) else T
            gates.insert(int(s), (int(op), int(rng.integers(n))))
    return gates
