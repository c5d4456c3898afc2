"""Seeded synthetic workloads (numpy PCG64, seed = 1000 + config id).

  cfg0  n=12, 200 Cliffords from {H, S, CX, CZ}, 8 T; 256 bitstrings;
        16 Paulis of weight 1-4
  cfg1  n=64, 2000 Cliffords, 24 T; 65536 bitstrings; full norm;
        Z-basis samples on all qubits
  cfg2  256 circuits, n in [16, 128], 30 n Cliffords, t in [4, 32];
        Z-basis samples and 64 Paulis of weight 2-6 per circuit
  cfg3  n=48, 1500 Cliffords, 30 T; 128 Paulis of weight 2-6 with N(0,1)
        weights; coefficients perturbed by N(0, 1e-3)
  cfg4  n=128, 6000 Cliffords, 40 T; 2^16 bitstrings of two words

Gates are (op, q0[, q1]) tuples with the op codes of include/necstar_hip.h.
"""
from __future__ import annotations

import numpy as np

H, X, Y, Z, S, SDG, SQRTX, SQRTXDG, CX, CZ, SWAP, T, TDG, CCX = range(14)
CLIFFORD_SET = (H, S, CX, CZ)


def random_circuit(rng, n, n_clifford, n_t, gate_set=CLIFFORD_SET):
    gates = []
    for _ in range(n_clifford):
        op = gate_set[int(rng.integers(len(gate_set)))]
        if op in (CX, CZ, SWAP):
            a, b = rng.choice(n, 2, replace=False)
            gates.append((int(op), int(a), int(b)))
        else:
            gates.append((int(op), int(rng.integers(n))))
    if n_t:
        slots = sorted(int(s) for s in rng.choice(n_clifford + n_t, n_t, replace=False))
        for s in slots:
            gates.insert(s, (T, int(rng.integers(n))))
    return gates


def random_paulis(rng, n, count, wmin, wmax):
    """Dense little-endian Pauli op lists (0=I, 1=X, 2=Y, 3=Z)."""
    out = []
    for _ in range(count):
        w = int(rng.integers(wmin, wmax + 1))
        ops = [0] * n
        for q in rng.choice(n, w, replace=False):
            ops[int(q)] = int(rng.integers(1, 4))
        out.append(ops)
    return out


def random_bitstrings(rng, n, count):
    words = max(1, (n + 63) // 64)
    bits = rng.integers(0, 2 ** 63, size=(count, words), dtype=np.uint64) * np.uint64(2)
    bits ^= rng.integers(0, 2, size=(count, words), dtype=np.uint64)
    for w in range(words):
        hi = n - 64 * w
        if hi < 64:
            bits[:, w] &= np.uint64((1 << max(hi, 0)) - 1)
    return bits


def config(cid: int):
    rng = np.random.default_rng(1000 + cid)
    if cid == 0:
        n = 12
        return dict(name="cfg0", n=n, gates=random_circuit(rng, n, 200, 8),
                    bitstrings=random_bitstrings(rng, n, 256), paulis=random_paulis(rng, n, 16, 1, 4))
    if cid == 1:
        n = 64
        return dict(name="cfg1", n=n, gates=random_circuit(rng, n, 2000, 24),
                    bitstrings=random_bitstrings(rng, n, 65536), shots=1024, sample_seed=1001)
    if cid == 2:
        circuits = []
        for _ in range(256):
            n = int(rng.integers(16, 129))
            t = int(rng.integers(4, 33))
            circuits.append(dict(n=n, gates=random_circuit(rng, n, 30 * n, t),
                                 paulis=random_paulis(rng, n, 64, 2, 6)))
        return dict(name="cfg2", circuits=circuits, shots=1024)
    if cid == 3:
        n = 48
        return dict(name="cfg3", n=n, gates=random_circuit(rng, n, 1500, 30),
                    paulis=random_paulis(rng, n, 128, 2, 6), weights=rng.standard_normal(128),
                    perturb_seed=2003)
    if cid == 4:
        n = 128
        return dict(name="cfg4", n=n, gates=random_circuit(rng, n, 6000, 40),
                    bitstrings=random_bitstrings(rng, n, 1 << 16))
    raise ValueError(cid)


def to_circuit(n, gates):
    """necstar.QuantumCircuit from gate tuples."""
    from necstar import QuantumCircuit
    from necstar.gate import QuantumGate
    qc = QuantumCircuit(n)
    qc._gates = [QuantumGate(g[0], g[1:]) for g in gates]
    return qc
