"""Random Clifford circuits, restated in Python (TEST INFRASTRUCTURE ONLY).

Restates crates/stabilizer-ch-form-rust/src/circuit/random_clifford.rs:7-218
(Bravyi-Maslov sampling, arXiv:2003.09412v2) with the random draws behind
two callables, so that tests can feed the reference's StdRng stream (Rng of
oracle.py) or enumerate every outcome of the draws:

    draw_f64()  -- rand 0.8 ``rng.gen::<f64>()``
    draw_bit()  -- rand 0.8 ``rng.gen_range(0..=1)`` for u8: UniformInt<u8>
                   widens to u32 and draws one next_u32; for range 2 the
                   rejection zone is the whole u32 range, so the result is
                   (v * 2) >> 32 = v >> 31.

Only tests/ may import this module.  Gates are (code, q0, q1) in the gate
numbering of oracle.py.
"""
from __future__ import annotations

import math

from . import oracle as O


def _f64_to_isize(x: float) -> int:
    """Rust's saturating ``f64 as isize``."""
    if math.isnan(x):
        return 0
    if math.isinf(x):
        return (1 << 63) - 1 if x > 0 else -(1 << 63)
    return max(-(1 << 63), min((1 << 63) - 1, int(x)))


def mallows_a(m: int, r: float) -> int:
    """random_clifford.rs:18-27: the index a in [1, 2m] drawn from r."""
    four_m = math.inf if m >= 512 else float(4 ** m)  # 4.0f64.powi(m)
    val = r * (four_m - 1.0) + 1.0
    if math.isnan(val):
        cl = 0
    elif math.isinf(val):
        cl = _f64_to_isize(math.inf)
    else:
        cl = _f64_to_isize(float(math.ceil(math.log2(val))))
    a = (2 * m + 1) - cl
    return 1 if a <= 0 else a


def sample_quantum_mallows(n, draw_f64):
    """random_clifford.rs:7-46"""
    h, s = [0] * n, [0] * n
    available = list(range(n))
    for i in range(n):
        m = n - i
        a = mallows_a(m, draw_f64())
        if a <= m:
            h[i], k1 = 1, a
        else:
            h[i], k1 = 0, 2 * m - a + 1
        k0 = k1 - 1
        if 0 <= k0 < len(available):
            s[i] = available.pop(k0)
        else:  # usize wrap of 0 - 1, or a boundary rounding: the last index
            s[i] = available.pop()
    return h, s


def clifford_params(n, draw_f64, draw_bit):
    """random_clifford.rs:63-119"""
    h, s = sample_quantum_mallows(n, draw_f64)
    g1 = [[0] * n for _ in range(n)]
    g2 = [[0] * n for _ in range(n)]
    d1 = [[int(i == j) for j in range(n)] for i in range(n)]
    d2 = [[int(i == j) for j in range(n)] for i in range(n)]
    pz = [draw_bit() for _ in range(n)]
    px = [draw_bit() for _ in range(n)]
    for i in range(n):
        g2[i][i] = draw_bit()
        if h[i]:
            g1[i][i] = draw_bit()
    for i in range(1, n):
        for j in range(i):
            b = draw_bit()
            g2[i][j] = g2[j][i] = b
            d2[i][j] = draw_bit()
            hi, hj, si, sj = h[i] == 1, h[j] == 1, s[i], s[j]
            if ((si < sj or hj) and hi) or (hj and sj < si):
                b = draw_bit()
                g1[i][j] = g1[j][i] = b
            if (si < sj or hj) and (si > sj or not hi) and (hj or not hi):
                d1[i][j] = draw_bit()
    return dict(h=h, s=s, pz=pz, px=px, gamma1=g1, delta1=d1, gamma2=g2, delta2=d2)


def _hadamard_free(gates, n, gamma, delta, pz=None, px=None):
    """random_clifford.rs:122-165"""
    for j in range(n):
        for i in range(j + 1, n):
            if delta[i][j]:
                gates.append((O.CX, j, i))
    for i in range(n):
        if gamma[i][i]:
            gates.append((O.S, i, 0))
        for j in range(i):
            if gamma[i][j]:
                gates.append((O.CZ, i, j))
    if pz is not None:
        for i in range(n):
            if pz[i] and px[i]:
                gates.append((O.Y, i, 0))
            elif pz[i]:
                gates.append((O.Z, i, 0))
            elif px[i]:
                gates.append((O.X, i, 0))


def _permutation(gates, s):
    """random_clifford.rs:168-182"""
    p = list(range(len(s)))
    for i, target in enumerate(s):
        if p[i] != target:
            j = p.index(target)
            gates.append((O.SWAP, i, j))
            p[i], p[j] = p[j], p[i]


def circuit_from_params(n, prm):
    """random_clifford.rs:189-218: F2 (with Paulis), permutation, H, F1."""
    gates = []
    _hadamard_free(gates, n, prm["gamma2"], prm["delta2"], prm["pz"], prm["px"])
    _permutation(gates, prm["s"])
    gates += [(O.H, i, 0) for i in range(n) if prm["h"][i]]
    _hadamard_free(gates, n, prm["gamma1"], prm["delta1"])
    return gates


def random_clifford(n, seed):
    """The reference's gate list for (n, seed) (seed as in oracle.seed_bytes)."""
    if n == 0:
        return []
    rng = O.Rng(seed)
    prm = clifford_params(n, rng.gen_f64, lambda: rng.next_u32() >> 31)
    return circuit_from_params(n, prm)
