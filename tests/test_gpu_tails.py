"""Exact leaf norms of sampling tails (nx_state_completion_norms): every
term restricted to the d completion qubits (tails.hip) must give, for each
completion x = base | y, A + sqrt2 B = |<x|psi>|^2 / |g|^2 as the amplitude
kernel evaluates it (FP64, 1e-12 of the scale) and, on small states, as the
oracle's amplitude (amplitude.rs:28-44) gives it.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import fp64
import workloads as WL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SQRT2 = 1.4142135623730951


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nx():
    import necstar
    from necstar import _lib
    _lib.check(_lib.lib().nx_init(0))
    return necstar


def completion_norms(st, base_words, qs):
    from necstar import _lib
    d = len(qs)
    A = np.zeros(1 << d)
    B = np.zeros(1 << d)
    base = np.ascontiguousarray(np.asarray(base_words, np.uint64))
    q = np.ascontiguousarray(np.asarray(qs, np.int32))
    _lib.check(_lib.lib().nx_state_completion_norms(
        st._h, base.ctypes.data_as(C.POINTER(C.c_uint64)), d, q.ctypes.data_as(C.POINTER(C.c_int32)),
        A.ctypes.data_as(C.POINTER(C.c_double)), B.ctypes.data_as(C.POINTER(C.c_double))))
    return A, B


def completions(base_words, qs, W):
    d = len(qs)
    out = np.zeros((1 << d, W), np.uint64)
    for y in range(1 << d):
        x = [int(w) for w in base_words]
        for i, q in enumerate(qs):
            if (y >> i) & 1:
                x[q >> 6] |= 1 << (q & 63)
        out[y] = x
    return out


CASES = [  # (seed, n, cliffords, t, d)
    (1, 6, 60, 6, 3),
    (2, 12, 200, 8, 7),
    (3, 20, 300, 10, 14),
    (4, 64, 2000, 12, 14),
    (5, 70, 700, 10, 10),
    (6, 98, 2940, 8, 14),     # a cfg2-like circuit: 30 n Cliffords
    (7, 128, 3840, 6, 16),
    (8, 113, 3390, 14, 14),
    (9, 40, 1200, 4, 1),
]


@pytest.mark.parametrize("seed,n,nc,nt,d", CASES)
def test_completion_norms_match_amplitudes(nx, seed, n, nc, nt, d):
    rng = np.random.default_rng(7700 + seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=tuple(range(11)))
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    W = max(1, (n + 63) // 64)
    g2 = abs(st.global_factor) ** 2
    for trial in range(3):
        qs = [int(q) for q in rng.choice(n, d, replace=False)]
        if trial == 0:  # a sampling tail: the lowest qubits, highest first
            qs = list(range(d - 1, -1, -1))
        base = [0] * W
        for q in range(n):
            if q not in qs and rng.integers(2):
                base[q >> 6] |= 1 << (q & 63)
        say(f"tail seed={seed} n={n} chi={st.stabilizer_rank} d={d} trial={trial} qs={qs[:6]}...")
        A, B = completion_norms(st, base, qs)
        amps = st.amplitudes(completions(base, qs, W))
        want = np.abs(amps) ** 2 / g2
        got = A + SQRT2 * B
        scale = max(float(np.max(want)), 1e-300)
        bad = np.abs(got - want) > 1e-12 * scale
        assert not np.any(bad), (np.nonzero(bad)[0][:5], got[bad][:5], want[bad][:5])
        if n <= 12:
            o = O.State.compile(n, gates)
            for y in range(1 << d):
                bits = np.array([(int(completions(base, qs, W)[y, 0]) >> q) & 1 for q in range(n)], np.uint8)
                fp64.check(got[y], abs(o.amplitude(bits)) ** 2 / g2, scale, f"tail oracle y={y}")
