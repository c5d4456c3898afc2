"""The amplitude kernel's per-term nibble tables have two builders: k_qform
writes them with the quadratic forms in one launch (one word per row, when
amplitudes are the first use of freshly computed forms), k_amp_tables
writes them afterwards (forms already computed for another call, and every
wider row).  Both must give the same amplitudes bit for bit, including terms
with more than eight constraints (support tested through the C^T tables),
and the amplitudes must match the oracle (amplitude.rs:28-44) on small
states.
"""
import sys

import numpy as np
import pytest

import fp64
import workloads as WL
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nx():
    import necstar
    from necstar import _lib
    _lib.check(_lib.lib().nx_init(0))
    return necstar


FEW_H = (WL.S, WL.CX, WL.CZ, WL.S, WL.CX, WL.CZ, WL.X, WL.S, WL.CX, WL.CZ, WL.H)

CASES = [  # (seed, n, cliffords, t, gate set)
    (1, 12, 150, 6, WL.CLIFFORD_SET),
    (2, 40, 400, 9, WL.CLIFFORD_SET),
    (3, 64, 2000, 12, WL.CLIFFORD_SET),
    (4, 50, 160, 8, FEW_H),     # most terms with more than eight constraints
    (5, 64, 220, 10, FEW_H),
    (6, 100, 600, 8, WL.CLIFFORD_SET),   # two words: k_amp_tables in both orders
]


@pytest.mark.parametrize("seed,n,nc,nt,gs", CASES)
def test_table_builders_agree(nx, seed, n, nc, nt, gs):
    from necstar import _lib
    L = _lib.lib()
    rng = np.random.default_rng(8800 + seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=gs)
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    bits = WL.random_bitstrings(rng, n, 1500)
    say(f"tables seed={seed} n={n} chi={st.stabilizer_rank}")
    _lib.check(L.nx_state_drop_cache(st._h))
    first = st.amplitudes(bits)          # forms and tables in one launch (n <= 64)
    _lib.check(L.nx_state_drop_cache(st._h))
    norm = st.norm()                     # forms alone
    later = st.amplitudes(bits)          # tables by k_amp_tables
    assert np.array_equal(first, later)
    assert st.norm() == norm
    if n <= 12:
        o = O.State.compile(n, gates)
        scale = max(float(np.max(np.abs(first))), 1e-300)
        for b in range(64):
            x = np.array([(int(bits[b, 0]) >> q) & 1 for q in range(n)], np.uint8)
            want = o.amplitude(x)
            fp64.check(first[b].real, want.real, scale, f"re b={b}")
            fp64.check(first[b].imag, want.imag, scale, f"im b={b}")
