"""Batched sampling (nx_batch_sample_exact, necstar.sample_many): many states
sampled in one call, each state's windows evaluated together with the other
states' in shared launches.

Every state's outcomes must equal what QuantumState.sample(..., exact=True)
draws for it alone with the same seed (same RNG order, same exact node
probabilities), bit for bit; on Clifford states (probabilities 0, 1/2^k, 1)
both equal the oracle's seeded outcomes (sampling.rs:80-189).
"""
import ctypes as C
import sys

import numpy as np
import pytest

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


def _state(nx, n, gates):
    return nx.QuantumState.from_circuit(WL.to_circuit(n, gates))


def _oracle_counts(o, qargs, shots, seed):
    return {"".join("1" if b else "0" for b in bits): c for bits, c in o.sample(qargs, shots, seed)}


def test_batch_equals_single_calls(nx):
    """Clifford+T states of 5..100 qubits (one and two row words, chi 1..~200),
    every qubit or a subset, 1..1024 shots: each state's batched outcomes are
    its own sample(exact=True) outcomes."""
    rng = np.random.default_rng(8101)
    specs = [(5, 6, None, 700), (12, 8, None, 1024), (30, 10, 17, 300), (64, 8, None, 200), (70, 9, None, 64),
             (100, 6, 40, 1), (16, 12, None, 513), (40, 0, None, 128), (9, 14, 5, 2000), (64, 4, None, 1024)]
    states, qargs, shots, seeds = [], [], [], []
    for i, (n, t, nq, sh) in enumerate(specs):
        gates = WL.random_circuit(rng, n, 10 * n, t, gate_set=tuple(range(11)))
        st = _state(nx, n, gates)
        qa = list(range(n)) if nq is None else [int(q) for q in rng.choice(n, nq, replace=False)]
        states.append(st)
        qargs.append(qa)
        shots.append(sh)
        seeds.append(9000 + i)
        say(f"state {i}: n={n} t={t} chi={st.stabilizer_rank} nq={len(qa)} shots={sh}")
    stats = {}
    got = nx.sample_many(states, qargs, shots, seeds, stats=stats)
    say(f"batch stats {stats}")
    for i, st in enumerate(states):
        want = st.sample(qargs[i], shots[i], seed=seeds[i], exact=True)
        assert got[i] == want, i
        assert sum(got[i].values()) == shots[i]
    assert stats["windows"] > 0 and stats["rounds"] > 0
    # the same call again: the same outcomes (no state carried between calls)
    assert nx.sample_many(states, qargs, shots, seeds) == got


def test_batch_many_small_tails(nx):
    """Many states of 3..9 qubits (every qubit sampled): their exact tails
    have fewer than 64 completions, so one launch's wavefronts straddle
    several tails; with larger states beside them."""
    rng = np.random.default_rng(8404)
    states, qargs, shots, seeds = [], [], [], []
    for i in range(40):
        n = int(rng.integers(3, 10)) if i % 8 else int(rng.integers(20, 70))
        gates = WL.random_circuit(rng, n, 10 * n, int(rng.integers(0, 9)), gate_set=tuple(range(11)))
        states.append(_state(nx, n, gates))
        qargs.append(list(range(n)))
        shots.append(int(rng.integers(1, 400)))
        seeds.append(6100 + i)
    got = nx.sample_many(states, qargs, shots, seeds)
    for i, st in enumerate(states):
        assert got[i] == st.sample(qargs[i], shots[i], seed=seeds[i], exact=True), i


def test_batch_clifford_states_equal_oracle(nx):
    """Clifford states: the batched outcomes equal the oracle's seeded
    outcomes (the reference's FP64 sums are exact on dyadic probabilities)."""
    rng = np.random.default_rng(8202)
    cases = [(14, 14, 2000), (20, 13, 300), (70, 10, 64), (66, 66, 40), (8, 8, 1000)]
    states, os_, qargs, shots, seeds = [], [], [], [], []
    for i, (n, nq, sh) in enumerate(cases):
        gates = WL.random_circuit(rng, n, 10 * n, 0, gate_set=tuple(range(11)))
        states.append(_state(nx, n, gates))
        os_.append(O.State.compile(n, gates))
        qargs.append([int(q) for q in rng.choice(n, nq, replace=False)])
        shots.append(sh)
        seeds.append(700 + i)
    got = nx.sample_many(states, qargs, shots, seeds)
    for i in range(len(cases)):
        say(f"clifford batch state {i}: n={cases[i][0]}")
        assert got[i] == _oracle_counts(os_[i], qargs[i], shots[i], seeds[i]), i


def test_batch_edge_cases(nx):
    """Zero shots and an empty batch; output capacity reported
    when the buffer is too small; bad qubit indices rejected."""
    rng = np.random.default_rng(8303)
    gates = WL.random_circuit(rng, 10, 100, 6, gate_set=tuple(range(11)))
    st = _state(nx, 10, gates)
    got = nx.sample_many([st, st], [list(range(10)), [3, 1]], [0, 50], [1, 3])
    assert got[0] == {}
    assert got[1] == st.sample([3, 1], 50, seed=3, exact=True)
    assert nx.sample_many([], [], []) == []
    from necstar import _lib
    L = _lib.lib()
    nq = np.array([10, 10], np.int32)
    qa = np.array(list(range(10)) * 2, np.int32)
    sh = np.array([100, 100], np.uint64)
    seeds = (C.c_uint8 * 64)()
    handles = (C.c_void_p * 2)(st._h, st._h)
    nout = np.zeros(2, np.int64)
    bits = np.zeros(2, np.uint64)
    counts = np.zeros(1, np.uint64)
    code = L.nx_batch_sample_exact(2, handles, nq.ctypes.data_as(C.POINTER(C.c_int32)),
                                   qa.ctypes.data_as(C.POINTER(C.c_int32)), sh.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   C.cast(seeds, C.POINTER(C.c_uint8)), 1, bits.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   counts.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   nout.ctypes.data_as(C.POINTER(C.c_int64)), None)
    assert code != 0
    assert b"too small" in L.nx_last_error()
    single = st.sample(list(range(10)), 100, seed=bytes(32), exact=True)
    assert nout[0] == nout[1] == len(single)
    with pytest.raises(ValueError):
        nx.sample_many([st], [[]], [5])  # empty qubit list, as sample() rejects it
    with pytest.raises(ValueError):
        nx.sample_many([st], [[10]], [5])
    with pytest.raises(ValueError):
        nx.sample_many([st], [[1, 1]], [5])
