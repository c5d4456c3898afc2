"""Pauli-variant pair sums and batches on the GPU.

QuantumState.exp_values evaluates all of its Pauli strings per (bra, ket)
pair in one elimination (k_pairs_var); necstar.Batch does that for many
states of uneven n and chi in one launch plan (optionally autotuned and
captured into a HIP graph).  Both are checked against the oracle's
exp_value / norm (the reference's exp_value.rs:7-61 and norm.rs:8-35
restated on the CPU) and against the single-string GPU path, within
1e-12 relative (fp64.check).  Graph replays are checked one by one, with
eager library work between them."""
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


def scales(st, ops_list):
    """Term scales of st.exp_value for each dense op list."""
    return np.array([fp64.exp_scale(st, ops) for ops in ops_list])


def pauli_str(ops):
    return "".join("IXYZ"[x] for x in reversed(ops))


def random_ops(rng, n, count, full=4):
    ops = WL.random_paulis(rng, n, count - full - 1, 1, min(6, n))
    for _ in range(full):  # dense strings touching every qubit
        ops.append([int(x) for x in rng.integers(1, 4, size=n)])
    ops.append([0] * n)  # identity: the norm squared
    return ops


def informative_ops(st, nx, rng, n, count):
    """Strings with nonzero expectation first (found among 400 random
    low-weight candidates), then random ones, dense ones and the identity,
    so that wrong variant phases or right-hand sides change the sums."""
    cands = WL.random_paulis(rng, n, 400, 1, 3)
    ev = st.exp_values([nx.PauliString.from_str(pauli_str(p)) for p in cands])
    good = [c for c, e in zip(cands, ev) if abs(e) > 1e-3][: count // 2]
    return good, good + random_ops(rng, n, count - len(good))


VAR_CASES = [  # (seed, n, cliffords, t, number of strings, strings checked against the oracle)
    (41, 5, 20, 5, 70, 70),
    (42, 12, 40, 8, 64, 64),
    (43, 30, 90, 10, 129, 40),
    (44, 64, 150, 8, 65, 24),
    (45, 70, 180, 9, 33, 12),
    (46, 130, 300, 5, 20, 6),
    (47, 240, 300, 12, 20, 6),
]


@pytest.mark.parametrize("seed,n,nc,nt,npauli,ncheck", VAR_CASES)
def test_exp_values_variants_vs_oracle(nx, seed, n, nc, nt, npauli, ncheck):
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=tuple(range(11)))
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    o = O.State.compile(n, gates)
    good, ops = informative_ops(st, nx, rng, n, npauli)
    assert len(ops) == npauli
    paulis = [nx.PauliString.from_str(pauli_str(p)) for p in ops]
    say(f"variants seed={seed} n={n} chi={o.chi} strings={npauli} informative={len(good)}")
    ev = st.exp_values(paulis)
    sc = scales(st, ops)
    # identity string = norm^2
    fp64.check(ev[-1], o.norm_squared(), sc[-1], "identity string = norm^2")
    idx = list(range(ncheck - 1)) + [npauli - 1]
    ref = np.array([o.exp_value(("dense", ops[k])) for k in idx])
    fp64.check(ev[idx], ref, sc[idx], "variants vs oracle")
    assert np.sum(np.abs(ref[:-1]) > 1e-3) >= 3, "the check needs strings with nonzero expectation"
    # every string against the single-string GPU path
    single = [st.exp_value(p) for p in paulis]
    fp64.check(ev, single, sc, "variants vs single-string path")
    say(f"  worst relative error so far {fp64.worst}")


def test_exp_values_complex_coefficients(nx):
    rng = np.random.default_rng(47)
    n = 9
    gates = WL.random_circuit(rng, n, 90, 6, gate_set=tuple(range(11)))
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    o = O.State.compile(n, gates)
    c = st.coefficients() * np.exp(1j * rng.uniform(0, 2 * np.pi, st.stabilizer_rank))
    c += 1e-2 * (rng.standard_normal(len(c)) + 1j * rng.standard_normal(len(c)))
    st.set_coefficients(c)
    o.set_complex_coeffs(c)
    ops = random_ops(rng, n, 40)
    ev = st.exp_values([nx.PauliString.from_str(pauli_str(p)) for p in ops])
    fp64.check(ev, [o.exp_value(("dense", p)) for p in ops], scales(st, ops), "complex coefficients")


def _mixed_states(nx, rng):
    specs = [(16, 4), (40, 7), (64, 8), (65, 6), (100, 5), (128, 4), (20, 9)]
    states, olist = [], []
    for n, t in specs:
        gates = WL.random_circuit(rng, n, 10 * n, t)
        states.append(nx.QuantumState.from_circuit(WL.to_circuit(n, gates)))
        olist.append((n, gates))
    return states, olist


def test_batch_matches_plain_calls_and_graph_replays(nx):
    rng = np.random.default_rng(48)
    states, olist = _mixed_states(nx, rng)
    counts = [0, 1, 5, 63, 64, 65, 130]
    paulis, opl = [], []
    for st, k in zip(states, counts):
        ops = WL.random_paulis(rng, st.num_qubits, k, 2, 6)
        paulis.append([nx.PauliString.from_str(pauli_str(p)) for p in ops])
        opl.append(ops)
    # a state with no terms left (impossible projection)
    z = nx.QuantumState.from_circuit(nx.QuantumCircuit(3))
    z.project_unnormalized(0, True)
    assert z.stabilizer_rank == 0
    states.append(z)
    paulis.append([nx.PauliString.from_str("XIZ")])
    opl.append([[1, 0, 3]])
    say("batch plain reference")
    ref_norm = [st.norm() for st in states]
    ref_vals = [np.array([st.exp_value(p) for p in pl]) for st, pl in zip(states, paulis)]
    norm_sc = [np.sqrt(fp64.state_scale(st, st)) if st.stabilizer_rank else 0.0 for st in states]
    val_sc = [scales(st, ops) for st, ops in zip(states, opl)]
    # oracle on the small states
    for i in (0, 1, 6):
        n, gates = olist[i]
        o = O.State.compile(n, gates)
        fp64.check(ref_norm[i], o.norm(), norm_sc[i], f"plain norm {i} vs oracle")
    for autotune, graph in ((False, False), (True, False), (False, True), (True, True)):
        say(f"batch autotune={autotune} graph={graph}")
        b = nx.Batch(states, paulis, autotune=autotune, graph=graph)
        J, waves, vw = b.shape
        assert J in (8, 16, 32) and waves >= 1 and vw == 2
        for rep in range(3):
            norms, vals = b.run()
            fp64.check(norms, ref_norm, norm_sc, f"batch norms {autotune} {graph} {rep}")
            for i, (v, r) in enumerate(zip(vals, ref_vals)):
                assert v.shape == r.shape, (autotune, graph, rep, i)
                if len(r):
                    fp64.check(v, r, val_sc[i], f"batch values {autotune} {graph} {rep} state {i}")
            states[2].norm()  # eager work on the same stream between replays
        del b


def test_batch_copies_states(nx):
    rng = np.random.default_rng(49)
    n = 10
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, WL.random_circuit(rng, n, 100, 5)))
    p = [nx.PauliString.from_str("Z0"), nx.PauliString.from_str("X1 Y2")]
    ref = [st.exp_value(q) for q in p]
    sc = scales(st, [[3] + [0] * (n - 1), [0, 1, 2] + [0] * (n - 3)])
    b = nx.Batch([st], [p], autotune=False, graph=True)
    st.apply_h(0)
    st.apply_cx(0, 1)
    norms, vals = b.run()
    fp64.check(vals[0], ref, sc, "batch holds a copy")


def test_batch_errors(nx):
    st = nx.QuantumState.from_circuit(nx.QuantumCircuit(3))
    with pytest.raises(ValueError, match="Invalid Pauli string length"):
        nx.Batch([st], [[nx.PauliString.from_str("XX")]])
    with pytest.raises(ValueError):
        nx.Batch([st], [])
