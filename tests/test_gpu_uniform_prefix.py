"""QuantumState.uniform_prefix (nx_state_uniform_prefix) and its use in
sample().

One symbolic pass over the pairs decides, for every t != 0 supported on the
first D qubits of qargs at once, whether some pair value
<phi_i| Z^t |phi_j> is nonzero; the answer is the largest D for which none
is.  Checked against brute force over every pair and every such t from the
oracle's per-term state vectors (exact equality), against the dense state's
marginals (the certified uniformity must hold), on closed-form states with
two-word rows and more than 64 coordinates, and by seeded sample() runs that
skip the certified levels yet give the oracle's outcomes."""
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


def pairwise_depth(o, qargs):
    """Largest D such that <phi_i| Z^t |phi_j> = 0 for every pair and every
    t != 0 on qargs[:D] (brute force)."""
    T = np.array([np.asarray(o.term(i).statevector()) for i in range(o.chi)])
    idx = np.arange(T.shape[1])
    for D in range(1, len(qargs) + 1):
        top = qargs[D - 1]
        for mask in range(1 << (D - 1)):
            tm = 1 << top
            for k in range(D - 1):
                if (mask >> k) & 1:
                    tm |= 1 << qargs[k]
            sign = 1.0 - 2.0 * (np.array([bin(x).count("1") for x in (idx & tm)]) & 1)
            G = T.conj() @ (T * sign).T
            if np.any(np.abs(G) > 1e-9):
                return D - 1
    return len(qargs)


def marginal_uniform_depth(sv, qargs):
    p = np.abs(np.asarray(sv)) ** 2
    tot = p.sum()
    idx = np.arange(len(p))
    best = 0
    for D in range(1, len(qargs) + 1):
        key = np.zeros(len(p), np.int64)
        for k in range(D):
            key |= ((idx >> qargs[k]) & 1) << k
        m = np.bincount(key, weights=p, minlength=1 << D)
        if np.all(np.abs(m - tot / (1 << D)) <= 1e-12 * max(tot, 1e-300)):
            best = D
        else:
            break
    return best


CASES = [(701, 4, 2), (702, 5, 3), (703, 6, 4), (704, 7, 3), (705, 8, 5), (706, 9, 2), (707, 6, 0)]


@pytest.mark.parametrize("seed,n,nt", CASES)
def test_uniform_prefix_vs_brute_force(nx, seed, n, nt):
    rng = np.random.default_rng(seed)
    for trial in range(3):
        gates = WL.random_circuit(rng, n, 6 * n, nt, gate_set=tuple(range(11)))
        st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
        o = O.State.compile(n, gates)
        k = int(rng.integers(1, n + 1))
        qargs = [int(q) for q in rng.permutation(n)[:k]]
        d = st.uniform_prefix(qargs)
        ref = pairwise_depth(o, qargs)
        true = marginal_uniform_depth(o.statevector(), qargs)
        say(f"uniform prefix seed={seed} trial={trial} n={n} chi={o.chi} qargs={qargs} d={d} ref={ref} true={true}")
        assert d == ref
        assert d <= true


def test_uniform_prefix_closed_forms(nx):
    n = 70
    qc = nx.QuantumCircuit(n)
    for q in range(n):
        qc.apply_h(q)
    qc.apply_s(5)
    qc.apply_cz(3, 67)
    st = nx.QuantumState.from_circuit(qc)
    order = list(range(n - 1, -1, -1))
    # |+>^70 with diagonal gates: every Z marginal is uniform (two-word rows,
    # 70 coordinates)
    assert st.uniform_prefix(order) == n
    qc = nx.QuantumCircuit(n)
    for q in range(n):
        if q != 66:
            qc.apply_h(q)
    qc.apply_cx(3, 66)
    st = nx.QuantumState.from_circuit(qc)
    # qubit 66 copies qubit 3: coordinates 3 (qubit 66) and 66 (qubit 3)
    assert st.uniform_prefix(order) == 66
    assert st.uniform_prefix([66]) == 1
    assert st.uniform_prefix([66, 3]) == 1
    assert st.uniform_prefix([65, 64, 2, 66, 3]) == 4
    with pytest.raises(ValueError):
        st.uniform_prefix([70])


@pytest.mark.parametrize("seed,n,nq,shots", [(721, 12, 12, 800), (722, 16, 10, 300)])
def test_sample_skipping_uniform_levels_same_stream(nx, seed, n, nq, shots):
    """Clifford states (dyadic probabilities): the levels certified uniform
    are drawn with p = 1/2 and no norms, and the outcomes equal the
    oracle's, which evaluates every node."""
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, 8 * n, 0, gate_set=tuple(range(11)))
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    o = O.State.compile(n, gates)
    qargs = [int(q) for q in rng.choice(n, nq, replace=False)]
    order = sorted(qargs, reverse=True)
    d = st.uniform_prefix(order)
    say(f"sample with uniform prefix seed={seed} n={n} qargs={qargs} d={d}")
    assert d >= 2
    res_g = st.sample(qargs, shots, seed=seed)
    res_o = {"".join("1" if b else "0" for b in bits): c for bits, c in o.sample(qargs, shots, seed)}
    assert res_g == res_o


def _masked_terms(o, fixed):
    T = np.array([np.asarray(o.term(i).statevector()) for i in range(o.chi)])
    idx = np.arange(T.shape[1])
    keep = np.ones(len(idx), bool)
    for q, b in fixed:
        keep &= ((idx >> q) & 1) == b
    return T * keep


@pytest.mark.parametrize("seed,n,nt", [(731, 6, 3), (732, 7, 4), (733, 8, 2), (734, 9, 3)])
def test_uniform_prefix_given_vs_brute_force(nx, seed, n, nt):
    """Conditional version: pairs of the terms projected onto the fixed
    outcomes (checked on the masked term vectors), and the conditional
    marginal of the dense state."""
    rng = np.random.default_rng(seed)
    for trial in range(3):
        gates = WL.random_circuit(rng, n, 6 * n, nt, gate_set=tuple(range(11)))
        st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
        o = O.State.compile(n, gates)
        perm = [int(q) for q in rng.permutation(n)]
        nf = int(rng.integers(1, n - 1))
        fixed = [(q, int(rng.integers(2))) for q in perm[:nf]]
        qargs = perm[nf:]
        d = st.uniform_prefix(qargs, fixed)
        T = _masked_terms(o, fixed)
        idx = np.arange(T.shape[1])
        ref = len(qargs)
        for D in range(1, len(qargs) + 1):
            hit = False
            for mask in range(1 << (D - 1)):
                tm = 1 << qargs[D - 1]
                for k in range(D - 1):
                    if (mask >> k) & 1:
                        tm |= 1 << qargs[k]
                sign = 1.0 - 2.0 * (np.array([bin(x).count("1") for x in (idx & tm)]) & 1)
                if np.any(np.abs(T.conj() @ (T * sign).T) > 1e-9):
                    hit = True
                    break
            if hit:
                ref = D - 1
                break
        sv = np.asarray(o.statevector())
        keep = np.ones(len(sv), bool)
        for q, b in fixed:
            keep &= ((np.arange(len(sv)) >> q) & 1) == b
        true = marginal_uniform_depth(sv * keep, qargs) if np.abs(sv[keep]).sum() > 0 else len(qargs)
        say(f"uniform prefix given seed={seed} trial={trial} n={n} fixed={fixed} qargs={qargs} d={d} ref={ref} true={true}")
        assert d == ref
        assert d <= true
