"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Integer results are compared bit for bit: per-term CH-form tables (G, F,
M, gamma, v, s, phase), exact coefficient scalars, exact pair scalars
<phi_i|phi_j>, exact per-term amplitude phases.  FP64 sums are compared with
the relative bound of fp64.check: |gpu - oracle| <= 1e-12 |oracle| wherever
the value is at least 1e-3 of the sum of its terms' magnitudes.
"""
import sys

import numpy as np
import pytest

import dense as D
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


def compile_both(n, gates):
    st = WL.to_circuit(n, gates)
    from necstar import QuantumState
    return QuantumState.from_circuit(st), O.State.compile(n, gates)


def assert_tables_equal(g, o):
    assert g.stabilizer_rank == o.chi
    assert g.num_qubits == o.n
    z, k, r = g.scalar_coefficients()
    for i in range(o.chi):
        tg, to = g.export_term(i), o.term(i).tables()
        for key in ("G", "F", "M", "gamma", "v", "s"):
            assert np.array_equal(tg[key], to[key]), (i, key)
        assert tg["phase"] == to["phase"], i
        oz, ok, orr = o.coeff_scalar(i)
        assert (bool(z[i]), int(k[i]), int(r[i])) == (oz, ok, orr), i


CASES = [  # (seed, n, cliffords, t, gate set)
    (1, 3, 30, 3, tuple(range(11))),
    (2, 5, 60, 5, tuple(range(11))),
    (3, 8, 100, 7, tuple(range(11))),
    (4, 12, 200, 8, WL.CLIFFORD_SET),
    (5, 30, 300, 10, tuple(range(11))),
    (6, 63, 400, 6, WL.CLIFFORD_SET),
    (7, 64, 500, 8, tuple(range(11))),
    (8, 70, 500, 9, tuple(range(11))),
    (9, 130, 700, 5, WL.CLIFFORD_SET),
    (10, 200, 600, 12, WL.CLIFFORD_SET),   # four-word rows
    (11, 240, 700, 16, tuple(range(11))),  # 240 + 16 ancillas: the 256-qubit maximum while compiling
]
# pair sums at one to four row words (the 70-qubit case is covered by the amplitude and table tests)
PAIR_CASES = CASES[:7] + CASES[8:]


@pytest.mark.parametrize("seed,n,nc,nt,gs", CASES)
def test_compile_tables_bit_exact(nx, seed, n, nc, nt, gs):
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=gs)
    say(f"compile seed={seed} n={n} cliffords={nc} t={nt}")
    g, o = compile_both(n, gates)
    assert_tables_equal(g, o)


@pytest.mark.parametrize("seed,n,nc,nt,gs", PAIR_CASES)
def test_pair_scalars_exact(nx, seed, n, nc, nt, gs):
    rng = np.random.default_rng(100 + seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=gs)
    g, o = compile_both(n, gates)
    say(f"pairs seed={seed} n={n} chi={o.chi}")
    tab = g.pair_table()
    chi = o.chi
    idx = [(i, j) for i in range(chi) for j in range(chi)]
    if len(idx) > 400:
        pick = rng.choice(len(idx), 400, replace=False)
        idx = [idx[p] for p in pick]
    for i, j in idx:
        (z, k, r), _ = o.term(i).inner_product(o.term(j))
        v = int(tab[i, j])
        if z:
            assert v == 0xFFFFFFFF, (i, j, v)
        else:
            assert v == (k | (r << 8)), (i, j, v, k, r)


@pytest.mark.parametrize("seed,n,nc,nt,gs", PAIR_CASES)
def test_norm_exp_value_inner_product(nx, seed, n, nc, nt, gs):
    rng = np.random.default_rng(200 + seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=gs)
    g, o = compile_both(n, gates)
    say(f"norm/exp seed={seed} n={n} chi={o.chi}")
    fp64.check(g.norm() ** 2, o.norm_squared(), fp64.state_scale(g, g), "norm^2")
    fp64.check(g.norm(), o.norm(), np.sqrt(fp64.state_scale(g, g)), "norm")
    from necstar import PauliString
    for ops in WL.random_paulis(rng, n, 4, 1, min(4, n)):
        s = "".join("IXYZ"[x] for x in reversed(ops))
        ev_g = g.exp_value(PauliString.from_str(s))
        ev_o = o.exp_value(("dense", ops))
        fp64.check(ev_g, ev_o, fp64.exp_scale(g, ops), f"exp_value {s}")
    gates2 = WL.random_circuit(rng, n, nc, nt, gate_set=gs)
    g2, o2 = compile_both(n, gates2)
    ip_g, ip_o = g.inner_product(g2), o.inner_product(o2)
    fp64.check(ip_g, ip_o, fp64.state_scale(g, g2), "inner_product")
    say(f"  worst relative error so far {fp64.worst}")


@pytest.mark.parametrize("seed,n,nc,nt,gs", CASES)
def test_amplitudes_exact(nx, seed, n, nc, nt, gs):
    rng = np.random.default_rng(300 + seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=gs)
    g, o = compile_both(n, gates)
    bits = WL.random_bitstrings(rng, n, 48)
    # a few strings inside the support: basis states of nonzero amplitude
    say(f"amplitudes seed={seed} n={n} chi={o.chi}")
    vals, pt, rr = g.amplitudes(bits, per_term=True)
    for b in range(len(bits)):
        x = np.array([(int(bits[b, q // 64]) >> (q % 64)) & 1 for q in range(n)], np.uint8)
        val_o, (zo, ko, ro) = o.amplitude(x, per_term=True)
        for i in range(o.chi):
            if zo[i]:
                assert pt[i, b] == 255, (b, i)
            else:
                assert pt[i, b] == ko[i] and rr[i] == ro[i], (b, i, pt[i, b], ko[i], rr[i], ro[i])
        fp64.check(vals[b], val_o, fp64.amp_scale(o, (zo, ko, ro)), f"amplitude {b}")


def _support_circuit(rng, n, k, nt):
    """H on k qubits, then CX / CZ / S / X gates and nt T gates: the support
    of the state is {c + sum_i h_i v_i} for the columns v_i and constant c
    returned (bit j of a column = qubit j)."""
    hq = [int(q) for q in rng.choice(n, k, replace=False)]
    gates = [(O.H, q, 0) for q in hq]
    cols = np.zeros((k, n), np.uint8)
    for i, q in enumerate(hq):
        cols[i, q] = 1
    const = np.zeros(n, np.uint8)
    tpos = set(int(p) for p in rng.choice(3 * n, nt, replace=False))
    for step in range(3 * n):
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        r = rng.random()
        if r < 0.5:
            gates.append((O.CX, a, b))
            cols[:, b] ^= cols[:, a]
            const[b] ^= const[a]
        elif r < 0.7:
            gates.append((O.CZ, a, b))
        elif r < 0.85:
            gates.append((O.S, a, 0))
        else:
            gates.append((O.X, a, 0))
            const[a] ^= 1
        if step in tpos:
            gates.append((O.T, a, 0))
    return gates, cols, const


def _pack(xbits, n):
    words = (n + 63) // 64
    out = np.zeros(words, np.uint64)
    for q in np.nonzero(xbits)[0]:
        out[q // 64] |= np.uint64(1) << np.uint64(q % 64)
    return out


@pytest.mark.parametrize("seed,n,k", [(901, 64, 3), (902, 64, 57), (903, 64, 58), (904, 64, 60),
                                      (905, 100, 10), (906, 100, 93), (907, 100, 94), (908, 130, 40)])
def test_amplitudes_constraint_paths(nx, seed, n, k):
    """Terms with many support constraints (tested against a table of C^T
    rows when there are more than 8) and with few (tested row by row),
    including exactly 8 and 9, evaluated at strings inside the support and
    at random strings; per-term phases and zero flags bit-exact against the
    oracle, sums within 1e-12."""
    rng = np.random.default_rng(seed)
    gates, cols, const = _support_circuit(rng, n, k, 4)
    g, o = compile_both(n, gates)
    ncs = sorted({n - int(o.term(i).tables()["v"].sum()) for i in range(o.chi)})
    say(f"amplitudes constraints seed={seed} n={n} k={k} chi={o.chi} constraint counts={ncs}")
    inside = [(const ^ (rng.integers(0, 2, k).astype(np.uint8) @ cols % 2).astype(np.uint8)) for _ in range(40)]
    rand = [rng.integers(0, 2, n).astype(np.uint8) for _ in range(16)]
    xs = inside + rand
    bits = np.array([_pack(x, n) for x in xs], np.uint64)
    vals, pt, rr = g.amplitudes(bits, per_term=True)
    nonzero = 0
    for b, x in enumerate(xs):
        val_o, (zo, ko, ro) = o.amplitude(x, per_term=True)
        for i in range(o.chi):
            if zo[i]:
                assert pt[i, b] == 255, (b, i)
            else:
                nonzero += 1
                assert pt[i, b] == ko[i] and rr[i] == ro[i], (b, i, pt[i, b], ko[i], rr[i], ro[i])
        fp64.check(vals[b], val_o, fp64.amp_scale(o, (zo, ko, ro)), f"amplitude {b}")
    assert nonzero > 0
    # both support tests are exercised where the case is meant to
    if n - k > 8:
        assert max(ncs) > 8
    if n - k <= 8:
        assert min(ncs) <= 8


def test_statevector_vs_dense(nx):
    rng = np.random.default_rng(77)
    for trial in range(6):
        n = int(rng.integers(1, 9))
        gates = D.random_circuit(rng, n, 40, int(rng.integers(0, 7)),
                                 gate_set=tuple(range(11)) if n > 1 else tuple(range(8)), tdg=True)
        say(f"statevector trial={trial} n={n}")
        from necstar import QuantumState
        g = QuantumState.from_circuit(WL.to_circuit(n, gates))
        o = O.State.compile(n, gates)
        sv = g.statevector_array()
        # |<x|phi_i>| <= 1 bounds every entry's term scale by |g| sum |c_i|
        fp64.check(sv, o.statevector(), abs(o.global_factor) * np.abs(o.coeffs()).sum(), "statevector")
        assert np.abs(sv - D.run(n, gates)).max() < 1e-9


def test_projection_discard_gates_parity(nx):
    rng = np.random.default_rng(404)
    for trial in range(8):
        n = int(rng.integers(3, 10))
        gates = WL.random_circuit(rng, n, 60, int(rng.integers(2, 7)), gate_set=tuple(range(11)))
        say(f"project/discard trial={trial} n={n}")
        g, o = compile_both(n, gates)
        cl = WL.random_circuit(rng, n, 20, 0, gate_set=tuple(range(11)))
        from necstar.gate import QuantumGate
        g.apply_gates([QuantumGate(x[0], x[1:]) for x in cl])
        for x in cl:
            o.apply_gate(*x)
        assert_tables_equal(g, o)
        q, outc = int(rng.integers(n)), bool(rng.integers(2))
        g.project_unnormalized(q, outc)
        o.project_unnormalized(q, outc)
        assert_tables_equal(g, o)
        if o.chi:
            fp64.check(g.norm() ** 2, o.norm_squared(), fp64.state_scale(g, g), "projected norm^2")
        g.project_unnormalized(q, False)
        o.project_unnormalized(q, False)
        if o.chi:
            g.discard(q)
            o.discard(q)
            assert_tables_equal(g, o)


def test_measure_same_stream_as_oracle(nx):
    """Same seed -> same measure() outcomes as the oracle (both restate rand
    0.8's StdRng; measure projects and renormalises per qubit as
    measurement.rs:34-103 does), and the same global factor bit for bit
    afterwards: the outcome norms are the reference's FP64 sums, so the
    renormalisation 1/sqrt(p) is too.  H and S after the circuit make the
    pair values complex."""
    rng = np.random.default_rng(505)
    for trial in range(6):
        n = int(rng.integers(3, 9))
        gates = WL.random_circuit(rng, n, 10 * n, 6 + trial, gate_set=tuple(range(11)))
        for q in range(n):
            if rng.integers(2):
                gates += [(WL.H, q), (WL.S, q)]
        g, o = compile_both(n, gates)
        say(f"measure trial={trial} n={n} chi={o.chi}")
        qargs = [int(q) for q in rng.permutation(n)]
        gc, oc = g.clone(), o.clone()
        m_g = gc.measure(qargs, seed=trial + 7)
        m_o = oc.measure(qargs, trial + 7)
        assert m_g == m_o
        assert gc.global_factor == oc.global_factor, (gc.global_factor, oc.global_factor)


def _oracle_counts(o, qargs, shots, seed):
    return {"".join("1" if b else "0" for b in bits): c for bits, c in o.sample(qargs, shots, seed)}


@pytest.mark.parametrize("seed,n,nt,nq,shots", [
    (501, 6, 12, 6, 500),     # one window, every qubit (chi 27)
    (502, 8, 16, 8, 800),     # two windows, every qubit (chi 36)
    (503, 10, 14, 7, 1000),   # subset of qubits, unsorted, two windows (chi 45)
    (504, 9, 18, 9, 300),     # two windows, chi 96
    (505, 7, 12, 4, 2000),    # subset, one window (chi 24)
    (506, 12, 14, 12, 200),   # every qubit, two windows (chi 16)
])
def test_sample_same_stream_as_oracle(nx, seed, n, nt, nq, shots):
    """sample() on Clifford+T states with generic (irrational, tied)
    probabilities: every node probability is the reference's FP64 value
    (norm.rs:8-30 summed over the projected terms in its order), so a seed
    gives the oracle's outcomes exactly (sampling.rs:80-189).  Single-qubit
    H and S after the circuit make the pair values complex, so the order of
    each pair's operands matters to the rounding.  The exact mode
    (exact=True) may differ on these states; its outcomes are reported."""
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, 12 * n, nt, gate_set=tuple(range(11)))
    for q in range(n):
        if rng.integers(2):
            gates += [(WL.H, q), (WL.S, q)]
    g, o = compile_both(n, gates)
    qargs = [int(q) for q in rng.permutation(n)[:nq]]
    say(f"sample seed={seed} n={n} chi={o.chi} qargs={qargs} shots={shots}")
    res_o = _oracle_counts(o, qargs, shots, seed)
    res_g = g.sample(qargs, shots, seed=seed)
    assert res_g == res_o
    assert res_g == g.sample(qargs, shots, seed=seed)
    res_x = g.sample(qargs, shots, seed=seed, exact=True)
    assert sum(res_x.values()) == shots
    say(f"  exact mode {'equal' if res_x == res_o else 'differs'}")


@pytest.mark.parametrize("seed,n,nq,shots", [
    (611, 14, 14, 2000),   # three 6-qubit look-ahead windows
    (612, 20, 13, 300),    # subset of qubits, unsorted
    (613, 70, 10, 64),     # two-word rows
    (614, 66, 66, 40),     # every qubit sampled: certified levels, then an exact-amplitude tail (two words)
    (615, 200, 20, 40),    # four-word rows
    (616, 196, 128, 6),    # the 128-qubit sampling maximum, four-word rows
])
def test_sample_clifford_windows_same_stream(nx, seed, n, nq, shots):
    """Clifford states (probabilities 0, 1/2^k, 1): seeded outcomes equal
    the oracle's across several look-ahead windows, in both modes (the
    reference's sums are exact here, so the exact mode agrees too)."""
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, 10 * n, 0, gate_set=tuple(range(11)))
    g, o = compile_both(n, gates)
    qargs = [int(q) for q in rng.choice(n, nq, replace=False)]
    say(f"clifford sample seed={seed} n={n} qargs={qargs}")
    res_o = _oracle_counts(o, qargs, shots, seed)
    assert g.sample(qargs, shots, seed=seed) == res_o
    assert g.sample(qargs, shots, seed=seed, exact=True) == res_o


def test_sample_dyadic_clifford_t_exact_mode(nx):
    """The exact mode on Clifford+T states whose outcome probabilities are
    dyadic (the reference's FP64 sums are exact): the same seed gives the
    oracle's outcomes."""
    rng = np.random.default_rng(505)
    checked = 0
    for trial in range(8):
        n = int(rng.integers(3, 8))
        gates = WL.random_circuit(rng, n, 50, 4, gate_set=tuple(range(11)))
        g, o = compile_both(n, gates)
        qargs = [int(q) for q in rng.permutation(n)]
        p = np.abs(np.asarray(o.statevector())) ** 2
        if not np.all(np.abs(p - np.round(p * 2.0 ** 24) / 2.0 ** 24) <= 1e-12):
            continue
        say(f"dyadic sample trial={trial} n={n} chi={o.chi}")
        assert g.sample(qargs, 500, seed=trial, exact=True) == _oracle_counts(o, qargs, 500, trial)
        checked += 1
    assert checked >= 1


def test_projection_norms_vs_oracle(nx):
    """projection_norms (projections as constraints of the pair sums, all
    2^k look-ahead outcomes in one pass) against the oracle's project +
    norm_squared (projection.rs:10-31, norm.rs:8-30) per outcome."""
    rng = np.random.default_rng(701)
    for trial, (n, nfix, nlook) in enumerate([(3, 0, 1), (4, 1, 0), (5, 0, 2), (6, 2, 3), (9, 3, 6), (12, 4, 7),
                                              (70, 5, 4), (200, 3, 5)]):
        gates = WL.random_circuit(rng, n, 12 * n, 5, gate_set=tuple(range(11)))
        g, o = compile_both(n, gates)
        qs = [int(q) for q in rng.choice(n, nfix + nlook, replace=False)]
        fixed = [(q, int(rng.integers(2))) for q in qs[:nfix]]
        look = qs[nfix:]
        say(f"projection norms trial={trial} n={n} chi={o.chi} fixed={fixed} look={look}")
        got = g.projection_norms(look, fixed)
        ref = []
        for k in range(1 << nlook):
            c = o.clone()
            for q, b in fixed:
                c.project_unnormalized(q, b)
            for i, q in enumerate(look):
                c.project_unnormalized(q, (k >> i) & 1)
            ref.append(c.norm_squared())
        ref = np.array(ref)
        # pair magnitudes are at most 1: (|g| sum |c_i|)^2 bounds the term scale
        fp64.check(got, ref, (abs(o.global_factor) * np.abs(o.coeffs()).sum()) ** 2, f"projection norms {trial}")
        # exact sums: an outcome of probability zero gives exactly 0.0, as in
        # the oracle (sampling's binomial draws depend on it)
        assert np.all(got[ref == 0.0] == 0.0), (trial, got, ref)


@pytest.mark.parametrize("seed,n,nc,nt,nq,shots", [
    (601, 9, 90, 6, 9, 20000),     # two windows, all qubits
    (602, 12, 120, 5, 8, 20000),   # subset of qubits, unsorted
])
def test_sample_distribution_clifford_t(nx, seed, n, nc, nt, nq, shots):
    """Clifford+T states, exact mode: the sampled distribution matches the exact one
    (from the oracle's state vector; total variation within three times its
    expected sampling noise) and a seed reproduces its outcomes."""
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=tuple(range(11)))
    g, o = compile_both(n, gates)
    qargs = [int(q) for q in rng.choice(n, nq, replace=False)]
    say(f"sample distribution seed={seed} n={n} chi={o.chi} qargs={qargs}")
    res = g.sample(qargs, shots, seed=seed, exact=True)
    assert res == g.sample(qargs, shots, seed=seed, exact=True)
    assert sum(res.values()) == shots
    p = np.abs(np.asarray(o.statevector())) ** 2
    exact = {}
    for x in range(1 << n):
        if p[x] > 0:
            key = "".join("1" if (x >> q) & 1 else "0" for q in qargs)
            exact[key] = exact.get(key, 0.0) + p[x]
    assert set(res) <= set(k for k, v in exact.items() if v > 1e-12)
    tv = 0.5 * sum(abs(res.get(k, 0) / shots - v) for k, v in exact.items())
    # E[tv] <= 0.5 sum_k sqrt(v_k (1 - v_k) / shots); allow three times that
    bound = 0.5 * sum(np.sqrt(v * (1 - v) / shots) for v in exact.values())
    assert tv <= 3 * bound, (tv, bound)


def test_errors_match_oracle(nx):
    from necstar import PauliString, QuantumCircuit, QuantumState
    k = [(0, 0), (11, 0), (8, 0, 1)]
    with pytest.raises(ValueError, match="Qubit index 3 is out of bounds for 3 qubits"):
        QuantumState.from_circuit(WL.to_circuit(2, [(0, 0), (11, 3), (8, 0, 1)]))
    st = QuantumState.from_circuit(WL.to_circuit(2, k))
    with pytest.raises(ValueError, match="Direct application of non-Clifford gate T is not supported."):
        st.apply_gates([__import__("necstar").QuantumGate.t(0)])
    with pytest.raises(ValueError, match="Qubit index 5 is out of bounds for 2 qubits."):
        st.apply_h(5)
    with pytest.raises(ValueError, match="Two-qubit gate requires two different qubit indices"):
        st.apply_cx(1, 1)
    with pytest.raises(ValueError, match="Invalid Pauli string length: expected 2, found 3."):
        st.exp_value(PauliString.from_str("ZZZ"))
    with pytest.raises(ValueError, match="Invalid Pauli string length: expected 2, found 5."):
        st.exp_value(PauliString.from_str("Z4"))
    with pytest.raises(ValueError, match="Qubit index list is empty."):
        st.sample([], 10)
    with pytest.raises(ValueError, match="Duplicate qubit index found: 0."):
        st.measure([0, 0])
    with pytest.raises(ValueError, match="Gate CCX is not supported."):
        c = QuantumCircuit(3)
        c.apply_ccx(0, 1, 2)
        QuantumState.from_circuit(c)
    big = QuantumState.from_circuit(QuantumCircuit(3))
    with pytest.raises(ValueError, match="The qubit counts of the two states must match"):
        st.inner_product(big)
    wide = QuantumState.from_circuit(QuantumCircuit(130))
    with pytest.raises(ValueError, match="Sampling more than 128 qubits is not supported."):
        wide.sample(list(range(129)), 10)
    with pytest.raises(O.OracleError, match="Sampling more than 128 qubits"):
        O.State.compile(130, []).sample(list(range(129)), 10, 0)


def test_qubit_limit(nx):
    """The device tables hold at most 256 qubits (four row words) while
    compiling, ancillas of the T gadgets included; a circuit above that is
    rejected on the host before any launch, and one exactly at it compiles."""
    from necstar import QuantumState
    with pytest.raises(ValueError, match="above 256 qubits"):
        QuantumState.from_circuit(WL.to_circuit(250, [(0, 0)] + [(11, q) for q in range(7)]))
    with pytest.raises(ValueError, match="above 256 qubits"):
        QuantumState.from_circuit(WL.to_circuit(257, [(0, 0)]))
    g, o = compile_both(250, [(0, q) for q in range(6)] + [(11, q) for q in range(6)])
    assert_tables_equal(g, o)
    fp64.check(g.norm(), o.norm(), np.sqrt(fp64.state_scale(g, g)), "norm at 256 compile qubits")


def test_config0_workload(nx):
    cfg = WL.config(0)
    n = cfg["n"]
    g, o = compile_both(n, cfg["gates"])
    say(f"cfg0 chi={o.chi}")
    assert_tables_equal(g, o)
    fp64.check(g.norm(), o.norm(), np.sqrt(fp64.state_scale(g, g)), "cfg0 norm")
    assert abs(g.norm() - 1.0) <= 1e-12
    vals = g.amplitudes(cfg["bitstrings"])
    for b in range(len(cfg["bitstrings"])):
        x = np.array([(int(cfg["bitstrings"][b, 0]) >> q) & 1 for q in range(n)], np.uint8)
        # 12 terms: one term chunk, so the sum runs in the reference's order
        assert vals[b] == o.amplitude(x)
    from necstar import PauliString
    for ops in cfg["paulis"]:
        s = "".join("IXYZ"[x] for x in reversed(ops))
        fp64.check(g.exp_value(PauliString.from_str(s)), o.exp_value(("dense", ops)), fp64.exp_scale(g, ops),
                   f"cfg0 exp_value {s}")


def test_config1_full_size(nx):
    """cfg1 at its real shape (n=64, 2000 Cliffords, 24 T): chi and sampled
    tables, pair scalars and per-term amplitudes bit-exact against the
    oracle, 256 amplitude sums within 1e-12 relative (cfg1 amplitudes are
    of order 2^-32 with terms cancelling by up to ~400x); the norm of a
    compiled unitary circuit is 1 (size-independent property)."""
    cfg = WL.config(1)
    n = cfg["n"]
    g, o = compile_both(n, cfg["gates"])
    say(f"cfg1 chi={o.chi}")
    assert g.stabilizer_rank == o.chi
    rng = np.random.default_rng(11)
    z, k, r = g.scalar_coefficients()
    for i in rng.choice(o.chi, 24, replace=False):
        i = int(i)
        tg, to = g.export_term(i), o.term(i).tables()
        for key in ("G", "F", "M", "gamma", "v", "s"):
            assert np.array_equal(tg[key], to[key]), (i, key)
        assert tg["phase"] == to["phase"]
        assert (bool(z[i]), int(k[i]), int(r[i])) == o.coeff_scalar(i)
    # (the oracle's norm at cfg1, 266,085 pairs, takes minutes: the unit
    # norm of a compiled unitary circuit is the check here)
    assert abs(g.norm() - 1.0) <= 1e-12
    tab = g.pair_table()
    for _ in range(60):
        i, j = (int(v) for v in rng.integers(0, o.chi, 2))
        (zz, kk, rr), _ = o.term(i).inner_product(o.term(j))
        assert int(tab[i, j]) == (0xFFFFFFFF if zz else (kk | (rr << 8))), (i, j)
    # the first 256 strings of the workload, evaluated as part of all
    # 65,536 (the launch shape, hence the term chunking, of the benchmark)
    allv = g.amplitudes(cfg["bitstrings"])
    bits = cfg["bitstrings"][:256]
    vals, pt, rr = g.amplitudes(bits, per_term=True)
    rels = []
    for b in range(len(bits)):
        x = np.array([(int(bits[b, 0]) >> q) & 1 for q in range(n)], np.uint8)
        val_o, (zo, ko, ro) = o.amplitude(x, per_term=True)
        assert np.array_equal(pt[:, b] == 255, zo)
        assert np.array_equal(pt[~zo, b], ko[~zo]) and np.array_equal(rr[~zo], ro[~zo])
        sc = fp64.amp_scale(o, (zo, ko, ro))
        rels.append(fp64.check(vals[b], val_o, sc, f"cfg1 amplitude {b}"))
        rels.append(fp64.check(allv[b], val_o, sc, f"cfg1 amplitude {b} (65,536-string launch)"))
    say(f"cfg1 amplitudes: largest relative error {max(rels):.3g} over {len(bits)} strings")


def test_config4_full_size(nx):
    """cfg4 at its real shape (n=128, 6000 Cliffords, 40 T: chi = 78,732 on
    two-word rows).  The oracle cannot compile it in test time, so two
    independent device paths are compared: the amplitude kernel's <x|psi>
    against inner_product(|x>, psi) evaluated by the pair kernel (the
    reference defines the amplitude through the same CH-form algebra,
    amplitude.rs:28-44 vs inner_product.rs:14-48)."""
    cfg = WL.config(4)
    n = cfg["n"]
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    chi = st.stabilizer_rank
    say(f"cfg4 chi={chi}")
    assert chi == 78732
    # random 128-bit strings: amplitudes of order 2^-64, compared relative
    # to their own size (zero amplitudes must be zero on both paths)
    bits = cfg["bitstrings"][:6]
    amp = st.amplitudes(bits)
    for b in range(len(bits)):
        qc = nx.QuantumCircuit(n)
        for w in range(2):
            word = int(bits[b, w])
            for q in range(64):
                if (word >> q) & 1:
                    qc.apply_x(64 * w + q)
        basis = nx.QuantumState.from_circuit(qc)
        ip = basis.inner_product(st)
        scale = max(abs(ip), abs(amp[b]), 1e-300)
        assert abs(ip - amp[b]) <= 1e-12 * scale, (b, ip, amp[b])


def test_amplitudes_ignore_bits_beyond_n(nx):
    """Ragged input: strings given with more words than the state needs, and
    set bits above qubit n - 1, read as the string of their first n bits
    (the reference's bitstrings have exactly n entries)."""
    rng = np.random.default_rng(1201)
    n = 70
    gates = WL.random_circuit(rng, n, 300, 6, gate_set=tuple(range(11)))
    g, o = compile_both(n, gates)
    bits = WL.random_bitstrings(rng, n, 32)
    wide = np.concatenate([bits, rng.integers(0, 2**63, (32, 1), dtype=np.uint64)], axis=1)
    wide[:, 1] |= rng.integers(0, 2**57, 32, dtype=np.uint64) << np.uint64(6)
    assert np.array_equal(g.amplitudes(wide), g.amplitudes(bits))
    x = np.array([(int(bits[0, q // 64]) >> (q % 64)) & 1 for q in range(n)], np.uint8)
    val_o, pt = o.amplitude(x, per_term=True)
    fp64.check(g.amplitudes(wide)[0], val_o, fp64.amp_scale(o, pt), "ragged amplitude")
