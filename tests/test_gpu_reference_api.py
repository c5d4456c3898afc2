"""The reference's Python state tests (tests/test_state_initialization.py,
tests/test_state_calculation.py, tests/test_state_simulation.py) run against
this necstar module on the GPU, with the same assertions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nc():
    import necstar
    return necstar


def make_toffoli_state(nc):
    circuit = nc.QuantumCircuit(3)
    circuit.apply_h(0)
    circuit.apply_h(1)
    circuit.apply_h(2)
    circuit.apply_cx(1, 2)
    circuit.apply_tdg(2)
    circuit.apply_cx(0, 2)
    circuit.apply_t(2)
    circuit.apply_cx(1, 2)
    circuit.apply_tdg(2)
    circuit.apply_cx(0, 2)
    circuit.apply_t(1)
    circuit.apply_t(2)
    circuit.apply_cx(0, 1)
    circuit.apply_h(2)
    circuit.apply_t(0)
    circuit.apply_tdg(1)
    circuit.apply_cx(0, 1)
    return nc.QuantumState.from_circuit(circuit)


def test_state_initialization(nc):
    circuit = nc.QuantumCircuit(2)
    circuit.apply_h(0)
    circuit.apply_t(0)
    circuit.apply_cx(0, 1)
    state = nc.QuantumState.from_circuit(circuit)
    assert state.num_qubits == 2
    assert state.stabilizer_rank == 2
    ref = [1 / 2 ** 0.5, 0.0, 0.0, 0.5 + 0.5j]
    for amp, r in zip(state.to_statevector(), ref):
        assert abs(amp - r) < 1e-10
    assert abs(state.norm() - 1.0) < 1e-10
    assert str(state) == "QuantumState(num_qubits=2, stabilizer_rank=2)"


def test_state_initialization_invalid_circuit(nc):
    circuit = nc.QuantumCircuit(2)
    circuit.apply_h(0)
    circuit.apply_t(3)
    circuit.apply_cx(0, 1)
    with pytest.raises(ValueError):
        nc.QuantumState.from_circuit(circuit)


def test_state_expectation_value(nc):
    st = make_toffoli_state(nc)
    assert np.isclose(st.exp_value(nc.PauliString.from_str("ZII")), 0.5)
    for bad in ["ZZIIII", "Z0 Y2 X4", "ZI"]:
        with pytest.raises(ValueError):
            st.exp_value(nc.PauliString.from_str(bad))


def test_state_inner_product(nc):
    s1, s2 = make_toffoli_state(nc), make_toffoli_state(nc)
    assert np.isclose(s1.inner_product(s2), 1.0)
    for n in (4, 2):
        c = nc.QuantumCircuit(n)
        c.apply_h(0)
        with pytest.raises(ValueError):
            s1.inner_product(nc.QuantumState.from_circuit(c))


def test_state_sample(nc):
    st = make_toffoli_state(nc)
    shots = 10000
    counts = st.sample([0, 1, 2], shots, seed=42)
    for outcome in {"000", "100", "010", "111"}:
        assert outcome in counts
        assert sum(counts.values()) == shots
        assert abs(counts[outcome] / shots - 0.25) < 0.05
    rev = st.sample([2, 1, 0], shots, seed=40)
    for outcome in {"000", "001", "010", "111"}:
        assert outcome in rev
    with pytest.raises(ValueError):
        st.sample([0, 1, 2, 3], shots)
    with pytest.raises(ValueError):
        st.sample([0, 0], shots)
    with pytest.raises(ValueError):
        st.sample([], shots)


def test_state_sample_determinism(nc):
    st = make_toffoli_state(nc)
    c1 = st.sample([0, 1, 2], 1000, seed=123)
    c2 = st.sample([0, 1, 2], 1000, seed=123)
    assert c1 == c2
    assert c1 != st.sample([0, 1, 2], 1000, seed=456)


def _one_hot(sv):
    sv = np.asarray(sv)
    nz = np.where(np.abs(sv) > 1e-10)[0]
    return len(nz) == 1 and np.isclose(np.abs(sv[nz[0]]), 1.0)


def test_state_measure(nc):
    for i in range(20):
        st = make_toffoli_state(nc)
        bits = st.measure([0, 1, 2], seed=i)
        assert "".join(map(str, map(int, bits))) in {"000", "100", "010", "111"}
        assert _one_hot(st.to_statevector())
        assert list(bits) == list(st.measure([0, 1, 2]))
    with pytest.raises(ValueError):
        make_toffoli_state(nc).measure([0, 0])
    with pytest.raises(ValueError):
        make_toffoli_state(nc).measure([0, 1, 10])


def test_state_measure_all(nc):
    for i in range(20):
        st = make_toffoli_state(nc)
        bits = st.measure_all(seed=i)
        assert "".join(map(str, map(int, bits))) in {"000", "100", "010", "111"}
        assert _one_hot(st.to_statevector())
        assert list(bits) == list(st.measure_all())


def test_state_measure_determinism(nc):
    b1 = make_toffoli_state(nc).measure([0, 1, 2], seed=123)
    b2 = make_toffoli_state(nc).measure([0, 1, 2], seed=123)
    assert list(b1) == list(b2)
    assert list(b1) != list(make_toffoli_state(nc).measure([0, 1, 2], seed=456))


def test_state_project_normalized(nc):
    st = make_toffoli_state(nc)
    st.project_normalized(0, True)
    assert np.isclose(st.norm(), 1.0)
    assert st.num_qubits == 3
    assert st.measure([0])[0]
    with pytest.raises(ValueError):
        st.project_normalized(0, False)
    with pytest.raises(ValueError):
        make_toffoli_state(nc).project_normalized(3, True)


def test_state_project_unnormalized(nc):
    st = make_toffoli_state(nc)
    st.project_unnormalized(0, True)
    assert st.num_qubits == 3
    assert st.measure([0])[0]
    st.project_unnormalized(0, False)
    with pytest.raises(ValueError):
        make_toffoli_state(nc).project_unnormalized(3, True)


def test_empty_inputs(nc):
    """Zero strings, zero Pauli strings, zero shots and no projected qubits:
    empty results (or the norm), no launch, as the reference's loops give
    for empty input (sampling.rs:91-94 returns at zero shots)."""
    import numpy as np
    from necstar import Batch, QuantumState
    qc = nc.QuantumCircuit(3)
    qc.apply_h(0)
    qc.apply_t(0)
    qc.apply_cx(0, 1)
    st = QuantumState.from_circuit(qc)
    assert st.amplitudes(np.zeros((0, 1), np.uint64)).shape == (0,)
    vals, pt, rr = st.amplitudes(np.zeros((0, 1), np.uint64), per_term=True)
    assert vals.shape == (0,) and pt.shape == (st.stabilizer_rank, 0) and rr.shape == (st.stabilizer_rank,)
    assert st.exp_values([]).shape == (0,)
    assert st.sample([0, 1], 0, seed=1) == {}
    assert abs(st.projection_norms([], [])[0] - st.norm() ** 2) <= 1e-12
    loss, grad = st.grad_exp_values([], [])
    assert loss == 0.0 and not np.any(grad)
    norms, values = Batch([st], [[]]).run()
    assert abs(norms[0] - st.norm()) <= 1e-12 and len(values[0]) == 0
