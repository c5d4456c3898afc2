"""The synthetic workloads at their real shapes on the GPU.

  cfg1  sample(range(64), 1024, seed=1001) of the 64-qubit, 24-T state in
        both modes (reference order, exact): sampled outcomes have a nonzero
        amplitude (oracle), and every qubit's sampled marginal agrees with
        exp_value(Z_q) within 4 sigma;
  cfg2  all 256 circuits through necstar.Batch (autotuned, HIP graph) against
        the plain per-call norm() / exp_value() and, on the smallest
        circuits, the oracle, within 1e-12 relative;
  cfg3  the 1,500-Clifford, 30-T circuit with its 128 weight-2..6 strings:
        every expectation value and every gradient entry is exactly 0 because
        every pair value <phi_j|P_k phi_i> vanishes (each string anticommutes
        with a stabilizer all terms share); the pair values are checked
        against the oracle's exact pair scalars on sampled (i, j, k) and in
        full on the GPU for a few strings.
Reference semantics: sampling.rs:13-189, norm.rs:8-35, exp_value.rs:7-61.
"""
import ctypes as C
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


def pauli_str(ops):
    return "".join("IXYZ"[x] for x in reversed(ops))


@pytest.mark.timeout(600)
def test_cfg1_sampling_support_and_marginals(nx):
    cfg = WL.config(1)
    n = cfg["n"]
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    say(f"cfg1 sampling chi={st.stabilizer_rank}")
    qs = list(range(n))
    counts_x = st.sample(qs, cfg["shots"], seed=cfg["sample_seed"], exact=True)
    assert counts_x == st.sample(qs, cfg["shots"], seed=cfg["sample_seed"], exact=True)
    say("cfg1 reference-order sampling")
    counts_r = st.sample(qs, cfg["shots"], seed=cfg["sample_seed"])
    o = O.State.compile(n, cfg["gates"])
    zs = [nx.PauliString.from_str(f"Z{q}") for q in range(n)]
    ez = st.exp_values(zs)
    rng = np.random.default_rng(17)
    for mode, counts in (("exact", counts_x), ("reference order", counts_r)):
        assert sum(counts.values()) == cfg["shots"]
        # (a) sampled outcomes lie in the support: nonzero amplitude by the oracle
        keys = sorted(counts)
        for key in [keys[int(i)] for i in rng.choice(len(keys), min(32, len(keys)), replace=False)]:
            x = np.array([1 if ch == "1" else 0 for ch in key], np.uint8)  # char q <-> qubit q
            assert abs(o.amplitude(x)) > 0.0, (mode, key)
        # (b) single-qubit marginals against <Z_q> (the state has norm 1)
        shots = cfg["shots"]
        ones = np.zeros(n)
        for key, c in counts.items():
            ones += c * np.array([ch == "1" for ch in key])
        p1 = (1.0 - ez) / 2.0
        sigma = np.sqrt(p1 * (1.0 - p1) / shots)
        dev = np.abs(ones / shots - p1)
        say(f"cfg1 {mode}: largest marginal deviation {np.max(dev / np.maximum(sigma, 1e-300)):.2f} sigma, "
            f"{len(counts)} distinct outcomes")
        assert np.all(dev <= 4.0 * sigma + 1e-12), [(mode, q, ones[q] / shots, p1[q])
                                                     for q in np.nonzero(dev > 4 * sigma)[0]]


def test_sample_reports_needed_capacity(nx):
    """nx_state_sample with too small an output buffer fails, reports the
    number of distinct outcomes, and succeeds with that capacity; the result
    equals QuantumState.sample's for the same seed."""
    from necstar import _lib
    L = _lib.lib()
    rng = np.random.default_rng(1501)
    n = 10
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, WL.random_circuit(rng, n, 100, 4)))
    qargs = (C.c_int32 * n)(*range(n))
    seed = (C.c_uint8 * 32)(*([7] + [0] * 31))
    nout = C.c_int64(-1)
    bits = np.zeros(2, np.uint64)
    counts = np.zeros(1, np.uint64)
    P64 = C.POINTER(C.c_uint64)
    code = L.nx_state_sample(st._h, n, qargs, 400, seed, 1, bits.ctypes.data_as(P64), counts.ctypes.data_as(P64),
                             C.byref(nout))
    assert code == _lib.NX_E_VALUE and "too small" in L.nx_last_error().decode()
    k = nout.value
    assert k > 1
    bits = np.zeros(2 * k, np.uint64)
    counts = np.zeros(k, np.uint64)
    _lib.check(L.nx_state_sample(st._h, n, qargs, 400, seed, k, bits.ctypes.data_as(P64),
                                 counts.ctypes.data_as(P64), C.byref(nout)))
    assert nout.value == k and int(counts.sum()) == 400
    ref = st.sample(list(range(n)), 400, seed=7)
    got = {"".join("1" if (int(bits[2 * i]) >> j) & 1 else "0" for j in range(n)): int(counts[i]) for i in range(k)}
    assert got == ref


@pytest.mark.timeout(900)
def test_cfg2_batch_vs_plain_and_oracle(nx):
    cfg = WL.config(2)
    circs = cfg["circuits"]
    states = []
    for i, c in enumerate(circs):
        states.append(nx.QuantumState.from_circuit(WL.to_circuit(c["n"], c["gates"])))
        if i % 64 == 63:
            say(f"cfg2 compiled {i + 1} circuits")
    chis = [s.stabilizer_rank for s in states]
    say(f"cfg2: chi max {max(chis)}, mean {np.mean(chis):.0f}")
    paulis = [[nx.PauliString.from_str(pauli_str(p)) for p in c["paulis"]] for c in circs]
    b = nx.Batch(states, paulis, autotune=True, graph=True)
    J, waves, vw = b.shape
    say(f"cfg2 batch shape J={J} waves={waves} variant words={vw}")
    norms, vals = b.run()
    norms2, vals2 = b.run()  # a second graph replay gives the same values
    assert norms == norms2 and all(np.array_equal(a, c) for a, c in zip(vals, vals2))
    K = 4
    for i, (st, c) in enumerate(zip(states, circs)):
        nrm = st.norm()
        fp64.check_lazy(norms[i], nrm, lambda _k: np.sqrt(fp64.state_scale(st, st)), f"cfg2 norm {i}")
        plain = np.array([st.exp_value(p) for p in paulis[i][:K]])
        fp64.check_lazy(vals[i][:K], plain, lambda k: fp64.exp_scale(st, c["paulis"][k]), f"cfg2 values {i}")
        assert abs(nrm - 1.0) <= 1e-12  # compiled unitary circuits
        if i % 64 == 63:
            say(f"cfg2 plain checks {i + 1}; worst relative error {fp64.worst}")
    # the oracle on the circuits with the fewest T gates (smallest chi), all 64 strings
    small = sorted(range(len(circs)), key=lambda i: (chis[i], circs[i]["n"]))[:3]
    for i in small:
        c = circs[i]
        o = O.State.compile(c["n"], c["gates"])
        assert o.chi == chis[i]
        say(f"cfg2 oracle circuit {i}: n={c['n']} chi={o.chi}")
        fp64.check(norms[i], o.norm(), np.sqrt(fp64.state_scale(states[i], states[i])), f"cfg2 oracle norm {i}")
        ref = np.array([o.exp_value(("dense", p)) for p in c["paulis"]])
        sc = np.array([fp64.exp_scale(states[i], p) for p in c["paulis"]])
        fp64.check(vals[i], ref, sc, f"cfg2 oracle values {i}")


@pytest.mark.timeout(900)
def test_cfg3_workload_zero_pairs(nx):
    cfg = WL.config(3)
    n = cfg["n"]
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, cfg["gates"]))
    chi = st.stabilizer_rank
    rng = np.random.default_rng(cfg["perturb_seed"])
    c = st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi))
    st.set_coefficients(c)
    paulis = [nx.PauliString.from_str(pauli_str(p)) for p in cfg["paulis"]]
    say(f"cfg3 chi={chi}, {len(paulis)} strings")
    ev = st.exp_values(paulis)
    assert np.all(ev == 0.0), ev[ev != 0.0]
    loss, grad = st.grad_exp_values(paulis, cfg["weights"])
    assert loss == 0.0 and np.all(grad == 0.0)
    # every pair value of a few strings vanishes (exact tables on the GPU)
    for k in (0, 1, 2, 3):
        tab = st.cross_pair_table(fp64.pauli_image(st, cfg["paulis"][k]))
        assert np.all(tab == np.uint32(0xFFFFFFFF)), (k, int(np.sum(tab != np.uint32(0xFFFFFFFF))))
    # ... and for sampled (i, j, k) in the oracle's exact pair scalars
    o = O.State.compile(n, cfg["gates"])
    assert o.chi == chi
    for _ in range(96):
        i, j = (int(v) for v in rng.integers(0, chi, 2))
        k = int(rng.integers(len(paulis)))
        pi = o.term(i).clone()
        for q, op in enumerate(cfg["paulis"][k]):
            if op:
                pi.apply(op, q)
        (zero, _, _), _ = o.term(j).inner_product(pi)
        assert zero, (i, j, k)
    # the norm is not degenerate: the forms do overlap
    assert abs(st.norm() - 1.0) < 1e-2
