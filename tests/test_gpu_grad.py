"""Coefficient backward (one pass over tiles of pairs, y = G c) on the GPU.

L(c) = sum_k w_k exp_value(P_k) with complex128 coefficients c (the
reference's StabilizerDecomposedState<Complex64> semantics).  The gradient in
PyTorch's convention, 2 |g|^2 sum_k w_k G_k c with (G_k)_ji = <phi_j|P_k phi_i>,
is checked against (1) a CPU FP64 Gram gradient built from the oracle's exact
pair scalars, (2) central finite differences of the GPU forward (h = 1e-6; L is
quadratic in (Re c, Im c), so the difference quotient is exact up to rounding),
(3) torch.autograd through necstar.autograd.ExpValueLoss."""
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


def _setup(nx, seed, n, nc, nt, npauli):
    rng = np.random.default_rng(seed)
    gates = WL.random_circuit(rng, n, nc, nt, gate_set=tuple(range(11)))
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    o = O.State.compile(n, gates)
    chi = st.stabilizer_rank
    c = st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi))
    st.set_coefficients(c)
    o.set_complex_coeffs(c)
    ops = WL.random_paulis(rng, n, npauli, 1, min(6, n))
    paulis = [nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in ops]
    w = rng.standard_normal(npauli)
    return st, o, c, ops, paulis, w


def test_grad_vs_oracle_gram(nx):
    st, o, c, ops, paulis, w = _setup(nx, 31, 6, 60, 5, 6)
    chi = o.chi
    say(f"grad vs oracle chi={chi}")
    loss, grad = st.grad_exp_values(paulis, w)
    y = np.zeros(chi, complex)
    for k, p in enumerate(ops):
        G = np.zeros((chi, chi), complex)
        for i in range(chi):
            pi = o.term(i).clone()
            for q, op in enumerate(p):
                if op:
                    pi.apply(op, q)
            for j in range(chi):
                G[j, i] = o.term(j).inner_product(pi)[1]
        y += w[k] * (G @ c)
    g2 = abs(o.global_factor) ** 2
    ref = 2 * g2 * y
    assert np.abs(ref).max() > 1e-3, "the check needs a non-trivial gradient"
    # normwise relative: every entry within 1e-12 of the largest
    assert np.abs(grad - ref).max() <= 1e-12 * np.abs(ref).max()
    ev_o = np.array([o.exp_value(("dense", p)) for p in ops])
    sc = np.array([fp64.exp_scale(st, p) for p in ops])
    fp64.check(loss, float(np.dot(w, ev_o)), float(np.abs(w) @ sc), "loss")
    fp64.check(st.exp_values(paulis), ev_o, sc, "forward values")


def test_grad_vs_central_differences(nx):
    st, o, c, ops, paulis, w = _setup(nx, 32, 8, 80, 6, 4)
    chi = st.stabilizer_rank
    say(f"grad vs FD chi={chi}")
    _, grad = st.grad_exp_values(paulis, w)
    assert np.abs(grad).max() > 1e-3, "the check needs a non-trivial gradient"
    h = 1e-6

    def L(cc):
        s2 = st.clone()
        s2.set_coefficients(cc)
        return float(np.dot(w, s2.exp_values(paulis)))

    for i in range(min(chi, 8)):
        for d in (1.0, 1j):
            e = np.zeros(chi, complex)
            e[i] = d
            fd = (L(c + h * e) - L(c - h * e)) / (2 * h)
            an = grad[i].real if d == 1.0 else grad[i].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (i, d, fd, an)


def test_torch_autograd_function(nx):
    torch = pytest.importorskip("torch")
    from necstar.autograd import exp_value_loss
    st, o, c, ops, paulis, w = _setup(nx, 33, 7, 70, 5, 5)
    # strings with nonzero expectation, so that the gradient is not zero
    rng = np.random.default_rng(34)
    cands = WL.random_paulis(rng, st.num_qubits, 60, 1, 3)
    ev = st.exp_values([nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in cands])
    ops = [p for p, e in zip(cands, ev) if abs(e) > 1e-3][:5]
    assert len(ops) >= 2
    paulis = [nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in ops]
    w = w[: len(ops)]
    coeffs = torch.tensor(c, dtype=torch.complex128, requires_grad=True)
    loss = exp_value_loss(coeffs, st, paulis, w)
    loss.backward()
    ref_loss, ref_grad = st.grad_exp_values(paulis, w)
    sc = float(np.abs(w) @ np.array([fp64.exp_scale(st, p) for p in ops]))
    fp64.check(loss.item(), ref_loss, sc, "autograd loss")
    assert np.abs(ref_grad).max() > 1e-3
    assert np.abs(coeffs.grad.cpu().numpy() - ref_grad).max() <= 1e-12 * np.abs(ref_grad).max()


def test_config3_shape_gradient_fd(nx):
    """cfg3 scale (n=48, 30 T gates): the Gram gradient matches central
    differences on sampled coordinates.  On the cfg3 circuit itself (1500
    scrambling Cliffords) every low-weight Pauli has zero expectation: it
    anticommutes with a stabilizer shared by all terms, so every pair value
    vanishes.  The check therefore uses a shallower circuit (120 Cliffords,
    same n and t) and single-qubit strings with nonzero expectation."""
    cfg = WL.config(3)
    n = cfg["n"]
    gates = WL.random_circuit(np.random.default_rng(3003), n, 120, 30)
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, gates))
    chi = st.stabilizer_rank
    rng = np.random.default_rng(cfg["perturb_seed"])
    c = st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi))
    st.set_coefficients(c)
    cands = [nx.PauliString.from_str(f"{op}{q}") for q in range(n) for op in "ZXY"]
    ev = st.exp_values(cands)
    paulis = [p for p, e in zip(cands, ev) if abs(e) > 1e-3][:8]
    assert len(paulis) >= 2
    w = cfg["weights"][: len(paulis)]
    say(f"cfg3 chi={chi} paulis={[str(p) for p in paulis]}")
    loss, grad = st.grad_exp_values(paulis, w)
    assert np.abs(grad).max() > 1e-3
    fp64.check(loss, float(np.dot(w, st.exp_values(paulis))), abs(loss) + 1.0, "cfg3-shape loss")
    h = 1e-6

    def L(cc):
        s2 = st.clone()
        s2.set_coefficients(cc)
        return float(np.dot(w, s2.exp_values(paulis)))

    for i in [int(np.argmax(np.abs(grad.real))), int(rng.integers(chi))]:
        e = np.zeros(chi, complex)
        e[int(i)] = 1.0
        fd = (L(c + h * e) - L(c - h * e)) / (2 * h)
        assert abs(fd - grad[int(i)].real) <= 1e-6 * max(1.0, abs(grad[int(i)].real))


def test_torch_autograd_on_device(nx):
    """The autograd function with CUDA coefficient tensors (device entry
    point on PyTorch's stream): the same loss and gradient as the host entry
    point, and the gradient of a loss built on top of it by torch."""
    torch = pytest.importorskip("torch")
    from necstar.autograd import exp_value_loss
    st, o, c, ops, paulis, w = _setup(nx, 35, 7, 70, 5, 5)
    rng = np.random.default_rng(36)
    cands = WL.random_paulis(rng, st.num_qubits, 60, 1, 3)
    ev = st.exp_values([nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in cands])
    ops = [p for p, e in zip(cands, ev) if abs(e) > 1e-3][:6]
    assert len(ops) >= 2
    paulis = [nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in ops]
    w = rng.standard_normal(len(ops))
    ref_loss, ref_grad = st.grad_exp_values(paulis, w)
    coeffs = torch.tensor(c, dtype=torch.complex128, device="cuda", requires_grad=True)
    loss = exp_value_loss(coeffs, st, paulis, w)
    assert loss.device.type == "cuda"
    (3.0 * loss).backward()
    assert abs(loss.item() - ref_loss) <= 1e-12 * abs(ref_loss)
    assert np.abs(ref_grad).max() > 1e-3
    assert np.abs(coeffs.grad.cpu().numpy() - 3.0 * ref_grad).max() <= 1e-12 * 3.0 * np.abs(ref_grad).max()


def _euler_check(c, loss, grad, what):
    """L = |g|^2 c^H G c is homogeneous of degree 2 in (c, conj c), so
    sum_j conj(c_j) grad_j = 2 L exactly (G Hermitian: no imaginary part)."""
    z = np.vdot(c, grad)
    scale = float(np.sum(np.abs(c) * np.abs(grad)))
    assert abs(z.real - 2.0 * loss) <= 1e-11 * scale, (what, z, loss, scale)
    assert abs(z.imag) <= 1e-11 * scale, (what, z, scale)


@pytest.mark.timeout(900)
def test_grad_beyond_dense_limit(nx):
    """chi above 32,768 (where a dense chi x chi Gram matrix stops fitting):
    n = 64 with 40 T gates.  Checked by size-independent properties: the
    loss equals sum_k w_k exp_value(P_k) from the forward kernel, the Euler
    identity of the quadratic form, and a central difference along a random
    direction (exact for a quadratic up to rounding)."""
    rng = np.random.default_rng(3101)
    n = 64
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, WL.random_circuit(rng, n, 300, 40)))
    chi = st.stabilizer_rank
    say(f"grad beyond the dense limit: chi={chi}")
    assert chi > 32768
    c = st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi))
    st.set_coefficients(c)
    cands = [nx.PauliString.from_str(f"Z{q}") for q in range(n)] + [nx.PauliString.from_str(f"X{q}") for q in range(n)]
    ev = st.exp_values(cands)
    sel = [k for k in np.argsort(-np.abs(ev)) if abs(ev[k]) > 1e-4][:3]
    assert len(sel) >= 1, "the check needs strings with nonzero expectation"
    paulis = [cands[k] for k in sel]
    w = rng.standard_normal(len(sel))
    loss, grad = st.grad_exp_values(paulis, w)
    assert np.abs(grad).max() > 1e-6
    # two FP64 sums of ~chi^2 / 2 = 9e8 complex-coefficient pair terms in
    # different orders (tiles vs units): consistent to 1e-10 relative
    ref = float(np.dot(w, ev[sel]))
    assert abs(loss - ref) <= 1e-10 * abs(ref), (loss, ref)
    _euler_check(c, loss, grad, "chi > 32768")
    d = rng.standard_normal(chi) + 1j * rng.standard_normal(chi)
    d /= np.linalg.norm(d)
    h = 1e-3

    def L(cc):
        s2 = st.clone()
        s2.set_coefficients(cc)
        return float(np.dot(w, s2.exp_values(paulis)))

    fd = (L(c + h * d) - L(c - h * d)) / (2 * h)
    an = float(np.real(np.vdot(d, grad)))
    assert abs(fd - an) <= 1e-7 * max(abs(an), np.abs(grad).max()), (fd, an)


def test_cfg3_gradient_properties(nx):
    """cfg3's shape (n=48, 30 T gates) on a non-degenerate circuit (150
    Cliffords, strings with nonzero expectation, 128 of them): the Euler
    identity and the loss of the forward pass, with more than 64 strings
    (two variant words)."""
    cfg = WL.config(3)
    n = cfg["n"]
    rng = np.random.default_rng(3201)
    st = nx.QuantumState.from_circuit(WL.to_circuit(n, WL.random_circuit(rng, n, 150, 30)))
    chi = st.stabilizer_rank
    c = st.coefficients() + 1e-3 * (rng.standard_normal(chi) + 1j * rng.standard_normal(chi))
    st.set_coefficients(c)
    cands = WL.random_paulis(rng, n, 600, 1, 3)
    ev = st.exp_values([nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in cands])
    good = [p for p, e in zip(cands, ev) if abs(e) > 1e-3]
    ops = (good + WL.random_paulis(rng, n, 128, 2, 6))[:128]
    paulis = [nx.PauliString.from_str("".join("IXYZ"[x] for x in reversed(p))) for p in ops]
    w = rng.standard_normal(128)
    say(f"cfg3 shape: chi={chi}, {len(good)} strings with nonzero expectation")
    assert len(good) >= 8
    loss, grad = st.grad_exp_values(paulis, w)
    ev2 = st.exp_values(paulis)
    assert np.abs(grad).max() > 1e-3
    sc = float(np.abs(w) @ np.array([fp64.exp_scale(st, p) for p in ops[:8]] + [0.0] * 120)) + abs(loss)
    fp64.check(loss, float(np.dot(w, ev2)), sc, "cfg3-shape loss vs forward")
    _euler_check(c, loss, grad, "cfg3 shape")
