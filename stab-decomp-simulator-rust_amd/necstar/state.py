"""QuantumState (python/src/state.rs:12-222) over the HIP C-ABI."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib, loaded
from .gate import QuantumGate, _usize
from .pauli_string import PauliString


def parse_seed(seed):
    """python/src/utils.rs:4-39: None, an int (< 2^64 packed into 8
    little-endian bytes, larger ints into 32), or exactly 32 bytes."""
    if seed is None:
        return None
    if isinstance(seed, int):
        if 0 <= seed < 2 ** 64:
            return seed.to_bytes(8, "little") + bytes(24)
        return seed.to_bytes(32, "little")
    if isinstance(seed, (bytes,)):
        if len(seed) != 32:
            raise ValueError(f"Seed must be 32 bytes, got {len(seed)}")
        return bytes(seed)
    raise TypeError("Seed must be an integer or 32 bytes")


def _seed_ptr(seed):
    b = parse_seed(seed)
    if b is None:
        return None, None
    buf = (C.c_uint8 * 32).from_buffer_copy(b)
    return buf, C.cast(buf, C.POINTER(C.c_uint8))


def _i32_array(values):
    arr = np.ascontiguousarray(np.asarray(values, dtype=np.int64))
    if arr.size and (arr.min() < -(2 ** 31) or arr.max() >= 2 ** 31):
        raise OverflowError("qubit index out of range")
    arr = arr.astype(np.int32)
    if arr.size == 0:
        arr = np.zeros(1, np.int32)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_int32))


def _gate_rows(gates):
    rows = np.zeros((max(len(gates), 1), 4), np.int32)
    for i, g in enumerate(gates):
        rows[i] = g._row()
    return rows


class QuantumState:
    __slots__ = ("_h",)

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                L = loaded()
            except Exception:  # interpreter shutdown: module state is gone
                return
            if L is not None:
                L.nx_state_free(h)

    # -- construction --------------------------------------------------------
    @staticmethod
    def from_circuit(circuit) -> "QuantumState":
        rows = _gate_rows(circuit.gates)
        out = C.c_void_p()
        check(lib().nx_state_compile(circuit.num_qubits, circuit.num_gates,
                                     rows.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(out)))
        return QuantumState(out.value)

    def clone(self) -> "QuantumState":
        out = C.c_void_p()
        check(lib().nx_state_clone(self._h, C.byref(out)))
        return QuantumState(out.value)

    # -- getters -------------------------------------------------------------
    @property
    def num_qubits(self) -> int:
        v = C.c_int32()
        check(lib().nx_state_num_qubits(self._h, C.byref(v)))
        return v.value

    @property
    def stabilizer_rank(self) -> int:
        v = C.c_int64()
        check(lib().nx_state_stabilizer_rank(self._h, C.byref(v)))
        return v.value

    def __str__(self):
        return f"QuantumState(num_qubits={self.num_qubits}, stabilizer_rank={self.stabilizer_rank})"

    # -- read-only computations ----------------------------------------------
    def to_statevector(self):
        n = self.num_qubits
        if n > 28:
            out = np.zeros(2, np.float64)
        else:
            out = np.zeros(2 << n, np.float64)
        check(lib().nx_state_statevector(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return [complex(a, b) for a, b in zip(out[0::2], out[1::2])]

    def statevector_array(self) -> np.ndarray:
        """to_statevector as a numpy complex128 array (no list conversion)."""
        n = self.num_qubits
        out = np.zeros(2 << min(n, 28), np.float64)
        check(lib().nx_state_statevector(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.view(np.complex128)

    def inner_product(self, other: "QuantumState") -> complex:
        re, im = C.c_double(), C.c_double()
        check(lib().nx_state_inner_product(self._h, other._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def norm(self) -> float:
        v = C.c_double()
        check(lib().nx_state_norm(self._h, C.byref(v)))
        return v.value

    def exp_value(self, pauli_string: PauliString) -> float:
        ops, opp = _i32_array(pauli_string._ops)
        qs, qsp = _i32_array(pauli_string._qubits if not pauli_string._dense else [])
        v = C.c_double()
        check(lib().nx_state_exp_value(self._h, 1 if pauli_string._dense else 0, len(pauli_string._ops),
                                       opp, qsp, C.byref(v)))
        return v.value

    def _dense_ops(self, paulis):
        """Dense little-endian op rows [len(paulis), n] with the length checks
        of exp_value.rs:9-27."""
        n = self.num_qubits
        rows = np.zeros((max(len(paulis), 1), max(n, 1)), np.int32)
        for k, p in enumerate(paulis):
            if p._dense:
                if len(p._ops) != n:
                    raise ValueError(f"Invalid Pauli string length: expected {n}, found {len(p._ops)}.")
                rows[k, :n] = p._ops
            else:
                mx = max(p._qubits) if p._qubits else 0
                if mx >= n:
                    raise ValueError(f"Invalid Pauli string length: expected {n}, found {mx + 1}.")
                for o, q in zip(p._ops, p._qubits):
                    rows[k, q] = o
        return rows

    def projection_norms(self, look, fixed=()):
        """norm_squared() after projecting onto outcome b of qubit q for each
        (q, b) in `fixed` and onto every outcome of the qubits in `look`
        (API addition; k-th entry: outcome (k >> i) & 1 on look[i])."""
        look = [int(q) for q in look]
        fq = np.array([int(q) for q, _ in fixed] + [0], np.int32)
        fb = np.array([1 if b else 0 for _, b in fixed] + [0], np.int32)
        lq = np.array(look + [0], np.int32)
        out = np.zeros(1 << len(look), np.float64)
        P = C.POINTER(C.c_int32)
        check(lib().nx_state_projection_norms(self._h, len(fixed), fq.ctypes.data_as(P), fb.ctypes.data_as(P),
                                              len(look), lq.ctypes.data_as(P),
                                              out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def uniform_prefix(self, qargs, fixed=()) -> int:
        """The largest D such that the Z-basis marginal of qargs[:D] is
        provably uniform, given outcome b of qubit q for each (q, b) in
        `fixed` (API addition; sample() skips those levels)."""
        arr, ptr = _i32_array(qargs)
        out = C.c_int32()
        if not fixed:
            check(lib().nx_state_uniform_prefix(self._h, len(qargs), ptr, C.byref(out)))
        else:
            fq, fqp = _i32_array([q for q, _ in fixed])
            fb, fbp = _i32_array([1 if b else 0 for _, b in fixed])
            check(lib().nx_state_uniform_prefix_given(self._h, len(fixed), fqp, fbp, len(qargs), ptr,
                                                      C.byref(out)))
        return out.value

    def exp_values(self, paulis) -> np.ndarray:
        """exp_value for a list of PauliStrings (one call through the C-ABI)."""
        paulis = list(paulis)
        rows = self._dense_ops(paulis)
        out = np.zeros(max(len(paulis), 1), np.float64)
        check(lib().nx_state_exp_values(self._h, len(paulis), rows.ctypes.data_as(C.POINTER(C.c_int32)),
                                        out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[: len(paulis)]

    def grad_exp_values(self, paulis, weights):
        """L = sum_k w_k exp_value(P_k) and dL/dc in PyTorch's convention for
        complex leaves (2 dL/d conj(c) = 2 |g|^2 sum_k w_k G_k c)."""
        paulis = list(paulis)
        w = np.ascontiguousarray(np.asarray(weights, np.float64))
        if w.shape != (len(paulis),):
            raise ValueError("one weight per Pauli string is required")
        rows = self._dense_ops(paulis)
        chi = self.stabilizer_rank
        grad = np.zeros(2 * max(chi, 1), np.float64)
        loss = C.c_double()
        check(lib().nx_state_grad_exp_values(self._h, len(paulis), rows.ctypes.data_as(C.POINTER(C.c_int32)),
                                             w.ctypes.data_as(C.POINTER(C.c_double)), C.byref(loss),
                                             grad.ctypes.data_as(C.POINTER(C.c_double))))
        return loss.value, grad[: 2 * chi].view(np.complex128).copy()

    def amplitudes(self, bitstrings, per_term: bool = False):
        """<x|psi> for a batch of bitstrings (API addition).  bitstrings:
        uint64 array [B] (n <= 64) or [B, words], little-endian (qubit j =
        bit j).  Returns complex128 [B]; with per_term, also the exact
        per-term phases (uint8 [chi, B], 255 = zero) and exponents r [chi]."""
        n = self.num_qubits
        words = max(1, (n + 63) // 64)
        b = np.ascontiguousarray(np.asarray(bitstrings, dtype=np.uint64))
        if b.ndim == 1:
            b = b.reshape(-1, 1)
        if b.shape[1] < words:
            b = np.concatenate([b, np.zeros((b.shape[0], words - b.shape[1]), np.uint64)], axis=1)
        b = np.ascontiguousarray(b)
        nb = b.shape[0]
        out = np.zeros(2 * max(nb, 1), np.float64)
        chi = self.stabilizer_rank
        pt = np.zeros((max(chi, 1), max(nb, 1)), np.uint8) if per_term else None
        rr = np.zeros(max(chi, 1), np.int32) if per_term else None
        check(lib().nx_state_amplitudes(
            self._h, nb, b.ctypes.data_as(C.POINTER(C.c_uint64)), b.shape[1],
            out.ctypes.data_as(C.POINTER(C.c_double)),
            pt.ctypes.data_as(C.POINTER(C.c_uint8)) if per_term else None,
            rr.ctypes.data_as(C.POINTER(C.c_int32)) if per_term else None))
        vals = out[: 2 * nb].view(np.complex128)
        if per_term:
            return vals, pt[:chi, :nb], rr[:chi]
        return vals

    # -- measurement ---------------------------------------------------------
    def measure(self, qargs, seed=None):
        qargs = [_usize(q) for q in qargs]
        arr, p = _i32_array(qargs)
        _buf, sp = _seed_ptr(seed)
        out = (C.c_uint8 * max(len(qargs), 1))()
        check(lib().nx_state_measure(self._h, len(qargs), p, sp, out))
        return [bool(out[i]) for i in range(len(qargs))]

    def measure_all(self, seed=None):
        return self.measure(list(range(self.num_qubits)), seed)

    def sample(self, qargs, shots, seed=None, exact=False):
        """sampling.rs:13-189.  exact=False: node probabilities as the
        reference's FP64 sums (a seed gives the reference's outcomes);
        exact=True: exact probabilities, certified uniform levels and exact
        tails (same distribution, much faster on large states; seeded
        outcomes can differ where the reference's sums round)."""
        qargs = [_usize(q) for q in qargs]
        shots = _usize(shots)
        arr, p = _i32_array(qargs)
        _buf, sp = _seed_ptr(seed)
        if sp is None:  # one entropy seed for both attempts below
            _buf, sp = _seed_ptr(os.urandom(32))
        # a first capacity that fits most calls; when more distinct outcomes
        # occur, the library reports how many and the call is repeated with
        # the same seed (the reference's ShotCount has no size limit)
        cap = min(shots, 1 << min(len(qargs), 20)) + 1
        nout = C.c_int64()
        for _attempt in range(2):
            bits = np.zeros(2 * cap, np.uint64)
            counts = np.zeros(cap, np.uint64)
            nout.value = 0
            fn = lib().nx_state_sample_exact if exact else lib().nx_state_sample
            code = fn(self._h, len(qargs), p, shots, sp, cap,
                                         bits.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         counts.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(nout))
            if code == 0 or nout.value <= cap:
                break
            cap = nout.value
        check(code)
        res = {}
        k = len(qargs)
        for i in range(nout.value):
            word = int(bits[2 * i]) | (int(bits[2 * i + 1]) << 64)
            key = "".join("1" if (word >> j) & 1 else "0" for j in range(k))
            res[key] = int(counts[i])
        return res

    # -- state updates -------------------------------------------------------
    def project_normalized(self, qubit: int, outcome: bool) -> None:
        check(lib().nx_state_project(self._h, _usize(qubit), int(bool(outcome)), 1))

    def project_unnormalized(self, qubit: int, outcome: bool) -> None:
        check(lib().nx_state_project(self._h, _usize(qubit), int(bool(outcome)), 0))

    def discard(self, qubit: int) -> None:
        check(lib().nx_state_discard(self._h, _usize(qubit)))

    def apply_gate(self, gate: QuantumGate) -> None:
        self.apply_gates([gate])

    def apply_gates(self, gates) -> None:
        gates = list(gates)
        rows = _gate_rows(gates)
        check(lib().nx_state_apply_gates(self._h, len(gates), rows.ctypes.data_as(C.POINTER(C.c_int32))))

    def apply_x(self, qubit): self.apply_gate(QuantumGate.x(qubit))

    def apply_y(self, qubit): self.apply_gate(QuantumGate.y(qubit))

    def apply_z(self, qubit): self.apply_gate(QuantumGate.z(qubit))

    def apply_h(self, qubit): self.apply_gate(QuantumGate.h(qubit))

    def apply_s(self, qubit): self.apply_gate(QuantumGate.s(qubit))

    def apply_sdg(self, qubit): self.apply_gate(QuantumGate.sdg(qubit))

    def apply_sqrt_x(self, qubit): self.apply_gate(QuantumGate.sqrt_x(qubit))

    def apply_sqrt_xdg(self, qubit): self.apply_gate(QuantumGate.sqrt_xdg(qubit))

    def apply_cx(self, control, target): self.apply_gate(QuantumGate.cx(control, target))

    def apply_cz(self, qarg1, qarg2): self.apply_gate(QuantumGate.cz(qarg1, qarg2))

    def apply_swap(self, qarg1, qarg2): self.apply_gate(QuantumGate.swap(qarg1, qarg2))

    # -- parity / coefficient access (additions) -------------------------------
    def export_term(self, i: int):
        n = self.num_qubits
        G = np.zeros((n, n), np.uint8); F = np.zeros((n, n), np.uint8); M = np.zeros((n, n), np.uint8)
        g = np.zeros(max(n, 1), np.uint8); v = np.zeros(max(n, 1), np.uint8); s = np.zeros(max(n, 1), np.uint8)
        ph = np.zeros(1, np.uint8)
        P = C.POINTER(C.c_uint8)
        check(lib().nx_state_export_term(self._h, i, *(a.ctypes.data_as(P) for a in (G, F, M, g, v, s, ph))))
        return dict(G=G, F=F, M=M, gamma=g[:n], v=v[:n], s=s[:n], phase=int(ph[0]))

    def scalar_coefficients(self):
        chi = self.stabilizer_rank
        z = np.zeros(max(chi, 1), np.uint8); k = np.zeros(max(chi, 1), np.uint8); r = np.zeros(max(chi, 1), np.int64)
        check(lib().nx_state_get_scalar_coefficients(self._h, z.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                     k.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                     r.ctypes.data_as(C.POINTER(C.c_int64))))
        return z[:chi].astype(bool), k[:chi], r[:chi]

    def coefficients(self) -> np.ndarray:
        chi = self.stabilizer_rank
        out = np.zeros(2 * max(chi, 1), np.float64)
        check(lib().nx_state_get_coefficients(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[: 2 * chi].view(np.complex128).copy()

    def set_coefficients(self, c) -> None:
        c = np.ascontiguousarray(np.asarray(c, np.complex128))
        if c.shape != (self.stabilizer_rank,):
            raise ValueError("one coefficient per stabilizer term is required")
        buf = c.view(np.float64)
        check(lib().nx_state_set_coefficients(self._h, buf.ctypes.data_as(C.POINTER(C.c_double))))

    @property
    def global_factor(self) -> complex:
        re, im = C.c_double(), C.c_double()
        check(lib().nx_state_global_factor(self._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def pair_table(self) -> np.ndarray:
        chi = self.stabilizer_rank
        out = np.zeros((max(chi, 1), max(chi, 1)), np.uint32)
        check(lib().nx_state_pair_table(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out[:chi, :chi]

    def cross_pair_table(self, other: "QuantumState") -> np.ndarray:
        """Exact <self_i|other_j> (uint32: 0xFFFFFFFF zero, else k | r << 8)."""
        ca, cb = self.stabilizer_rank, other.stabilizer_rank
        out = np.zeros((max(ca, 1), max(cb, 1)), np.uint32)
        check(lib().nx_state_cross_pair_table(self._h, other._h, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out.reshape(-1)[: ca * cb].reshape(ca, cb)
