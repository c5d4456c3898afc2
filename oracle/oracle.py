"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates the reference algorithms of
mutekichi/stab-decomp-simulator-rust in plain C (necstar_oracle.c).  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker / the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libnecstar_oracle.so")

# gate op codes (same numbering as include/necstar_hip.h)
H, X, Y, Z, S, SDG, SQRTX, SQRTXDG, CX, CZ, SWAP, T, TDG, CCX = range(14)
GATE_NAMES = ["H", "X", "Y", "Z", "S", "Sdg", "SqrtX", "SqrtXdg", "CX", "CZ",
              "Swap", "T", "Tdg", "CCX"]

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        IP, DP = C.POINTER(C.c_int), C.POINTER(C.c_double)
        U8P, U64P = C.POINTER(C.c_uint8), C.POINTER(C.c_uint64)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_chf_new": (P, [I, IP]),
            "orc_chf_free": (None, [P]),
            "orc_chf_clone": (P, [P]),
            "orc_chf_num_qubits": (I, [P]),
            "orc_chf_apply_gate": (I, [P, I, I, I]),
            "orc_chf_project": (I, [P, I, I, IP]),
            "orc_chf_discard": (I, [P, I]),
            "orc_chf_kron": (P, [P, P, IP]),
            "orc_chf_amplitude_at_zero": (I, [P, IP, IP, IP]),
            "orc_chf_amplitude": (I, [P, U8P, IP, IP, IP]),
            "orc_chf_inner_product": (I, [P, P, IP, IP, IP, DP, DP]),
            "orc_chf_export": (None, [P, U8P, U8P, U8P, U8P, U8P, U8P, U8P]),
            "orc_chf_statevector": (I, [P, DP]),
            "orc_compile": (P, [I, I, IP, IP]),
            "orc_state_free": (None, [P]),
            "orc_state_clone": (P, [P]),
            "orc_state_num_qubits": (I, [P]),
            "orc_state_rank": (I, [P]),
            "orc_state_is_complex": (I, [P]),
            "orc_state_term": (P, [P, I]),
            "orc_state_coeff_scalar": (None, [P, I, IP, IP, C.POINTER(C.c_int64)]),
            "orc_state_coeff_complex": (None, [P, I, DP, DP]),
            "orc_state_set_complex_coeffs": (None, [P, DP]),
            "orc_state_global_factor": (None, [P, DP, DP]),
            "orc_state_apply_gate": (I, [P, I, I, I, I]),
            "orc_state_norm_squared": (I, [P, DP]),
            "orc_state_norm": (I, [P, DP]),
            "orc_state_exp_value": (I, [P, I, I, IP, IP, DP]),
            "orc_state_inner_product": (I, [P, P, DP, DP]),
            "orc_state_project_unnormalized": (I, [P, I, I]),
            "orc_state_project_normalized": (I, [P, I, I]),
            "orc_state_discard": (I, [P, I]),
            "orc_state_statevector": (I, [P, DP]),
            "orc_state_amplitude": (I, [P, U8P, DP, DP, IP, IP, IP]),
            "orc_state_measure": (I, [P, I, IP, U8P, U8P]),
            "orc_state_sample": (I, [P, I, IP, C.c_uint64, U8P, I, U64P, U64P, IP]),
            "orc_rng_new": (P, [U8P]),
            "orc_rng_free": (None, [P]),
            "orc_rng_next_u32": (C.c_uint32, [P]),
            "orc_rng_next_u64": (C.c_uint64, [P]),
            "orc_rng_gen_f64": (D, [P]),
            "orc_rng_binomial": (C.c_uint64, [P, C.c_uint64, D]),
            "orc_chacha_block": (None, [C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64, I,
                                        C.POINTER(C.c_uint32)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class OracleError(ValueError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def _check(code: int):
    if code:
        raise OracleError(code, lib().orc_last_error().decode())


def _ip(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def _u8p(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8))


def seed_bytes(seed) -> bytes:
    """python/src/utils.rs:4-39 seed parsing (int or 32 bytes)."""
    if isinstance(seed, (bytes, bytearray)):
        if len(seed) != 32:
            raise ValueError(f"Seed must be 32 bytes, got {len(seed)}")
        return bytes(seed)
    if isinstance(seed, int) and 0 <= seed < 2 ** 64:
        return seed.to_bytes(8, "little") + bytes(24)
    if isinstance(seed, int):
        return seed.to_bytes(32, "little")
    raise TypeError("Seed must be an integer or 32 bytes")


class CHForm:
    """A single reference CH-form (stabilizer-ch-form-rust StabilizerCHForm)."""

    def __init__(self, ptr, owned=True):
        self.ptr = ptr
        self.owned = owned

    @classmethod
    def new(cls, n):
        err = C.c_int(0)
        p = lib().orc_chf_new(n, C.byref(err))
        _check(err.value)
        return cls(p)

    def __del__(self):
        if getattr(self, "owned", False) and self.ptr:
            lib().orc_chf_free(self.ptr)
            self.ptr = None

    def clone(self):
        return CHForm(lib().orc_chf_clone(self.ptr))

    @property
    def n(self):
        return lib().orc_chf_num_qubits(self.ptr)

    def apply(self, op, q0, q1=0):
        _check(lib().orc_chf_apply_gate(self.ptr, op, q0, q1))
        return self

    def project(self, q, outcome):
        det = C.c_int(0)
        _check(lib().orc_chf_project(self.ptr, q, int(outcome), C.byref(det)))
        return bool(det.value)

    def discard(self, q):
        _check(lib().orc_chf_discard(self.ptr, q))

    def kron(self, other):
        err = C.c_int(0)
        p = lib().orc_chf_kron(self.ptr, other.ptr, C.byref(err))
        _check(err.value)
        return CHForm(p)

    def amplitude(self, bits):
        b, bp = _u8p(bits)
        z, k, r = C.c_int(), C.c_int(), C.c_int()
        _check(lib().orc_chf_amplitude(self.ptr, bp, C.byref(z), C.byref(k), C.byref(r)))
        return (bool(z.value), k.value, r.value)

    def inner_product(self, other):
        z, k, r = C.c_int(), C.c_int(), C.c_int()
        re, im = C.c_double(), C.c_double()
        _check(lib().orc_chf_inner_product(self.ptr, other.ptr, C.byref(z), C.byref(k),
                                           C.byref(r), C.byref(re), C.byref(im)))
        return (bool(z.value), k.value, r.value), complex(re.value, im.value)

    def tables(self):
        n = self.n
        G = np.zeros((n, n), np.uint8); F = np.zeros((n, n), np.uint8)
        M = np.zeros((n, n), np.uint8); g = np.zeros(n, np.uint8)
        v = np.zeros(n, np.uint8); s = np.zeros(n, np.uint8); ph = np.zeros(1, np.uint8)
        P8 = C.POINTER(C.c_uint8)
        lib().orc_chf_export(self.ptr, *(a.ctypes.data_as(P8) for a in (G, F, M, g, v, s, ph)))
        return dict(G=G, F=F, M=M, gamma=g, v=v, s=s, phase=int(ph[0]))

    def statevector(self):
        out = np.zeros(2 << self.n, np.float64)
        _check(lib().orc_chf_statevector(self.ptr, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[0::2] + 1j * out[1::2]


def gates_array(gates):
    """gates: iterable of (op, q0[, q1[, q2]]) -> int32 [ng, 4]."""
    arr = np.zeros((max(len(gates), 1), 4), np.int32)
    for i, g in enumerate(gates):
        for j, v in enumerate(g):
            arr[i, j] = v
    return arr


def pauli_args(pauli):
    """pauli: ('dense', [ops little-endian]) or ('sparse', [(op, q), ...])."""
    kind, data = pauli
    if kind == "dense":
        ops = np.asarray(data, np.int32)
        return 1, len(ops), ops, np.zeros(max(len(ops), 1), np.int32)
    ops = np.asarray([o for o, _ in data] or [0], np.int32)
    qs = np.asarray([q for _, q in data] or [0], np.int32)
    return 0, len(data), ops, qs


class State:
    """A reference chi-term state (necstar StabilizerDecomposedState)."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def compile(cls, n, gates):
        arr = gates_array(gates)
        err = C.c_int(0)
        p = lib().orc_compile(n, len(gates), arr.ctypes.data_as(C.POINTER(C.c_int)), C.byref(err))
        _check(err.value)
        return cls(p)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_state_free(self.ptr)
            self.ptr = None

    def clone(self):
        return State(lib().orc_state_clone(self.ptr))

    @property
    def n(self):
        return lib().orc_state_num_qubits(self.ptr)

    @property
    def chi(self):
        return lib().orc_state_rank(self.ptr)

    def term(self, i):
        return CHForm(lib().orc_state_term(self.ptr, i), owned=False)

    def coeff_scalar(self, i):
        z, k, r = C.c_int(), C.c_int(), C.c_int64()
        lib().orc_state_coeff_scalar(self.ptr, i, C.byref(z), C.byref(k), C.byref(r))
        return (bool(z.value), k.value, r.value)

    def coeffs(self):
        out = np.zeros(self.chi, np.complex128)
        re, im = C.c_double(), C.c_double()
        for i in range(self.chi):
            lib().orc_state_coeff_complex(self.ptr, i, C.byref(re), C.byref(im))
            out[i] = complex(re.value, im.value)
        return out

    def set_complex_coeffs(self, c):
        c = np.ascontiguousarray(np.asarray(c, np.complex128))
        buf = np.empty(2 * len(c), np.float64)
        buf[0::2] = c.real
        buf[1::2] = c.imag
        lib().orc_state_set_complex_coeffs(self.ptr, buf.ctypes.data_as(C.POINTER(C.c_double)))

    @property
    def global_factor(self):
        re, im = C.c_double(), C.c_double()
        lib().orc_state_global_factor(self.ptr, C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def apply_gate(self, op, q0, q1=0, q2=0):
        _check(lib().orc_state_apply_gate(self.ptr, op, q0, q1, q2))

    def norm_squared(self):
        out = C.c_double()
        _check(lib().orc_state_norm_squared(self.ptr, C.byref(out)))
        return out.value

    def norm(self):
        out = C.c_double()
        _check(lib().orc_state_norm(self.ptr, C.byref(out)))
        return out.value

    def exp_value(self, pauli):
        dense, ln, ops, qs = pauli_args(pauli)
        out = C.c_double()
        _check(lib().orc_state_exp_value(self.ptr, dense, ln, ops.ctypes.data_as(C.POINTER(C.c_int)),
                                         qs.ctypes.data_as(C.POINTER(C.c_int)), C.byref(out)))
        return out.value

    def inner_product(self, other):
        re, im = C.c_double(), C.c_double()
        _check(lib().orc_state_inner_product(self.ptr, other.ptr, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def project_unnormalized(self, q, outcome):
        _check(lib().orc_state_project_unnormalized(self.ptr, q, int(outcome)))

    def project_normalized(self, q, outcome):
        _check(lib().orc_state_project_normalized(self.ptr, q, int(outcome)))

    def discard(self, q):
        _check(lib().orc_state_discard(self.ptr, q))

    def statevector(self):
        out = np.zeros(2 << self.n, np.float64)
        _check(lib().orc_state_statevector(self.ptr, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[0::2] + 1j * out[1::2]

    def amplitude(self, bits, per_term=False):
        b, bp = _u8p(bits)
        re, im = C.c_double(), C.c_double()
        chi = max(self.chi, 1)
        z = np.zeros(chi, np.int32); k = np.zeros(chi, np.int32); r = np.zeros(chi, np.int32)
        IP = C.POINTER(C.c_int)
        _check(lib().orc_state_amplitude(self.ptr, bp, C.byref(re), C.byref(im),
                                         z.ctypes.data_as(IP), k.ctypes.data_as(IP),
                                         r.ctypes.data_as(IP)))
        val = complex(re.value, im.value)
        if per_term:
            n = self.chi
            return val, (z[:n].astype(bool), k[:n], r[:n])
        return val

    def measure(self, qargs, seed):
        qa, qp = _ip(qargs)
        sd, sp = _u8p(np.frombuffer(seed_bytes(seed), np.uint8))
        out = np.zeros(max(len(qargs), 1), np.uint8)
        _check(lib().orc_state_measure(self.ptr, len(qargs), qp, sp,
                                       out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return [bool(b) for b in out[:len(qargs)]]

    def sample(self, qargs, shots, seed, max_out=None):
        """Returns a list of (outcome bit list, count) in DFS order."""
        qa, qp = _ip(qargs)
        sd, sp = _u8p(np.frombuffer(seed_bytes(seed), np.uint8))
        if max_out is None:
            max_out = min(shots, 1 << min(len(qargs), 20)) + 1
        bits = np.zeros(2 * max_out, np.uint64)
        counts = np.zeros(max_out, np.uint64)
        nout = C.c_int(0)
        U64P = C.POINTER(C.c_uint64)
        _check(lib().orc_state_sample(self.ptr, len(qargs), qp, shots, sp, max_out,
                                      bits.ctypes.data_as(U64P), counts.ctypes.data_as(U64P),
                                      C.byref(nout)))
        res = []
        for i in range(nout.value):
            lo, hi = int(bits[2 * i]), int(bits[2 * i + 1])
            word = lo | (hi << 64)
            res.append(([bool((word >> b) & 1) for b in range(len(qargs))], int(counts[i])))
        return res


class Rng:
    """rand 0.8 StdRng restatement (ChaCha12)."""

    def __init__(self, seed):
        sd, sp = _u8p(np.frombuffer(seed_bytes(seed), np.uint8))
        self.ptr = lib().orc_rng_new(sp)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_rng_free(self.ptr)
            self.ptr = None

    def next_u32(self):
        return lib().orc_rng_next_u32(self.ptr)

    def next_u64(self):
        return lib().orc_rng_next_u64(self.ptr)

    def gen_f64(self):
        return lib().orc_rng_gen_f64(self.ptr)

    def binomial(self, n, p):
        return lib().orc_rng_binomial(self.ptr, n, p)


def chacha_block(key_words, counter, stream, rounds):
    key = (C.c_uint32 * 8)(*key_words)
    out = (C.c_uint32 * 16)()
    lib().orc_chacha_block(key, counter, stream, rounds, out)
    return list(out)
