"""QuantumCircuit (python/src/circuit.rs:9-235, crates/necstar/src/circuit/mod.rs,
OpenQASM 2 text I/O of crates/necstar/src/circuit/parser.rs:17-230)."""
from __future__ import annotations

import re

from . import gate as G
from .gate import QuantumGate, _usize

_ID = r"[a-zA-Z][a-zA-Z0-9_]*"
_QREG_RE = re.compile(r"qreg\s+(" + _ID + r")\s*\[\s*(\d+)\s*\]\s*;")
_G1_RE = re.compile(r"([a-z_]+)\s+(" + _ID + r")\[(\d+)\]\s*;")
_G2_RE = re.compile(r"([a-z_]+)\s+(" + _ID + r")\[(\d+)\],\s*(" + _ID + r")\[(\d+)\]\s*;")
_G3_RE = re.compile(r"([a-z_]+)\s+(" + _ID + r")\[(\d+)\],\s*(" + _ID + r")\[(\d+)\],\s*(" + _ID + r")\[(\d+)\]\s*;")
_G1 = {"h": G.H, "x": G.X, "y": G.Y, "z": G.Z, "s": G.S, "sdg": G.SDG, "sx": G.SQRTX, "sxdg": G.SQRTXDG,
       "t": G.T, "tdg": G.TDG}
_G2 = {"cx": G.CX, "cz": G.CZ, "swap": G.SWAP}
_G3 = {"ccx": G.CCX}


class _QasmError(ValueError):
    pass


def _parse_qasm(text: str):
    """crates/necstar/src/circuit/parser.rs:17-158"""
    num_qubits = None
    gates = []
    for line_content in text.splitlines():
        line = line_content.strip()
        if not line or line.startswith("//"):
            continue
        if line.startswith("OPENQASM") or line.startswith("include"):
            continue
        m = _QREG_RE.search(line)
        if m:
            if num_qubits is not None:
                raise _QasmError("QASM parsing error: Multiple qreg declarations are not supported.")
            num_qubits = int(m.group(2))
            continue
        if line.startswith("measure"):
            continue
        matched = False
        m = _G3_RE.search(line)
        if m and m.group(1) in _G3:
            gates.append(QuantumGate(_G3[m.group(1)], (int(m.group(3)), int(m.group(5)), int(m.group(7)))))
            matched = True
        if not matched:
            m = _G2_RE.search(line)
            if m and m.group(1) in _G2:
                gates.append(QuantumGate(_G2[m.group(1)], (int(m.group(3)), int(m.group(5)))))
                matched = True
        if not matched:
            m = _G1_RE.search(line)
            if m and m.group(1) in _G1:
                gates.append(QuantumGate(_G1[m.group(1)], (int(m.group(3)),)))
                matched = True
        if not matched:
            raise _QasmError(f"QASM parsing error: Unrecognized or malformed line: {line}")
    if num_qubits is None:
        raise _QasmError("QASM parsing error: qreg declaration not found in QASM string.")
    return num_qubits, gates


class QuantumCircuit:
    def __init__(self, num_qubits: int):
        self._n = _usize(num_qubits)
        self._gates = []

    @property
    def num_qubits(self) -> int:
        return self._n

    @property
    def num_gates(self) -> int:
        return len(self._gates)

    @property
    def gates(self):
        return list(self._gates)

    @staticmethod
    def from_qasm_file(path: str) -> "QuantumCircuit":
        try:
            with open(path) as f:
                text = f.read()
        except OSError as e:
            raise FileNotFoundError(
                f"Failed to read QASM file: QASM parsing error: Failed to read file '{path}': {e}") from None
        try:
            n, gates = _parse_qasm(text)
        except _QasmError as e:
            raise FileNotFoundError(f"Failed to read QASM file: {e}") from None
        qc = QuantumCircuit(n)
        qc._gates = gates
        return qc

    @staticmethod
    def from_qasm_str(qasm: str) -> "QuantumCircuit":
        try:
            n, gates = _parse_qasm(qasm)
        except _QasmError as e:
            raise ValueError(f"Failed to parse QASM string: {e}") from None
        qc = QuantumCircuit(n)
        qc._gates = gates
        return qc

    def to_qasm_str(self, reg_name: str) -> str:
        lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', f"qreg {reg_name}[{self._n}];"]
        lines += [g._qasm(reg_name) for g in self._gates]
        return "\n".join(lines)

    def to_qasm_file(self, path: str, reg_name: str) -> None:
        try:
            with open(path, "w") as f:
                f.write(self.to_qasm_str(reg_name))
        except OSError as e:
            raise ValueError(f"Failed to write QASM file: IO error: {e}") from None

    @staticmethod
    def random_clifford(n: int, seed=None) -> "QuantumCircuit":
        """python/src/circuit.rs:73-79: a uniformly random Clifford circuit
        (Bravyi-Maslov; crates/stabilizer-ch-form-rust/src/circuit/
        random_clifford.rs:7-218), drawn on the host by the library from the
        seed's StdRng stream."""
        import ctypes as C

        import numpy as np

        from ._lib import check, lib
        from .state import _seed_ptr

        n = _usize(n)
        _buf, sp = _seed_ptr(seed)
        cap = 2 * n * n + 3 * n + 1 if n <= 1024 else 1  # larger n is rejected by the library
        rows = np.zeros((cap, 4), np.int32)
        count = C.c_int64()
        check(lib().nx_random_clifford(min(n, 2 ** 31 - 1), sp, cap,
                                       rows.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(count)))
        qc = QuantumCircuit(n)
        qc._gates = [QuantumGate(int(r[0]), tuple(int(q) for q in r[1:1 + G.ARITY[r[0]]]))
                     for r in rows[: count.value]]
        return qc

    def append(self, other: "QuantumCircuit") -> None:
        self._gates.extend(other._gates)

    def tensor(self, other: "QuantumCircuit") -> "QuantumCircuit":
        qc = QuantumCircuit(self._n + other._n)
        qc._gates = list(self._gates) + [g._shifted(self._n) for g in other._gates]
        return qc

    def apply_gate(self, gate_name: str, qubits) -> None:
        """python/src/circuit.rs:90-187"""
        qubits = [_usize(q) for q in qubits]
        name = gate_name.lower()
        one = {"h": (G.H, "H"), "x": (G.X, "X"), "y": (G.Y, "Y"), "z": (G.Z, "Z"), "s": (G.S, "S"),
               "sdg": (G.SDG, "Sdg"), "sqrtx": (G.SQRTX, "SqrtX"), "sqrtxdg": (G.SQRTXDG, "SqrtXdg"),
               "t": (G.T, "T"), "tdg": (G.TDG, "Tdg")}
        if name in one:
            op, label = one[name]
            if len(qubits) != 1:
                raise ValueError(f"{label} gate requires 1 qubit index.")
            self._gates.append(QuantumGate(op, qubits))
            return
        if name in ("cx", "cnot"):
            if len(qubits) != 2:
                raise ValueError("CX gate requires 2 qubit indices (control, target).")
            self._gates.append(QuantumGate(G.CX, qubits))
            return
        if name == "cz":
            if len(qubits) != 2:
                raise ValueError("CZ gate requires 2 qubit indices.")
            self._gates.append(QuantumGate(G.CZ, qubits))
            return
        if name == "swap":
            if len(qubits) != 2:
                raise ValueError("SWAP gate requires 2 qubit indices.")
            self._gates.append(QuantumGate(G.SWAP, qubits))
            return
        raise ValueError(f"Unknown gate name: '{gate_name}'")

    def apply_h(self, target): self._gates.append(QuantumGate.h(target))

    def apply_x(self, target): self._gates.append(QuantumGate.x(target))

    def apply_y(self, target): self._gates.append(QuantumGate.y(target))

    def apply_z(self, target): self._gates.append(QuantumGate.z(target))

    def apply_s(self, target): self._gates.append(QuantumGate.s(target))

    def apply_sdg(self, target): self._gates.append(QuantumGate.sdg(target))

    def apply_sqrt_x(self, target): self._gates.append(QuantumGate.sqrt_x(target))

    def apply_sqrt_xdg(self, target): self._gates.append(QuantumGate.sqrt_xdg(target))

    def apply_t(self, target): self._gates.append(QuantumGate.t(target))

    def apply_tdg(self, target): self._gates.append(QuantumGate.tdg(target))

    def apply_cx(self, control, target): self._gates.append(QuantumGate.cx(control, target))

    def apply_cz(self, qarg1, qarg2): self._gates.append(QuantumGate.cz(qarg1, qarg2))

    def apply_swap(self, qarg1, qarg2): self._gates.append(QuantumGate.swap(qarg1, qarg2))

    def apply_ccx(self, control1, control2, target):
        self._gates.append(QuantumGate.ccx(control1, control2, target))

    def __str__(self):
        return f"QuantumCircuit(num_qubits={self._n}) [" + ", ".join(str(g) for g in self._gates) + "]"

    def _gate_rows(self):
        return [g._row() for g in self._gates]
