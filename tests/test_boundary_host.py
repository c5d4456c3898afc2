"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every symbol include/necstar_hip.h declares, and the host-side mirror of the
reference's Python API (gate, circuit, Pauli parsing, seeds) behaves like
python/src/{gate,circuit,pauli_string,utils}.rs.  No GPU calls here."""
import ctypes
import os
import re

import pytest

from necstar import PauliString, QuantumCircuit, QuantumGate
from necstar import _lib
from necstar.state import parse_seed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "necstar_hip.h")).read()
    return sorted(set(re.findall(r"\b(nx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), "Python signatures must cover the header exactly"


def test_library_is_gfx950_code_object():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


# tests/test_gate.py of the reference
def test_gate_strings():
    g = QuantumGate.x(0)
    assert g.name == "X" and g.qubits == [0] and str(g) == "X(0)" and repr(g) == "<QuantumGate: X(0)>"
    g = QuantumGate.cx(1, 2)
    assert g.name == "CX" and g.qubits == [1, 2] and str(g) == "CX(1, 2)"
    cl = [QuantumGate.h(0), QuantumGate.x(1), QuantumGate.y(2), QuantumGate.z(3), QuantumGate.s(4),
          QuantumGate.sdg(5), QuantumGate.sqrt_x(6), QuantumGate.sqrt_xdg(7), QuantumGate.cx(8, 9),
          QuantumGate.cz(10, 11), QuantumGate.swap(12, 13)]
    assert all(g.is_clifford() for g in cl) and not any(g.is_t_type() for g in cl)
    tt = [QuantumGate.t(0), QuantumGate.tdg(1)]
    assert all(g.is_t_type() for g in tt) and not any(g.is_clifford() for g in tt)
    assert str(QuantumGate.ccx(0, 1, 2)) == "CCX(0, 1, 2)"


# tests/test_pauli_string.py of the reference
def test_pauli_dense_sparse_identity():
    p = PauliString.from_str("IXYZ")
    assert str(p) == "IXYZ" and repr(p) == "PauliString('IXYZ')" and not p.is_identity
    for bad in ["IXYZA", "IXY1", "I X Y Z", "xyz"]:
        with pytest.raises(ValueError):
            PauliString.from_str(bad)
    p = PauliString.from_str("X1 Y3 Z0")
    assert str(p) == "X1 Y3 Z0" and repr(p) == "PauliString('X1 Y3 Z0')"
    assert str(PauliString.from_str("x1 y3 z0")) == "X1 Y3 Z0"
    for bad in ["X1 Y3 A0", "X1 Y3 Z", "X1 Y3 Z-1", "X1 Y3 Z1.5", "X1 Y3 Z1"]:
        with pytest.raises(ValueError):
            PauliString.from_str(bad)
    for s in ["", "I"]:
        p = PauliString.from_str(s)
        assert str(p) == "I" and repr(p) == "PauliString('I')" and p.is_identity
    # little-endian storage: "IX" puts X on qubit 0
    assert PauliString.from_str("IX")._ops == [1, 0]


# tests/test_circuit.py of the reference (QASM and builders)
def test_circuit_builders_and_strings():
    c = QuantumCircuit(2)
    assert c.num_qubits == 2 and len(c.gates) == 0
    c.apply_h(0)
    c.apply_cx(0, 1)
    c.apply_t(1)
    assert str(c) == "QuantumCircuit(num_qubits=2) [H(0), CX(0, 1), T(1)]"
    c2 = QuantumCircuit(2)
    c2.apply_gate("x", [1])
    c2.apply_gate("cz", [0, 1])
    assert [g.name for g in c2.gates] == ["X", "CZ"] and c2.gates[1].qubits == [0, 1]
    with pytest.raises(ValueError):
        c2.apply_gate("invalid_gate", [0])
    with pytest.raises(ValueError):
        c2.apply_gate("cx", [0])
    with pytest.raises(ValueError):
        c2.apply_gate("h", [0, 1])
    a, b = QuantumCircuit(2), QuantumCircuit(3)
    a.apply_h(0)
    b.apply_cx(0, 1)
    b.apply_t(2)
    t = a.tensor(b)
    assert t.num_qubits == 5 and [g.qubits for g in t.gates] == [[0], [2, 3], [4]]
    a.append(b)
    assert a.num_gates == 3


def test_qasm_roundtrip(tmp_path):
    q = QuantumCircuit.from_qasm_str('\nOPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\nh q[0];\ncx q[0], q[1];\nt q[1];\n')
    assert q.num_qubits == 2 and [g.name for g in q.gates] == ["H", "CX", "T"]
    assert q.to_qasm_str("q") == 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\nh q[0];\ncx q[0], q[1];\nt q[1];'
    with pytest.raises(ValueError):
        QuantumCircuit.from_qasm_str('OPENQASM 2.0;\nqreg q[2];\nt q[1];\nrx(pi/4) q[0];\n')
    with pytest.raises(ValueError):
        QuantumCircuit.from_qasm_str('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2]\nh q[0]\n')
    path = str(tmp_path / "c.qasm")
    q.to_qasm_file(path, reg_name="q")
    r = QuantumCircuit.from_qasm_file(path)
    assert [(g.name, g.qubits) for g in r.gates] == [(g.name, g.qubits) for g in q.gates]


def test_seed_parsing():
    assert parse_seed(None) is None
    assert parse_seed(5) == (5).to_bytes(8, "little") + bytes(24)
    assert parse_seed(2 ** 70) == (2 ** 70).to_bytes(32, "little")
    assert parse_seed(bytes(range(32))) == bytes(range(32))
    with pytest.raises(ValueError, match="Seed must be 32 bytes, got 3"):
        parse_seed(b"abc")
    with pytest.raises(TypeError):
        parse_seed(1.5)
    with pytest.raises(OverflowError):
        parse_seed(1 << 300)
    with pytest.raises(OverflowError):
        parse_seed(-1)


def test_pow_half_identity():
    """The pair kernel forms 2^{-r/2} as ldexp(pow(2, -1/2) or 1, -floor(r/2))
    in place of the host table entry pow(2, -r/2): the two agree bit for bit
    for every |r| <= 1024 (the table range), and the constant it uses is
    pow(2, -1/2)."""
    import math
    assert math.pow(2.0, -0.5) == float.fromhex("0x1.6a09e667f3bcdp-1")
    for r in range(-1024, 1025):
        assert math.pow(2.0, -r / 2.0) == math.ldexp(math.pow(2.0, -0.5) if r & 1 else 1.0, -(r >> 1))
