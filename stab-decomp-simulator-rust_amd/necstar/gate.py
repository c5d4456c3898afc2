"""QuantumGate (python/src/gate.rs:4-122, crates/necstar/src/circuit/gates.rs)."""
from __future__ import annotations

H, X, Y, Z, S, SDG, SQRTX, SQRTXDG, CX, CZ, SWAP, T, TDG, CCX = range(14)
NAMES = ["H", "X", "Y", "Z", "S", "Sdg", "SqrtX", "SqrtXdg", "CX", "CZ", "Swap", "T", "Tdg", "CCX"]
QASM = ["h", "x", "y", "z", "s", "sdg", "sx", "sxdg", "cx", "cz", "swap", "t", "tdg", "ccx"]
ARITY = [1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 3]


def _usize(q):
    q = int(q)
    if q < 0:
        raise OverflowError("can't convert negative int to unsigned")
    return q


class QuantumGate:
    __slots__ = ("_op", "_qubits")

    def __init__(self, op: int, qubits):
        self._op = op
        self._qubits = tuple(_usize(q) for q in qubits)

    @staticmethod
    def h(target): return QuantumGate(H, (target,))

    @staticmethod
    def x(target): return QuantumGate(X, (target,))

    @staticmethod
    def y(target): return QuantumGate(Y, (target,))

    @staticmethod
    def z(target): return QuantumGate(Z, (target,))

    @staticmethod
    def s(target): return QuantumGate(S, (target,))

    @staticmethod
    def sdg(target): return QuantumGate(SDG, (target,))

    @staticmethod
    def sqrt_x(target): return QuantumGate(SQRTX, (target,))

    @staticmethod
    def sqrt_xdg(target): return QuantumGate(SQRTXDG, (target,))

    @staticmethod
    def cx(control, target): return QuantumGate(CX, (control, target))

    @staticmethod
    def cz(qarg1, qarg2): return QuantumGate(CZ, (qarg1, qarg2))

    @staticmethod
    def swap(qarg1, qarg2): return QuantumGate(SWAP, (qarg1, qarg2))

    @staticmethod
    def t(target): return QuantumGate(T, (target,))

    @staticmethod
    def tdg(target): return QuantumGate(TDG, (target,))

    @staticmethod
    def ccx(control1, control2, target): return QuantumGate(CCX, (control1, control2, target))

    @property
    def name(self) -> str:
        return NAMES[self._op]

    @property
    def qubits(self):
        return list(self._qubits)

    def __str__(self):
        return f"{NAMES[self._op]}({', '.join(str(q) for q in self._qubits)})"

    def __repr__(self):
        return f"<QuantumGate: {self}>"

    def __eq__(self, other):
        return isinstance(other, QuantumGate) and (self._op, self._qubits) == (other._op, other._qubits)

    def __hash__(self):
        return hash((self._op, self._qubits))

    def is_clifford(self) -> bool:
        return self._op <= SWAP

    def is_t_type(self) -> bool:
        return self._op in (T, TDG)

    def _row(self):
        q = list(self._qubits) + [0] * (3 - len(self._qubits))
        return (self._op, q[0], q[1], q[2])

    def _shifted(self, offset):
        return QuantumGate(self._op, tuple(q + offset for q in self._qubits))

    def _qasm(self, reg):
        return f"{QASM[self._op]} " + ", ".join(f"{reg}[{q}]" for q in self._qubits) + ";"
