"""PauliString (python/src/pauli_string.rs:8-35,
crates/stabilizer-ch-form-rust/src/types/pauli/mod.rs:19-168)."""
from __future__ import annotations

import re

_TERM_RE = re.compile(r"^(?i:([XYZ]))(\d+)$")
_CODE = {"I": 0, "X": 1, "Y": 2, "Z": 3}
_CHAR = "IXYZ"


class PauliString:
    """Dense strings are stored little-endian (index 0 = qubit 0); the
    textual form puts qubit 0 rightmost.  Sparse strings keep term order."""

    __slots__ = ("_dense", "_ops", "_qubits")

    def __init__(self, dense: bool, ops, qubits):
        self._dense = dense
        self._ops = list(ops)
        self._qubits = list(qubits)

    @staticmethod
    def from_str(s: str) -> "PauliString":
        try:
            return PauliString._parse(s)
        except ValueError as e:
            raise ValueError(f"Invalid Pauli string format: Pauli string parsing error: {e}") from None

    @staticmethod
    def _parse(s: str) -> "PauliString":
        trimmed = s.strip()
        if not trimmed or trimmed.lower() == "i":
            return PauliString(False, [], [])
        if any(c.isdigit() for c in trimmed):  # sparse (mod.rs:43-87)
            ops, qubits, seen = [], [], set()
            for term in trimmed.split():
                m = _TERM_RE.match(term)
                if not m:
                    raise ValueError(f"invalid sparse Pauli term: '{term}' in string: '{trimmed}'")
                q = int(m.group(2))
                if q in seen:
                    raise ValueError(f"duplicate Pauli operator on qubit {q} in string: '{trimmed}'")
                seen.add(q)
                ops.append(_CODE[m.group(1).upper()])
                qubits.append(q)
            return PauliString(False, ops, qubits)
        ops = []  # dense (mod.rs:19-41)
        for i, ch in enumerate(trimmed):
            if ch not in _CODE:
                raise ValueError(f"invalid character '{ch}' at position {i} in dense PauliString")
            ops.append(_CODE[ch])
        ops.reverse()
        return PauliString(True, ops, [])

    @property
    def is_identity(self) -> bool:
        return all(o == 0 for o in self._ops)

    def __str__(self):
        if self._dense:
            s = "".join(_CHAR[o] for o in reversed(self._ops))
            return s if s else "I"
        if not self._ops:
            return "I"
        return " ".join(f"{_CHAR[o]}{q}" for o, q in zip(self._ops, self._qubits))

    def __repr__(self):
        return f"PauliString('{self}')"
