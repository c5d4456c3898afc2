"""necstar -- MI355X-native drop-in for the reference Python module.

Same classes, methods, getters, strings and error types as the PyO3 module
of mutekichi/stab-decomp-simulator-rust (python/src/lib.rs:14-22,
python/src/{gate,circuit,pauli_string,state,utils}.rs).  Every QuantumState
operation runs through the C-ABI of include/necstar_hip.h on a gfx950 GPU.
"""
from .gate import QuantumGate
from .circuit import QuantumCircuit
from .pauli_string import PauliString
from .state import QuantumState
from .batch import Batch, sample_many

__all__ = ["QuantumGate", "QuantumCircuit", "QuantumState", "PauliString", "Batch", "sample_many"]
