"""Batched norms and Pauli expectation values over many states (API
addition: the reference evaluates QuantumState.norm and exp_value one call
at a time, crates/necstar/src/state/stabilizer_decomposed_state/
{norm.rs:8-35, exp_value.rs:7-61}).

    b = Batch([st0, st1, ...], [[P, Q, ...], [...], ...], autotune=True, graph=True)
    norms, values = b.run()      # norms[i], values[i][k] = st_i.exp_value(P_k)

Each state is copied when the batch is created.  All pairs of all states
and all Pauli strings of a state run in one persistent pair launch per
row-word width; with graph=True the launches are captured once into a HIP
graph and every run() replays it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, loaded

NX_BATCH_AUTOTUNE = 1
NX_BATCH_GRAPH = 2


class Batch:
    def __init__(self, states, paulis, autotune: bool = True, graph: bool = True):
        states = list(states)
        paulis = [list(p) for p in paulis]
        if len(paulis) != len(states):
            raise ValueError("one Pauli list per state is required")
        rows = [st._dense_ops(pl)[: len(pl)] for st, pl in zip(states, paulis)]
        flat = np.concatenate([r.reshape(-1) for r in rows] + [np.zeros(1, np.int32)]).astype(np.int32)
        npauli = np.array([len(p) for p in paulis] + [0], np.int32)
        handles = (C.c_void_p * max(len(states), 1))(*[st._h for st in states])
        flags = (NX_BATCH_AUTOTUNE if autotune else 0) | (NX_BATCH_GRAPH if graph else 0)
        h = C.c_void_p()
        check(lib().nx_batch_create(len(states), handles, npauli.ctypes.data_as(C.POINTER(C.c_int32)),
                                    flat.ctypes.data_as(C.POINTER(C.c_int32)), flags, C.byref(h)))
        self._h = h
        self._counts = [len(p) for p in paulis]
        nout = C.c_int64()
        check(lib().nx_batch_num_outputs(self._h, C.byref(nout)))
        self._nout = nout.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                L = loaded()
            except Exception:  # interpreter shutdown: module state is gone
                return
            if L is not None:
                L.nx_batch_free(h)

    @property
    def shape(self):
        """(pair-chunk length, total waves, variant words) of the launch plan."""
        J, waves, vw = C.c_int32(), C.c_int64(), C.c_int32()
        check(lib().nx_batch_shape(self._h, C.byref(J), C.byref(waves), C.byref(vw)))
        return J.value, waves.value, vw.value

    def run_flat(self) -> np.ndarray:
        """Raw output: for each state, its norm then its expectation values."""
        out = np.zeros(max(self._nout, 1), np.float64)
        check(lib().nx_batch_run(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[: self._nout]

    def run(self):
        flat = self.run_flat()
        norms, values, pos = [], [], 0
        for k in self._counts:
            norms.append(float(flat[pos]))
            values.append(flat[pos + 1: pos + 1 + k].copy())
            pos += 1 + k
        return norms, values
